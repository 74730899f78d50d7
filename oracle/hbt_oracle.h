/*
 * hbt_oracle.h -- CPU oracle for the HBT pair-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's HBT_correlation pair loops (src/HBT_correlation.cpp), its
 * mode-10 sample reader (src/particleSamples.cpp), its random stream
 * (src/Random.{h,cpp}, libstdc++ 11 <random>) and its text writers.  It is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
 * leg to check results against; the product (libhbt_hip.so) never links or calls it.
 *
 * Pinning status of each function is recorded in the header comment of
 * hbt_oracle.c.
 */
#ifndef HBT_ORACLE_H
#define HBT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of HBT_correlation::HBT_correlation (HBT_correlation.cpp:22-61). */
typedef struct {
    int qnpts;
    double q_min, q_max;
    int n_KT, n_Kphi;
    double KT_min, KT_max;
    int azimuthal_flag;
    int invariant_radius_flag;
    int long_comoving_boost;
    double HBTrap_min, HBTrap_max;
    double needed_number_of_pairs; /* as read by getVal (double), converted like the reference */
    int number_of_mixed_events;    /* <= 0: reference rule floor(mixed_nev/2)+1 (HBT_correlation.cpp:200) */
} orc_params;

/* Derived constants, computed exactly as the reference does. */
typedef struct {
    double delta_q;  /* (q_max - q_min)/(qnpts - 1)            :28 */
    double dKT;      /* (KT_max - KT_min)/(n_KT - 1)           :48 */
    double dKphi;    /* 2*M_PI/n_Kphi                          :49 */
    double q_base;   /* q_min - delta_q/2.                     :369 */
    double q_lo_cut; /* q_min - delta_q/2. + 1e-8              :363 */
    double q_hi_cut; /* q_max + delta_q/2. - 1e-8              :364 */
    double KT_min_sq, KT_max_sq; /*                            :289-290 */
    double hbarC_inv;            /* 1./hbarC                   :253 */
    double rap_cut_min, rap_cut_max; /* tanh(HBTrap_*)         :255-256 */
    uint64_t needed;             /* (ull) needed_number_of_pairs :36 */
    uint64_t needed_qinv;        /* 50*needed in ull arithmetic :341 */
} orc_derived;

/* Stage reach counters (for the FLOP model; not part of the reference). */
typedef struct {
    uint64_t enumerated, kt, out, side, binned;
} orc_stages;

typedef struct {
    orc_params p;
    orc_derived d;
    double psi_ref;
    int nphi;       /* 1 when azimuthal_flag==0, n_Kphi otherwise */
    long nbins3d;   /* n_KT*nphi*qnpts^3 */
    /* 3D histograms, flat [kt][kphi][iqout][iqside][iqlong] (iqlong fastest) */
    double *num_count, *q_out_sum, *q_side_sum, *q_long_sum, *num, *den;
    /* per-bin |.| sums, used as the floor of the tolerance metric */
    double *abs_num, *abs_q_out, *abs_q_side, *abs_q_long;
    /* q_inv histograms [kt][iq] */
    double *qinv_sum, *qinv_num, *qinv_count, *qinv_den;
    /* pair counters: [kt] or [kt][kphi] */
    uint64_t *cnt_num_kt, *cnt_den_kt, *cnt_num_qinv, *cnt_den_qinv;
    uint64_t *cnt_num_ktphi, *cnt_den_ktphi;
    orc_stages same_st, mixed_st;
    uint64_t kphi_skipped;
} orc_state;

/* A buffer of events in structure-of-arrays form with event offsets. */
typedef struct {
    int nev;
    const int64_t *ev_off; /* nev+1 offsets into the particle arrays */
    const double *px, *py, *pz, *E, *x, *y, *z, *t;
} orc_buffer;

int orc_init(orc_state *s, const orc_params *p);
void orc_free(orc_state *s);
void orc_derive(const orc_params *p, orc_derived *d);

/* psi_ref over all particles of the buffer (HBT_correlation.cpp:233-249). */
double orc_flow_event_plane_angle(const orc_buffer *b, int n_order);

/* Same-event pairs of the whole buffer (HBT_correlation.cpp:251-462). */
void orc_combine_same(orc_state *s, const orc_buffer *b);

/* Mixed pairs of event iev of `b` with the M partners of `mix`
 * (HBT_correlation.cpp:464-692).  cs[2k], cs[2k+1] = cos, sin of partner k's
 * rotation angle. */
void orc_combine_mixed(orc_state *s, const orc_buffer *b, int iev,
                       const orc_buffer *mix, int M, const int *partner_ids,
                       const double *cs);

/* ---------------- random stream (Random.h:11-25, libstdc++ 11) ---------- */
typedef struct {
    uint32_t mt[624];
    int idx;
} orc_rng;
void orc_rng_seed(orc_rng *r, uint32_t seed);
uint32_t orc_rng_next(orc_rng *r);
int orc_rand_int_uniform(orc_rng *r); /* uniform_int_distribution<int>() */
double orc_rand_uniform(orc_rng *r);  /* uniform_real_distribution<double>(0,1) */

/* Draws the mixing plan of one buffer in the reference's order
 * (HBT_correlation.cpp:197-216 and :493-497): for each event, M partner ids
 * then M rotation angles.  ids has nev*M entries, cs 2*nev*M.  Returns M. */
int orc_draw_mixing_plan(orc_rng *r, int nev, int mixed_nev,
                         int number_of_mixed_events, int *ids, double *cs);

/* Whole calculate_HBT_correlation_function on one buffer
 * (HBT_correlation.cpp:177-218).  `mix` may equal `b` (aliased mixed list,
 * particleSamples.cpp:528-531). */
void orc_calculate_buffer(orc_state *s, orc_rng *r, const orc_buffer *b,
                          const orc_buffer *mix);

/* Derived counters: number_of_pairs_*_KTdiff etc. are kept exactly. */

/* ---------------- writers (HBT_correlation.cpp:694-855) ---------------- */
int orc_write_outputs(const orc_state *s, const char *path, int eco_output);

/* ---------------- mode-10 reader (particleSamples.cpp:1247-1346) -------- */
typedef struct orc_reader orc_reader;
orc_reader *orc_reader_open(const char *filename, int event_buffer_size,
                            int particle_monval);
/* Reads the next buffer.  Returns number of events (may be 0), fills
 * owned arrays accessible through orc_reader_buffer(). */
int orc_reader_next(orc_reader *rd);
int orc_reader_eof(const orc_reader *rd);
const orc_buffer *orc_reader_buffer(const orc_reader *rd);
/* number of events before filtering (all species) in last buffer */
long orc_reader_total_particles(const orc_reader *rd);
void orc_reader_close(orc_reader *rd);

/* ---------------- species table (particle_decay.cpp:23-129) ---------------- */
/* Charge, baryon number and strangeness per monval, sorted by monval. */
typedef struct {
    int n;
    int *monval, *charge, *baryon, *strange;
} orc_species;
/* read_resonances_list: the PDG table file, anti-baryon entries added. */
int orc_species_read(const char *path, orc_species *sp);
void orc_species_free(orc_species *sp);

/* ---------------- balance function (BalanceFunction.cpp) ---------------- */
typedef struct {
    int alpha, beta;             /* particle_alpha, particle_beta */
    double BpT_min, BpT_max, Brap_max;
    int Bnpts, rap_type;         /* rap_type 0: pseudo-rapidity, 1: rapidity */
} orc_bf_params;

typedef struct {
    orc_bf_params p;
    int Bnphi;
    double dphi, Bphi_min, drap, Brap_min;
    double *hist;                /* [8][Bnpts][Bnphi]: C_ab, C_abarbbar, C_abbar,
                                    C_abarb, then the four mixed-event ones */
    long N_b, N_bbar;
    const orc_species *sp;       /* for the species classes (may be NULL) */
    int bad_monval_seen;         /* a class was evaluated for a monval not in sp */
} orc_bf_state;

/* One buffer of events with all species (monval, mass and momenta). */
typedef struct {
    int nev;
    const int64_t *ev_off;
    const int *monval;
    const double *mass, *px, *py, *pz, *E;
} orc_bf_buffer;

int orc_bf_init(orc_bf_state *s, const orc_bf_params *p);
void orc_bf_free(orc_bf_state *s);
/* calculate_balance_function on one buffer (BalanceFunction.cpp:63-101);
 * mix may equal b (the aliased mixed list). */
void orc_bf_calculate(orc_bf_state *s, orc_rng *r, const orc_bf_buffer *b,
                      const orc_bf_buffer *mix);
/* output_balance_function (BalanceFunction.cpp:211-318). */
int orc_bf_write_outputs(const orc_bf_state *s, const char *path);

#ifdef __cplusplus
}
#endif
#endif
