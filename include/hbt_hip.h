/*
 * hbt_hip.h -- C ABI of the MI355X (gfx950) HBT pair-correlation engine.
 *
 * The reference has no C ABI; its boundary for this path is the C++ class
 * HBT_correlation (src/HBT_correlation.h:14-104), called per particle
 * buffer by Analysis::HBTAnalysis (src/Analysis.cpp:817-835).  Every entry
 * point below replaces one piece of that class; the comment on each names
 * the reference member it stands in for.  The drop-in C++ class
 * (hadronic_afterburner_toolkit_amd/csrc/host/HBT_correlation.h) keeps the
 * reference's signatures and forwards to these functions.
 *
 * Conventions: plain pointers and sizes only; every function returns 0 on
 * success or a negative HBT_ERR_* code, with a message in
 * hbt_last_error(ctx).  A context is single-owner and not thread-safe; all
 * its device work runs on its own HIP stream.  Host arrays passed in are
 * only read during the call.
 */
#ifndef HBT_HIP_H
#define HBT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HBT_ABI_VERSION 1

enum {
    HBT_OK = 0,
    HBT_ERR_ARG = -1,        /* invalid argument or parameter */
    HBT_ERR_HIP = -2,        /* a HIP runtime call failed */
    HBT_ERR_NOMEM = -3,      /* device or host allocation failed */
    HBT_ERR_STATE = -4,      /* call out of order (e.g. run before upload) */
    HBT_ERR_CAPACITY = -5,   /* request exceeds a stated limit */
    HBT_ERR_NODEVICE = -6    /* no usable gfx950 device */
};

/* Arithmetic of the pair cascade. */
enum {
    HBT_PREC_FP64 = 0,  /* plain mode: FP64, bit-exact bins/counts with the reference */
    HBT_PREC_FP32C = 1  /* precision form: FP32 + compensation, FP64 final reduction */
};

/* Kernel family. */
enum {
    HBT_KERNEL_PLAIN = 0, /* LDS-tiled direct cascade, one thread per i */
    HBT_KERNEL_FAST = 1   /* prefilter + per-wave LDS compaction, exact cascade on survivors */
};

/* Analysis parameters: the keys HBT_correlation reads in its constructor
 * (HBT_correlation.cpp:22-46), same names and meaning, plus the engine's
 * own choices at the end. */
typedef struct {
    int32_t qnpts;
    double q_min, q_max;
    int32_t n_KT;            /* number of K_T EDGES, as in the reference (:48) */
    double KT_min, KT_max;
    int32_t n_Kphi;
    int32_t azimuthal_flag;
    int32_t invariant_radius_flag;
    int32_t long_comoving_boost;
    double HBTrap_min, HBTrap_max;
    double needed_number_of_pairs;   /* converted to unsigned long long like :36 */
    int32_t number_of_mixed_events;  /* <=0: reference rule floor(mixed_nev/2)+1 (:200) */
    /* engine choices */
    int32_t precision;       /* HBT_PREC_* */
    int32_t kernel;          /* HBT_KERNEL_* */
    int32_t device;          /* HIP device ordinal */
    /* K_phi angles within this many 1e-12 rad of a bin edge (or of 0 and
     * 2 pi) are decided on the host with glibc's atan2, as the reference
     * computes them; the device's atan2 may differ from it in the last bits.
     * 0: 1 (1e-12 rad).  Larger values only move more pairs to the host. */
    int32_t kphi_guard_prad;
    /* histogram accumulation of the fast kernel: 0 = every binned pair added
     * to the histogram with device-scope atomics; 1 = slab binning where it
     * applies (no q_inv, qnpts^2 <= 2048, and n_KT * qnpts <= 512, or with
     * K_phi bins n_KT * n_Kphi * qnpts <= 4096): binned same-event pairs
     * written as records per (block, K_T x K_phi x q_out plane)
     * and summed in LDS, without device-scope atomics (mixed pairs still add
     * their u32 directly); 2 = the one hbt_autotune measures faster on the
     * prepared batch, timing every launch shape both ways (atomics until
     * then).  Counts are the same; sums agree within each form's tolerance.
     * hbt_slab_info reports which one the last launch used. */
    int32_t binning;
    int32_t reserved[3];
} hbt_config;

typedef struct hbt_ctx hbt_ctx;
typedef struct hbt_rng hbt_rng;

/* Replaces HBT_correlation::HBT_correlation (HBT_correlation.cpp:16-143):
 * validates the parameters, derives the grids and allocates zeroed
 * histograms on the device. */
int hbt_create(const hbt_config *cfg, hbt_ctx **out);

/* Replaces HBT_correlation::~HBT_correlation (HBT_correlation.cpp:145-175). */
int hbt_destroy(hbt_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed
 * hbt_create when ctx is NULL). */
const char *hbt_last_error(const hbt_ctx *ctx);

/* Zeroes histograms (no reference counterpart: a fresh HBT_correlation). */
int hbt_reset(hbt_ctx *ctx);

/* Replaces the particle gathers of combine_and_bin_particle_pairs and
 * combine_and_bin_particle_pairs_mixed_events (HBT_correlation.cpp:253-281,
 * :466-489, :498-551): takes one buffer of nev events (already filtered to
 * one species, particleSamples.cpp:1329-1346) as host SoA arrays with
 * nev+1 event offsets, applies the tanh(HBTrap) cut on pz/E, and makes the
 * kept particles resident in HBM.  slot 0 = same-event list, slot 1 = the
 * mixed-event list when read_in_real_mixed_events=1 (otherwise slot 1
 * aliases slot 0, particleSamples.cpp:528-531).  soa = {px,py,pz,E,x,y,z,t}.
 * The copies go through pinned host buffers that the context keeps (sized
 * to its largest upload) until hbt_destroy; the call returns once they are
 * done, so the caller may reuse its arrays at once. */
int hbt_upload(hbt_ctx *ctx, int32_t slot, int32_t nev, const int64_t *ev_off,
               const double *const soa[8]);

/* Overrides psi_ref for azimuthal (azimuthal_flag=1) pairs of all groups.
 * By default each group's psi_ref is computed at upload exactly as
 * calculate_flow_event_plane_angle(2) does (HBT_correlation.cpp:181-183,
 * :233-249), over all particles passed in, before the rapidity cut. */
int hbt_set_psi_ref(hbt_ctx *ctx, double psi_ref);

/* Replaces combine_and_bin_particle_pairs (HBT_correlation.cpp:251-462):
 * all i<j pairs of the slot-0 buffer, binned into the resident histograms. */
int hbt_run_same(hbt_ctx *ctx);

/* Replaces the per-event loop of combine_and_bin_particle_pairs_mixed_events
 * (HBT_correlation.cpp:464-692) for every event of the slot-0 buffer:
 * partner_ids[nev*M] index events of the mixed list, cos_sin[2*nev*M] hold
 * cos and sin of each partner's rotation angle, drawn on the host in the
 * reference's order (hbt_rng_draw_plan). */
int hbt_run_mixed(hbt_ctx *ctx, int32_t M, const int32_t *partner_ids,
                  const double *cos_sin);

/* Raw results, host arrays owned by the caller, sized for n_KT (edges),
 * nphi = (azimuthal_flag ? n_Kphi : 1) and qnpts^3 bins per (K_T, K_phi).
 * Layout [kt][kphi][iqout][iqside][iqlong], iqlong fastest.  Any pointer
 * may be NULL to skip it.  The pair counters are the reference's
 * number_of_pairs_*_KTdiff (HBT_correlation.h:42-48). */
typedef struct {
    uint64_t *num_count;     /* correl_3d_num_count */
    uint64_t *den;           /* correl_3d_denorm */
    double *q_out_sum, *q_side_sum, *q_long_sum; /* q_*_mean before division */
    double *num;             /* correl_3d_num (sum of cos) */
    uint64_t *qinv_count, *qinv_den; /* [kt][q] */
    double *qinv_sum, *qinv_num;
    uint64_t *cnt_num_kt, *cnt_den_kt, *cnt_num_qinv, *cnt_den_qinv; /* [kt] */
    uint64_t *cnt_num_ktphi, *cnt_den_ktphi;                         /* [kt][kphi] */
} hbt_result;

/* Replaces reading the member arrays in the output_* writers
 * (HBT_correlation.cpp:694-855).  Synchronises the context's stream. */
int hbt_read(hbt_ctx *ctx, hbt_result *out);

/* Stage reach counters of the last runs since reset (enumerated, K_T cut
 * passed, q_out passed, q_side passed, binned), same and mixed; the inputs
 * of the FLOP model.  Only the plain kernel counts every stage; the fast
 * kernel reports enumerated and binned. */
int hbt_stage_counts(hbt_ctx *ctx, uint64_t same[5], uint64_t mixed[5]);

/* Pairs the context decided on the host so far: every window passed and
 * the K_phi angle lay within kphi_guard_prad of a bin edge, where the
 * device's atan2 may differ from the reference's glibc atan2 (no counterpart
 * in the reference). */
int hbt_deferred_pairs(hbt_ctx *ctx, uint64_t *n);

/* Replaces RandomUtil::Random (Random.h:11-25, Random.cpp:7-14):
 * std::mt19937(seed) with the reference's two distributions.  seed -1
 * draws from std::random_device like the reference. */
int hbt_rng_create(int32_t seed, hbt_rng **out);
int hbt_rng_destroy(hbt_rng *rng);
int32_t hbt_rng_int(hbt_rng *rng);    /* rand_int_uniform() */
double hbt_rng_uniform(hbt_rng *rng); /* rand_uniform() */

/* Draws the mixing plan of one buffer in the reference's order
 * (HBT_correlation.cpp:200-216 for ids, :495-497 for angles): per event,
 * M partner ids (redrawn while equal to the event unless mixed_nev==1),
 * then M angles theta = U*2*pi, returned as cos, sin.  Returns M (>0) or an
 * error code.  number_of_mixed_events <= 0 selects floor(mixed_nev/2)+1. */
int32_t hbt_rng_draw_plan(hbt_rng *rng, int32_t nev, int32_t mixed_nev,
                          int32_t number_of_mixed_events, int32_t *partner_ids,
                          double *cos_sin);

/* Replaces calculate_HBT_correlation_function (HBT_correlation.cpp:177-218)
 * on the buffer(s) already uploaded: psi_ref (when azimuthal) of each group
 * computed at upload over all its particles (:181-183), the mixing plan of
 * each group drawn from rng in group order (:200-216, :495-497), then all
 * same-event and mixed pairs as one batch.  Histograms accumulate across
 * calls like the reference's.  Returns once the batch is launched: the next
 * hbt_upload does its host work while it runs, and hbt_read, hbt_sync,
 * hbt_stage_counts and the next upload wait for it. */
int hbt_calculate_buffer(hbt_ctx *ctx, hbt_rng *rng);

/* ---------------- batched form ----------------
 * A slot may hold G groups at once, each one of the reference's particle
 * buffers (Analysis.cpp:821-833 processes them one after another).  The
 * batch runs every same-event pair of every group and every mixed pair of
 * every event in ONE persistent launch of the fast kernel (a work queue of
 * tiles of uneven size), captured once in a hipGraph and replayed by
 * hbt_launch_batch.  Results equal G consecutive
 * calculate_HBT_correlation_function calls. */

/* hbt_upload for G groups: grp_ev[G+1] are event offsets of the groups
 * (grp_ev[0] = 0, grp_ev[G] = nev).  The mixed list (slot 1), when used,
 * must have the same number of groups. */
int hbt_upload_groups(hbt_ctx *ctx, int32_t slot, int32_t ngroups, const int32_t *grp_ev,
                      int32_t nev, const int64_t *ev_off, const double *const soa[8]);

/* Uploads the mixing plans of all groups and prepares the batch: M_per_group[G]
 * partners per event, partner_ids and cos_sin laid out event after event
 * (each event's M rows), ids relative to the event's own group, as
 * hbt_rng_draw_plan returns them. */
int hbt_prepare_batch(hbt_ctx *ctx, const int32_t *M_per_group, const int32_t *partner_ids,
                      const double *cos_sin);

/* Enqueues the prepared batch on the context's stream (asynchronous).
 * Launches may be queued back to back.  An upload, hbt_prepare_batch,
 * hbt_run_same/mixed, hbt_set_psi_ref or a tuning change issued before
 * hbt_sync first waits for the queued launches, which therefore see the
 * data they were launched on. */
int hbt_launch_batch(hbt_ctx *ctx);

/* Waits for the context's stream; *batch_ms (may be NULL) receives the
 * device time of the last hbt_launch_batch. */
int hbt_sync(hbt_ctx *ctx, double *batch_ms);

/* Pairs enumerated by one launch of the prepared batch (same-event
 * sum N_g(N_g-1)/2, mixed sum N_e * sum N_partner) and its work items. */
int hbt_batch_info(hbt_ctx *ctx, uint64_t *enum_same, uint64_t *enum_mixed, int32_t *nitems);

/* How hbt_launch_batch issues the prepared batch: *graph = 1 when it
 * replays the hipGraph captured by hbt_prepare_batch (counter fill +
 * persistent fast kernel), 0 when it launches directly (plain kernel, or a
 * runtime that refused the capture, e.g. under a counter-collecting
 * profiler).  HBT_ERR_STATE without a prepared batch.  No counterpart in the
 * reference (its loops run on the host). */
int hbt_batch_graph(hbt_ctx *ctx, int32_t *graph);

/* Work queues of the prepared batch: *nqueues = 1, 2, 4 or 8 (one per XCD at
 * most; every item of a group in one queue), qoff[0..*nqueues] (when not
 * NULL, room for 9) the queues' first items.  HBT_ERR_STATE without a
 * prepared batch.  No counterpart in the reference. */
int hbt_batch_queues(hbt_ctx *ctx, int32_t *nqueues, int32_t *qoff);

/* Slab binning of the last fast launch (hbt_config.binning): *on = 1 when
 * its binned same-event pairs went through slab records, *direct = how many
 * of them took the direct atomic path instead (their (block, slab) region
 * was full, or their stored values round across a bin edge), *capacity =
 * records the regions held.  No counterpart in the reference. */
int hbt_slab_info(hbt_ctx *ctx, int32_t *on, uint64_t *direct, uint64_t *capacity);

/* ---- allocation-free entry point ------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same-event
 * and mixed-event pairs of one buffer with the plain FP64 kernel -- what
 * hbt_upload + hbt_run_same + hbt_run_mixed do for a context created with
 * HBT_KERNEL_PLAIN, HBT_PREC_FP64 -- with every array in device memory owned
 * by the caller, launched on the caller's stream, and no allocation:
 *   ev_off[nev+1] (int64, device): event offsets into the soa columns;
 *   soa[8] (host array of device pointers): px py pz E x y z t of the
 *     buffer, already filtered to one species; the tanh(HBTrap) rapidity cut
 *     is applied here, keeping the reference's gather order;
 *   n_particles = ev_off[nev] - ev_off[0], max_event >= the largest event
 *     (they size the grids);
 *   partner_ids[nev*M] (int32, device, ids of events of this buffer, which
 *     is also the mixed-event list), cos_sin[2*nev*M] (FP64, device): the
 *     mixing plan as hbt_rng_draw_plan returns it;
 *   bins (device): the histogram, n_KT * qnpts^3 bins of 8 doubles
 *     {count (an exact FP64 integer), sum q_out, sum q_side, sum q_long,
 *     sum cos, den (u64 bit pattern), -, -}, bin [kt][iq_out][iq_side]
 *     [iq_long] (iq_long fastest); pairs are added to what it holds;
 *   stream: a hipStream_t (NULL: the null stream).
 * Limits: azimuthal_flag = invariant_radius_flag = 0, the pair cap off
 * (needed_number_of_pairs >= 1e18), kernel HBT_KERNEL_PLAIN, precision
 * HBT_PREC_FP64; at most HBT_NOALLOC_MAX_PARTICLES particles and
 * HBT_NOALLOC_MAX_EVENTS events (HBT_ERR_CAPACITY beyond).  The kept
 * particles, their offsets and the stage counters live in a fixed device
 * array of about 16 MB per GPU, held until the process ends and shared by
 * every call on that GPU: calls that overlap on one GPU race.  Returns once
 * the work is queued; the histogram is complete when the stream is.  Errors:
 * hbt_noalloc_last_error(). */
#define HBT_NOALLOC_MAX_PARTICLES (1 << 18)
#define HBT_NOALLOC_MAX_EVENTS (1 << 14)
int hbt_calculate_noalloc(const hbt_config *cfg, int32_t nev, int64_t n_particles,
                          int64_t max_event, const int64_t *ev_off, const double *const soa[8],
                          int32_t M, const int32_t *partner_ids, const double *cos_sin,
                          double *bins, void *stream);
const char *hbt_noalloc_last_error(void);

/* Launch shape of the fast kernel: R i-particles per lane (1, 2, 4), j
 * chunk of same-event tiles, partner windows per mixed work item (one per
 * plan row whose partner has particles in the tile's |pT| window),
 * persistent blocks per CU; ms is the measured batch time when set by
 * hbt_autotune. */
typedef struct {
    int32_t R;
    int32_t jchunk;
    int32_t kchunk;
    int32_t blocks_per_cu;
    double ms;
} hbt_tuning;

int hbt_get_tuning(hbt_ctx *ctx, hbt_tuning *t);
int hbt_set_tuning(hbt_ctx *ctx, const hbt_tuning *t);

/* Autotuner: times the prepared batch over a grid of launch shapes (one
 * launch each after a warm-up), times its four fastest again (five launches
 * each) and keeps the lowest median (written to *best when not NULL, ms =
 * that median).  Clears the histograms. */
int hbt_autotune(hbt_ctx *ctx, hbt_tuning *best);

/* ---------------- request server (continuous batching) ----------------
 *
 * Serves independent HBT analyses as they arrive.  The reference runs one
 * analysis per process (Analysis::HBTAnalysis, Analysis.cpp:817-835, one
 * HBT_correlation per run); a request here is that whole run: its events,
 * grouped into buffers of events_per_group consecutive events (what the
 * reference's event_buffer_size produces), each buffer processed as by
 * calculate_HBT_correlation_function (HBT_correlation.cpp:177-218) with the
 * mixing plan drawn from RandomUtil::Random(seed) buffer after buffer.  The
 * server shares one configuration (q grid, K_T bins, flags, precision) among
 * its requests.
 *
 * Requests are admitted, in arrival order among those that fit, when their
 * HBM footprint (particles, plan, histograms) fits the device arena and a
 * resident slot is free; their particle records and histograms then stay in
 * HBM until the result is read back.  A large request goes in parts of whole
 * groups (part_particles): each part is admitted, uploaded and launched as
 * soon as it is prepared, while the next one is prepared, all adding to the
 * request's histograms.  Every launch of the fast kernel
 * carries work items of all resident requests (each with its own view of
 * particles and histograms), a bounded number of pairs per launch, shared
 * among them smallest-remaining-first, so requests that arrive while others
 * run join the next launch.  Results equal those of the same analysis run
 * alone (counts bit-identical; sums to FP64 rounding of the addition order).
 *
 * Calls are thread-safe.  A finished request keeps its result slot until
 * hbt_server_release. */
typedef struct hbt_server hbt_server;

typedef struct {
    int64_t hbm_bytes;       /* device arena for resident requests */
    int32_t max_resident;    /* requests resident at once (admitted, not released) */
    int32_t max_step_items;  /* work items per launch */
    double step_pairs;       /* enumerated pairs per launch (scheduling quantum) */
    hbt_tuning tune;         /* launch shape of the fast kernel */
    /* a request is prepared, uploaded and launched in parts of whole groups
     * of about this many input particles, which overlap (0: 1,000,000) */
    int32_t part_particles;
    int32_t reserved[7];
} hbt_server_config;

enum { HBT_REQ_QUEUED = 0, HBT_REQ_LOADING = 1, HBT_REQ_ACTIVE = 2, HBT_REQ_DRAINING = 3,
       HBT_REQ_DONE = 4, HBT_REQ_FAILED = 5 };

typedef struct {
    int32_t state;            /* HBT_REQ_* */
    int32_t nev;              /* events */
    int64_t n_kept;           /* particles kept by the rapidity cut */
    uint64_t enum_same;       /* enumerated same-event pairs */
    uint64_t enum_mixed;      /* enumerated mixed pairs */
    int64_t footprint;        /* bytes of HBM while resident */
    double t_submit;          /* seconds since hbt_server_create: submit called */
    double t_queued;          /* host preparation done, queued for admission */
    double t_admit;           /* admitted (HBM allocated) */
    double t_active;          /* particles resident, first launch possible */
    double t_first_launch;    /* first launch carrying its work */
    double t_last_launch;     /* launch carrying its last work items */
    double t_done;            /* result in host memory */
} hbt_request_info;

typedef struct {
    int64_t submitted, completed, failed;
    int64_t launches;         /* kernel launches */
    int64_t items;            /* work items launched */
    double pairs;             /* enumerated pairs launched */
    int64_t arena_bytes, arena_used, arena_peak;
    int32_t resident, resident_peak, queued;
    int32_t deferred;  /* pairs decided on the host (K_phi at a bin edge) */
    double busy_s;            /* summed device time of the launches (HIP events) */
} hbt_server_stats;

/* cfg must select the fast kernel (either precision); sc may be NULL for
 * defaults (hbm_bytes 16 GiB, max_resident 16, max_step_items 262144,
 * step_pairs 2e9, tune {2, 1024, 4, 8}). */
int hbt_server_create(const hbt_config *cfg, const hbt_server_config *sc, hbt_server **out);
/* Stops the server: queued requests are dropped, launched work finishes. */
int hbt_server_destroy(hbt_server *srv);
const char *hbt_server_last_error(const hbt_server *srv);

/* Copies one analysis (SoA {px,py,pz,E,x,y,z,t} of nev events, offsets
 * nev+1) and returns at once with its id.  Fails with HBT_ERR_CAPACITY when
 * it can never fit the arena or when the pair cap could bind. */
int hbt_server_submit(hbt_server *srv, int32_t nev, const int64_t *ev_off,
                      const double *const soa[8], int32_t events_per_group, int32_t seed,
                      int64_t *req_id);
int hbt_server_info(hbt_server *srv, int64_t req_id, hbt_request_info *info);
/* Waits until the request is done or failed, at most timeout_s seconds
 * (negative: no limit).  HBT_OK when done, HBT_ERR_STATE on timeout. */
int hbt_server_wait(hbt_server *srv, int64_t req_id, double timeout_s);
/* Histograms and counters of a finished request (fields may be NULL). */
int hbt_server_result(hbt_server *srv, int64_t req_id, hbt_result *out);
/* Frees a finished or failed request and its result slot. */
int hbt_server_release(hbt_server *srv, int64_t req_id);
int hbt_server_get_stats(hbt_server *srv, hbt_server_stats *st);

/* ---------------- host-side pieces (no GPU needed) ---------------- */

/* Replaces output_HBTcorrelation and the three writers it calls
 * (HBT_correlation.cpp:220-228, 694-855): writes
 * HBT_correlation_function[_inv]_KT_<a>_<b>[_Kphi_<c>].dat into path from
 * raw results (hbt_read), byte-identical to the reference's files for the
 * same values.  eco_output = the reference's ecoOutput key.  what selects
 * the writers: HBT_WRITE_INV (output_correlation_function_inv),
 * HBT_WRITE_3D (output_correlation_function), HBT_WRITE_KPHI
 * (output_correlation_function_Kphi_differential), or 0 for what
 * output_HBTcorrelation writes under the configuration's flags. */
enum { HBT_WRITE_INV = 1, HBT_WRITE_3D = 2, HBT_WRITE_KPHI = 4 };
int hbt_write_outputs(const hbt_config *cfg, const hbt_result *r, const char *path,
                      int32_t eco_output, int32_t what);

/* Replaces HBT_correlation::calculate_flow_event_plane_angle
 * (HBT_correlation.cpp:233-249) on n particles (px, py). */
double hbt_flow_event_plane_angle(int32_t n, const double *px, const double *py,
                                  int32_t n_order);

/* Replaces particleSamples' mode-10 reader (particleSamples.cpp:1247-1286,
 * gz_readline :2209-2218) with the species filter (:1329-1346): each
 * hbt_reader_next reads whole events while the running particle total (all
 * species) is below event_buffer_size and returns the number of events
 * (possibly 0 at end of file); hbt_reader_buffer exposes the kept
 * particles (SoA {px,py,pz,E,x,y,z,t}, event offsets) until the next call.
 * particle_monval selects one Monte-Carlo id, 99999 (all species) or a
 * species class (9999, +-9998, +-9997, +-9996, decide_to_pick_OSCAR,
 * particleSamples.cpp:625-678), which needs hbt_reader_set_species; then a
 * particle whose monval is not in the table makes hbt_reader_next return
 * HBT_ERR_ARG (the reference's lookup throws), and hbt_reader_bad_monval
 * names it.  Numbers
 * are read as strtod reads them (bit-identical values); lines are parsed on
 * several host threads. */
typedef struct hbt_reader hbt_reader;
int hbt_reader_open(const char *filename, int32_t event_buffer_size, int32_t particle_monval,
                    hbt_reader **out);
/* The same for read_in_mode 10 (gzip or plain text, as above) or 9
 * (particleSamples::read_in_particle_samples_binary, particleSamples.cpp:
 * 1203-1245: per event an int count, then per particle an int Monte-Carlo id
 * and 9 floats mass t x y z E px py pz, widened to double).  The reference's
 * mode-9 reader files every particle under event 0; this one keeps the
 * file's events. */
int hbt_reader_open_mode(const char *filename, int32_t read_in_mode, int32_t event_buffer_size,
                         int32_t particle_monval, hbt_reader **out);
int32_t hbt_reader_next(hbt_reader *rd);
int32_t hbt_reader_eof(const hbt_reader *rd);
int hbt_reader_buffer(const hbt_reader *rd, const int64_t **ev_off, const double *soa[8],
                      const double **mass, int64_t *total_read);
int hbt_reader_close(hbt_reader *rd);
/* The species table (hbt_species_read, below) for a species-class
 * particle_monval: the particle_decay lookups of particleSamples::
 * decide_to_pick_OSCAR; copied. */
typedef struct hbt_species_table hbt_species_table;
int hbt_reader_set_species(hbt_reader *rd, const hbt_species_table *t);
int64_t hbt_reader_bad_monval(const hbt_reader *rd);
/* The Monte-Carlo id of each particle of the last buffer (particle_info::
 * monval, particleSamples.cpp:1275), valid until the next call. */
int hbt_reader_monval(const hbt_reader *rd, const int32_t **monval);

/* Wall time (ms, HIP events on the context's stream) of the last
 * hbt_run_same / hbt_run_mixed kernel launches. */
int hbt_last_kernel_ms(hbt_ctx *ctx, double *same_ms, double *mixed_ms);

/* Number of particles kept by the rapidity cut in a slot, and of events. */
int hbt_slot_info(hbt_ctx *ctx, int32_t slot, int32_t *nev, int64_t *n_kept);

/* ---------------- balance function (BalanceFunction.cpp) ----------------
 * The pair histograms of BalanceFunction (BalanceFunction.cpp:15-197) on the
 * GPU: for a = particle_alpha, b = particle_beta and their antiparticles
 * (ids -a, -b), same-event (a,b), (abar,bbar), (a,bbar), (abar,b) and the
 * four mixed-event histograms, each [Bnpts][20] over (Delta y or Delta eta,
 * Delta phi), counts in doubles as the reference keeps them.  Particles
 * outside [BpT_min, BpT_max] are skipped.  An id names one Monte-Carlo id,
 * 99999 (all species) or a species class of particleSamples::
 * decide_to_pick_OSCAR (particleSamples.cpp:625-678): 9999 charged, 9998 /
 * -9998 positive / negative, 9997 / -9997 baryons / anti-baryons, 9996 /
 * -9996 positive / negative strangeness; the antiparticle list of an id is
 * that of -id.  Classes need the species table (hbt_bf_set_species). */
typedef struct {
    int32_t particle_alpha, particle_beta;
    double BpT_min, BpT_max, Brap_max;
    int32_t Bnpts;     /* Delta y bins, 2..817 */
    int32_t rap_type;  /* 0: pseudo-rapidity, 1: rapidity */
    int32_t device;
    int32_t kernel;    /* HBT_BF_KERNEL_FAST (default) or HBT_BF_KERNEL_EXACT */
} hbt_bf_config;

/* Per-pair forms of the balance-function kernel; both give the reference's
 * counts.  EXACT evaluates the reference's FP64 bin arithmetic for every
 * pair; FAST takes the bins from FP32 differences of per-particle values in
 * bin units and sends each pair within a guard band of a bin edge (or of
 * Delta y = 0) to the FP64 arithmetic. */
#define HBT_BF_KERNEL_FAST 0
#define HBT_BF_KERNEL_EXACT 1

/* What the last hbt_bf_calculate did: host_ms = species lists, draws and
 * uploads on the host; prep_ms / kernel_ms = the per-particle preparation
 * and the pair kernels (HIP events on the context's stream); pairs =
 * enumerated pairs over the eight histograms; fallback = pairs of the FAST
 * form that took the FP64 arithmetic; kernel = the form that ran; launches
 * = pair-kernel launches (each below 2^32 pairs per histogram). */
typedef struct {
    double host_ms, prep_ms, kernel_ms;
    uint64_t pairs, fallback;
    int32_t kernel, launches;
} hbt_bf_info;

typedef struct hbt_bf_ctx hbt_bf_ctx;
/* Replaces the BalanceFunction constructor (BalanceFunction.cpp:15-61). */
int hbt_bf_create(const hbt_bf_config *cfg, hbt_bf_ctx **out);
int hbt_bf_destroy(hbt_bf_ctx *ctx);
const char *hbt_bf_last_error(const hbt_bf_ctx *ctx);
/* Zeroes the histograms and the N_b / N_bbar counts. */
int hbt_bf_reset(hbt_bf_ctx *ctx);
/* Replaces calculate_balance_function on one buffer (BalanceFunction.cpp:
 * 63-101): nev events (ev_off nev+1) with all species, cols = {mass, px,
 * py, pz, E}; the mixed list is a separate buffer or, with mix_off NULL,
 * the same one (particleSamples.cpp:528-531).  The mixing draws come from
 * rng in the reference's order: per mixed histogram, per event a partner
 * rand_int_uniform() % mix_nev and a rotation 2 pi rand_uniform(). */
int hbt_bf_calculate(hbt_bf_ctx *ctx, int32_t nev, const int64_t *ev_off,
                     const int32_t *monval, const double *const cols[5], int32_t mix_nev,
                     const int64_t *mix_off, const int32_t *mix_monval,
                     const double *const mix_cols[5], hbt_rng *rng);
/* The same with the mixing draws made by the caller from its own stream (the
 * reference's shared RandomUtil::Random): mix_ev[h * nev + e] and
 * mix_rot[h * nev + e] for mixed histogram h = 0..3 (C_ab, C_abarbbar,
 * C_abbar, C_abarb) and event e, drawn in that order as
 * rand_int_uniform() % mix_nev and rand_uniform() * 2. * M_PI
 * (BalanceFunction.cpp:165-169). */
int hbt_bf_calculate_planned(hbt_bf_ctx *ctx, int32_t nev, const int64_t *ev_off,
                             const int32_t *monval, const double *const cols[5],
                             int32_t mix_nev, const int64_t *mix_off,
                             const int32_t *mix_monval, const double *const mix_cols[5],
                             const int32_t *mix_ev, const double *mix_rot);
/* hist: 8 * Bnpts * 20 doubles (C_ab, C_abarbbar, C_abbar, C_abarb, then the
 * mixed four), N_b / N_bbar: b / bbar particles inside the pT cut. */
int hbt_bf_read(hbt_bf_ctx *ctx, double *hist, int64_t *N_b, int64_t *N_bbar);
/* Timing and counts of the last hbt_bf_calculate (no counterpart in the
 * reference, which has no timing). */
int hbt_bf_last_info(hbt_bf_ctx *ctx, hbt_bf_info *info);
/* The charge, baryon number and strangeness of n species, for the species
 * classes; replaces the particle_decay lookups of decide_to_pick_OSCAR
 * (particle_decay.cpp:181-196).  A particle whose monval is not in the table
 * makes hbt_bf_calculate fail when a class is evaluated for it, where the
 * reference's map lookup throws. */
int hbt_bf_set_species(hbt_bf_ctx *ctx, int32_t n, const int32_t *monval, const int32_t *charge,
                       const int32_t *baryon, const int32_t *strange);

/* ---------------- species table (particle_decay.cpp:23-129) ----------------
 * A PDG table file in the reference's format (EOS/pdg.dat): per species
 * "monval name mass width gspin baryon strange charm bottom gisospin charge
 * decays" and one 8-number record per decay channel; every baryon (baryon
 * number 1) gets its anti-baryon entry, as the reference adds it. */
int hbt_species_read(const char *path, hbt_species_table **out);
int32_t hbt_species_count(const hbt_species_table *t);
int hbt_species_get(const hbt_species_table *t, int32_t i, int32_t *monval, int32_t *charge,
                    int32_t *baryon, int32_t *strange);
int hbt_species_free(hbt_species_table *t);
/* Replaces output_balance_function (BalanceFunction.cpp:211-318): writes
 * Balance_function_<a>_<b>_Delta_y.dat, _Delta_phi.dat and
 * Correlation_function_<a>_<b>_2D.dat into path. */
int hbt_bf_write_outputs(const hbt_bf_config *cfg, const double *hist, const char *path);

#ifdef __cplusplus
}
#endif
#endif
