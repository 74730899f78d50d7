/*
 * hbt_oracle.c -- CPU oracle (test infrastructure only) for the HBT
 * pair-correlation path.  Plain C restatement of the reference; every
 * function cites the reference lines it follows.  Only tests/, the smoke
 * check and bench.py's CPU-baseline leg may load it, and only to check results against.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math: strict
 * IEEE double, no contraction, like the reference's g++ -O3 on x86-64).
 *
 * Pinning (what checks each part against something other than itself):
 *  - orc_combine_same / orc_combine_mixed / orc_calculate_buffer:
 *      bit-exact (every count, every FP64 sum) against golden dumps of the
 *      reference's own HBT_correlation compiled from its sources by
 *      oracle/Makefile (tests/golden/make_golden.py; checked by
 *      tests/test_oracle_golden.py): the reference fixture, seeded synthetic
 *      buffers, binding caps, and hand-placed pairs exactly on the window
 *      edges lo+1e-8 / hi-1e-8, one ulp around them, on interior bin edges,
 *      at K_perp = KT_min / KT_max and on K_phi bin edges
 *      (tests/golden/edge_inputs.py).  The mixed comparisons (>= on hi),
 *      which the reference reaches only through random rotations, are pinned
 *      pair by pair on the same designed pairs against the reference's
 *      rules (tests/test_edge_known_answers.py).
 *  - orc_flow_event_plane_angle: HBT_unittest.cc:39,43,47 golden psi_ref
 *      (tests/test_oracle_golden.py).
 *  - orc_rng_* / orc_draw_mixing_plan: first draws from seed 0 against the
 *      values the reference's libstdc++ produces (1178568022,
 *      0.71518936510962539) and draw for draw against std::mt19937 +
 *      std::uniform_*_distribution compiled by g++ (tests/test_rng_plan.py);
 *      whole plans pinned indirectly through the mixed-event histograms of
 *      the reference runs.
 *  - orc_reader_*: particleSamples_unittest.cc:199-206 counts (619/635 pi+)
 *      on the reference fixture, and the reference's buffer sizes on every
 *      golden case (tests/test_oracle_golden.py).
 *  - orc_write_outputs: sha256 of every file the reference wrote on the
 *      golden cases (tests/test_host_io.py).
 */
#define _GNU_SOURCE
#include "hbt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static const double ORC_HBARC = 0.197327053; /* parameters.h:4 */

/* HBT_correlation.cpp:25-61, :253-256, :289-290; the constant expressions
 * are evaluated exactly as the reference writes them. */
void orc_derive(const orc_params *p, orc_derived *d) {
    d->delta_q = (p->q_max - p->q_min) / (p->qnpts - 1);
    d->dKT = (p->KT_max - p->KT_min) / (p->n_KT - 1);
    d->dKphi = 2 * M_PI / p->n_Kphi;
    d->q_base = p->q_min - d->delta_q / 2.;
    d->q_lo_cut = p->q_min - d->delta_q / 2. + 1e-8;
    d->q_hi_cut = p->q_max + d->delta_q / 2. - 1e-8;
    d->KT_min_sq = p->KT_min * p->KT_min;
    d->KT_max_sq = p->KT_max * p->KT_max;
    d->hbarC_inv = 1. / ORC_HBARC;
    d->rap_cut_max = tanh(p->HBTrap_max);
    d->rap_cut_min = tanh(p->HBTrap_min);
    d->needed = (uint64_t)(unsigned long long)p->needed_number_of_pairs;
    d->needed_qinv = (uint64_t)50 * d->needed; /* wraps like ull in :341 */
}

int orc_init(orc_state *s, const orc_params *p) {
    memset(s, 0, sizeof(*s));
    s->p = *p;
    orc_derive(p, &s->d);
    s->nphi = p->azimuthal_flag ? p->n_Kphi : 1;
    long q = p->qnpts;
    s->nbins3d = (long)p->n_KT * s->nphi * q * q * q;
    double **arrs[] = {&s->num_count, &s->q_out_sum, &s->q_side_sum,
                       &s->q_long_sum, &s->num, &s->den, &s->abs_num,
                       &s->abs_q_out, &s->abs_q_side, &s->abs_q_long};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) {
        *arrs[k] = (double *)calloc((size_t)s->nbins3d, sizeof(double));
        if (!*arrs[k]) return -1;
    }
    long nq = (long)p->n_KT * q;
    double **qa[] = {&s->qinv_sum, &s->qinv_num, &s->qinv_count, &s->qinv_den};
    for (size_t k = 0; k < 4; k++) {
        *qa[k] = (double *)calloc((size_t)nq, sizeof(double));
        if (!*qa[k]) return -1;
    }
    s->cnt_num_kt = (uint64_t *)calloc(p->n_KT, 8);
    s->cnt_den_kt = (uint64_t *)calloc(p->n_KT, 8);
    s->cnt_num_qinv = (uint64_t *)calloc(p->n_KT, 8);
    s->cnt_den_qinv = (uint64_t *)calloc(p->n_KT, 8);
    s->cnt_num_ktphi = (uint64_t *)calloc((size_t)p->n_KT * p->n_Kphi, 8);
    s->cnt_den_ktphi = (uint64_t *)calloc((size_t)p->n_KT * p->n_Kphi, 8);
    return 0;
}

void orc_free(orc_state *s) {
    double *arrs[] = {s->num_count, s->q_out_sum, s->q_side_sum,
                      s->q_long_sum, s->num, s->den, s->abs_num,
                      s->abs_q_out, s->abs_q_side, s->abs_q_long,
                      s->qinv_sum, s->qinv_num, s->qinv_count, s->qinv_den};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) free(arrs[k]);
    free(s->cnt_num_kt); free(s->cnt_den_kt);
    free(s->cnt_num_qinv); free(s->cnt_den_qinv);
    free(s->cnt_num_ktphi); free(s->cnt_den_ktphi);
    memset(s, 0, sizeof(*s));
}

/* HBT_correlation.cpp:233-249 */
double orc_flow_event_plane_angle(const orc_buffer *b, int n_order) {
    double vn_real = 0.0, vn_imag = 0.0;
    for (int iev = 0; iev < b->nev; iev++) {
        for (int64_t i = b->ev_off[iev]; i < b->ev_off[iev + 1]; i++) {
            double temp_phi = atan2(b->py[i], b->px[i]);
            /* g++ -O3 merges the reference's cos()/sin() of one argument
             * into glibc sincos(); spelled out here so the result does not
             * depend on the compiler doing the same */
            double sn, cn;
            sincos(n_order * temp_phi, &sn, &cn);
            vn_real += cn;
            vn_imag += sn;
        }
    }
    return atan2(vn_imag, vn_real) / n_order;
}

typedef struct {
    long n, cap;
    double *px, *py, *pz, *E, *x, *y, *z, *t;
} plist;

static void pl_reserve(plist *l, long n) {
    if (n <= l->cap) return;
    long c = l->cap ? l->cap : 256;
    while (c < n) c *= 2;
    double **a[] = {&l->px, &l->py, &l->pz, &l->E, &l->x, &l->y, &l->z, &l->t};
    for (int k = 0; k < 8; k++) *a[k] = (double *)realloc(*a[k], c * sizeof(double));
    l->cap = c;
}
static void pl_push(plist *l, double px, double py, double pz, double E,
                    double x, double y, double z, double t) {
    pl_reserve(l, l->n + 1);
    long k = l->n++;
    l->px[k] = px; l->py[k] = py; l->pz[k] = pz; l->E[k] = E;
    l->x[k] = x; l->y[k] = y; l->z[k] = z; l->t[k] = t;
}
static void pl_free(plist *l) {
    free(l->px); free(l->py); free(l->pz); free(l->E);
    free(l->x); free(l->y); free(l->z); free(l->t);
    memset(l, 0, sizeof(*l));
}

static inline long bin3(const orc_state *s, int kt, int kphi, int io, int is, int il) {
    long q = s->p.qnpts;
    return ((((long)kt * s->nphi + kphi) * q + io) * q + is) * q + il;
}

/* K_phi index with the reference's wrap loops (:408-423 / :657-672).
 * Returns -1 when the pair is skipped. */
static int kphi_index(const orc_state *s, double K_x, double K_y) {
    double local_K_phi = atan2(K_y, K_x);
    double delta_phi = local_K_phi - s->psi_ref;
    while (delta_phi < 0.) delta_phi += 2. * M_PI;
    while (delta_phi > 2. * M_PI) delta_phi -= 2. * M_PI;
    int idx = (int)(delta_phi / s->d.dKphi);
    if (idx < 0 || idx >= s->p.n_Kphi) return -1;
    return idx;
}

/* Same-event pairs: gather with rapidity cut (HBT_correlation.cpp:253-281),
 * then the i<j loop (:289-460). */
void orc_combine_same(orc_state *s, const orc_buffer *b) {
    const orc_derived *d = &s->d;
    const orc_params *p = &s->p;
    plist L = {0};
    for (int ev = 0; ev < b->nev; ev++) {
        for (int64_t i = b->ev_off[ev]; i < b->ev_off[ev + 1]; i++) {
            double temp_ratio = b->pz[i] / b->E[i];
            if (temp_ratio > d->rap_cut_min && temp_ratio < d->rap_cut_max)
                pl_push(&L, b->px[i], b->py[i], b->pz[i], b->E[i], b->x[i],
                        b->y[i], b->z[i], b->t[i]);
        }
    }
    const long n = L.n;
    const int qnpts = p->qnpts;
    for (long i = 0; i < n; i++) {
        const double p1x = L.px[i], p1y = L.py[i], p1z = L.pz[i], p1E = L.E[i];
        const double x1 = L.x[i], y1 = L.y[i], z1 = L.z[i], t1 = L.t[i];
        for (long j = i + 1; j < n; j++) {
            const double p2x = L.px[j], p2y = L.py[j], p2z = L.pz[j], p2E = L.E[j];
            s->same_st.enumerated++;
            double K_z = 0.5 * (p1z + p2z);
            double K_E = 0.5 * (p1E + p2E);
            double K_z_over_K_E = K_z / K_E;
            double K_x = 0.5 * (p1x + p2x);
            double K_y = 0.5 * (p1y + p2y);
            double K_perp_sq = K_x * K_x + K_y * K_y;
            if (K_perp_sq < d->KT_min_sq || K_perp_sq > d->KT_max_sq) continue;
            s->same_st.kt++;
            double K_perp = sqrt(K_perp_sq);
            int Kperp_idx = (int)((K_perp - p->KT_min) / d->dKT);
            double q_x = p1x - p2x;
            double q_y = p1y - p2y;
            double q_z = p1z - p2z;
            double q_E = p1E - p2E;
            double t_diff = t1 - L.t[j];
            double x_diff = x1 - L.x[j];
            double y_diff = y1 - L.y[j];
            double z_diff = z1 - L.z[j];
            if (p->invariant_radius_flag == 1) { /* :339-356 */
                double local_q_inv = sqrt(-(q_E * q_E - q_x * q_x - q_y * q_y - q_z * q_z));
                if (s->cnt_num_qinv[Kperp_idx] < d->needed_qinv) {
                    if (local_q_inv > d->q_lo_cut && local_q_inv < d->q_hi_cut) {
                        int qinv_idx = (int)((local_q_inv - d->q_base) / d->delta_q);
                        /* the reference has no upper check here; an index
                         * of qnpts cannot occur for dq >> 1e-8 */
                        if (qinv_idx < qnpts) {
                            s->cnt_num_qinv[Kperp_idx]++;
                            double cos_qx = cos(d->hbarC_inv *
                                (q_E * t_diff - q_x * x_diff - q_y * y_diff - q_z * z_diff));
                            long k = (long)Kperp_idx * qnpts + qinv_idx;
                            s->qinv_count[k]++;
                            s->qinv_sum[k] += local_q_inv;
                            s->qinv_num[k] += cos_qx;
                        }
                    }
                }
            }
            double cos_K_phi = K_x / K_perp;
            double sin_K_phi = K_y / K_perp;
            double local_q_out = q_x * cos_K_phi + q_y * sin_K_phi;
            if (local_q_out < d->q_lo_cut || local_q_out > d->q_hi_cut) continue;
            int qout_idx = (int)((local_q_out - d->q_base) / d->delta_q);
            if (qout_idx >= qnpts) continue;
            s->same_st.out++;
            double local_q_side = q_y * cos_K_phi - q_x * sin_K_phi;
            if (local_q_side < d->q_lo_cut || local_q_side > d->q_hi_cut) continue;
            int qside_idx = (int)((local_q_side - d->q_base) / d->delta_q);
            if (qside_idx >= qnpts) continue;
            s->same_st.side++;
            double local_q_long = q_z;
            if (p->long_comoving_boost) {
                double Mt = sqrt(K_E * K_E - K_z * K_z);
                double boost_gamma = K_E / Mt;
                double boost_beta = K_z_over_K_E;
                local_q_long = boost_gamma * (q_z - boost_beta * q_E);
            }
            if (local_q_long < d->q_lo_cut || local_q_long > d->q_hi_cut) continue;
            int qlong_idx = (int)((local_q_long - d->q_base) / d->delta_q);
            if (qlong_idx >= qnpts) continue;
            int Kphi_idx = 0;
            if (p->azimuthal_flag == 0) {
                if (s->cnt_num_kt[Kperp_idx] > d->needed) continue;
                s->cnt_num_kt[Kperp_idx]++;
            } else {
                Kphi_idx = kphi_index(s, K_x, K_y);
                if (Kphi_idx < 0) { s->kphi_skipped++; continue; }
                uint64_t *c = &s->cnt_num_ktphi[(long)Kperp_idx * p->n_Kphi + Kphi_idx];
                if (*c > d->needed) continue;
                (*c)++;
            }
            s->same_st.binned++;
            double cos_qx = cos(d->hbarC_inv *
                (q_E * t_diff - q_x * x_diff - q_y * y_diff - q_z * z_diff));
            long k = bin3(s, Kperp_idx, Kphi_idx, qout_idx, qside_idx, qlong_idx);
            s->num_count[k]++;
            s->q_out_sum[k] += local_q_out;
            s->q_side_sum[k] += local_q_side;
            s->q_long_sum[k] += local_q_long;
            s->num[k] += cos_qx;
            s->abs_num[k] += fabs(cos_qx);
            s->abs_q_out[k] += fabs(local_q_out);
            s->abs_q_side[k] += fabs(local_q_side);
            s->abs_q_long[k] += fabs(local_q_long);
        }
    }
    pl_free(&L);
}

/* Mixed pairs of event iev (HBT_correlation.cpp:464-692). */
void orc_combine_mixed(orc_state *s, const orc_buffer *b, int iev,
                       const orc_buffer *mix, int M, const int *partner_ids,
                       const double *cs) {
    const orc_derived *d = &s->d;
    const orc_params *p = &s->p;
    plist L1 = {0}, L2 = {0};
    for (int64_t i = b->ev_off[iev]; i < b->ev_off[iev + 1]; i++) { /* :471-489 */
        double temp_ratio = b->pz[i] / b->E[i];
        if (temp_ratio > d->rap_cut_min && temp_ratio < d->rap_cut_max)
            pl_push(&L1, b->px[i], b->py[i], b->pz[i], b->E[i], 0, 0, 0, 0);
    }
    for (int k = 0; k < M; k++) { /* :493-552 */
        const double cos_phi = cs[2 * k], sin_phi = cs[2 * k + 1];
        int id = partner_ids[k];
        for (int64_t i = mix->ev_off[id]; i < mix->ev_off[id + 1]; i++) {
            double temp_ratio = mix->pz[i] / mix->E[i];
            if (temp_ratio > d->rap_cut_min && temp_ratio < d->rap_cut_max) {
                double px_temp = mix->px[i], py_temp = mix->py[i];
                pl_push(&L2, px_temp * cos_phi - py_temp * sin_phi,
                        px_temp * sin_phi + py_temp * cos_phi, mix->pz[i],
                        mix->E[i], 0, 0, 0, 0);
            }
        }
    }
    const int qnpts = p->qnpts;
    for (long i = 0; i < L1.n; i++) {
        const double p1z = L1.pz[i], p1x = L1.px[i], p1y = L1.py[i], p1E = L1.E[i];
        for (long j = 0; j < L2.n; j++) {
            const double p2z = L2.pz[j], p2x = L2.px[j], p2y = L2.py[j], p2E = L2.E[j];
            s->mixed_st.enumerated++;
            double K_z = 0.5 * (p1z + p2z);
            double K_E = 0.5 * (p1E + p2E);
            double K_z_over_K_E = K_z / K_E;
            double K_x = 0.5 * (p1x + p2x);
            double K_y = 0.5 * (p1y + p2y);
            double K_perp_sq = K_x * K_x + K_y * K_y;
            if (K_perp_sq < d->KT_min_sq || K_perp_sq > d->KT_max_sq) continue;
            s->mixed_st.kt++;
            double K_perp = sqrt(K_perp_sq);
            int Kperp_idx = (int)((K_perp - p->KT_min) / d->dKT);
            double q_z = p1z - p2z;
            double q_E = p1E - p2E;
            double q_x = p1x - p2x;
            double q_y = p1y - p2y;
            if (p->invariant_radius_flag == 1) { /* :592-607 */
                double local_q_inv = sqrt(-(q_E * q_E - q_x * q_x - q_y * q_y - q_z * q_z));
                if (local_q_inv > d->q_lo_cut && local_q_inv < d->q_hi_cut) {
                    int qinv_idx = (int)((local_q_inv - d->q_base) / d->delta_q);
                    if (qinv_idx < qnpts && s->cnt_den_qinv[Kperp_idx] < d->needed_qinv) {
                        s->cnt_den_qinv[Kperp_idx]++;
                        s->qinv_den[(long)Kperp_idx * qnpts + qinv_idx] += 1.0;
                    }
                }
            }
            double cos_K_phi = K_x / K_perp;
            double sin_K_phi = K_y / K_perp;
            double local_q_out = q_x * cos_K_phi + q_y * sin_K_phi;
            if (local_q_out < d->q_lo_cut || local_q_out >= d->q_hi_cut) continue;
            int qout_idx = (int)((local_q_out - d->q_base) / d->delta_q);
            if (qout_idx >= qnpts) continue;
            s->mixed_st.out++;
            double local_q_side = q_y * cos_K_phi - q_x * sin_K_phi;
            if (local_q_side < d->q_lo_cut || local_q_side >= d->q_hi_cut) continue;
            int qside_idx = (int)((local_q_side - d->q_base) / d->delta_q);
            if (qside_idx >= qnpts) continue;
            s->mixed_st.side++;
            double local_q_long = q_z;
            if (p->long_comoving_boost) {
                double Mt = sqrt(K_E * K_E - K_z * K_z);
                double boost_gamma = K_E / Mt;
                double boost_beta = K_z_over_K_E;
                local_q_long = boost_gamma * (q_z - boost_beta * q_E);
            }
            if (local_q_long < d->q_lo_cut || local_q_long >= d->q_hi_cut) continue;
            int qlong_idx = (int)((local_q_long - d->q_base) / d->delta_q);
            if (qlong_idx >= qnpts) continue;
            int Kphi_idx = 0;
            if (p->azimuthal_flag == 0) {
                if (s->cnt_den_kt[Kperp_idx] > d->needed) continue;
                s->cnt_den_kt[Kperp_idx]++;
            } else {
                Kphi_idx = kphi_index(s, K_x, K_y);
                if (Kphi_idx < 0) { s->kphi_skipped++; continue; }
                uint64_t *c = &s->cnt_den_ktphi[(long)Kperp_idx * p->n_Kphi + Kphi_idx];
                if (*c > d->needed) continue;
                (*c)++;
            }
            s->mixed_st.binned++;
            s->den[bin3(s, Kperp_idx, Kphi_idx, qout_idx, qside_idx, qlong_idx)] += 1.0;
        }
    }
    pl_free(&L1);
    pl_free(&L2);
}

/* ---------------- random stream ----------------
 * std::mt19937 (Matsumoto & Nishimura 1998, the parameters fixed by
 * [rand.predef]) and libstdc++ 11.4's distributions:
 *  - uniform_int_distribution<int>() over [0, INT_MAX]: bits/uniform_int_dist.h
 *    takes Lemire's nearly-divisionless path for a 32-bit engine
 *    (64-bit product, threshold = -range % range).
 *  - uniform_real_distribution<double>(0,1): generate_canonical<double,53>
 *    with two engine calls: (u1 + u2*2^32)/2^64, clamped below 1
 *    (bits/random.tcc), then *(b-a)+a.
 * Random.h:11-25, Random.cpp:7-14. */
void orc_rng_seed(orc_rng *r, uint32_t seed) {
    r->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->idx = 624;
}

uint32_t orc_rng_next(orc_rng *r) {
    if (r->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
            r->mt[i] = r->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        r->idx = 0;
    }
    uint32_t y = r->mt[r->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

int orc_rand_int_uniform(orc_rng *r) {
    const uint32_t range = 0x80000000u; /* INT_MAX - 0 + 1 */
    uint64_t product = (uint64_t)orc_rng_next(r) * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        uint32_t threshold = (uint32_t)(-range) % range;
        while (low < threshold) {
            product = (uint64_t)orc_rng_next(r) * (uint64_t)range;
            low = (uint32_t)product;
        }
    }
    return (int)(product >> 32);
}

double orc_rand_uniform(orc_rng *r) {
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; k++) {
        sum += (double)orc_rng_next(r) * tmp;
        tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret * (1.0 - 0.0) + 0.0;
}

/* HBT_correlation.cpp:200-216 (ids) interleaved with :495-497 (angles). */
int orc_draw_mixing_plan(orc_rng *r, int nev, int mixed_nev,
                         int number_of_mixed_events, int *ids, double *cs) {
    int M = number_of_mixed_events > 0 ? number_of_mixed_events : mixed_nev / 2 + 1;
    for (int iev = 0; iev < nev; iev++) {
        for (int count = 0; count < M; count++) {
            int id = orc_rand_int_uniform(r) % mixed_nev;
            while (iev == id && mixed_nev != 1) id = orc_rand_int_uniform(r) % mixed_nev;
            ids[(long)iev * M + count] = id;
        }
        for (int k = 0; k < M; k++) {
            const double random_rotation = orc_rand_uniform(r) * 2 * M_PI;
            double sn, cn;
            sincos(random_rotation, &sn, &cn); /* see orc_flow_event_plane_angle */
            cs[2 * ((long)iev * M + k)] = cn;
            cs[2 * ((long)iev * M + k) + 1] = sn;
        }
    }
    return M;
}

/* HBT_correlation.cpp:177-218 */
void orc_calculate_buffer(orc_state *s, orc_rng *r, const orc_buffer *b,
                          const orc_buffer *mix) {
    if (s->p.azimuthal_flag == 1) s->psi_ref = orc_flow_event_plane_angle(b, 2);
    orc_combine_same(s, b);
    int nev = b->nev, mixed_nev = mix->nev;
    if (nev <= 0) return;
    int M = s->p.number_of_mixed_events > 0 ? s->p.number_of_mixed_events
                                            : mixed_nev / 2 + 1;
    int *ids = (int *)malloc(sizeof(int) * (size_t)M);
    double *cs = (double *)malloc(sizeof(double) * 2 * (size_t)M);
    for (int iev = 0; iev < nev; iev++) {
        /* per event: M ids, then M angles inside the mixed call */
        for (int count = 0; count < M; count++) {
            int id = orc_rand_int_uniform(r) % mixed_nev;
            while (iev == id && mixed_nev != 1) id = orc_rand_int_uniform(r) % mixed_nev;
            ids[count] = id;
        }
        for (int k = 0; k < M; k++) {
            const double random_rotation = orc_rand_uniform(r) * 2 * M_PI;
            double sn, cn;
            sincos(random_rotation, &sn, &cn);
            cs[2 * k] = cn;
            cs[2 * k + 1] = sn;
        }
        orc_combine_mixed(s, b, iev, mix, M, ids, cs);
    }
    free(ids);
    free(cs);
}

/* ---------------- writers ----------------
 * Formatting of the reference's iostreams: `std::scientific <<
 * std::setw(18) << std::setprecision(8)` on the first field only, then
 * "    " separators (printf "%18.8e" / "%.8e"); file names print KT and
 * K_phi values with the default ostream format ("%g"). */
static int write_3d(const orc_state *s, const char *path, int eco) { /* :726-783 */
    const orc_params *p = &s->p;
    const int q = p->qnpts;
    const double dq = s->d.delta_q;
    for (int iK = 0; iK < p->n_KT - 1; iK++) {
        double npair_ratio = (double)s->cnt_num_kt[iK] / (double)s->cnt_den_kt[iK];
        char fn[4096];
        snprintf(fn, sizeof(fn), "%s/HBT_correlation_function_KT_%g_%g.dat", path,
                 p->KT_min + iK * s->d.dKT, p->KT_min + (iK + 1) * s->d.dKT);
        FILE *f = fopen(fn, "w");
        if (!f) return -1;
        for (int il = 0; il < q; il++)
            for (int io = 0; io < q; io++)
                for (int is = 0; is < q; is++) {
                    long k = bin3(s, iK, 0, io, is, il);
                    int npart_num = (int)s->num_count[k];
                    int npart_denorm = (int)s->den[k];
                    double qo, qs, ql, cnum, cden;
                    if (npart_num < 2 || npart_denorm < 2) {
                        qo = p->q_min + io * dq;
                        qs = p->q_min + is * dq;
                        ql = p->q_min + il * dq;
                        cnum = 0.0;
                        cden = npart_denorm;
                    } else {
                        qo = s->q_out_sum[k] / npart_num;
                        qs = s->q_side_sum[k] / npart_num;
                        ql = s->q_long_sum[k] / npart_num;
                        cnum = s->num[k];
                        cden = npair_ratio * s->den[k];
                    }
                    if (eco == 1)
                        fprintf(f, "%18.8e    %.8e\n", cnum, cden);
                    else
                        fprintf(f, "%18.8e    %.8e    %.8e    %.8e    %.8e\n", qo, qs,
                                ql, cnum, cden);
                }
        fclose(f);
    }
    return 0;
}

static int write_3d_kphi(const orc_state *s, const char *path, int eco) { /* :785-855 */
    const orc_params *p = &s->p;
    const int q = p->qnpts;
    const double dq = s->d.delta_q;
    for (int iK = 0; iK < p->n_KT - 1; iK++) {
        for (int iphi = 0; iphi < p->n_Kphi; iphi++) {
            double npair_ratio =
                (double)s->cnt_num_ktphi[(long)iK * p->n_Kphi + iphi] /
                (double)s->cnt_den_ktphi[(long)iK * p->n_Kphi + iphi];
            char fn[4096];
            snprintf(fn, sizeof(fn),
                     "%s/HBT_correlation_function_KT_%g_%g_Kphi_%g.dat", path,
                     p->KT_min + iK * s->d.dKT, p->KT_min + (iK + 1) * s->d.dKT,
                     iphi * s->d.dKphi);
            FILE *f = fopen(fn, "w");
            if (!f) return -1;
            for (int il = 0; il < q; il++)
                for (int io = 0; io < q; io++)
                    for (int is = 0; is < q; is++) {
                        long k = bin3(s, iK, iphi, io, is, il);
                        int npart_num = (int)s->num_count[k];
                        int npart_denorm = (int)s->den[k];
                        double qo, qs, ql, cnum, cden;
                        if (npart_num < 2 || npart_denorm < 2) {
                            qo = p->q_min + io * dq;
                            qs = p->q_min + is * dq;
                            ql = p->q_min + il * dq;
                            cnum = 0.0;
                            cden = npart_denorm;
                        } else {
                            qo = s->q_out_sum[k] / npart_num;
                            qs = s->q_side_sum[k] / npart_num;
                            ql = s->q_long_sum[k] / npart_num;
                            cnum = s->num[k];
                            cden = npair_ratio * s->den[k];
                        }
                        if (eco == 1)
                            fprintf(f, "%18.8e    %.8e\n", cnum, cden);
                        else
                            fprintf(f, "%18.8e    %.8e    %.8e    %.8e    %.8e\n",
                                    qo, qs, ql, cnum, cden);
                    }
            fclose(f);
        }
    }
    return 0;
}

static int write_inv(const orc_state *s, const char *path, int eco) { /* :694-724 */
    const orc_params *p = &s->p;
    for (int iK = 0; iK < p->n_KT - 1; iK++) {
        double npair_ratio = (double)s->cnt_num_qinv[iK] / (double)s->cnt_den_qinv[iK];
        char fn[4096];
        snprintf(fn, sizeof(fn), "%s/HBT_correlation_function_inv_KT_%g_%g.dat", path,
                 p->KT_min + iK * s->d.dKT, p->KT_min + (iK + 1) * s->d.dKT);
        FILE *f = fopen(fn, "w");
        if (!f) return -1;
        for (int iq = 0; iq < p->qnpts; iq++) {
            long k = (long)iK * p->qnpts + iq;
            double q_inv_local = s->qinv_sum[k] / s->qinv_count[k];
            double cnum = s->qinv_num[k];
            double cden = s->qinv_den[k] * npair_ratio;
            if (eco == 1)
                fprintf(f, "%18.8e    %.8e\n", cnum, cden);
            else
                fprintf(f, "%18.8e    %.8e    %.8e\n", q_inv_local, cnum, cden);
        }
        fclose(f);
    }
    return 0;
}

/* HBT_correlation.cpp:220-228 */
int orc_write_outputs(const orc_state *s, const char *path, int eco_output) {
    int rc = 0;
    if (s->p.invariant_radius_flag == 1) rc |= write_inv(s, path, eco_output);
    if (s->p.azimuthal_flag == 0)
        rc |= write_3d(s, path, eco_output);
    else
        rc |= write_3d_kphi(s, path, eco_output);
    return rc;
}

/* ---------------- mode-10 reader ---------------- */
struct orc_reader {
    gzFile gz;
    int event_buffer_size;
    int monval;
    int eof;
    long total;
    plist L;
    int64_t *off;
    int off_cap;
    orc_buffer buf;
};

orc_reader *orc_reader_open(const char *filename, int event_buffer_size,
                            int particle_monval) {
    gzFile gz = gzopen(filename, "rb");
    if (!gz) return NULL;
    orc_reader *rd = (orc_reader *)calloc(1, sizeof(orc_reader));
    rd->gz = gz;
    rd->event_buffer_size = event_buffer_size;
    rd->monval = particle_monval;
    return rd;
}

/* One line without its '\n' (particleSamples.cpp:2209-2218).  Sets *eof
 * like gzeof(): only when a read found no more data. */
static int read_line(orc_reader *rd, char *buf, int cap) {
    int n = 0;
    for (;;) {
        int c = gzgetc(rd->gz);
        if (c < 0) { rd->eof = 1; break; }
        if (c == '\n') break;
        if (n < cap - 1) buf[n++] = (char)c;
    }
    buf[n] = 0;
    return n;
}

/* particleSamples.cpp:1247-1286 (read whole events while the running
 * particle total is below event_buffer_size), :1329-1346 (species filter by
 * exact monval; 99999 keeps all). */
int orc_reader_next(orc_reader *rd) {
    rd->L.n = 0;
    rd->total = 0;
    int nev = 0;
    char line[4096];
    long num_particles = 0;
    if (rd->off_cap < 1) { rd->off_cap = 64; rd->off = (int64_t *)malloc(8 * 64); }
    rd->off[0] = 0;
    while (num_particles < rd->event_buffer_size) {
        read_line(rd, line, sizeof(line));
        if (rd->eof) break;
        int n_particle = 0;
        sscanf(line, "%d", &n_particle);
        for (int ip = 0; ip < n_particle; ip++) {
            read_line(rd, line, sizeof(line));
            int mon = 0;
            double mass, t, x, y, z, E, px, py, pz;
            mass = t = x = y = z = E = px = py = pz = 0.0;
            sscanf(line, "%d %lf %lf %lf %lf %lf %lf %lf %lf %lf", &mon, &mass, &t, &x,
                   &y, &z, &E, &px, &py, &pz);
            if (rd->monval == 99999 || mon == rd->monval)
                pl_push(&rd->L, px, py, pz, E, x, y, z, t);
        }
        num_particles += n_particle;
        nev++;
        if (nev + 1 > rd->off_cap) {
            rd->off_cap *= 2;
            rd->off = (int64_t *)realloc(rd->off, 8 * (size_t)rd->off_cap);
        }
        rd->off[nev] = rd->L.n;
    }
    rd->total = num_particles;
    rd->buf.nev = nev;
    rd->buf.ev_off = rd->off;
    rd->buf.px = rd->L.px; rd->buf.py = rd->L.py; rd->buf.pz = rd->L.pz;
    rd->buf.E = rd->L.E; rd->buf.x = rd->L.x; rd->buf.y = rd->L.y;
    rd->buf.z = rd->L.z; rd->buf.t = rd->L.t;
    return nev;
}

int orc_reader_eof(const orc_reader *rd) { return rd->eof; }
const orc_buffer *orc_reader_buffer(const orc_reader *rd) { return &rd->buf; }
long orc_reader_total_particles(const orc_reader *rd) { return rd->total; }

void orc_reader_close(orc_reader *rd) {
    if (!rd) return;
    gzclose(rd->gz);
    pl_free(&rd->L);
    free(rd->off);
    free(rd);
}

/* ======================================================================
 * Balance function (BalanceFunction.cpp).  Restated from
 *   constructor  :15-61   (bins: drap = 2|Brap_max|/(Bnpts-1),
 *                          Brap_min = -|Brap_max| - drap/2, 20 phi bins of
 *                          2 pi/20 from -pi/2)
 *   per buffer   :63-101  (four same-event histograms, then four mixed ones,
 *                          each mixed call drawing per event a partner
 *                          rand_int_uniform() % nev_mixed and a rotation
 *                          rand_uniform() * 2 * pi)
 *   pair loops   :120-197 (pT cut on both, phi bin floor(...) % Bnphi made
 *                          non-negative, |dy| < 1e-10 skipped, dy < Brap_min
 *                          skipped, y bin truncation, 0 <= bin < Bnpts)
 *   N_b, N_bbar  :68-69, :199-209
 *   writer       :211-318
 * with the per-particle quantities of particleSamples::boostParticles
 * (particleSamples.cpp:442-468, rap_shift 0): pT, phi = atan2(py, px),
 * y = asinh(pz / mT), eta = asinh(pz / pT), and the species lists of
 * filter_particles (:1329-1460): exact Monte-Carlo id, 99999 = all.
 * ====================================================================== */

int orc_bf_init(orc_bf_state *s, const orc_bf_params *p) {
    memset(s, 0, sizeof(*s));
    s->p = *p;
    if (p->Bnpts < 2) return -1;
    s->drap = 2. * fabs(p->Brap_max) / (p->Bnpts - 1);
    s->Brap_min = -fabs(p->Brap_max) - 0.5 * s->drap;
    s->Bnphi = 20;
    s->dphi = 2. * M_PI / s->Bnphi;
    s->Bphi_min = -M_PI / 2.;
    s->hist = (double *)calloc((size_t)8 * p->Bnpts * s->Bnphi, sizeof(double));
    return s->hist ? 0 : -1;
}

void orc_bf_free(orc_bf_state *s) {
    free(s->hist);
    s->hist = NULL;
}

typedef struct {
    int n;
    double *pT, *phi, *y;
    int64_t *off; /* nev+1 */
} bf_list;

/* particle_decay::get_particle_charge & co. (particle_decay.cpp:181-196):
 * the map lookup; NULL when monval is not in the table (the reference's
 * std::map::at throws). */
static const int *sp_find(const orc_species *sp, int monval, int *idx) {
    if (!sp) return NULL;
    int lo = 0, hi = sp->n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (sp->monval[mid] == monval) {
            *idx = mid;
            return &sp->monval[mid];
        }
        if (sp->monval[mid] < monval)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return NULL;
}

/* particleSamples::decide_to_pick_OSCAR (particleSamples.cpp:625-678). */
static int bf_pick(orc_bf_state *s, int poi, int monval) {
    if (poi == 99999) return 1;
    const int cls = poi == 9999 || poi == 9998 || poi == -9998 || poi == 9997 || poi == -9997 ||
                    poi == 9996 || poi == -9996;
    if (!cls) return monval == poi;
    int i = 0;
    if (!sp_find(s->sp, monval, &i)) {
        s->bad_monval_seen = 1;
        return 0;
    }
    const int charge = s->sp->charge[i], baryon = s->sp->baryon[i], strange = s->sp->strange[i];
    switch (poi) {
        case 9999: return charge != 0;
        case 9998: return charge > 0;
        case -9998: return charge < 0;
        case 9997: return baryon > 0;
        case -9997: return baryon < 0;
        case 9996: return strange > 0;
        default: return strange < 0; /* -9996 */
    }
}

static void bf_list_build(orc_bf_state *s, bf_list *L, const orc_bf_buffer *b, int poi,
                          int rap_type) {
    const int64_t n = b->nev > 0 ? b->ev_off[b->nev] - b->ev_off[0] : 0;
    L->pT = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    L->phi = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    L->y = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    L->off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(b->nev + 1));
    int k = 0;
    L->off[0] = 0;
    for (int e = 0; e < b->nev; e++) {
        for (int64_t i = b->ev_off[e]; i < b->ev_off[e + 1]; i++) {
            if (!bf_pick(s, poi, b->monval[i])) continue;
            const double px = b->px[i], py = b->py[i], pz = b->pz[i];
            const double pT = sqrt(px * px + py * py);
            const double mT = sqrt(pT * pT + b->mass[i] * b->mass[i]);
            L->pT[k] = pT;
            L->phi[k] = atan2(py, px);
            L->y[k] = rap_type == 0 ? asinh(pz / pT) : asinh(pz / mT);
            k++;
        }
        L->off[e + 1] = k;
    }
    L->n = k;
}

static void bf_list_free(bf_list *L) {
    free(L->pT);
    free(L->phi);
    free(L->y);
    free(L->off);
}

static void bf_pair(const orc_bf_state *s, double *hist, double phi_a, double y_a, double phi_b,
                    double y_b, double rot, int mixed) {
    const double delta_phi_local = mixed ? (phi_a - phi_b + rot) : (phi_a - phi_b);
    int phi_idx = ((int)floor((delta_phi_local - s->Bphi_min) / s->dphi)) % s->Bnphi;
    if (phi_idx < 0) phi_idx += s->Bnphi;
    const double delta_y_local = y_a - y_b;
    if (fabs(delta_y_local) < 1e-10) return;
    if (delta_y_local < s->Brap_min) return;
    const int y_bin_idx = (int)((delta_y_local - s->Brap_min) / s->drap);
    if (y_bin_idx >= 0 && y_bin_idx < s->p.Bnpts) hist[(size_t)y_bin_idx * s->Bnphi + phi_idx] += 1.;
}

static int bf_cut(const orc_bf_state *s, double pT) {
    return pT < s->p.BpT_min || pT > s->p.BpT_max;
}

static void bf_same(const orc_bf_state *s, double *hist, const bf_list *A, const bf_list *B,
                    int nev) {
    for (int e = 0; e < nev; e++)
        for (int64_t i = A->off[e]; i < A->off[e + 1]; i++) {
            if (bf_cut(s, A->pT[i])) continue;
            for (int64_t j = B->off[e]; j < B->off[e + 1]; j++) {
                if (bf_cut(s, B->pT[j])) continue;
                bf_pair(s, hist, A->phi[i], A->y[i], B->phi[j], B->y[j], 0.0, 0);
            }
        }
}

static void bf_mixed(const orc_bf_state *s, orc_rng *r, double *hist, const bf_list *A,
                     const bf_list *B, int nev, int nev_mixed) {
    for (int e = 0; e < nev; e++) {
        const int em = orc_rand_int_uniform(r) % nev_mixed;
        const double rot = orc_rand_uniform(r) * 2. * M_PI;
        for (int64_t i = A->off[e]; i < A->off[e + 1]; i++) {
            if (bf_cut(s, A->pT[i])) continue;
            for (int64_t j = B->off[em]; j < B->off[em + 1]; j++) {
                if (bf_cut(s, B->pT[j])) continue;
                bf_pair(s, hist, A->phi[i], A->y[i], B->phi[j], B->y[j], rot, 1);
            }
        }
    }
}

static long bf_count(const orc_bf_state *s, const bf_list *L) {
    long c = 0;
    for (int k = 0; k < L->n; k++) c += !bf_cut(s, L->pT[k]);
    return c;
}

void orc_bf_calculate(orc_bf_state *s, orc_rng *r, const orc_bf_buffer *b,
                      const orc_bf_buffer *mix) {
    const int rt = s->p.rap_type;
    bf_list a, abar, bb, bbar, bm, bbarm;
    bf_list_build(s, &a, b, s->p.alpha, rt);
    bf_list_build(s, &abar, b, -s->p.alpha, rt);
    bf_list_build(s, &bb, b, s->p.beta, rt);
    bf_list_build(s, &bbar, b, -s->p.beta, rt);
    bf_list_build(s, &bm, mix, s->p.beta, rt);
    bf_list_build(s, &bbarm, mix, -s->p.beta, rt);
    s->N_b += bf_count(s, &bb);
    s->N_bbar += bf_count(s, &bbar);
    const size_t H = (size_t)s->p.Bnpts * s->Bnphi;
    bf_same(s, s->hist + 0 * H, &a, &bb, b->nev);
    bf_same(s, s->hist + 1 * H, &abar, &bbar, b->nev);
    bf_same(s, s->hist + 2 * H, &a, &bbar, b->nev);
    bf_same(s, s->hist + 3 * H, &abar, &bb, b->nev);
    bf_mixed(s, r, s->hist + 4 * H, &a, &bm, b->nev, mix->nev);
    bf_mixed(s, r, s->hist + 5 * H, &abar, &bbarm, b->nev, mix->nev);
    bf_mixed(s, r, s->hist + 6 * H, &a, &bbarm, b->nev, mix->nev);
    bf_mixed(s, r, s->hist + 7 * H, &abar, &bm, b->nev, mix->nev);
    bf_list_free(&a);
    bf_list_free(&abar);
    bf_list_free(&bb);
    bf_list_free(&bbar);
    bf_list_free(&bm);
    bf_list_free(&bbarm);
}

int orc_bf_write_outputs(const orc_bf_state *s, const char *path) {
    const int np = s->p.Bnpts, nf = s->Bnphi;
    const double *C[8];
    for (int a = 0; a < 8; a++) C[a] = s->hist + (size_t)a * np * nf;
#define H(a, i, j) C[a][(size_t)(i) * nf + (j)]
    double N_OS = 0., N_OS_mixed = 0., N_SS = 0., N_SS_mixed = 0.;
    double *os_y = calloc(np, 8), *ss_y = calloc(np, 8), *osm_y = calloc(np, 8),
           *ssm_y = calloc(np, 8);
    double *os_p = calloc(nf, 8), *ss_p = calloc(nf, 8), *osm_p = calloc(nf, 8),
           *ssm_p = calloc(nf, 8);
    for (int i = 0; i < np; i++) {
        for (int j = 0; j < nf; j++) {
            os_y[i] += H(0, i, j) + H(1, i, j);
            ss_y[i] += H(2, i, j) + H(3, i, j);
            osm_y[i] += H(4, i, j) + H(5, i, j);
            ssm_y[i] += H(6, i, j) + H(7, i, j);
        }
        N_OS += os_y[i];
        N_SS += ss_y[i];
        N_OS_mixed += osm_y[i];
        N_SS_mixed += ssm_y[i];
    }
    for (int j = 0; j < nf; j++)
        for (int i = 0; i < np; i++) {
            os_p[j] += H(0, i, j) + H(1, i, j);
            ss_p[j] += H(2, i, j) + H(3, i, j);
            osm_p[j] += H(4, i, j) + H(5, i, j);
            ssm_p[j] += H(6, i, j) + H(7, i, j);
        }
    char fn[4096];
    snprintf(fn, sizeof fn, "%s/Balance_function_%d_%d_Delta_y.dat", path, s->p.alpha, s->p.beta);
    FILE *f = fopen(fn, "w");
    if (!f) return -1;
    fprintf(f, "# DeltaY  Delta_C2  C2(OS)  rho2(OS)  rho1^2(OS)  C2(SS) rho2(SS)  rho1^2(SS)\n");
    for (int i = 0; i < np; i++) {
        const double dy = s->Brap_min + (i + 0.5) * s->drap;
        const double c_os = os_y[i] / osm_y[i] * N_OS_mixed / N_OS;
        const double c_ss = ss_y[i] / ssm_y[i] * N_SS_mixed / N_SS;
        fprintf(f, "%18.8e   %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e\n", dy, c_os - c_ss, c_os,
                os_y[i], osm_y[i], c_ss, ss_y[i], ssm_y[i]);
    }
    fclose(f);
    snprintf(fn, sizeof fn, "%s/Balance_function_%d_%d_Delta_phi.dat", path, s->p.alpha,
             s->p.beta);
    f = fopen(fn, "w");
    if (!f) return -1;
    fprintf(f, "# Delta_phi  Delta_C2  C2(OS)  rho2(OS)  rho1^2(OS)  C2(SS) rho2(SS)  rho1^2(SS)\n");
    for (int j = 0; j < nf; j++) {
        const double dp = s->Bphi_min + (j + 0.5) * s->dphi;
        const double c_os = (os_p[j] / osm_p[j] * N_OS_mixed / N_OS);
        const double c_ss = (ss_p[j] / ssm_p[j] * N_SS_mixed / N_SS);
        fprintf(f, "%18.8e   %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e\n", dp, c_os - c_ss, c_os,
                os_p[j], osm_p[j], c_ss, ss_p[j], ssm_p[j]);
    }
    fclose(f);
    snprintf(fn, sizeof fn, "%s/Correlation_function_%d_%d_2D.dat", path, s->p.alpha, s->p.beta);
    f = fopen(fn, "w");
    if (!f) return -1;
    fprintf(f, "# DY  Dphi  C2(OS)  rho2(OS)  rho1^2(OS)  C2(SS)  rho2(SS)  rho1^2(SS)\n");
    for (int i = 0; i < np; i++)
        for (int j = 0; j < nf; j++) {
            const double dy = s->Brap_min + (i + 0.5) * s->drap;
            const double dp = s->Bphi_min + (j + 0.5) * s->dphi;
            const double c_os = ((H(0, i, j) + H(1, i, j)) / (H(4, i, j) + H(5, i, j) + 1e-15));
            const double c_ss = ((H(2, i, j) + H(3, i, j)) / (H(6, i, j) + H(7, i, j) + 1e-15));
            fprintf(f, "%18.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e  %.8e\n", dy, dp,
                    c_os * N_OS_mixed / N_OS, H(0, i, j) + H(1, i, j), H(4, i, j) + H(5, i, j),
                    c_ss * N_SS_mixed / N_SS, H(2, i, j) + H(3, i, j), H(6, i, j) + H(7, i, j));
        }
    fclose(f);
#undef H
    free(os_y); free(ss_y); free(osm_y); free(ssm_y);
    free(os_p); free(ss_p); free(osm_p); free(ssm_p);
    return 0;
}

/* ---------------- species table (particle_decay.cpp:23-129) ---------------- */
/* read_resonances_list: per species monval name mass width gspin baryon
 * strange charm bottom gisospin charge decays, then decays channel records
 * (dummy, Npart, branching ratio, five products).  A species with baryon
 * number 1 adds its anti-species (-monval, quantum numbers negated); before
 * that, the reference's table lookup of each decay product of that baryon
 * creates a zero entry for a product not yet in the table (:96-103). */
typedef struct {
    int n, cap;
    int *m, *q; /* q: 3 ints per entry */
} sp_build;

static int spb_find(const sp_build *b, int m) {
    for (int i = 0; i < b->n; i++)
        if (b->m[i] == m) return i;
    return -1;
}

static void spb_put(sp_build *b, int m, int c, int ba, int st) {
    int i = spb_find(b, m);
    if (i < 0) {
        if (b->n == b->cap) {
            b->cap = b->cap ? 2 * b->cap : 256;
            b->m = (int *)realloc(b->m, sizeof(int) * (size_t)b->cap);
            b->q = (int *)realloc(b->q, 3 * sizeof(int) * (size_t)b->cap);
        }
        i = b->n++;
        b->m[i] = m;
    }
    b->q[3 * i] = c;
    b->q[3 * i + 1] = ba;
    b->q[3 * i + 2] = st;
}

static int cmp_int_idx(const void *x, const void *y) {
    const int *a = (const int *)x, *b = (const int *)y;
    return (a[0] > b[0]) - (a[0] < b[0]);
}

int orc_species_read(const char *path, orc_species *sp) {
    memset(sp, 0, sizeof(*sp));
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    sp_build b = {0, 0, NULL, NULL};
    int monval, rc = 0;
    while (fscanf(f, "%d", &monval) == 1) {
        char name[256];
        double mass, width, br;
        int gspin, baryon, strange, charm, bottom, giso, charge, decays;
        if (fscanf(f, "%255s %lf %lf %d %d %d %d %d %d %d %d", name, &mass, &width, &gspin,
                   &baryon, &strange, &charm, &bottom, &giso, &charge, &decays) != 11 ||
            decays < 0) {
            rc = 2;
            break;
        }
        int *parts = (int *)malloc(sizeof(int) * 5 * (size_t)(decays + 1));
        for (int d = 0; d < decays && !rc; d++) {
            int dummy, npart;
            if (fscanf(f, "%d %d %lf %d %d %d %d %d", &dummy, &npart, &br, &parts[5 * d],
                       &parts[5 * d + 1], &parts[5 * d + 2], &parts[5 * d + 3],
                       &parts[5 * d + 4]) != 8)
                rc = 2;
        }
        if (!rc) {
            spb_put(&b, monval, charge, baryon, strange);
            if (baryon == 1) {
                for (int k = 0; k < 5 * decays; k++)
                    if (parts[k] != 0 && spb_find(&b, parts[k]) < 0) spb_put(&b, parts[k], 0, 0, 0);
                spb_put(&b, -monval, -charge, -baryon, -strange);
            }
        }
        free(parts);
        if (rc) break;
    }
    if (!rc && !feof(f)) rc = 2;
    fclose(f);
    if (rc) {
        free(b.m);
        free(b.q);
        return rc;
    }
    /* sort (monval, index) pairs */
    int *pairs = (int *)malloc(sizeof(int) * 2 * (size_t)(b.n + 1));
    for (int i = 0; i < b.n; i++) {
        pairs[2 * i] = b.m[i];
        pairs[2 * i + 1] = i;
    }
    qsort(pairs, (size_t)b.n, 2 * sizeof(int), cmp_int_idx);
    sp->n = b.n;
    sp->monval = (int *)malloc(sizeof(int) * (size_t)(b.n + 1));
    sp->charge = (int *)malloc(sizeof(int) * (size_t)(b.n + 1));
    sp->baryon = (int *)malloc(sizeof(int) * (size_t)(b.n + 1));
    sp->strange = (int *)malloc(sizeof(int) * (size_t)(b.n + 1));
    for (int i = 0; i < b.n; i++) {
        const int j = pairs[2 * i + 1];
        sp->monval[i] = b.m[j];
        sp->charge[i] = b.q[3 * j];
        sp->baryon[i] = b.q[3 * j + 1];
        sp->strange[i] = b.q[3 * j + 2];
    }
    free(pairs);
    free(b.m);
    free(b.q);
    return 0;
}

void orc_species_free(orc_species *sp) {
    free(sp->monval);
    free(sp->charge);
    free(sp->baryon);
    free(sp->strange);
    memset(sp, 0, sizeof(*sp));
}
