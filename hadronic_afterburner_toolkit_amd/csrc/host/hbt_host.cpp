// hbt_host.cpp -- host-side pieces of the C ABI that need no GPU: the
// reference's text writers, its mode-10 and mode-9 sample readers and
// psi_ref.
// Compiled into libhbt_hip.so next to the engine.
#include <math.h>
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <iomanip>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/hbt_hip.h"

namespace {

struct Grid {
    int q, nKT, nphi;
    double q_min, dq, KT_min, dKT, dKphi;
};

Grid grid_of(const hbt_config *c) {
    Grid g;
    g.q = c->qnpts;
    g.nKT = c->n_KT;
    g.nphi = c->azimuthal_flag ? c->n_Kphi : 1;
    g.q_min = c->q_min;
    g.dq = (c->q_max - c->q_min) / (c->qnpts - 1);     // HBT_correlation.cpp:28
    g.KT_min = c->KT_min;
    g.dKT = (c->KT_max - c->KT_min) / (c->n_KT - 1);   // :48
    g.dKphi = 2 * M_PI / (c->n_Kphi > 0 ? c->n_Kphi : 1);  // :49
    return g;
}

// One table (one K_T bin, one K_phi bin) in the reference's line format
// (HBT_correlation.cpp:735-777 / :798-849): q_long outermost, q_side
// innermost; a bin with fewer than 2 numerator or denominator pairs prints
// the grid q values, 0 and the raw denominator.
void write_table(std::ofstream &output, const Grid &g, const hbt_result *r, long base,
                 double npair_ratio, int eco) {
    const int q = g.q;
    for (int iqlong = 0; iqlong < q; iqlong++)
        for (int iqout = 0; iqout < q; iqout++)
            for (int iqside = 0; iqside < q; iqside++) {
                const long k = base + ((long)iqout * q + iqside) * q + iqlong;
                const int npart_num = (int)(double)r->num_count[k];
                const int npart_denorm = (int)(double)r->den[k];
                double q_out_local, q_side_local, q_long_local;
                double correl_fun_num, correl_fun_denorm;
                if (npart_num < 2 || npart_denorm < 2) {
                    q_out_local = g.q_min + iqout * g.dq;  // q_out[iqout] (:30-34)
                    q_side_local = g.q_min + iqside * g.dq;
                    q_long_local = g.q_min + iqlong * g.dq;
                    correl_fun_num = 0.0;
                    correl_fun_denorm = npart_denorm;
                } else {
                    q_out_local = r->q_out_sum[k] / npart_num;
                    q_side_local = r->q_side_sum[k] / npart_num;
                    q_long_local = r->q_long_sum[k] / npart_num;
                    correl_fun_num = r->num[k];
                    correl_fun_denorm = npair_ratio * (double)r->den[k];
                }
                if (eco == 1) {
                    output << std::scientific << std::setw(18) << std::setprecision(8)
                           << correl_fun_num << "    " << correl_fun_denorm << std::endl;
                } else {
                    output << std::scientific << std::setw(18) << std::setprecision(8)
                           << q_out_local << "    " << q_side_local << "    " << q_long_local
                           << "    " << correl_fun_num << "    " << correl_fun_denorm
                           << std::endl;
                }
            }
}

}  // namespace

extern "C" {

int hbt_write_outputs(const hbt_config *c, const hbt_result *r, const char *path,
                      int32_t eco_output, int32_t what) {
    if (!c || !r || !path) return HBT_ERR_ARG;
    if (!r->num_count || !r->den || !r->q_out_sum || !r->q_side_sum || !r->q_long_sum ||
        !r->num)
        return HBT_ERR_ARG;
    if (what == 0)  // output_HBTcorrelation, :220-228
        what = (c->invariant_radius_flag == 1 ? HBT_WRITE_INV : 0) |
               (c->azimuthal_flag == 0 ? HBT_WRITE_3D : HBT_WRITE_KPHI);
    const Grid g = grid_of(c);
    const long q3 = (long)g.q * g.q * g.q;
    if (what & HBT_WRITE_INV) {  // output_correlation_function_inv, :694-724
        if (!r->qinv_count || !r->qinv_den || !r->qinv_sum || !r->qinv_num || !r->cnt_num_qinv ||
            !r->cnt_den_qinv)
            return HBT_ERR_ARG;
        for (int iK = 0; iK < g.nKT - 1; iK++) {
            const double npair_ratio = (static_cast<double>(r->cnt_num_qinv[iK]) /
                                        static_cast<double>(r->cnt_den_qinv[iK]));
            std::ostringstream filename;
            filename << path << "/HBT_correlation_function_inv_KT_" << (g.KT_min + iK * g.dKT)
                     << "_" << (g.KT_min + (iK + 1) * g.dKT) << ".dat";
            std::ofstream output(filename.str().c_str());
            if (!output) return HBT_ERR_ARG;
            for (int iq = 0; iq < g.q; iq++) {
                const long k = (long)iK * g.q + iq;
                const double q_inv_local = r->qinv_sum[k] / (double)r->qinv_count[k];
                const double correl_fun_num = r->qinv_num[k];
                const double correl_fun_denorm = (double)r->qinv_den[k] * npair_ratio;
                if (eco_output == 1)
                    output << std::scientific << std::setw(18) << std::setprecision(8)
                           << correl_fun_num << "    " << correl_fun_denorm << std::endl;
                else
                    output << std::scientific << std::setw(18) << std::setprecision(8)
                           << q_inv_local << "    " << correl_fun_num << "    "
                           << correl_fun_denorm << std::endl;
            }
        }
    }
    if (what & HBT_WRITE_3D) {  // output_correlation_function, :726-783
        if (!r->cnt_num_kt || !r->cnt_den_kt) return HBT_ERR_ARG;
        for (int iK = 0; iK < g.nKT - 1; iK++) {
            const double npair_ratio = (static_cast<double>(r->cnt_num_kt[iK]) /
                                        static_cast<double>(r->cnt_den_kt[iK]));
            std::ostringstream filename;
            filename << path << "/HBT_correlation_function_KT_" << (g.KT_min + iK * g.dKT) << "_"
                     << (g.KT_min + (iK + 1) * g.dKT) << ".dat";
            std::ofstream output(filename.str().c_str());
            if (!output) return HBT_ERR_ARG;
            write_table(output, g, r, (long)iK * q3, npair_ratio, eco_output);
        }
    }
    if (what & HBT_WRITE_KPHI) {  // output_correlation_function_Kphi_differential, :785-855
        if (!r->cnt_num_ktphi || !r->cnt_den_ktphi) return HBT_ERR_ARG;
        for (int iK = 0; iK < g.nKT - 1; iK++)
            for (int iKphi = 0; iKphi < g.nphi; iKphi++) {
                const long kk = (long)iK * g.nphi + iKphi;
                const double npair_ratio = (static_cast<double>(r->cnt_num_ktphi[kk]) /
                                            static_cast<double>(r->cnt_den_ktphi[kk]));
                std::ostringstream filename;
                filename << path << "/HBT_correlation_function_KT_" << (g.KT_min + iK * g.dKT)
                         << "_" << (g.KT_min + (iK + 1) * g.dKT) << "_Kphi_"
                         << (iKphi * g.dKphi) << ".dat";
                std::ofstream output(filename.str().c_str());
                if (!output) return HBT_ERR_ARG;
                write_table(output, g, r, kk * q3, npair_ratio, eco_output);
            }
    }
    return HBT_OK;
}

double hbt_flow_event_plane_angle(int32_t n, const double *px, const double *py, int32_t n_order) {
    // HBT_correlation.cpp:233-249 over all particles of the buffer, in order.
    // The reference's build evaluates the cos/sin pair with glibc sincos().
    double vn_real = 0.0, vn_imag = 0.0;
    for (int32_t i = 0; i < n; i++) {
        const double temp_phi = atan2(py[i], px[i]);
        double sn, cn;
        sincos(n_order * temp_phi, &sn, &cn);
        vn_real += cn;
        vn_imag += sn;
    }
    return atan2(vn_imag, vn_real) / n_order;
}

}  // extern "C"

// ---------------------------------------------- mode-10 and mode-9 readers

// Charge, baryon number and strangeness per monval (the species table).
struct hbt_species_table {
    struct Q {
        int32_t charge, baryon, strange;
    };
    std::vector<int32_t> order;  // monval in order of first insertion
    std::map<int32_t, Q> q;
};

struct hbt_reader {
    int mode = 10;  // 10: text (gzip or plain), 9: binary records
    gzFile gz = nullptr;
    FILE *bin = nullptr;
    int event_buffer_size = 0;
    int monval = 0;
    bool eof = false;
    // decompressed text not yet consumed: win[pos, len), NUL after len;
    // the window grows by doubling (realloc, no zero fill)
    char *win = nullptr;
    size_t cap = 0, pos = 0, len = 0;
    bool no_more_data = false;
    // last buffer: filtered particles (SoA) and their event offsets
    std::vector<int64_t> off;
    std::vector<double> cols[8];  // px py pz E x y z t
    std::vector<double> mass;
    std::vector<int32_t> monv;  // Monte-Carlo id of each kept particle
    long long total = 0;
    // species classes (particle_monval 9999, +-9998, +-9997, +-9996)
    bool has_species = false;
    std::map<int32_t, hbt_species_table::Q> species;
    long long bad_monval = 0;
    bool bad = false;
};

namespace {

bool is_species_class(int id) {
    return id == 9999 || id == 9998 || id == -9998 || id == 9997 || id == -9997 || id == 9996 ||
           id == -9996;
}

// particleSamples::decide_to_pick_OSCAR (particleSamples.cpp:625-678); a
// class evaluated for a monval missing from the table is recorded (the
// reference's lookup throws) and the particle is not kept.
bool reader_pick(hbt_reader *rd, int mon) {
    const int poi = rd->monval;
    if (poi == 99999) return true;
    if (!is_species_class(poi)) return mon == poi;
    const auto it = rd->species.find(mon);
    if (it == rd->species.end()) {
        if (!rd->bad) rd->bad_monval = mon;
        rd->bad = true;
        return false;
    }
    const hbt_species_table::Q &q = it->second;
    switch (poi) {
        case 9999: return q.charge != 0;
        case 9998: return q.charge > 0;
        case -9998: return q.charge < 0;
        case 9997: return q.baryon > 0;
        case -9997: return q.baryon < 0;
        case 9996: return q.strange > 0;
        default: return q.strange < 0;  // -9996
    }
}

// ---- number parsing, bit-identical to strtod / strtol

// Exact powers of ten representable in double.
const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                           1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }
inline bool is_delim(char c) { return is_blank(c) || c == '\n' || c == '\0'; }

// One decimal number at *p (leading blanks skipped), as strtod reads it.
// Clinger's fast path (W. D. Clinger, "How to read floating point numbers
// accurately", PLDI 1990): when the significant digits form an integer
// m < 2^53 and the decimal exponent e lies in [-22, 22], m and 10^|e| are
// exact doubles and one IEEE multiply or divide gives the correctly rounded
// value, which is what glibc's strtod returns.  Anything else (more digits,
// larger exponents, hex, inf/nan, a malformed field) is handed to strtod.
inline double parse_double(const char *&p) {
    const char *s = p;
    while (is_blank(*s)) s++;
    const char *q = s;
    bool neg = false;
    if (*q == '-' || *q == '+') neg = *q++ == '-';
    uint64_t m = 0;
    int sig = 0, e10 = 0;
    bool any = false, fast = true;
    for (; *q >= '0' && *q <= '9'; q++) {
        any = true;
        if (m == 0 && *q == '0') continue;
        if (sig < 19) {
            m = m * 10 + (uint64_t)(*q - '0');
            sig++;
        } else {
            fast = false;
        }
    }
    if (*q == '.') {
        for (q++; *q >= '0' && *q <= '9'; q++) {
            any = true;
            if (m == 0 && *q == '0') {
                e10--;
                continue;
            }
            if (sig < 19) {
                m = m * 10 + (uint64_t)(*q - '0');
                sig++;
                e10--;
            } else {
                fast = false;
            }
        }
    }
    if (any && (*q == 'e' || *q == 'E')) {
        const char *r = q + 1;
        bool eneg = false;
        if (*r == '-' || *r == '+') eneg = *r++ == '-';
        if (*r >= '0' && *r <= '9') {
            int x = 0;
            for (; *r >= '0' && *r <= '9'; r++)
                if (x < 100000) x = x * 10 + (*r - '0');
            e10 += eneg ? -x : x;
            q = r;
        }
    }
    if (!any || !fast || !is_delim(*q) || m >= (1ull << 53) || e10 < -22 || e10 > 22) {
        char *end = nullptr;
        const double v = strtod(s, &end);
        p = end;
        return v;
    }
    double v = (double)m;
    v = e10 >= 0 ? v * kPow10[e10] : v / kPow10[-e10];
    p = q;
    return neg ? -v : v;
}

// One integer at *p as strtol(p, &end, 10) reads it (leading blanks, sign);
// long digit strings and malformed fields go to strtol.
inline long parse_long(const char *&p) {
    const char *s = p;
    while (is_blank(*s)) s++;
    const char *q = s;
    bool neg = false;
    if (*q == '-' || *q == '+') neg = *q++ == '-';
    long v = 0;
    int nd = 0;
    for (; *q >= '0' && *q <= '9' && nd < 18; q++, nd++) v = v * 10 + (*q - '0');
    if (nd == 0 || (*q >= '0' && *q <= '9')) {
        char *end = nullptr;
        const long r = strtol(s, &end, 10);
        p = end;
        return r;
    }
    p = q;
    return neg ? -v : v;
}

// ---- text window

constexpr size_t kChunk = size_t(4) << 20;

// Room for n more bytes and the NUL after them.
bool reserve_window(hbt_reader *rd, size_t n) {
    if (rd->len + n + 1 <= rd->cap) return true;
    size_t cap = std::max<size_t>(rd->cap * 2, rd->len + n + 1);
    char *w = (char *)realloc(rd->win, cap);
    if (!w) return false;
    rd->win = w;
    rd->cap = cap;
    return true;
}

// Appends up to kChunk decompressed bytes; false when the file has no more.
bool refill(hbt_reader *rd) {
    if (rd->no_more_data || !reserve_window(rd, kChunk)) {
        rd->no_more_data = true;
        return false;
    }
    const int got = gzread(rd->gz, rd->win + rd->len, (unsigned)kChunk);
    if (got <= 0) {
        rd->no_more_data = true;
        rd->win[rd->len] = '\0';
        return false;
    }
    rd->len += (size_t)got;
    rd->win[rd->len] = '\0';
    return true;
}

// The next line [start, end) of the window, without its '\n'.  As
// gz_readline + gzeof (particleSamples.cpp:2209-2218): a line that reaches
// the end of the data without '\n' sets eof (and is returned as read).
void next_line(hbt_reader *rd, size_t &start, size_t &end) {
    for (;;) {
        const char *b = rd->win + rd->pos;
        const void *nl = rd->len > rd->pos ? memchr(b, '\n', rd->len - rd->pos) : nullptr;
        if (nl) {
            start = rd->pos;
            end = (size_t)((const char *)nl - rd->win);
            rd->pos = end + 1;
            return;
        }
        if (!refill(rd)) {
            start = rd->pos;
            end = rd->len;
            rd->pos = rd->len;
            rd->eof = true;
            return;
        }
    }
}

// Runs fn(begin, end) over [0, n) on up to 16 host threads.
template <typename F>
void parallel_lines(size_t n, F fn) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nt = std::max<size_t>(1, std::min<size_t>(hw, n / 4096));
    if (nt <= 1) {
        fn((size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

// Mode 10 (particleSamples.cpp:1247-1286): whole events while the running
// particle total (all species) is below event_buffer_size.  The event
// structure is scanned first (line boundaries only), then the particle
// lines "monval mass t x y z E px py pz" are parsed on several threads and
// filtered by species (:1329-1346) in file order.
int32_t read_text_buffer(hbt_reader *rd) {
    // drop what earlier buffers consumed
    if (rd->pos > 0) {
        memmove(rd->win, rd->win + rd->pos, rd->len - rd->pos);
        rd->len -= rd->pos;
        rd->pos = 0;
        rd->win[rd->len] = '\0';
    }
    if (!rd->win && !reserve_window(rd, 0)) return HBT_ERR_NOMEM;
    rd->win[rd->len] = '\0';
    std::vector<size_t> ls, le;  // particle line bounds
    std::vector<int> nper;       // particles per event
    long long num_particles = 0;
    while (num_particles < rd->event_buffer_size) {
        size_t a, b;
        next_line(rd, a, b);
        if (rd->eof) break;
        rd->win[b] = '\0';  // the '\n' (or the end) terminates the line
        const char *h = rd->win + a;
        const int n_particle = (int)parse_long(h);
        for (int ip = 0; ip < n_particle; ip++) {
            next_line(rd, a, b);
            ls.push_back(a);
            le.push_back(b);
        }
        nper.push_back(n_particle);
        num_particles += n_particle;
    }
    const size_t nl = ls.size();
    for (size_t k = 0; k < nl; k++) rd->win[le[k]] = '\0';
    std::vector<int> mon(nl);
    std::vector<double> val(nl * 9);
    const char *base = rd->win;
    parallel_lines(nl, [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; k++) {
            const char *p = base + ls[k];
            mon[k] = (int)parse_long(p);
            double *v = &val[k * 9];  // mass t x y z E px py pz
            for (int f = 0; f < 9; f++) v[f] = parse_double(p);
        }
    });
    rd->off.assign(1, 0);
    for (auto &c : rd->cols) c.clear();
    rd->mass.clear();
    rd->monv.clear();
    size_t k = 0;
    for (int n : nper) {
        for (int ip = 0; ip < n; ip++, k++) {
            if (!reader_pick(rd, mon[k])) continue;
            const double *v = &val[k * 9];
            rd->cols[0].push_back(v[6]);
            rd->cols[1].push_back(v[7]);
            rd->cols[2].push_back(v[8]);
            rd->cols[3].push_back(v[5]);
            rd->cols[4].push_back(v[2]);
            rd->cols[5].push_back(v[3]);
            rd->cols[6].push_back(v[4]);
            rd->cols[7].push_back(v[1]);
            rd->mass.push_back(v[0]);
            rd->monv.push_back(mon[k]);
        }
        rd->off.push_back((int64_t)rd->cols[0].size());
    }
    rd->total = num_particles;
    return (int32_t)nper.size();
}

// Mode 9 (particleSamples.cpp:1203-1245): per event an int n, then n
// records of an int Monte-Carlo id and 9 floats (mass t x y z E px py pz),
// native byte order; the values widen to double.  The reference files every
// particle under event 0 (its event index never advances); here each record
// belongs to its own event, as the file lays them out.
int32_t read_binary_buffer(hbt_reader *rd) {
    rd->off.assign(1, 0);
    for (auto &c : rd->cols) c.clear();
    rd->mass.clear();
    rd->monv.clear();
    long long num_particles = 0;
    int nev = 0;
    std::vector<char> rec;
    while (num_particles < rd->event_buffer_size) {
        int32_t n_particle = 0;
        if (fread(&n_particle, sizeof(n_particle), 1, rd->bin) != 1) {
            rd->eof = true;
            break;
        }
        const size_t rb = 4 + 9 * sizeof(float);
        rec.assign((size_t)std::max(0, n_particle) * rb, 0);
        const size_t got = rec.empty() ? 0 : fread(rec.data(), 1, rec.size(), rd->bin);
        if (got < rec.size()) rd->eof = true;  // truncated: missing values read as 0
        for (int ip = 0; ip < n_particle; ip++) {
            const char *r = rec.data() + (size_t)ip * rb;
            int32_t mon;
            float a[9];
            memcpy(&mon, r, 4);
            memcpy(a, r + 4, sizeof(a));
            if (!reader_pick(rd, mon)) continue;
            rd->cols[0].push_back(a[6]);
            rd->cols[1].push_back(a[7]);
            rd->cols[2].push_back(a[8]);
            rd->cols[3].push_back(a[5]);
            rd->cols[4].push_back(a[2]);
            rd->cols[5].push_back(a[3]);
            rd->cols[6].push_back(a[4]);
            rd->cols[7].push_back(a[1]);
            rd->mass.push_back(a[0]);
            rd->monv.push_back(mon);
        }
        num_particles += std::max(0, n_particle);
        nev++;
        rd->off.push_back((int64_t)rd->cols[0].size());
        if (rd->eof) break;
    }
    rd->total = num_particles;
    return nev;
}

}  // namespace

extern "C" {

int hbt_reader_open_mode(const char *filename, int32_t read_in_mode, int32_t event_buffer_size,
                         int32_t particle_monval, hbt_reader **out) {
    if (!filename || !out) return HBT_ERR_ARG;
    *out = nullptr;
    if (read_in_mode != 9 && read_in_mode != 10) return HBT_ERR_ARG;
    hbt_reader *rd = new hbt_reader();
    rd->mode = read_in_mode;
    rd->event_buffer_size = event_buffer_size;
    rd->monval = particle_monval;
    if (read_in_mode == 10) {
        rd->gz = gzopen(filename, "rb");
        if (rd->gz) gzbuffer(rd->gz, 1u << 20);
    } else {
        rd->bin = fopen(filename, "rb");
    }
    if (!rd->gz && !rd->bin) {
        delete rd;
        return HBT_ERR_ARG;
    }
    *out = rd;
    return HBT_OK;
}

int hbt_reader_open(const char *filename, int32_t event_buffer_size, int32_t particle_monval,
                    hbt_reader **out) {
    return hbt_reader_open_mode(filename, 10, event_buffer_size, particle_monval, out);
}

int32_t hbt_reader_next(hbt_reader *rd) {
    if (!rd) return HBT_ERR_ARG;
    if (is_species_class(rd->monval) && !rd->has_species) return HBT_ERR_STATE;
    const int32_t n = rd->mode == 9 ? read_binary_buffer(rd) : read_text_buffer(rd);
    return rd->bad ? HBT_ERR_ARG : n;
}

int hbt_reader_set_species(hbt_reader *rd, const hbt_species_table *t) {
    if (!rd || !t) return HBT_ERR_ARG;
    rd->species = t->q;
    rd->has_species = true;
    return HBT_OK;
}

int64_t hbt_reader_bad_monval(const hbt_reader *rd) { return rd && rd->bad ? rd->bad_monval : 0; }

int hbt_reader_monval(const hbt_reader *rd, const int32_t **monval) {
    if (!rd || !monval) return HBT_ERR_ARG;
    *monval = rd->monv.data();
    return HBT_OK;
}

int32_t hbt_reader_eof(const hbt_reader *rd) { return rd && rd->eof ? 1 : 0; }

int hbt_reader_buffer(const hbt_reader *rd, const int64_t **ev_off, const double *soa[8],
                      const double **mass, int64_t *total_read) {
    if (!rd) return HBT_ERR_ARG;
    if (ev_off) *ev_off = rd->off.data();
    if (soa)
        for (int a = 0; a < 8; a++) soa[a] = rd->cols[a].data();
    if (mass) *mass = rd->mass.data();
    if (total_read) *total_read = rd->total;
    return HBT_OK;
}

// output_balance_function (BalanceFunction.cpp:211-318): the sums over
// Delta phi and over Delta y in the reference's loop order, the mixed-event
// normalization N_mixed / N per charge combination, and its three files.
int hbt_bf_write_outputs(const hbt_bf_config *cfg, const double *hist, const char *path) {
    if (!cfg || !hist || !path || cfg->Bnpts < 2) return HBT_ERR_ARG;
    const int np = cfg->Bnpts, nf = 20;
    const double drap = 2. * fabs(cfg->Brap_max) / (np - 1);
    const double Brap_min = -fabs(cfg->Brap_max) - 0.5 * drap;
    const double dphi = 2. * M_PI / nf, Bphi_min = -M_PI / 2.;
    auto H = [&](int a, int i, int j) { return hist[((size_t)a * np + i) * nf + j]; };
    double N_OS = 0., N_OS_mixed = 0., N_SS = 0., N_SS_mixed = 0.;
    std::vector<double> os_y(np, 0.), ss_y(np, 0.), osm_y(np, 0.), ssm_y(np, 0.);
    std::vector<double> os_p(nf, 0.), ss_p(nf, 0.), osm_p(nf, 0.), ssm_p(nf, 0.);
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
    const std::string base = std::string(path) + "/";
    const std::string ab = std::to_string(cfg->particle_alpha) + "_" +
                           std::to_string(cfg->particle_beta);
    // one line: the first field right-aligned in 18 characters, then the
    // rest, %.8e like std::scientific with precision 8
    auto line = [](FILE *f, const char *sep0, const double *v, int n) {
        fprintf(f, "%18.8e", v[0]);
        for (int k = 1; k < n; k++) fprintf(f, "%s%.8e", k == 1 ? sep0 : "  ", v[k]);
        fputc('\n', f);
    };
    FILE *f = fopen((base + "Balance_function_" + ab + "_Delta_y.dat").c_str(), "w");
    if (!f) return HBT_ERR_ARG;
    fputs("# DeltaY  Delta_C2  C2(OS)  rho2(OS)  rho1^2(OS)  C2(SS) rho2(SS)  rho1^2(SS)\n", f);
    for (int i = 0; i < np; i++) {
        const double c_os = os_y[i] / osm_y[i] * N_OS_mixed / N_OS;
        const double c_ss = ss_y[i] / ssm_y[i] * N_SS_mixed / N_SS;
        const double v[8] = {Brap_min + (i + 0.5) * drap, c_os - c_ss, c_os, os_y[i], osm_y[i],
                             c_ss, ss_y[i], ssm_y[i]};
        line(f, "   ", v, 8);
    }
    fclose(f);
    f = fopen((base + "Balance_function_" + ab + "_Delta_phi.dat").c_str(), "w");
    if (!f) return HBT_ERR_ARG;
    fputs("# Delta_phi  Delta_C2  C2(OS)  rho2(OS)  rho1^2(OS)  C2(SS) rho2(SS)  rho1^2(SS)\n",
          f);
    for (int j = 0; j < nf; j++) {
        const double c_os = (os_p[j] / osm_p[j] * N_OS_mixed / N_OS);
        const double c_ss = (ss_p[j] / ssm_p[j] * N_SS_mixed / N_SS);
        const double v[8] = {Bphi_min + (j + 0.5) * dphi, c_os - c_ss, c_os, os_p[j], osm_p[j],
                             c_ss, ss_p[j], ssm_p[j]};
        line(f, "   ", v, 8);
    }
    fclose(f);
    f = fopen((base + "Correlation_function_" + ab + "_2D.dat").c_str(), "w");
    if (!f) return HBT_ERR_ARG;
    fputs("# DY  Dphi  C2(OS)  rho2(OS)  rho1^2(OS)  C2(SS)  rho2(SS)  rho1^2(SS)\n", f);
    for (int i = 0; i < np; i++)
        for (int j = 0; j < nf; j++) {
            const double c_os = ((H(0, i, j) + H(1, i, j)) / (H(4, i, j) + H(5, i, j) + 1e-15));
            const double c_ss = ((H(2, i, j) + H(3, i, j)) / (H(6, i, j) + H(7, i, j) + 1e-15));
            const double v[8] = {Brap_min + (i + 0.5) * drap, Bphi_min + (j + 0.5) * dphi,
                                 c_os * N_OS_mixed / N_OS, H(0, i, j) + H(1, i, j),
                                 H(4, i, j) + H(5, i, j), c_ss * N_SS_mixed / N_SS,
                                 H(2, i, j) + H(3, i, j), H(6, i, j) + H(7, i, j)};
            line(f, "  ", v, 8);
        }
    fclose(f);
    return HBT_OK;
}

int hbt_reader_close(hbt_reader *rd) {
    if (!rd) return HBT_OK;
    if (rd->gz) gzclose(rd->gz);
    if (rd->bin) fclose(rd->bin);
    free(rd->win);
    delete rd;
    return HBT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Species table: the quantum numbers of a PDG table file in the reference's
// format (EOS/pdg.dat, read by particle_decay::read_resonances_list,
// particle_decay.cpp:23-129).  Per species: monval name mass width gspin
// baryon strange charm bottom gisospin charge decays, then `decays` channel
// records of 8 numbers.  As the reference does, every species with baryon
// number 1 gets an anti-species entry -monval with baryon, strangeness and
// charge negated, and a decay product of such a baryon that has no entry yet
// gets one with zero quantum numbers (the reference's map lookup creates it;
// a later entry of that species replaces it).


extern "C" {

int hbt_species_read(const char *path, hbt_species_table **out) {
    if (!path || !out) return HBT_ERR_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f) return HBT_ERR_ARG;
    std::unique_ptr<hbt_species_table> t(new hbt_species_table());
    auto put = [&](int32_t m, hbt_species_table::Q v) {
        if (t->q.find(m) == t->q.end()) t->order.push_back(m);
        t->q[m] = v;
    };
    long long monval;
    while (f >> monval) {
        std::string name;
        double mass, width;
        long long gspin, baryon, strange, charm, bottom, gisospin, charge, decays;
        if (!(f >> name >> mass >> width >> gspin >> baryon >> strange >> charm >> bottom >>
              gisospin >> charge >> decays) ||
            decays < 0)
            return HBT_ERR_ARG;
        std::vector<long long> parts;
        for (long long d = 0; d < decays; d++) {
            long long dummy, npart, p[5];
            double br;
            if (!(f >> dummy >> npart >> br >> p[0] >> p[1] >> p[2] >> p[3] >> p[4]))
                return HBT_ERR_ARG;
            parts.insert(parts.end(), p, p + 5);
        }
        put((int32_t)monval, {(int32_t)charge, (int32_t)baryon, (int32_t)strange});
        if (baryon == 1) {
            for (long long d : parts)
                if (d != 0 && t->q.find((int32_t)d) == t->q.end()) put((int32_t)d, {0, 0, 0});
            put((int32_t)-monval, {(int32_t)-charge, (int32_t)-baryon, (int32_t)-strange});
        }
    }
    if (!f.eof()) return HBT_ERR_ARG;  // a token that is not a number
    *out = t.release();
    return HBT_OK;
}

int32_t hbt_species_count(const hbt_species_table *t) { return t ? (int32_t)t->order.size() : 0; }

int hbt_species_get(const hbt_species_table *t, int32_t i, int32_t *monval, int32_t *charge,
                    int32_t *baryon, int32_t *strange) {
    if (!t || i < 0 || i >= (int32_t)t->order.size()) return HBT_ERR_ARG;
    const int32_t m = t->order[i];
    const hbt_species_table::Q &v = t->q.at(m);
    if (monval) *monval = m;
    if (charge) *charge = v.charge;
    if (baryon) *baryon = v.baryon;
    if (strange) *strange = v.strange;
    return HBT_OK;
}

int hbt_species_free(hbt_species_table *t) {
    delete t;
    return HBT_OK;
}

}  // extern "C"

