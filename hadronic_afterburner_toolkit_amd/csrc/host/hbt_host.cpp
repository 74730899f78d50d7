// hbt_host.cpp -- host-side pieces of the C ABI that need no GPU: the
// reference's text writers, its mode-10 sample reader and psi_ref.
// Compiled into libhbt_hip.so next to the engine.
#include <math.h>
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
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

// ------------------------------------------------------------ mode-10 reader

struct hbt_reader {
    gzFile gz = nullptr;
    int event_buffer_size = 0;
    int monval = 0;
    bool eof = false;
    // last buffer: filtered particles (SoA) and their event offsets
    std::vector<int64_t> off;
    std::vector<double> cols[8];  // px py pz E x y z t
    std::vector<double> mass;
    long long total = 0;
};

namespace {

// One line without its '\n', byte at a time like gz_readline
// (particleSamples.cpp:2209-2218); end of file is flagged only when a read
// finds no more data, as gzeof does.
bool read_line(hbt_reader *rd, std::string &line) {
    line.clear();
    for (;;) {
        const int ch = gzgetc(rd->gz);
        if (ch < 0) {
            rd->eof = true;
            return false;
        }
        if (ch == '\n') return true;
        line.push_back((char)ch);
    }
}

}  // namespace

int hbt_reader_open(const char *filename, int32_t event_buffer_size, int32_t particle_monval,
                    hbt_reader **out) {
    if (!filename || !out) return HBT_ERR_ARG;
    *out = nullptr;
    // species selections other than one Monte-Carlo id need the PDG tables
    // (particleSamples.cpp:625-678), which this reader does not load
    if (particle_monval == 9999 || particle_monval == 9998 || particle_monval == -9998 ||
        particle_monval == 9997 || particle_monval == -9997 || particle_monval == 9996 ||
        particle_monval == -9996)
        return HBT_ERR_ARG;
    gzFile gz = gzopen(filename, "rb");
    if (!gz) return HBT_ERR_ARG;
    hbt_reader *rd = new hbt_reader();
    rd->gz = gz;
    rd->event_buffer_size = event_buffer_size;
    rd->monval = particle_monval;
    *out = rd;
    return HBT_OK;
}

// particleSamples.cpp:1247-1286 (whole events while the running particle
// total of all species is below event_buffer_size) and :1329-1346 (species
// filter by exact Monte-Carlo id; 99999 keeps everything).
int32_t hbt_reader_next(hbt_reader *rd) {
    if (!rd) return HBT_ERR_ARG;
    rd->off.assign(1, 0);
    for (auto &c : rd->cols) c.clear();
    rd->mass.clear();
    rd->total = 0;
    long long num_particles = 0;
    int nev = 0;
    std::string line;
    while (num_particles < rd->event_buffer_size) {
        read_line(rd, line);
        if (rd->eof) break;
        const int n_particle = (int)strtol(line.c_str(), nullptr, 10);
        for (int ip = 0; ip < n_particle; ip++) {
            read_line(rd, line);
            char *p = const_cast<char *>(line.c_str());
            char *e = nullptr;
            const int mon = (int)strtol(p, &e, 10);
            double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // mass t x y z E px py pz
            p = e;
            for (int k = 0; k < 9; k++) {
                v[k] = strtod(p, &e);
                p = e;
            }
            if (rd->monval == 99999 || mon == rd->monval) {
                rd->cols[0].push_back(v[6]);
                rd->cols[1].push_back(v[7]);
                rd->cols[2].push_back(v[8]);
                rd->cols[3].push_back(v[5]);
                rd->cols[4].push_back(v[2]);
                rd->cols[5].push_back(v[3]);
                rd->cols[6].push_back(v[4]);
                rd->cols[7].push_back(v[1]);
                rd->mass.push_back(v[0]);
            }
        }
        num_particles += n_particle;
        nev++;
        rd->off.push_back((int64_t)rd->cols[0].size());
    }
    rd->total = num_particles;
    return nev;
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

int hbt_reader_close(hbt_reader *rd) {
    if (!rd) return HBT_OK;
    if (rd->gz) gzclose(rd->gz);
    delete rd;
    return HBT_OK;
}

}  // extern "C"
