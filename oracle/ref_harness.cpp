// ref_harness.cpp -- drives the reference's own HBT_correlation (compiled
// from its sources into oracle/_ref/libhbt_ref.so) the way
// Analysis::HBTAnalysis does (Analysis.cpp:49-52, :817-835), then dumps the
// raw histograms and counters so the oracle and the GPU path can be compared
// with it bit for bit.  Test infrastructure only; runs in the build
// container, where /root/reference exists.
//
//   ref_harness <workdir> <dumpfile> [key=value ...]
//
// <workdir> holds parameters.dat, EOS/ (particle_decay.cpp:28 opens
// EOS/pdg.dat relative to the working directory) and results/ with
// particle_samples.gz; the reference writes its .dat outputs to results/.
#include <stdint.h>
#include <time.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <utility>
#include <vector>
#include <zlib.h>

// The histograms are private members; the dump needs to read them.
#define private public
#include "HBT_correlation.h"
#undef private

static double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void put(FILE *f, const void *p, size_t n) {
    if (fwrite(p, 1, n, f) != n) {
        perror("fwrite");
        exit(2);
    }
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <workdir> <dumpfile> [key=value ...]\n", argv[0]);
        return 1;
    }
    std::string dump = argv[2];
    if (chdir(argv[1]) != 0) {
        perror("chdir");
        return 1;
    }
    ParameterReader paraRdr;
    paraRdr.readFromFile("parameters.dat");
    paraRdr.readFromArguments(argc, argv, "#", 3);

    int randomSeed = paraRdr.getVal("randomSeed", -1);
    auto rng = std::make_shared<RandomUtil::Random>(randomSeed);
    const std::string path = "results";
    auto plist = std::make_shared<particleSamples>(paraRdr, path, rng);

    HBT_correlation hbt(paraRdr, path, rng);
    if (paraRdr.getVal("harness_psi_only", 0) == 1) {
        // psi_ref for n = 1, 2, 3 on the first buffer (the computation that
        // HBT_unittest.cc:23-48 pins), plus the particles it was computed on.
        plist->read_in_particle_samples_and_filter();
        hbt.set_particle_list(plist);
        FILE *f = fopen(dump.c_str(), "wb");
        if (!f) return 1;
        for (int n = 1; n <= 3; n++) {
            hbt.calculate_flow_event_plane_angle(n);
            double psi = hbt.get_psi_ref();
            put(f, &psi, 8);
        }
        int32_t nev = plist->get_number_of_events();
        put(f, &nev, 4);
        for (int e = 0; e < nev; e++) {
            int32_t np = plist->get_number_of_particles(e);
            put(f, &np, 4);
            for (int i = 0; i < np; i++) {
                double pxy[2] = {plist->get_particle(e, i).px, plist->get_particle(e, i).py};
                put(f, pxy, 16);
            }
        }
        fclose(f);
        return 0;
    }
    double t_calc = 0.0;
    int nbuf = 0;
    std::vector<int> nev_per_buffer;
    while (!plist->end_of_file()) {
        plist->read_in_particle_samples_and_filter();
        plist->read_in_particle_samples_mixed_event_and_filter();
        nev_per_buffer.push_back(plist->get_number_of_events());
        double t0 = now_s();
        hbt.calculate_HBT_correlation_function(plist);
        t_calc += now_s() - t0;
        nbuf++;
    }
    hbt.output_HBTcorrelation();

    FILE *f = fopen(dump.c_str(), "wb");
    if (!f) {
        perror("fopen dump");
        return 1;
    }
    const int az = hbt.azimuthal_flag_, qi = hbt.invariant_radius_flag_;
    const int nKT = hbt.n_KT, nphi = hbt.n_Kphi, q = hbt.qnpts;
    int32_t hdr[8] = {0x48425431, nKT, nphi, q, az, qi, nbuf, 0};
    put(f, hdr, sizeof(hdr));
    double extra[2] = {hbt.psi_ref, t_calc};
    put(f, extra, sizeof(extra));
    // 3D arrays in flat [kt][kphi][iqout][iqside][iqlong] order
    const int np = az ? nphi : 1;
    std::vector<double> buf((size_t)nKT * np * q * q * q);
    for (int a = 0; a < 6; a++) {
        size_t k = 0;
        for (int iK = 0; iK < nKT; iK++)
            for (int ip = 0; ip < np; ip++)
                for (int io = 0; io < q; io++)
                    for (int is = 0; is < q; is++)
                        for (int il = 0; il < q; il++) {
                            double v;
                            if (!az) {
                                double ****arr[6] = {hbt.correl_3d_num_count, hbt.q_out_mean,
                                                     hbt.q_side_mean, hbt.q_long_mean,
                                                     hbt.correl_3d_num, hbt.correl_3d_denorm};
                                v = arr[a][iK][io][is][il];
                            } else {
                                double *****arr[6] = {
                                    hbt.correl_3d_Kphi_diff_num_count, hbt.q_out_diff_mean,
                                    hbt.q_side_diff_mean, hbt.q_long_diff_mean,
                                    hbt.correl_3d_Kphi_diff_num,
                                    hbt.correl_3d_Kphi_diff_denorm};
                                v = arr[a][iK][ip][io][is][il];
                            }
                            buf[k++] = v;
                        }
        put(f, buf.data(), buf.size() * sizeof(double));
    }
    // q_inv arrays [kt][iq]
    if (qi) {
        double **qarr[4] = {hbt.correl_1d_inv_num_count, hbt.q_inv_mean,
                            hbt.correl_1d_inv_num, hbt.correl_1d_inv_denorm};
        for (int a = 0; a < 4; a++)
            for (int iK = 0; iK < nKT; iK++) put(f, qarr[a][iK], sizeof(double) * q);
    }
    // counters
    put(f, hbt.number_of_pairs_numerator_KTdiff.data(), 8 * nKT);
    put(f, hbt.number_of_pairs_denormenator_KTdiff.data(), 8 * nKT);
    put(f, hbt.number_of_pairs_numerator_KTdiff_qinv_.data(), 8 * nKT);
    put(f, hbt.number_of_pairs_denormenator_KTdiff_qinv_.data(), 8 * nKT);
    if (az) {
        for (int iK = 0; iK < nKT; iK++)
            put(f, hbt.number_of_pairs_numerator_KTKphidiff[iK], 8 * nphi);
        for (int iK = 0; iK < nKT; iK++)
            put(f, hbt.number_of_pairs_denormenator_KTKphidiff[iK], 8 * nphi);
    }
    int32_t nb = (int32_t)nev_per_buffer.size();
    put(f, &nb, 4);
    put(f, nev_per_buffer.data(), 4 * nev_per_buffer.size());
    fclose(f);
    return 0;
}
