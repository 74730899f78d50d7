// ref_bf_harness.cpp -- drives the reference's own BalanceFunction (compiled
// from its sources into oracle/_ref/libhbt_ref.so) the way
// Analysis::BalanceFunctionAnalysis does (Analysis.cpp:856-874), then dumps
// the eight pair histograms and the b / bbar counts so the oracle and the
// GPU path can be compared with it bit for bit.  Test infrastructure only;
// runs in the build container, where /root/reference exists.
//
//   ref_bf_harness <workdir> <dumpfile> [key=value ...]
//
// <workdir> holds parameters.dat, EOS/ (particle_decay.cpp opens EOS/pdg.dat)
// and results/ with particle_samples.gz; the reference writes its
// Balance_function_*.dat and Correlation_function_*.dat files to results/.
#include <stdint.h>
#include <unistd.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include <zlib.h>

// The histograms are private members; the dump needs to read them.
#define private public
#include "BalanceFunction.h"
#undef private

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
    BalanceFunction bf(paraRdr, path, rng);
    std::vector<int32_t> nev_per_buffer;
    while (!plist->end_of_file()) {
        plist->read_in_particle_samples_and_filter();
        plist->read_in_particle_samples_mixed_event_and_filter();
        nev_per_buffer.push_back(plist->get_number_of_events());
        bf.calculate_balance_function(plist);
    }
    bf.output_balance_function();

    FILE *f = fopen(dump.c_str(), "wb");
    if (!f) {
        perror("fopen dump");
        return 1;
    }
    int32_t hdr[4] = {0x42465231, bf.Bnpts, bf.Bnphi, (int32_t)nev_per_buffer.size()};
    put(f, hdr, sizeof(hdr));
    int64_t nb[2] = {(int64_t)bf.N_b, (int64_t)bf.N_bbar};
    put(f, nb, sizeof(nb));
    const std::vector<std::vector<double>> *h[8] = {
        &bf.C_ab,       &bf.C_abarbbar,       &bf.C_abbar,       &bf.C_abarb,
        &bf.C_mixed_ab, &bf.C_mixed_abarbbar, &bf.C_mixed_abbar, &bf.C_mixed_abarb};
    for (auto *a : h)
        for (int i = 0; i < bf.Bnpts; i++) put(f, (*a)[i].data(), sizeof(double) * bf.Bnphi);
    put(f, nev_per_buffer.data(), 4 * nev_per_buffer.size());
    fclose(f);
    return 0;
}
