// BalanceFunction.cpp -- see BalanceFunction.h.  Keys and grids as the
// reference's constructor (BalanceFunction.cpp:15-61), one buffer per
// calculate_balance_function (:63-101) with the mixing draws taken from the
// shared random stream in the reference's order (:165-169), the files of
// output_balance_function (:211-318).
#include "BalanceFunction.h"

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <vector>

BalanceFunction::BalanceFunction(ParameterReader &paraRdr, std::string path,
                                 std::shared_ptr<RandomUtil::Random> ran_gen)
    : path_(path), ran_gen_ptr(ran_gen) {
    cfg_.particle_alpha = (int)paraRdr.getVal("particle_alpha");
    cfg_.particle_beta = (int)paraRdr.getVal("particle_beta");
    cfg_.BpT_min = paraRdr.getVal("BpT_min");
    cfg_.BpT_max = paraRdr.getVal("BpT_max");
    cfg_.Bnpts = (int)paraRdr.getVal("Bnpts");
    cfg_.Brap_max = paraRdr.getVal("Brap_max");
    cfg_.rap_type = (int)paraRdr.getVal("rap_type", 1);
    cfg_.device = (int)paraRdr.getVal("hbt_device", 0);
    cfg_.kernel = (int)paraRdr.getVal("hbt_bf_kernel", HBT_BF_KERNEL_FAST);
    check(hbt_bf_create(&cfg_, &ctx_), "hbt_bf_create");
    // species classes pick by charge, baryon number or strangeness from the
    // PDG table the reference's particle_decay reads (particle_decay.cpp:26-36)
    bool classes = false;
    for (int id : {cfg_.particle_alpha, cfg_.particle_beta})
        classes = classes || id == 9999 || id == 9998 || id == -9998 || id == 9997 ||
                  id == -9997 || id == 9996 || id == -9996;
    if (classes) {
        const std::string pdg = paraRdr.getVal("resonance_weak_feed_down_flag", 0) == 1
                                    ? "EOS/pdg_weak.dat"
                                    : "EOS/pdg.dat";
        hbt_species_table *t = nullptr;
        if (hbt_species_read(pdg.c_str(), &t) != HBT_OK) {
            std::cout << "can not find the file " << pdg << "!" << std::endl;
            exit(1);
        }
        const int n = hbt_species_count(t);
        std::vector<int32_t> m(n), q(n), b(n), s(n);
        for (int i = 0; i < n; i++) hbt_species_get(t, i, &m[i], &q[i], &b[i], &s[i]);
        hbt_species_free(t);
        check(hbt_bf_set_species(ctx_, n, m.data(), q.data(), b.data(), s.data()),
              "hbt_bf_set_species");
    }
}

BalanceFunction::~BalanceFunction() { hbt_bf_destroy(ctx_); }

void BalanceFunction::check(int rc, const char *what) const {
    if (rc != HBT_OK) {
        std::cout << "BalanceFunction: " << what << " failed (" << rc
                  << "): " << hbt_bf_last_error(ctx_) << std::endl;
        exit(1);
    }
}

void BalanceFunction::calculate_balance_function(
    std::shared_ptr<particleSamples> particle_list_in) {
    set_particle_list(particle_list_in);
    const particleSamples::Events &ev = particle_list->same_events();
    const int nev = (int)ev.off.size() - 1;
    const bool sep = particle_list->mixed_is_separate();
    const particleSamples::Events &mix = particle_list->mixed_events();
    const int mix_nev = (int)mix.off.size() - 1;
    if (nev > 0 && mix_nev < 1) {
        std::cout << "BalanceFunction: the mixed-event buffer is empty" << std::endl;
        exit(1);
    }
    // C_mixed_ab, C_mixed_abarbbar, C_mixed_abbar, C_mixed_abarb: per event
    // a partner, then a rotation (BalanceFunction.cpp:165-169)
    std::vector<int32_t> plan_ev((size_t)4 * nev);
    std::vector<double> plan_rot((size_t)4 * nev);
    for (int h = 0; h < 4; h++)
        for (int e = 0; e < nev; e++) {
            plan_ev[(size_t)h * nev + e] = ran_gen_ptr->rand_int_uniform() % mix_nev;
            plan_rot[(size_t)h * nev + e] = ran_gen_ptr->rand_uniform() * 2. * M_PI;
        }
    // cols = {mass, px, py, pz, E}; Events::cols = {px, py, pz, E, x, y, z, t}
    const double *cols[5] = {ev.mass.data(), ev.cols[0].data(), ev.cols[1].data(),
                             ev.cols[2].data(), ev.cols[3].data()};
    const double *mcols[5] = {mix.mass.data(), mix.cols[0].data(), mix.cols[1].data(),
                              mix.cols[2].data(), mix.cols[3].data()};
    check(hbt_bf_calculate_planned(ctx_, nev, ev.off.data(), ev.monval.data(), cols,
                                   sep ? mix_nev : 0, sep ? mix.off.data() : nullptr,
                                   sep ? mix.monval.data() : nullptr, sep ? mcols : nullptr,
                                   plan_ev.data(), plan_rot.data()),
          "hbt_bf_calculate_planned");
}

void BalanceFunction::output_balance_function() {
    std::vector<double> hist((size_t)8 * cfg_.Bnpts * 20);
    check(hbt_bf_read(ctx_, hist.data(), nullptr, nullptr), "hbt_bf_read");
    check(hbt_bf_write_outputs(&cfg_, hist.data(), path_.c_str()), "hbt_bf_write_outputs");
}
