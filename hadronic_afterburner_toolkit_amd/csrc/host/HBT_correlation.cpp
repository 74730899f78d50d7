// HBT_correlation.cpp -- see HBT_correlation.h.  Mirrors
// src/HBT_correlation.cpp of the reference: the constructor reads the same
// keys (:22-46); calculate_HBT_correlation_function keeps its steps and
// random-draw order (:177-218); the output methods write the same files
// (:694-855).
#include "HBT_correlation.h"

#include <cmath>
#include <cstdlib>
#include <iostream>

HBT_correlation::HBT_correlation(ParameterReader &paraRdr, std::string path,
                                 std::shared_ptr<RandomUtil::Random> ran_gen)
    : paraRdr_(paraRdr), path_(path), ran_gen_ptr_(ran_gen) {
    hbt_config &c = cfg_;
    c = hbt_config();
    c.long_comoving_boost = paraRdr_.getVal("long_comoving_boost") == 1 ? 1 : 0;
    c.qnpts = (int)paraRdr_.getVal("qnpts");
    c.q_min = paraRdr_.getVal("q_min");
    c.q_max = paraRdr_.getVal("q_max");
    c.needed_number_of_pairs = paraRdr_.getVal("needed_number_of_pairs");
    c.azimuthal_flag = (int)paraRdr_.getVal("azimuthal_flag");
    c.invariant_radius_flag = (int)paraRdr_.getVal("invariant_radius_flag");
    c.n_KT = (int)paraRdr_.getVal("n_KT");
    c.n_Kphi = (int)paraRdr_.getVal("n_Kphi");
    c.KT_min = paraRdr_.getVal("KT_min");
    c.KT_max = paraRdr_.getVal("KT_max");
    c.HBTrap_min = paraRdr_.getVal("HBTrap_min");
    c.HBTrap_max = paraRdr_.getVal("HBTrap_max");
    c.number_of_mixed_events = (int)paraRdr_.getVal("number_of_mixed_events", 0);
    c.precision = (int)paraRdr_.getVal("hbt_precision", HBT_PREC_FP64);
    c.kernel = (int)paraRdr_.getVal("hbt_kernel", HBT_KERNEL_FAST);
    c.device = (int)paraRdr_.getVal("hbt_device", 0);
    c.kphi_guard_prad = (int)paraRdr_.getVal("hbt_kphi_guard_prad", 0);
    c.binning = (int)paraRdr_.getVal("hbt_binning", 0);
    check(hbt_create(&c, &ctx_), "hbt_create");
}

HBT_correlation::~HBT_correlation() { hbt_destroy(ctx_); }

void HBT_correlation::check(int rc, const char *what) const {
    if (rc == HBT_OK) return;
    std::cerr << "[Error] HBT_correlation: " << what << " failed (" << rc
              << "): " << hbt_last_error(ctx_) << std::endl;
    exit(1);
}

// The mixed-event list is a separate file only with read_in_real_mixed_events
// = 1; otherwise the reference's particleSamples aliases it to the
// same-event list (particleSamples.cpp:528-531).
bool HBT_correlation::mixed_is_separate() const {
    return paraRdr_.getVal("read_in_real_mixed_events", 0) == 1;
}

// The listed events of the same-event (or mixed-event) list, gathered through
// the reference's getters into the SoA columns px py pz E x y z t.
void HBT_correlation::gather(bool mixed, const std::vector<int> &event_list) {
    particleSamples &ps = *particle_list;
    off_.assign(1, 0);
    size_t n = 0;
    for (int e : event_list)
        n += (size_t)(mixed ? ps.get_number_of_particles_mixed_event(e)
                            : ps.get_number_of_particles(e));
    for (auto &c : cols_) c.resize(n);
    size_t k = 0;
    for (int e : event_list) {
        const int ne = mixed ? ps.get_number_of_particles_mixed_event(e)
                             : ps.get_number_of_particles(e);
        for (int i = 0; i < ne; i++, k++) {
            const particle_info p =
                mixed ? ps.get_particle_from_mixed_event(e, i) : ps.get_particle(e, i);
            cols_[0][k] = p.px;
            cols_[1][k] = p.py;
            cols_[2][k] = p.pz;
            cols_[3][k] = p.E;
            cols_[4][k] = p.x;
            cols_[5][k] = p.y;
            cols_[6][k] = p.z;
            cols_[7][k] = p.t;
        }
        off_.push_back((int64_t)k);
    }
}

// The listed events as one buffer in an engine slot.
void HBT_correlation::upload_events(int slot, bool mixed, const std::vector<int> &event_list) {
    gather(mixed, event_list);
    const double *soa[8];
    for (int a = 0; a < 8; a++) soa[a] = cols_[a].data();
    check(hbt_upload(ctx_, slot, (int32_t)event_list.size(), off_.data(), soa), "hbt_upload");
}

static std::vector<int> all_events(int nev) {
    std::vector<int> v((size_t)nev);
    for (int e = 0; e < nev; e++) v[(size_t)e] = e;
    return v;
}

// :233-249, over all particles of the buffer (no rapidity cut)
void HBT_correlation::calculate_flow_event_plane_angle(int n_order) {
    gather(false, all_events(particle_list->get_number_of_events()));
    psi_ref = hbt_flow_event_plane_angle((int32_t)cols_[0].size(), cols_[0].data(),
                                         cols_[1].data(), n_order);
}

void HBT_correlation::calculate_HBT_correlation_function(
    std::shared_ptr<particleSamples> particle_list_in) {
    set_particle_list(particle_list_in);
    const int nev = particle_list->get_number_of_events();
    number_of_oversample_events_ = nev;
    const int mixed_nev = particle_list->get_number_of_mixed_events();
    // slot 1 is left empty when the mixed list aliases the same-event list;
    // one an earlier combine_and_bin_particle_pairs_mixed_events call loaded
    // is overwritten with the mixed list the getters give
    if (mixed_is_separate() || slot1_loaded_) {
        upload_events(1, true, all_events(mixed_nev));
        slot1_loaded_ = true;
    }
    upload_events(0, false, all_events(nev));
    if (cfg_.azimuthal_flag == 1) {  // :181-183, on the columns just gathered
        psi_ref = hbt_flow_event_plane_angle((int32_t)cols_[0].size(), cols_[0].data(),
                                             cols_[1].data(), 2);
        check(hbt_set_psi_ref(ctx_, psi_ref), "hbt_set_psi_ref");
    }
    check(hbt_run_same(ctx_), "hbt_run_same");  // :190-194
    if (nev == 0) return;
    // :197-216: M partners per event, then inside the mixed call M angles
    const int M = cfg_.number_of_mixed_events > 0 ? cfg_.number_of_mixed_events
                                                  : mixed_nev / 2 + 1;
    std::vector<int32_t> ids((size_t)nev * M);
    std::vector<double> cs((size_t)2 * nev * M);
    for (int iev = 0; iev < nev; iev++) {
        for (int count = 0; count < M; count++) {
            int id = ran_gen_ptr_->rand_int_uniform() % mixed_nev;
            while (iev == id && mixed_nev != 1) id = ran_gen_ptr_->rand_int_uniform() % mixed_nev;
            ids[(size_t)iev * M + count] = id;
        }
        for (int k = 0; k < M; k++) {  // :495-497
            const double random_rotation = ran_gen_ptr_->rand_uniform() * 2 * M_PI;
            double sn, cn;
            sincos(random_rotation, &sn, &cn);  // as the reference's build evaluates cos, sin
            cs[2 * ((size_t)iev * M + k)] = cn;
            cs[2 * ((size_t)iev * M + k) + 1] = sn;
        }
    }
    check(hbt_run_mixed(ctx_, M, ids.data(), cs.data()), "hbt_run_mixed");
}

// :251-462 on the listed events of the current particle list
void HBT_correlation::combine_and_bin_particle_pairs(std::vector<int> event_list) {
    upload_events(0, false, event_list);
    check(hbt_run_same(ctx_), "hbt_run_same");
}

// :464-692: event_id against the listed mixed events; draws the rotation
// angles from the shared stream like the reference does inside this call
void HBT_correlation::combine_and_bin_particle_pairs_mixed_events(
    int event_id, std::vector<int> mixed_event_list) {
    const int M = (int)mixed_event_list.size();
    if (M == 0) return;
    // the mixed list goes to slot 1 first: an upload of slot 1 alone never
    // aliases slot 0
    upload_events(1, true, all_events(particle_list->get_number_of_mixed_events()));
    slot1_loaded_ = true;
    upload_events(0, false, std::vector<int>{event_id});
    std::vector<double> cs(2 * (size_t)M);
    for (int k = 0; k < M; k++) {
        const double random_rotation = ran_gen_ptr_->rand_uniform() * 2 * M_PI;
        double sn, cn;
        sincos(random_rotation, &sn, &cn);
        cs[2 * k] = cn;
        cs[2 * k + 1] = sn;
    }
    check(hbt_run_mixed(ctx_, M, mixed_event_list.data(), cs.data()), "hbt_run_mixed");
}

// Reads the raw results and runs the selected writers (:694-855).
void HBT_correlation::write(int what) {
    const long q = cfg_.qnpts, nphi = cfg_.azimuthal_flag ? cfg_.n_Kphi : 1;
    const size_t nb = (size_t)cfg_.n_KT * nphi * q * q * q, nq = (size_t)cfg_.n_KT * q;
    std::vector<uint64_t> num_count(nb), den(nb), qc(nq), qd(nq);
    std::vector<double> qo(nb), qs(nb), ql(nb), num(nb), qsum(nq), qnum(nq);
    std::vector<uint64_t> c1(cfg_.n_KT), c2(cfg_.n_KT), c3(cfg_.n_KT), c4(cfg_.n_KT);
    std::vector<uint64_t> c5((size_t)cfg_.n_KT * nphi), c6((size_t)cfg_.n_KT * nphi);
    hbt_result r = {num_count.data(), den.data(), qo.data(), qs.data(), ql.data(), num.data(),
                    qc.data(), qd.data(), qsum.data(), qnum.data(), c1.data(), c2.data(),
                    c3.data(), c4.data(), c5.data(), c6.data()};
    check(hbt_read(ctx_, &r), "hbt_read");
    const int eco = (int)paraRdr_.getVal("ecoOutput", 0);
    check(hbt_write_outputs(&cfg_, &r, path_.c_str(), eco, what), "hbt_write_outputs");
}

void HBT_correlation::output_HBTcorrelation() { write(0); }  // :220-228
void HBT_correlation::output_correlation_function_inv() { write(HBT_WRITE_INV); }
void HBT_correlation::output_correlation_function() { write(HBT_WRITE_3D); }
void HBT_correlation::output_correlation_function_Kphi_differential() { write(HBT_WRITE_KPHI); }
