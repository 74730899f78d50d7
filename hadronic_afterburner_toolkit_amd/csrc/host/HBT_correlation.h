// HBT_correlation -- drop-in replacement of the reference's class
// (src/HBT_correlation.h:14-104): same constructor, methods and output
// files; the pair loops run on the MI355X through the C ABI
// (include/hbt_hip.h).  It reads the particle container only through the
// reference's particleSamples getters (src/particleSamples.h:213-253:
// get_number_of_events, get_number_of_particles, get_particle and their
// mixed-event counterparts), so it compiles against the reference's
// particleSamples as well as against this repository's.  Engine keys read
// from the parameters (optional):
//   hbt_precision  0 = FP64 (bit-exact counts, default), 1 = precision form
//   hbt_kernel     0 = plain LDS-tiled kernel, 1 = fast kernel (default)
//   hbt_device     HIP device ordinal (default 0)
//   hbt_kphi_guard_prad  K_phi guard band in 1e-12 rad (default 1): pairs
//                  that close to a K_phi bin edge are binned on the host
//   hbt_binning    0 = device-scope atomics (default), 1 = slab binning
//   number_of_mixed_events  partners per event (default: the reference's
//                           floor(nev/2)+1)
#ifndef HBT_AMD_HBT_CORRELATION_H
#define HBT_AMD_HBT_CORRELATION_H

#include <memory>
#include <string>
#include <vector>

#include "ParameterReader.h"
#include "Random.h"
#include "hbt_hip.h"
#include "particleSamples.h"

class HBT_correlation {
  public:
    HBT_correlation(ParameterReader &paraRdr, std::string path,
                    std::shared_ptr<RandomUtil::Random> ran_gen);
    ~HBT_correlation();

    double get_psi_ref() { return psi_ref; }
    void set_particle_list(std::shared_ptr<particleSamples> particle_list_in) {
        particle_list = particle_list_in;
    }

    void calculate_flow_event_plane_angle(int n_order);
    void calculate_HBT_correlation_function(std::shared_ptr<particleSamples> particle_list_in);
    void combine_and_bin_particle_pairs(std::vector<int> event_list);
    void combine_and_bin_particle_pairs_mixed_events(int event_id,
                                                     std::vector<int> mixed_event_list);

    void output_HBTcorrelation();
    void output_correlation_function_inv();
    void output_correlation_function();
    void output_correlation_function_Kphi_differential();

  private:
    const ParameterReader paraRdr_;
    const std::string path_;
    std::shared_ptr<particleSamples> particle_list;
    std::shared_ptr<RandomUtil::Random> ran_gen_ptr_;
    hbt_config cfg_;
    hbt_ctx *ctx_ = nullptr;
    double psi_ref = 0.0;
    int number_of_oversample_events_ = 0;

    // SoA staging of the buffer handed to the engine (reused across calls)
    std::vector<int64_t> off_;
    std::vector<double> cols_[8];
    // engine slot 1 holds a mixed list uploaded by an earlier call
    bool slot1_loaded_ = false;

    void check(int rc, const char *what) const;
    bool mixed_is_separate() const;
    void gather(bool mixed, const std::vector<int> &event_list);
    void upload_events(int slot, bool mixed, const std::vector<int> &event_list);
    void write(int which);
};

#endif
