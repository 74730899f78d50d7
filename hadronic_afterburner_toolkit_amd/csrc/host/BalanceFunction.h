// BalanceFunction -- drop-in replacement of the reference's class
// (src/BalanceFunction.h:15-75): same constructor, calculate and output
// methods, same files; the pair histograms are filled on the MI355X through
// the C ABI (hbt_bf_*, include/hbt_hip.h).  The particle container must hold
// all species (particle_monval 99999), as the reference's particleSamples
// keeps its full list for the balance-function filters.  Engine keys read
// from the parameters (optional): hbt_device (default 0), hbt_bf_kernel
// (0 fast form, default; 1 the FP64 arithmetic for every pair).
#ifndef HBT_AMD_BALANCEFUNCTION_H
#define HBT_AMD_BALANCEFUNCTION_H

#include <memory>
#include <string>

#include "ParameterReader.h"
#include "Random.h"
#include "hbt_hip.h"
#include "particleSamples.h"

class BalanceFunction {
  public:
    BalanceFunction(ParameterReader &paraRdr, std::string path,
                    std::shared_ptr<RandomUtil::Random> ran_gen);
    ~BalanceFunction();

    void set_particle_list(std::shared_ptr<particleSamples> particle_list_in) {
        particle_list = particle_list_in;
    }
    void calculate_balance_function(std::shared_ptr<particleSamples> particle_list_in);
    void output_balance_function();

  private:
    std::string path_;
    std::shared_ptr<RandomUtil::Random> ran_gen_ptr;
    std::shared_ptr<particleSamples> particle_list;
    hbt_bf_config cfg_{};
    hbt_bf_ctx *ctx_ = nullptr;
    void check(int rc, const char *what) const;
};

#endif
