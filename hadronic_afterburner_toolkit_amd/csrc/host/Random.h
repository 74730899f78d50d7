// RandomUtil::Random -- the reference's random stream (src/Random.h:11-25):
// std::mt19937(seed) (seed -1 draws from std::random_device),
// uniform_real_distribution<double>(min, max) and
// uniform_int_distribution<int>() over [0, INT_MAX].  The mixing plan's
// partner ids and rotation angles are drawn from it in the reference's order.
#ifndef HBT_AMD_RANDOM_H
#define HBT_AMD_RANDOM_H

#include <memory>
#include <random>

namespace RandomUtil {

class Random {
  public:
    Random(int seed_in, double min = 0.0, double max = 1.0);
    double rand_uniform() { return rand_uniform_dist(*ran_generator); }
    int rand_int_uniform() { return rand_int_uniform_dist(*ran_generator); }
    int get_seed() const { return seed; }

  private:
    int seed;
    std::unique_ptr<std::mt19937> ran_generator;
    std::uniform_real_distribution<double> rand_uniform_dist;
    std::uniform_int_distribution<int> rand_int_uniform_dist;
};

}  // namespace RandomUtil

#endif
