// Random.cpp -- see Random.h (Random.cpp:7-14 of the reference).
#include "Random.h"

namespace RandomUtil {

Random::Random(int seed_in, double min, double max) : seed(seed_in), rand_uniform_dist(min, max) {
    if (seed == -1) {
        std::random_device dev;
        seed = dev();
    }
    ran_generator.reset(new std::mt19937(seed));
}

}  // namespace RandomUtil
