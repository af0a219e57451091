// spk_master_rng.h -- the `seed S` command's master generator: Marsaglia and
// Zaman's universal generator ("RANMAR", Marsaglia, Zaman & Tsang 1990;
// F. James, Comput. Phys. Commun. 60 (1990) 329), seeded and stepped as
// SPPARKS' RanMars (src/random_mars.cpp:43-105).  Header-only, host code;
// shared by libspk_hip.so and the spk_hip front end.
#pragma once

namespace spk {

class MasterRng {
 public:
  explicit MasterRng(int seed) {
    while (seed > 900000000) seed -= 900000000;
    // lagged-Fibonacci table from two seed pairs (random_mars.cpp:56-76)
    int ij = (seed - 1) / 30082, kl = (seed - 1) - 30082 * ij;
    int i = (ij / 177) % 177 + 2, j = ij % 177 + 2;
    int k = (kl / 169) % 178 + 1, l = kl % 169;
    for (int slot = 1; slot <= 97; slot++) {
      double acc = 0.0, bit = 0.5;
      for (int b = 0; b < 24; b++) {
        const int m = ((i * j) % 179) * k % 179;
        i = j;
        j = k;
        k = m;
        l = (53 * l + 1) % 169;
        if ((l * m) % 64 >= 32) acc = acc + bit;
        bit = 0.5 * bit;
      }
      u_[slot] = acc;
    }
    c_ = 362436.0 / 16777216.0;
    cd_ = 7654321.0 / 16777216.0;
    cm_ = 16777213.0 / 16777216.0;
    i97_ = 97;
    j97_ = 33;
    uniform();   // init's discarded first draw
  }

  double uniform() {
    double uni = u_[i97_] - u_[j97_];
    if (uni < 0.0) uni += 1.0;
    u_[i97_] = uni;
    i97_ = i97_ == 1 ? 97 : i97_ - 1;
    j97_ = j97_ == 1 ? 97 : j97_ - 1;
    c_ -= cd_;
    if (c_ < 0.0) c_ += cm_;
    uni -= c_;
    if (uni < 0.0) uni += 1.0;
    return uni;
  }

 private:
  double u_[98];
  double c_, cd_, cm_;
  int i97_, j97_;
};

// RandomPark(double) seed of a generator started from a master draw
// (src/random_park.cpp:45-49)
inline int park_seed_from_draw(double r) {
  int s = static_cast<int>(r * 2147483647);
  return s == 0 ? 1 : s;
}

}  // namespace spk
