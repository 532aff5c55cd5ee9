// ORACLE — test infrastructure only.  Nothing in the product links this.
//
// CPU restatement of the reference's random-number semantics:
//   RNG = std::minstd_rand                         (src/BasicTypes.hpp:27)
// as implemented by libstdc++ 11 (GCC headers in this container):
//   seeding      x = s mod m, 0 -> 1                (bits/random.tcc:117-124)
//   step         x = 48271 x mod (2^31 - 1)        (bits/random.h:358-362)
//   canonical    2 draws: sum = (x1-1) + (x2-1)*r, ret = sum / fl(r*r),
//                r = 2147483646, ret >= 1 -> nextafter(1, 0)
//                                                   (bits/random.tcc:3348-3378)
//   uniform(a,b) U*(b-a)+a                          (bits/random.h:1870)
//   bernoulli(p) U < p                              (bits/random.h:3640)
//   exponential  -log(1-U)/lambda                   (bits/random.h:4766)
// Everything here is compiled with -ffp-contract=off so that no product is
// fused into an add.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

constexpr uint64_t kLcgM = 2147483647ULL;
constexpr uint64_t kLcgA = 48271ULL;

struct Lcg {
  uint64_t x;
  explicit Lcg(uint64_t seed = 1) { Seed(seed); }
  void Seed(uint64_t seed) {
    x = seed % kLcgM;
    if (x == 0) x = 1;
  }
  uint64_t operator()() {
    x = (kLcgA * x) % kLcgM;
    return x;
  }
};

inline double Canonical(Lcg& g) {
  const double r = 2147483646.0;
  double sum = 0.0;
  double tmp = 1.0;
  sum += double(g() - 1) * tmp;
  tmp *= r;
  sum += double(g() - 1) * tmp;
  tmp *= r;
  double ret = sum / tmp;
  if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
  return ret;
}

inline double Uniform(Lcg& g, double a, double b) {
  return Canonical(g) * (b - a) + a;
}

inline bool Bernoulli(Lcg& g, double p) { return Canonical(g) < p; }

inline double Exponential(Lcg& g, double lambda) {
  return -std::log(1.0 - Canonical(g)) / lambda;
}

}  // namespace orc
