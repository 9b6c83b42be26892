// gs_crand.h -- the C library's rand() stream, restated on the host.
#pragma once

#include <cstdint>

namespace gs {

// The C library's rand() as the reference draws it (StrainWeakening's
// initial damage, PCDVC's Inflow branch; glibc's additive feedback generator): r_0 = seed (0 taken as 1), r_i =
// 16807 r_{i-1} mod (2^31 - 1) for i < 31, r_i = r_{i-31} for 31 <= i < 34,
// then r_i = r_{i-3} + r_{i-31} mod 2^32; the k-th output is r_{k+344} >> 1.
class CRand {
 public:
  CRand() : CRand(1u) {}
  explicit CRand(unsigned seed) {
    int32_t w = (int32_t)(seed == 0 ? 1u : seed);
    r_[0] = (uint32_t)w;
    for (int i = 1; i < 31; i++) {
      const int64_t hi = w / 127773, lo = w % 127773;
      int64_t v = 16807 * lo - 2836 * hi;  // Schrage's method for 16807 w mod (2^31 - 1)
      if (v < 0) v += 2147483647;
      w = (int32_t)v;
      r_[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; i++) r_[i] = r_[i - 31];
    n_ = 34;
    for (int k = 0; k < 310; k++) step();
  }
  int32_t next() { return (int32_t)(step() >> 1); }

 private:
  uint32_t step() {  // r_n = r_{n-3} + r_{n-31}, over a ring of 34
    const uint32_t v = r_[(n_ - 3) % 34] + r_[(n_ - 31) % 34];
    r_[n_ % 34] = v;
    n_++;
    return v;
  }
  uint32_t r_[34];
  uint64_t n_;
};

}  // namespace gs
