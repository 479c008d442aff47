// Exact replay of the ray march's step additions.
//
// The march advances by rayCurrent += v/2 in float (RayCastRender.cu:384,390).  To
// skip samples known to be invalid, the march must land on exactly the value the
// reference's sequence of float additions reaches.  Within one binade [2^e, 2^(e+1))
// every value is an integer multiple R of ulp = 2^(e-23) with R in [2^23, 2^24), and
// fl(R*ulp + inc) = (R + S)*ulp with S = floor(inc/ulp) + [frac(inc/ulp) > 1/2], as
// long as the result stays in the binade and frac(inc/ulp) != 1/2 (a tie, rounded to
// even, would make S depend on R).  So the k-th value is (R + k*S)*ulp, which gives
// the first value >= a target in integer arithmetic; binade crossings and ties take
// one real addition.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace vsg {

// First element >= target of r, fl(r+inc), fl(fl(r+inc)+inc), ... (r itself when
// r >= target).  *stalled is set when an addition leaves the value unchanged (the
// reference's loop would never end there).
__host__ __device__ inline float replay_steps(float r, float target, float inc, bool* stalled) {
  *stalled = false;
  for (int guard = 0; guard < 4096 && r < target; ++guard) {
    if (r > 0.0f && inc > 0.0f && r >= 1.17549435e-38f) {
      int e;
      (void)frexpf(r, &e);                  // r = m * 2^e, m in [0.5, 1)
      const float ulp = ldexpf(1.0f, e - 24);
      const float qi = inc / ulp;           // exact: scaling by a power of two
      const float I = floorf(qi);
      const float phi = qi - I;
      if (phi != 0.5f && I < 16777216.0f) {
        const uint32_t S = (uint32_t)I + (phi > 0.5f ? 1u : 0u);
        const uint32_t R = (uint32_t)(r / ulp);  // exact integer in [2^23, 2^24)
        if (S > 0u && R + S < 16777216u) {
          const float G = target / ulp;          // >= R; may exceed the binade
          const uint32_t last = (16777215u - R) / S;  // steps that stay in the binade
          uint32_t k;
          if (G <= 16777215.0f) {
            const uint32_t Gi = (uint32_t)ceilf(G);
            k = (Gi - R + S - 1u) / S;
            if (k <= last) return (float)(R + k * S) * ulp;
          }
          k = last;
          r = (float)(R + k * S) * ulp;
          if (r >= target) return r;
          // the next addition leaves the binade: take it for real
        }
      }
    }
    const float nx = r + inc;
    if (nx == r) {
      *stalled = true;
      return r;
    }
    r = nx;
  }
  return r;
}

}  // namespace vsg
