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

// ceil(num / den) and floor(num / den) for num, den < 2^24, den > 0: the float
// quotient is within one of the true one, then corrected exactly in integers.
__host__ __device__ inline uint32_t ceil_div24(uint32_t num, uint32_t den) {
  uint32_t k = (uint32_t)((float)num / (float)den);
  if (k * den < num) ++k;
  if (k > 0u && (k - 1u) * den >= num) --k;
  return k;
}

__host__ __device__ inline uint32_t floor_div24(uint32_t num, uint32_t den) {
  uint32_t k = (uint32_t)((float)num / (float)den);
  if (k * den > num) --k;
  if ((k + 1u) * den <= num) ++k;
  return k;
}

// First element >= target of r, fl(r+inc), fl(fl(r+inc)+inc), ... (r itself when
// r >= target); *prev receives the element before it (r when none precedes it).
// *stalled is set when an addition leaves the value unchanged (the reference's loop
// would never end there).
__host__ __device__ inline float replay_steps(float r, float target, float inc, bool* stalled, float* prev) {
  *stalled = false;
  *prev = r;
  for (int guard = 0; guard < 4096 && r < target; ++guard) {
    if (r >= 1.17549435e-38f && inc > 0.0f) {  // r a positive normal number
      int e;
      (void)frexpf(r, &e);                    // r = m * 2^e, m in [0.5, 1): ulp = 2^(e-24)
      const int sc = 24 - e;                  // scaling by 1/ulp is exact (a power of two)
      const float qi = ldexpf(inc, sc);       // inc / ulp (inf if it overflows)
      const float I = floorf(qi);
      const float phi = qi - I;
      if (phi != 0.5f && I < 16777216.0f) {
        const uint32_t S = (uint32_t)I + (phi > 0.5f ? 1u : 0u);
        const uint32_t R = (uint32_t)ldexpf(r, sc);  // exact integer in [2^23, 2^24)
        if (S > 0u && R + S < 16777216u) {
          const float G = ldexpf(target, sc);        // > R; may exceed the binade
          if (G <= 16777215.0f) {
            const uint32_t k = ceil_div24((uint32_t)ceilf(G) - R, S);  // >= 1
            if (R + k * S < 16777216u) {
              *prev = ldexpf((float)(R + (k - 1u) * S), -sc);
              return ldexpf((float)(R + k * S), -sc);
            }
          }
          // the target lies beyond the binade: its last element, then one real addition
          const uint32_t m = floor_div24(16777215u - R, S);
          if (m > 0u) {
            *prev = ldexpf((float)(R + (m - 1u) * S), -sc);
            r = ldexpf((float)(R + m * S), -sc);
          }
          if (r >= target) return r;
        }
      }
    }
    const float nx = r + inc;
    if (nx == r) {
      *stalled = true;
      return r;
    }
    *prev = r;
    r = nx;
  }
  return r;
}

__host__ __device__ inline float replay_steps(float r, float target, float inc, bool* stalled) {
  float prev;
  return replay_steps(r, target, inc, stalled, &prev);
}

// The value after k additions r += inc (the same closed form, by binade).  *stalled is
// set, and the value returned, at the first addition that leaves it unchanged.
__host__ __device__ inline float advance_steps(float r, float inc, uint32_t k, bool* stalled) {
  *stalled = false;
  while (k > 0u) {
    if (r >= 1.17549435e-38f && inc > 0.0f) {  // r a positive normal number
      int e;
      (void)frexpf(r, &e);
      const int sc = 24 - e;
      const float qi = ldexpf(inc, sc);
      const float I = floorf(qi);
      const float phi = qi - I;
      if (phi != 0.5f && I < 16777216.0f) {
        const uint32_t S = (uint32_t)I + (phi > 0.5f ? 1u : 0u);
        const uint32_t R = (uint32_t)ldexpf(r, sc);
        if (S > 0u && R + S < 16777216u) {
          // additions that stay in the binade: R + m S <= 2^24 - 1
          const uint32_t room = floor_div24(16777215u - R, S);
          const uint32_t m = k < room ? k : room;
          r = ldexpf((float)(R + m * S), -sc);
          k -= m;
          if (k == 0u) break;
        }
      }
    }
    const float nx = r + inc;
    if (nx == r) {
      *stalled = true;
      return r;
    }
    r = nx;
    --k;
  }
  return r;
}

}  // namespace vsg
