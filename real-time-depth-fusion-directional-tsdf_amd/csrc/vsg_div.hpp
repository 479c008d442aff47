// Correctly rounded division by a fixed divisor without the division macro.
//
// For a divisor b and rb = RN(1/b), two Newton-Raphson corrections of q0 = RN(a*rb)
// with exact FMA remainders return RN(a/b): the first makes the quotient faithful
// (within one ulp), and for a faithful quotient q and a reciprocal within half an ulp
// RN(q + (a - q*b)*rb) is the correctly rounded quotient (Markstein's theorem; the
// remainder a - q*b of a faithful quotient is exact).  No underflow or overflow may
// occur: the march divides positions of at most a few kilometres by a voxel size of
// a fraction of a metre.  Five FMA-class instructions instead of the ten of the IEEE
// division sequence (and its VCC wait states).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace vsg {

__host__ __device__ inline float div_fixed(float a, float b, float rb) {
  const float q0 = a * rb;
  const float e0 = fmaf(-q0, b, a);
  const float q1 = fmaf(e0, rb, q0);
  const float e1 = fmaf(-q1, b, a);
  return fmaf(e1, rb, q1);
}

}  // namespace vsg
