// The real root of a reduced monic cubic, shared by the quadratic-epigraph
// projections (proj.hip) and the fused lifting kernels (lift.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace opy {

// Real root of y^3 + a y + b = 0 by Cardano's formula (one real root) or the
// trigonometric form (three real roots; the largest is returned).
template <typename T>
__device__ inline T reduced_monic_cubic(T a, T b) {
  if (a == T(0)) return b == T(0) ? T(0) : cbrt(-b);
  if (b == T(0) && a > T(0)) return T(0);
  const T Q = a / T(3);
  const T R = -b / T(2);
  const T Q3 = Q * Q * Q;
  const T D = Q3 + R * R;
  if (D <= T(0)) {
    const T th = acos(R / sqrt(-Q3));
    return T(2) * sqrt(-Q) * cos(th / T(3));
  }
  T A = cbrt(fabs(R) + sqrt(D));
  if (R < T(0)) A = -A;
  return A - Q / A;
}

// (v, x2) projected onto {x2 >= 0.5*alph*v^2} (one-dimensional x1, no
// truncation): the same steps as the general projection for M = 1.
template <typename T>
__device__ inline void quadepi_project1(T& v, T& x2, T alph) {
  const T n2 = v * v;
  if (n2 > T(0)) {
    if (T(0.5) * alph * n2 > x2) {
      const T xn = sqrt(n2);
      const T l = T(2) / (alph * alph);
      const T yn = reduced_monic_cubic(l * (T(1) - x2 * alph), -l * xn);
      v *= yn / xn;
      x2 = T(0.5) * alph * yn * yn;
    }
  } else if (x2 < T(0)) {
    x2 = T(0);
  }
}

}  // namespace opy
