// Projections used as prox maps of conjugate functionals:
//  - L1NormsProj: per-point projection onto the Frobenius or spectral ball
//    (reference operators/proj.py:170-226, proj.cu:3-107)
//  - QuadEpiProj: projection onto the epigraph of a (shifted, optionally
//    truncated) paraboloid (reference proj.py:278-374, proj.cu:110-219)
#include "opy_common.h"
#include "quadepi.h"

#include <cmath>

namespace opy {
namespace {

// Per point x_i holds an (M1 x M2) matrix stored column by column
// (entry (r, c) at r + c*M1).  With M = (C, D) this is exactly the Jacobian
// of the gradient's (D, C) per-point block, which is how the reference's GPU
// kernel indexes it (proj.cu:11-21).
template <typename T>
__global__ void __launch_bounds__(kBlock) l1proj_frob_kernel(T* __restrict__ x, int64_t N, int M, T lbd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T* xi = x + i * M;
  T n2 = T(0);
  for (int m = 0; m < M; ++m) n2 += xi[m] * xi[m];
  if (n2 > lbd * lbd) {
    const T s = lbd / sqrt(n2);
    for (int m = 0; m < M; ++m) xi[m] *= s;
  }
}

// Specialised Frobenius projection for blocks of 2 (gradient of a scalar
// image): one 8-byte load and store per point.
template <typename T, typename V2>
__global__ void __launch_bounds__(kBlock) l1proj_frob2_kernel(V2* __restrict__ x, int64_t N, T lbd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  V2 v = x[i];
  const T n2 = v.x * v.x + v.y * v.y;
  if (n2 > lbd * lbd) {
    const T s = lbd / sqrt(n2);
    v.x *= s;
    v.y *= s;
    x[i] = v;
  }
}

// Spectral-ball projection via the 2x2 Gram matrix (one side of length 2).
// The Gram matrix and its eigen-decomposition are formed in FP64 for both
// element types: in FP32 the eigenvector formula (lmax - c11, c12) cancels
// badly for nearly equal singular values.
template <typename T>
__global__ void __launch_bounds__(kBlock) l1proj_spec_kernel(T* __restrict__ x, int64_t N, int M1, int M2, double lbd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T* xi = x + i * (M1 * M2);
  // Two "lines" of length lim: rows of a (2 x M2) matrix or columns of an (M1 x 2) matrix.
  const bool two_rows = (M1 <= M2);
  const int lim = two_rows ? M2 : M1;
  const int step_m = two_rows ? M1 : 1;   // stride between consecutive entries of a line
  const int step_j = two_rows ? 1 : M1;   // stride between the two lines
  double c11 = 0, c12 = 0, c22 = 0;
  for (int m = 0; m < lim; ++m) {
    const double a0 = xi[m * step_m];
    const double a1 = xi[m * step_m + step_j];
    c11 += a0 * a0;
    c12 += a0 * a1;
    c22 += a1 * a1;
  }
  const double tr = c11 + c22;
  const double d = sqrt(fmax(0.0, 0.25 * tr * tr - (c11 * c22 - c12 * c12)));
  const double lmax = fmax(0.0, 0.5 * tr + d);
  const double lmin = fmax(0.0, 0.5 * tr - d);
  const double smax = sqrt(lmax), smin = sqrt(lmin);
  if (!(smax > lbd)) return;
  double v11, v12, v21, v22;
  if (c12 == 0.0) {
    if (c11 >= c22) {
      v11 = 1; v12 = 0; v21 = 0; v22 = 1;
    } else {
      v11 = 0; v12 = 1; v21 = 1; v22 = 0;
    }
  } else {
    v11 = c12;
    v12 = c12;
    v21 = lmax - c11;
    v22 = lmin - c11;
    double nv = hypot(v11, v21);
    v11 /= nv;
    v21 /= nv;
    nv = hypot(v12, v22);
    v12 /= nv;
    v22 /= nv;
  }
  const double s1 = fmin(smax, lbd) / smax;
  const double s2 = smin > 0.0 ? fmin(smin, lbd) / smin : 0.0;
  const double m11 = s1 * v11 * v11 + s2 * v12 * v12;
  const double m12 = s1 * v11 * v21 + s2 * v12 * v22;
  const double m21 = s1 * v21 * v11 + s2 * v22 * v12;
  const double m22 = s1 * v21 * v21 + s2 * v22 * v22;
  for (int m = 0; m < lim; ++m) {
    const double a0 = xi[m * step_m];
    const double a1 = xi[m * step_m + step_j];
    xi[m * step_m] = (T)(a0 * m11 + a1 * m21);
    xi[m * step_m + step_j] = (T)(a0 * m12 + a1 * m22);
  }
}

// x_i = (x1 in R^M, x2) projected onto {x2 >= 0.5*alph*|x1 - s1|^2 + s2},
// optionally truncated at |x1 - s1| <= lbd (lbd <= 0 means no truncation).
template <typename T>
__global__ void __launch_bounds__(kBlock) quadepi_kernel(T* __restrict__ x, int64_t N, int M, T alph,
                                                         const T* __restrict__ shift, T lbd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T* xi = x + i * (M + 1);
  const T* si = shift ? shift + i * (M + 1) : nullptr;
  T v[OPY_QUADEPI_MAXM + 1];
  for (int m = 0; m <= M; ++m) v[m] = si ? xi[m] - si[m] : xi[m];
  T n2 = T(0);
  for (int m = 0; m < M; ++m) n2 += v[m] * v[m];
  T x2 = v[M];
  if (n2 > T(0)) {
    T xn = sqrt(n2);
    bool done = false;
    if (lbd > T(0)) {
      const T l = -xn / (lbd * alph) + T(0.5) * alph * lbd * lbd + T(1) / alph;
      if (l <= x2) {
        if (xn > lbd) {
          const T s = lbd / xn;
          for (int m = 0; m < M; ++m) v[m] *= s;
          const T top = T(0.5) * alph * lbd * lbd;
          if (top > x2) x2 = top;
        }
        done = true;
      }
    }
    if (!done && T(0.5) * alph * n2 > x2) {
      const T l = T(2) / (alph * alph);
      const T yn = reduced_monic_cubic(l * (T(1) - x2 * alph), -l * xn);
      const T s = yn / xn;
      for (int m = 0; m < M; ++m) v[m] *= s;
      x2 = T(0.5) * alph * yn * yn;
    }
  } else if (x2 < T(0)) {
    x2 = T(0);
  }
  v[M] = x2;
  for (int m = 0; m <= M; ++m) xi[m] = si ? v[m] + si[m] : v[m];
}

template <typename T>
int l1proj_impl(T* x, int64_t N, int M1, int M2, double lbd, int spectral, void* s) {
  if (N < 0 || M1 < 1 || M2 < 1) return fail(kInvalid, "l1norms_proj: bad shape");
  if (N > 0 && !x) return fail(kInvalid, "l1norms_proj: null pointer");
  if (!(lbd >= 0)) return fail(kInvalid, "l1norms_proj: lbd=%g must be >= 0", lbd);
  if (N == 0) return kOk;
  hipStream_t st = as_stream(s);
  const bool vec = (M1 == 1 || M2 == 1);
  if (spectral && !vec) {
    if (M1 != 2 && M2 != 2) return fail(kInvalid, "l1norms_proj: spectral ball needs M1 or M2 == 2");
    hipLaunchKernelGGL((l1proj_spec_kernel<T>), dim3(grid_for(N)), dim3(kBlock), 0, st, x, N, M1, M2, lbd);
  } else if (M1 * M2 == 2 && (reinterpret_cast<uintptr_t>(x) % (2 * sizeof(T)) == 0)) {
    using V2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;
    hipLaunchKernelGGL((l1proj_frob2_kernel<T, V2>), dim3(grid_for(N)), dim3(kBlock), 0, st,
                       reinterpret_cast<V2*>(x), N, (T)lbd);
  } else {
    hipLaunchKernelGGL((l1proj_frob_kernel<T>), dim3(grid_for(N)), dim3(kBlock), 0, st, x, N, M1 * M2, (T)lbd);
  }
  return check_launch("l1norms_proj");
}

template <typename T>
int quadepi_impl(T* x, int64_t N, int M, double alph, const T* shift, double lbd, void* s) {
  if (N < 0 || M < 1 || M > OPY_QUADEPI_MAXM) return fail(kInvalid, "quadepi_proj: M=%d outside [1,%d]", M, OPY_QUADEPI_MAXM);
  if (N > 0 && !x) return fail(kInvalid, "quadepi_proj: null pointer");
  if (!(alph > 0)) return fail(kInvalid, "quadepi_proj: alph=%g must be positive", alph);
  if (N == 0) return kOk;
  const T l = std::isfinite(lbd) ? (T)lbd : T(0);
  if (std::isfinite(lbd) && !(lbd > 0)) return fail(kInvalid, "quadepi_proj: lbd=%g must be positive", lbd);
  hipLaunchKernelGGL((quadepi_kernel<T>), dim3(grid_for(N)), dim3(kBlock), 0, as_stream(s), x, N, M, (T)alph, shift,
                     l);
  return check_launch("quadepi_proj");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_l1norms_proj_f32(float* x, int64_t N, int M1, int M2, double lbd, int spectral, void* s) {
  return l1proj_impl<float>(x, N, M1, M2, lbd, spectral, s);
}
OPY_API int opy_l1norms_proj_f64(double* x, int64_t N, int M1, int M2, double lbd, int spectral, void* s) {
  return l1proj_impl<double>(x, N, M1, M2, lbd, spectral, s);
}
OPY_API int opy_quadepi_proj_f32(float* x, int64_t N, int M, double alph, const float* shift, double lbd,
                                 void* s) {
  return quadepi_impl<float>(x, N, M, alph, shift, lbd, s);
}
OPY_API int opy_quadepi_proj_f64(double* x, int64_t N, int M, double alph, const double* shift, double lbd,
                                 void* s) {
  return quadepi_impl<double>(x, N, M, alph, shift, lbd, s);
}
