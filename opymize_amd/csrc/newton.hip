// Kernels of the semismooth Newton continuation (reference
// opymize/solvers/semismooth_newton.py): vector updates of its conjugate
// gradient and Armijo steps, and the Jacobians of the prox maps that make
// up its Newton system.  FP64 dot products are in reduce.hip.
#include "opy_common.h"

namespace opy {
namespace {

// z = a*x + b*y; z may alias x or y (each element is read before it is written)
template <typename T>
__global__ void axpby_kernel(const T* x, const T* y, T* z, int64_t n, T a, T b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  z[i] = a * x[i] + b * y[i];
}

// Jacobian of PosProj / NegProj at xbar applied to x: y (+)= x where xbar > 0
// (POS) or xbar < 0, else 0 (reference proj.py:33-50, :66-85: NihilOp with
// keep = xbar > 0 / xbar < 0)
template <typename T, bool POS>
__global__ void signproj_jac_kernel(const T* __restrict__ xbar, const T* x, T* y, int64_t n, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool keep = POS ? xbar[i] > T(0) : xbar[i] < T(0);
  const T v = keep ? x[i] : T(0);
  y[i] = add ? y[i] + v : v;
}

// Jacobian of the Frobenius-ball projection (radius lbd) at xbar applied to
// x, one point (M entries) per thread (reference proj.py:228-275): inside the
// ball y = x; outside, with n = |xbar| and u = xbar / n,
// y = (lbd / n) (x - <u, x> u).  The exterior test n^2 > lbd^2 is that of the
// projection kernel, so the Jacobian belongs to the map that was applied.
// y may be x (in place): a point's entries are all read before any is written.
template <typename T>
__global__ void __launch_bounds__(kBlock) l12proj_jac_kernel(const T* __restrict__ xbar, const T* x, T* y, int64_t N,
                                                             int M, T lbd, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const T* b = xbar + i * M;
  const T* xi = x + i * M;
  T* yi = y + i * M;
  T n2 = T(0);
  for (int m = 0; m < M; ++m) n2 += b[m] * b[m];
  if (n2 > lbd * lbd) {
    const T nrm = sqrt(n2);
    const T inv = T(1) / nrm;
    T d = T(0);
    for (int m = 0; m < M; ++m) d += (b[m] * inv) * xi[m];
    const T s = lbd * inv;
    for (int m = 0; m < M; ++m) {
      const T v = s * (xi[m] - d * (b[m] * inv));
      yi[m] = add ? yi[m] + v : v;
    }
  } else {
    for (int m = 0; m < M; ++m) yi[m] = add ? yi[m] + xi[m] : xi[m];
  }
}

template <typename T>
int axpby_impl(const T* x, const T* y, T* z, int64_t n, double a, double b, void* s) {
  if (n < 0) return fail(kInvalid, "axpby: negative size");
  if (n == 0) return kOk;
  if (!x || !y || !z) return fail(kInvalid, "axpby: null pointer");
  hipLaunchKernelGGL((axpby_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, z, n, (T)a, (T)b);
  return check_launch("axpby");
}

template <typename T>
int signproj_jac_impl(const T* xbar, const T* x, T* y, int64_t n, int positive, int add, void* s) {
  if (n < 0) return fail(kInvalid, "signproj_jac: negative size");
  if (n == 0) return kOk;
  if (!xbar || !x || !y) return fail(kInvalid, "signproj_jac: null pointer");
  if (positive)
    hipLaunchKernelGGL((signproj_jac_kernel<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), xbar, x, y,
                       n, add);
  else
    hipLaunchKernelGGL((signproj_jac_kernel<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), xbar, x, y,
                       n, add);
  return check_launch("signproj_jac");
}

template <typename T>
int l12proj_jac_impl(const T* xbar, const T* x, T* y, int64_t N, int M, double lbd, int add, void* s) {
  if (N < 0 || M < 1) return fail(kInvalid, "l12proj_jac: bad shape (N=%lld, M=%d)", (long long)N, M);
  if (!(lbd >= 0)) return fail(kInvalid, "l12proj_jac: lbd=%g must be >= 0", lbd);
  if (N == 0) return kOk;
  if (!xbar || !x || !y) return fail(kInvalid, "l12proj_jac: null pointer");
  hipLaunchKernelGGL((l12proj_jac_kernel<T>), dim3(grid_for(N)), dim3(kBlock), 0, as_stream(s), xbar, x, y, N, M,
                     (T)lbd, add);
  return check_launch("l12proj_jac");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_axpby_f32(const float* x, const float* y, float* z, int64_t n, double a, double b, void* s) {
  return axpby_impl<float>(x, y, z, n, a, b, s);
}
OPY_API int opy_axpby_f64(const double* x, const double* y, double* z, int64_t n, double a, double b, void* s) {
  return axpby_impl<double>(x, y, z, n, a, b, s);
}
OPY_API int opy_signproj_jac_f32(const float* xbar, const float* x, float* y, int64_t n, int positive, int add,
                                 void* s) {
  return signproj_jac_impl<float>(xbar, x, y, n, positive, add, s);
}
OPY_API int opy_signproj_jac_f64(const double* xbar, const double* x, double* y, int64_t n, int positive, int add,
                                 void* s) {
  return signproj_jac_impl<double>(xbar, x, y, n, positive, add, s);
}
OPY_API int opy_l12proj_jac_f32(const float* xbar, const float* x, float* y, int64_t N, int M, double lbd, int add,
                                void* s) {
  return l12proj_jac_impl<float>(xbar, x, y, N, M, lbd, add, s);
}
OPY_API int opy_l12proj_jac_f64(const double* xbar, const double* x, double* y, int64_t N, int M, double lbd, int add,
                                void* s) {
  return l12proj_jac_impl<double>(xbar, x, y, N, M, lbd, add, s);
}
