// Projections of the multi-label lifting model that the reference does not
// have: the box [lo, hi] (clip) and the probability simplex per point, plus
// the block-constant diagonal preconditioner for steps whose prox is a
// projection onto a non-separable set.
#include "opy_common.h"

namespace opy {
namespace {

template <typename T>
__global__ void clip_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, T lo, T hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = x[i];
  y[i] = v < lo ? lo : (v > hi ? hi : v);
}

// Euclidean projection of each point's K values onto {v >= 0, sum v = 1}:
// Michelot's finite algorithm (Michelot 1986).  Start with every entry
// active; theta = (sum of the active entries - 1) / #active; drop the entries
// with v <= theta; repeat until nothing is dropped (at most K rounds, theta
// only increases).  The result is max(v - theta, 0).  One thread per point,
// the values in registers (KMAX unrolled, entries from K on inert).
template <typename T, int KMAX>
__global__ void __launch_bounds__(kBlock) simplex_kernel(T* __restrict__ x, int64_t N, int K) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T* xi = x + i * K;
  T v[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) v[k] = k < K ? xi[k] : T(0);
  uint32_t act = (K >= 32) ? 0xffffffffu : ((1u << K) - 1u);
  T theta = T(0);
  for (int round = 0; round < KMAX; ++round) {
    T sum = T(0);
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      if (act & (1u << k)) sum += v[k];
    theta = (sum - T(1)) / (T)__popc(act);
    uint32_t keep = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      if ((act & (1u << k)) && v[k] > theta) keep |= 1u << k;
    if (keep == act) break;
    act = keep;
  }
#pragma unroll
  for (int k = 0; k < KMAX; ++k)
    if (k < K) {
      const T d = v[k] - theta;
      xi[k] = d > T(0) ? d : T(0);
    }
}

// Every block of `width` consecutive step sizes takes the smallest positive
// one of the block (0 if it has none).
template <typename T>
__global__ void block_min_kernel(T* __restrict__ x, int64_t nblocks, int width) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  T* xb = x + b * width;
  T m = T(0);
  for (int k = 0; k < width; ++k) {
    const T v = xb[k];
    if (v > T(0) && (m == T(0) || v < m)) m = v;
  }
  for (int k = 0; k < width; ++k) xb[k] = m;
}

template <typename T>
int clip_impl(const T* x, T* y, int64_t n, double lo, double hi, void* s) {
  if (n < 0) return fail(kInvalid, "clip: negative size");
  if (n > 0 && (!x || !y)) return fail(kInvalid, "clip: null pointer");
  if (!(lo <= hi)) return fail(kInvalid, "clip: lo=%g > hi=%g", lo, hi);
  if (n > 0)
    hipLaunchKernelGGL((clip_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, n, (T)lo, (T)hi);
  return check_launch("clip");
}

template <typename T>
int simplex_impl(T* x, int64_t N, int K, void* s) {
  if (N < 0 || K < 1 || K > 32) return fail(kInvalid, "simplex projection: N=%lld, K=%d (1 <= K <= 32)",
                                            (long long)N, K);
  if (N > 0 && !x) return fail(kInvalid, "simplex projection: null pointer");
  if (N == 0) return kOk;
  const dim3 g(grid_for(N)), b(kBlock);
  if (K <= 8) hipLaunchKernelGGL((simplex_kernel<T, 8>), g, b, 0, as_stream(s), x, N, K);
  else if (K <= 16) hipLaunchKernelGGL((simplex_kernel<T, 16>), g, b, 0, as_stream(s), x, N, K);
  else hipLaunchKernelGGL((simplex_kernel<T, 32>), g, b, 0, as_stream(s), x, N, K);
  return check_launch("simplex projection");
}

template <typename T>
int block_min_impl(T* x, int64_t nblocks, int width, void* s) {
  if (nblocks < 0 || width < 1) return fail(kInvalid, "block_min: bad shape");
  if (nblocks > 0 && !x) return fail(kInvalid, "block_min: null pointer");
  if (nblocks > 0)
    hipLaunchKernelGGL((block_min_kernel<T>), dim3(grid_for(nblocks)), dim3(kBlock), 0, as_stream(s), x, nblocks,
                       width);
  return check_launch("block_min");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_clip_f32(const float* x, float* y, int64_t n, double lo, double hi, void* s) {
  return clip_impl<float>(x, y, n, lo, hi, s);
}
OPY_API int opy_clip_f64(const double* x, double* y, int64_t n, double lo, double hi, void* s) {
  return clip_impl<double>(x, y, n, lo, hi, s);
}
OPY_API int opy_simplex_proj_f32(float* x, int64_t N, int K, void* s) { return simplex_impl<float>(x, N, K, s); }
OPY_API int opy_simplex_proj_f64(double* x, int64_t N, int K, void* s) { return simplex_impl<double>(x, N, K, s); }
OPY_API int opy_block_min_f32(float* x, int64_t nblocks, int width, void* s) {
  return block_min_impl<float>(x, nblocks, width, s);
}
OPY_API int opy_block_min_f64(double* x, int64_t nblocks, int width, void* s) {
  return block_min_impl<double>(x, nblocks, width, s);
}
