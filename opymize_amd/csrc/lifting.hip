// Per-point linear operators of functional-lifting models: indexed products
// (reference linear/indexed.py:34-94, indexed.cu:1-48) and the batched and
// "tangled" per-point matrix products (reference linear/einsum.py:130-214,
// einsum.cu:58-123).  All of them are a few multiply-adds per output entry
// over small coefficient arrays (served from cache), so each kernel is one
// streaming pass over x and y: one thread per output entry, the point
// index in the middle of the layout, consecutive threads on consecutive
// entries.
#include "opy_common.h"

namespace opy {
namespace {

// IndexedMult: y[j, i, m] (+)= -sum_l B[j, m, l] x[i, P[j, l]]
// x (N, K), y (J, N, M), B (J, M, L), P (J, L).  Indices of P outside
// [0, K) are skipped (the Python operator refuses them on construction).
template <typename T>
__global__ void __launch_bounds__(kBlock) indexedmult_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                             const T* __restrict__ B, const int32_t* __restrict__ P,
                                                             int64_t N, int J, int M, int L, int K, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)J * N * M) return;
  const int m = (int)(idx % M);
  const int64_t t = idx / M;
  const int64_t i = t % N;
  const int j = (int)(t / N);
  const T* xi = x + i * K;
  const T* Bjm = B + ((int64_t)j * M + m) * L;
  const int32_t* Pj = P + (int64_t)j * L;
  T acc = add ? y[idx] : T(0);
  for (int l = 0; l < L; ++l) {
    const int k = Pj[l];
    if ((unsigned)k < (unsigned)K) acc -= Bjm[l] * xi[k];
  }
  y[idx] = acc;
}

// IndexedMultAdj: y[i, P[j, l]] (+)= -sum_m B[j, m, l] x[j, i, m]
// x (J, N, M), y (N, K).  One thread per point i: it owns y[i, :], so the
// scatter needs no atomics; it clears its row first unless add.
template <typename T>
__global__ void __launch_bounds__(kBlock) indexedmult_adj_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                                 const T* __restrict__ B,
                                                                 const int32_t* __restrict__ P, int64_t N, int J,
                                                                 int M, int L, int K, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T* yi = y + i * K;
  if (!add)
    for (int k = 0; k < K; ++k) yi[k] = T(0);
  for (int j = 0; j < J; ++j) {
    const T* xji = x + ((int64_t)j * N + i) * M;
    for (int l = 0; l < L; ++l) {
      const int k = P[(int64_t)j * L + l];
      if ((unsigned)k >= (unsigned)K) continue;
      T v = yi[k];
      for (int m = 0; m < M; ++m) v -= B[((int64_t)j * M + m) * L + l] * xji[m];
      yi[k] = v;
    }
  }
}

// TangledMatrixMultR: y[m, i, k] (+)= sum_{j,l} x[j, i, l] A(j, l, m, k)
// x (J, N, L), y (M, N, K); A(j, l, m, k) = A[((j L + l) M + m) K + k]
// (A is J x L x M x K) or, transposed, A[((m K + k) J + j) L + l]
// (A is M x K x J x L).
template <typename T>
__global__ void __launch_bounds__(kBlock) tangled_matmul_r_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                                  const T* __restrict__ A, int64_t N, int J, int L,
                                                                  int M, int K, int trans, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * N * K) return;
  const int k = (int)(idx % K);
  const int64_t t = idx / K;
  const int64_t i = t % N;
  const int m = (int)(t / N);
  const int mk = m * K + k;
  T acc = add ? y[idx] : T(0);
  for (int j = 0; j < J; ++j) {
    const T* xji = x + ((int64_t)j * N + i) * L;
    for (int l = 0; l < L; ++l) {
      const T a = trans ? A[((int64_t)mk * J + j) * L + l] : A[((int64_t)j * L + l) * (M * K) + mk];
      acc += xji[l] * a;
    }
  }
  y[idx] = acc;
}

// MatrixMultRBatched: y[j, i, k] (+)= sum_m x[j, i, m] A(j, m, k)
// x (J, N, M), y (J, N, K); A(j, m, k) = A[(j M + m) K + k] (A is J x M x K)
// or, transposed, A[(j K + k) M + m] (A is J x K x M).
template <typename T>
__global__ void __launch_bounds__(kBlock) matmul_r_batched_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                                  const T* __restrict__ A, int64_t N, int J, int M,
                                                                  int K, int trans, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)J * N * K) return;
  const int k = (int)(idx % K);
  const int64_t t = idx / K;
  const int j = (int)(t / N);
  const T* xji = x + t * M;  // t = j N + i
  T acc = add ? y[idx] : T(0);
  for (int m = 0; m < M; ++m) {
    const T a = trans ? A[((int64_t)j * K + k) * M + m] : A[((int64_t)j * M + m) * K + k];
    acc += xji[m] * a;
  }
  y[idx] = acc;
}

template <typename T>
int indexedmult_impl(const T* x, T* y, const T* B, const int32_t* P, int64_t N, int J, int M, int L, int K, int adj,
                     int add, void* s) {
  if (N < 0 || J < 1 || M < 1 || L < 1 || K < 1) return fail(kInvalid, "indexedmult: bad shape");
  if (N > 0 && (!x || !y || !B || !P)) return fail(kInvalid, "indexedmult: null pointer");
  if (N == 0) return kOk;
  if (adj)
    hipLaunchKernelGGL((indexedmult_adj_kernel<T>), dim3(grid_for(N)), dim3(kBlock), 0, as_stream(s), x, y, B, P, N,
                       J, M, L, K, add);
  else
    hipLaunchKernelGGL((indexedmult_kernel<T>), dim3(grid_for((int64_t)J * N * M)), dim3(kBlock), 0, as_stream(s), x,
                       y, B, P, N, J, M, L, K, add);
  return check_launch("indexedmult");
}

template <typename T>
int tangled_impl(const T* x, T* y, const T* A, int64_t N, int J, int L, int M, int K, int trans, int add, void* s) {
  if (N < 0 || J < 1 || L < 1 || M < 1 || K < 1) return fail(kInvalid, "tangled_matmul_r: bad shape");
  if (N > 0 && (!x || !y || !A)) return fail(kInvalid, "tangled_matmul_r: null pointer");
  const int64_t n = (int64_t)M * N * K;
  if (n > 0)
    hipLaunchKernelGGL((tangled_matmul_r_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, A, N, J,
                       L, M, K, trans, add);
  return check_launch("tangled_matmul_r");
}

template <typename T>
int batched_impl(const T* x, T* y, const T* A, int64_t N, int J, int M, int K, int trans, int add, void* s) {
  if (N < 0 || J < 1 || M < 1 || K < 1) return fail(kInvalid, "matmul_r_batched: bad shape");
  if (N > 0 && (!x || !y || !A)) return fail(kInvalid, "matmul_r_batched: null pointer");
  const int64_t n = (int64_t)J * N * K;
  if (n > 0)
    hipLaunchKernelGGL((matmul_r_batched_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, A, N, J,
                       M, K, trans, add);
  return check_launch("matmul_r_batched");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_indexedmult_f32(const float* x, float* y, const float* B, const int32_t* P, int64_t N, int J, int M,
                                int L, int K, int adj, int add, void* s) {
  return indexedmult_impl<float>(x, y, B, P, N, J, M, L, K, adj, add, s);
}
OPY_API int opy_indexedmult_f64(const double* x, double* y, const double* B, const int32_t* P, int64_t N, int J, int M,
                                int L, int K, int adj, int add, void* s) {
  return indexedmult_impl<double>(x, y, B, P, N, J, M, L, K, adj, add, s);
}
OPY_API int opy_tangled_matmul_r_f32(const float* x, float* y, const float* A, int64_t N, int J, int L, int M, int K,
                                     int trans, int add, void* s) {
  return tangled_impl<float>(x, y, A, N, J, L, M, K, trans, add, s);
}
OPY_API int opy_tangled_matmul_r_f64(const double* x, double* y, const double* A, int64_t N, int J, int L, int M, int K,
                                     int trans, int add, void* s) {
  return tangled_impl<double>(x, y, A, N, J, L, M, K, trans, add, s);
}
OPY_API int opy_matmul_r_batched_f32(const float* x, float* y, const float* A, int64_t N, int J, int M, int K, int trans,
                                     int add, void* s) {
  return batched_impl<float>(x, y, A, N, J, M, K, trans, add, s);
}
OPY_API int opy_matmul_r_batched_f64(const double* x, double* y, const double* A, int64_t N, int J, int M, int K,
                                     int trans, int add, void* s) {
  return batched_impl<double>(x, y, A, N, J, M, K, trans, add, s);
}
