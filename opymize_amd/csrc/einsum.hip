// Dense per-point matrix products (reference linear/einsum.py:17-128,
// einsum.cu:3-57) and CSR sparse products for operators that only exist as
// a matrix.
#include "opy_common.h"

#include <algorithm>
#include <type_traits>

namespace opy {
namespace {

// MatrixMult, label-major: y[j, i] (+)= sum_k A(j, k) x[k, i]
// A(j, k) = A[j*K + k] (non-transposed, A is J x K) or A[k*J + j] (transposed, A is K x J).
template <typename T>
__global__ void __launch_bounds__(kBlock) matmul_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                        const T* __restrict__ A, int64_t N, int J, int K, int trans,
                                                        int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * J) return;
  const int j = (int)(idx / N);
  const int64_t i = idx - (int64_t)j * N;
  T acc = add ? y[idx] : T(0);
  for (int k = 0; k < K; ++k) {
    const T a = trans ? A[k * J + j] : A[j * K + k];
    acc += a * x[(int64_t)k * N + i];
  }
  y[idx] = acc;
}

// MatrixMultR, point-major: y[i, j] (+)= sum_k x[i, k] A(k, j)
// A(k, j) = A[k*J + j] (non-transposed, A is K x J) or A[j*K + k] (transposed, A is J x K).
// The matrix is staged in LDS; each thread owns one output entry.
template <typename T>
__global__ void __launch_bounds__(kBlock) matmul_r_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                          const T* __restrict__ A, int64_t N, int J, int K,
                                                          int trans, int add, int use_lds) {
  extern __shared__ unsigned char smem_raw[];
  T* As = reinterpret_cast<T*>(smem_raw);
  const T* Ar = A;
  if (use_lds) {
    for (int t = threadIdx.x; t < J * K; t += blockDim.x) As[t] = A[t];
    __syncthreads();
    Ar = As;
  }
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * J) return;
  const int64_t i = idx / J;
  const int j = (int)(idx - i * J);
  const T* xi = x + i * K;
  T acc = add ? y[idx] : T(0);
  for (int k = 0; k < K; ++k) {
    const T a = trans ? Ar[j * K + k] : Ar[k * J + j];
    acc += xi[k] * a;
  }
  y[idx] = acc;
}

// ---- FP32 per-point products on the matrix cores -------------------------
// Three kernels (the launcher picks by layout and size): this direct one, 16
// points at a time (point-major on small N or K > J); `matmul_mfma_lm_kernel`
// (label-major) and `matmul_mfma_pm_kernel` (point-major on large N), which
// stage 64 points per round through LDS.
//
// One wavefront forms the products of 16 points at a time as 16 x 16 tiles of
// v_mfma_f32_16x16x4f32 (FP32 operands, FP32 accumulation), in K/4 steps.
// Lane l of the wavefront holds the first operand's element (row l%16,
// k = l/16), the second operand's element (k = l/16, column l%16) and the
// result's rows 4*(l/16) + r, r = 0..3, of column l%16.
//   point-major (MatrixMultR, PM): rows = points, columns = outputs j; the
//     x rows are the first operand, the matrix the second;
//   label-major (MatrixMult): rows = outputs j, columns = points; the matrix
//     is the first operand, x the second.
// In both, lane l reads x at point p0 + l%16 and k = 4s + l/16, and holds the
// coefficient of x[., k = 4s + l/16] in y[., j = 16t + l%16] in register
// m[s][t] for the whole launch.  Entries beyond K, J or N are zero operands
// and skipped stores.
typedef float opy_f32x4 __attribute__((ext_vector_type(4)));

template <int KS, int JT, bool PM>
__global__ void __launch_bounds__(kBlock) matmul_mfma_kernel(const float* __restrict__ x, float* __restrict__ y,
                                                             const float* __restrict__ A, int64_t N, int J, int K,
                                                             int trans, int add) {
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;
  float m[KS][JT];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int k = 4 * s + lq, j = 16 * t + lr;
      float c = 0.0f;
      if (k < K && j < J) {
        // PM: y[i, j] = sum_k x[i, k] A(k, j), A(k, j) = A[k*J + j] (A is K x J) or A[j*K + k] (trans)
        // label-major: y[j, i] = sum_k A(j, k) x[k, i], A(j, k) = A[j*K + k] (A is J x K) or A[k*J + j] (trans)
        const bool kj = PM ? !trans : (bool)trans;
        c = kj ? A[k * J + j] : A[j * K + k];
      }
      m[s][t] = c;
    }
  }
  const int64_t ntiles = (N + 15) >> 4;
  const int64_t waves_per_block = blockDim.x >> 6;
  const int64_t nwaves = (int64_t)gridDim.x * waves_per_block;
  for (int64_t tile = (int64_t)blockIdx.x * waves_per_block + (threadIdx.x >> 6); tile < ntiles; tile += nwaves) {
    const int64_t p0 = tile << 4;
    const int64_t px = p0 + lr;
    float xv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + lq;
      xv[s] = 0.0f;
      if (px < N && k < K) xv[s] = PM ? x[px * K + k] : x[(int64_t)k * N + px];
    }
    opy_f32x4 acc[JT];
#pragma unroll
    for (int t = 0; t < JT; ++t) acc[t] = opy_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int t = 0; t < JT; ++t) {
        acc[t] = PM ? __builtin_amdgcn_mfma_f32_16x16x4f32(xv[s], m[s][t], acc[t], 0, 0, 0)
                    : __builtin_amdgcn_mfma_f32_16x16x4f32(m[s][t], xv[s], acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < JT; ++t) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t p, idx;
        int j;
        if (PM) {
          p = p0 + 4 * lq + r;
          j = 16 * t + lr;
          idx = p * J + j;
        } else {
          p = px;
          j = 16 * t + 4 * lq + r;
          idx = (int64_t)j * N + p;
        }
        if (p < N && j < J) y[idx] = add ? y[idx] + acc[t][r] : acc[t][r];
      }
    }
  }
}

// ---- label-major products with LDS staging ------------------------------
// MatrixMult (y[j, i] = sum_k A(j, k) x[k, i]): a wavefront takes 64 points at
// a time, loads their K rows of x as runs of 64 consecutive floats (one round
// ahead, while the previous round computes and stores), stages them in LDS,
// forms four 16 x 16 tiles per 16 outputs with v_mfma_f32_16x16x4f32 (the
// matrix as the first operand, held in registers for the launch), stages the
// 64 x J results in LDS over the dead x block and stores J runs of 64 floats.
// Rows k >= K of the staged x are zero, like their matrix entries.
constexpr int kMmPts = 64;  // points per wavefront and round
template <int KS, int JT>
constexpr int mm_lds_floats() {  // per wavefront: the larger of the staged x (KP rows) and y (JP rows), pitch 65
  return (4 * KS > 16 * JT ? 4 * KS : 16 * JT) * (kMmPts + 1);
}

template <int KS, int JT>
__global__ void __launch_bounds__(kBlock) matmul_mfma_lm_kernel(const float* __restrict__ x, float* __restrict__ y,
                                                                const float* __restrict__ A, int64_t N, int J, int K,
                                                                int trans, int add) {
  constexpr int KP = 4 * KS, JP = 16 * JT;
  constexpr int XP = kMmPts + 1;  // row pitch of the staged blocks (+1: spreads a column's reads over the banks)
  __shared__ float sbuf[kBlock / 64][mm_lds_floats<KS, JT>()];
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;
  float* sw = sbuf[threadIdx.x >> 6];
  float m[KS][JT];  // A(j, k) at k = 4s + lq, j = 16t + lr: A[j*K + k] (A is J x K) or A[k*J + j] (trans)
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int k = 4 * s + lq, j = 16 * t + lr;
      m[s][t] = (k < K && j < J) ? (trans ? A[k * J + j] : A[j * K + k]) : 0.0f;
    }
  const int64_t nrounds = (N + kMmPts - 1) / kMmPts;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  auto load_round = [&](int64_t rd, float (&xr)[KP]) {
    const int64_t p0 = rd * kMmPts;
    const bool in = p0 + lane < N;
#pragma unroll
    for (int k = 0; k < KP; ++k) xr[k] = (k < K && in) ? x[(int64_t)k * N + p0 + lane] : 0.0f;
  };
  int64_t rd = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  float xr[KP];
  if (rd < nrounds) load_round(rd, xr);
  for (; rd < nrounds; rd += nwaves) {
    const int64_t p0 = rd * kMmPts;
    // staged x: (k, point pt) at sw[k*XP + pt]
#pragma unroll
    for (int k = 0; k < KP; ++k) sw[k * XP + lane] = xr[k];
    if (rd + nwaves < nrounds) load_round(rd + nwaves, xr);
    opy_f32x4 acc[4][JT];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      float xv[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) xv[s] = sw[(4 * s + lq) * XP + 16 * tt + lr];
#pragma unroll
      for (int t = 0; t < JT; ++t) acc[tt][t] = opy_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int t = 0; t < JT; ++t)
          acc[tt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(m[s][t], xv[s], acc[tt][t], 0, 0, 0);
    }
    // staged y: (output j, point pt) at sw[j*XP + pt]; the staged x is dead
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int t = 0; t < JT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) sw[(16 * t + 4 * lq + r) * XP + 16 * tt + lr] = acc[tt][t][r];
    if (p0 + lane < N) {
      // in groups of 8 rows: the group's LDS reads (and y's, with add) go out before its stores
#pragma unroll 1
      for (int g = 0; g < J; g += 8) {
        float yr[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) yr[u] = g + u < J ? sw[(g + u) * XP + lane] : 0.0f;
        if (add) {
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (g + u < J) yr[u] += y[(int64_t)(g + u) * N + p0 + lane];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (g + u < J) y[(int64_t)(g + u) * N + p0 + lane] = yr[u];
      }
    }
  }
}

// ---- point-major products with LDS staging ------------------------------
// MatrixMultR (y[i, j] = sum_k x[i, k] A(k, j)): the 64 points of a round are
// one run of 64*K floats of x and one of 64*J floats of y.  A wavefront loads
// the run (one round ahead), copies it to LDS in the same order, reads the
// operands at stride K (k >= K read as 0), stages the results at stride J over
// the dead x run and stores the y run; no index arithmetic per element.
template <int KS, int JT>
constexpr int mm_pm_lds_floats() {  // per wavefront: the larger of the two runs
  return kMmPts * (4 * KS > 16 * JT ? 4 * KS : 16 * JT);
}

template <int KS, int JT>
__global__ void __launch_bounds__(kBlock) matmul_mfma_pm_kernel(const float* __restrict__ x, float* __restrict__ y,
                                                                const float* __restrict__ A, int64_t N, int J, int K,
                                                                int trans, int add) {
  constexpr int KP = 4 * KS, JP = 16 * JT;
  __shared__ float sbuf[kBlock / 64][mm_pm_lds_floats<KS, JT>()];
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;
  float* sw = sbuf[threadIdx.x >> 6];
  float m[KS][JT];  // A(k, j) at k = 4s + lq, j = 16t + lr: A[k*J + j] (A is K x J) or A[j*K + k] (trans)
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int k = 4 * s + lq, j = 16 * t + lr;
      m[s][t] = (k < K && j < J) ? (trans ? A[j * K + k] : A[k * J + j]) : 0.0f;
    }
  const int64_t nrounds = (N + kMmPts - 1) / kMmPts;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  // the round's x run in registers: element it of lane = x[p0*K + lane + 64*it], it < K
  auto load_round = [&](int64_t rd, float (&xr)[KP]) {
    const int64_t p0 = rd * kMmPts;
    const int64_t nv = min<int64_t>(kMmPts, N - p0) * K;
#pragma unroll
    for (int it = 0; it < KP; ++it) {
      const int e = lane + kMmPts * it;
      xr[it] = (it < K && e < nv) ? x[p0 * K + e] : 0.0f;
    }
  };
  int64_t rd = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  float xr[KP];
  if (rd < nrounds) load_round(rd, xr);
  for (; rd < nrounds; rd += nwaves) {
    const int64_t p0 = rd * kMmPts;
    const int np = (int)min<int64_t>(kMmPts, N - p0);
#pragma unroll
    for (int it = 0; it < KP; ++it)
      if (it < K) sw[lane + kMmPts * it] = xr[it];
    if (rd + nwaves < nrounds) load_round(rd + nwaves, xr);
    opy_f32x4 acc[4][JT];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      float xv[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + lq;
        xv[s] = k < K ? sw[(16 * tt + lr) * K + k] : 0.0f;
      }
#pragma unroll
      for (int t = 0; t < JT; ++t) acc[tt][t] = opy_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int t = 0; t < JT; ++t)
          acc[tt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv[s], m[s][t], acc[tt][t], 0, 0, 0);
    }
    // the y run of the round: (point pt, output j) at sw[pt*J + j]
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int t = 0; t < JT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = 16 * t + lr;
          if (j < J) sw[(16 * tt + 4 * lq + r) * J + j] = acc[tt][t][r];
        }
    float* yb = y + p0 * J;
    const int nv = np * J;
#pragma unroll 1
    for (int g = 0; g < J; g += 8) {
      float yr[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = lane + kMmPts * (g + u);
        yr[u] = (g + u < J && e < nv) ? sw[e] : 0.0f;
      }
      if (add) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = lane + kMmPts * (g + u);
          if (g + u < J && e < nv) yr[u] += yb[e];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = lane + kMmPts * (g + u);
        if (g + u < J && e < nv) yb[e] = yr[u];
      }
    }
  }
}

// the matrix-core path covers K, J <= 64 in FP32; returns false if the shape is outside it
template <bool PM>
bool launch_matmul_mfma(const float* x, float* y, const float* A, int64_t N, int J, int K, int trans, int add,
                        hipStream_t st) {
  if (K > 64 || J > 64) return false;
  // label-major: the staged kernel (64 points per wavefront and round), a wavefront per round.
  // Point-major: from 2^20 points the staged kernel on at most 2048 blocks, so that each
  // wavefront takes several rounds, with the next round's loads in flight and the matrix loaded
  // once (at 4M points 0.58-0.64 of HBM peak, against 0.48-0.54 for the direct kernel and for the
  // staged one with a wavefront per round); below that the direct kernel (16 points at a time,
  // whose finer grain keeps more wavefronts busy on small N; 0.44-0.47 at 512^2 for 15 x 30 and
  // 30 x 15, against 0.35-0.36 staged)
  const bool pm_staged = PM && N >= (int64_t(1) << 20);
  const int64_t units = (PM && !pm_staged) ? (N + 15) / 16 : (N + kMmPts - 1) / kMmPts;
  const int64_t blocks = std::min<int64_t>((units + 3) / 4, pm_staged ? 2048 : (1 << 16));
  const dim3 g((unsigned)std::max<int64_t>(blocks, 1)), b(kBlock);
  const int ks = K <= 4 ? 1 : K <= 8 ? 2 : K <= 16 ? 4 : K <= 32 ? 8 : 16;
  const int jt = J <= 16 ? 1 : J <= 32 ? 2 : 4;
#define OPY_MM(KS_, JT_)                                                                                      \
  do {                                                                                                      \
    if constexpr (PM) {                                                                                     \
      if (pm_staged)                                                                                        \
        hipLaunchKernelGGL((matmul_mfma_pm_kernel<KS_, JT_>), g, b, 0, st, x, y, A, N, J, K, trans, add);  \
      else                                                                                                  \
        hipLaunchKernelGGL((matmul_mfma_kernel<KS_, JT_, true>), g, b, 0, st, x, y, A, N, J, K, trans, add); \
    }                                                                                                       \
    else                                                                                                    \
      hipLaunchKernelGGL((matmul_mfma_lm_kernel<KS_, JT_>), g, b, 0, st, x, y, A, N, J, K, trans, add);    \
  } while (0)
#define OPY_MM_J(KS_)                 \
  switch (jt) {                       \
    case 1: OPY_MM(KS_, 1); break;    \
    case 2: OPY_MM(KS_, 2); break;    \
    default: OPY_MM(KS_, 4); break;   \
  }
  switch (ks) {
    case 1: OPY_MM_J(1); break;
    case 2: OPY_MM_J(2); break;
    case 4: OPY_MM_J(4); break;
    case 8: OPY_MM_J(8); break;
    default: OPY_MM_J(16); break;
  }
#undef OPY_MM_J
#undef OPY_MM
  return true;
}

// y (+)= A x for a CSR matrix; one thread per row.
template <typename T>
__global__ void __launch_bounds__(kBlock) csr_spmv_kernel(int64_t nrows, const int64_t* __restrict__ indptr,
                                                          const int64_t* __restrict__ indices,
                                                          const T* __restrict__ data, const T* __restrict__ x,
                                                          T* __restrict__ y, int add) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  T acc = T(0);
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) acc += data[e] * x[indices[e]];
  y[r] = add ? y[r] + acc : acc;
}

// row-wise l^p norms of a CSR matrix
template <typename T>
__global__ void __launch_bounds__(kBlock) csr_rowwise_kernel(int64_t nrows, const int64_t* __restrict__ indptr,
                                                             const T* __restrict__ data, double p,
                                                             T* __restrict__ y, int add) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double acc = 0.0;
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) acc += pow(fabs((double)data[e]), p);
  y[r] = add ? (T)((double)y[r] + acc) : (T)acc;
}

template <typename T>
int matmul_impl(const T* x, T* y, const T* A, int64_t N, int J, int K, int trans, int add, void* s) {
  if (N < 0 || J < 1 || K < 1) return fail(kInvalid, "matmul: bad shape");
  if (N > 0 && (!x || !y || !A)) return fail(kInvalid, "matmul: null pointer");
  const int64_t n = N * J;
  if (n > 0) {
    if constexpr (std::is_same<T, float>::value) {
      if (launch_matmul_mfma<false>(x, y, A, N, J, K, trans, add, as_stream(s))) return check_launch("matmul");
    }
    hipLaunchKernelGGL((matmul_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, A, N, J, K, trans,
                       add);
  }
  return check_launch("matmul");
}

template <typename T>
int matmul_r_impl(const T* x, T* y, const T* A, int64_t N, int J, int K, int trans, int add, void* s) {
  if (N < 0 || J < 1 || K < 1) return fail(kInvalid, "matmul_r: bad shape");
  if (N > 0 && (!x || !y || !A)) return fail(kInvalid, "matmul_r: null pointer");
  const int64_t n = N * J;
  const size_t abytes = (size_t)J * K * sizeof(T);
  const int use_lds = abytes <= 32 * 1024;
  if constexpr (std::is_same<T, float>::value) {
    if (n > 0 && launch_matmul_mfma<true>(x, y, A, N, J, K, trans, add, as_stream(s))) return check_launch("matmul_r");
  }
  if (n > 0)
    hipLaunchKernelGGL((matmul_r_kernel<T>), dim3(grid_for(n)), dim3(kBlock), use_lds ? abytes : 0, as_stream(s), x,
                       y, A, N, J, K, trans, add, use_lds);
  return check_launch("matmul_r");
}

template <typename T>
int csr_impl(int64_t nrows, const int64_t* indptr, const int64_t* indices, const T* data, const T* x, T* y, int add,
             void* s) {
  if (nrows < 0) return fail(kInvalid, "csr_spmv: bad shape");
  if (nrows > 0 && (!indptr || !y)) return fail(kInvalid, "csr_spmv: null pointer");
  if (nrows > 0)
    hipLaunchKernelGGL((csr_spmv_kernel<T>), dim3(grid_for(nrows)), dim3(kBlock), 0, as_stream(s), nrows, indptr,
                       indices, data, x, y, add);
  return check_launch("csr_spmv");
}

template <typename T>
int csr_rowwise_impl(int64_t nrows, const int64_t* indptr, const T* data, double p, T* y, int add, void* s) {
  if (nrows < 0) return fail(kInvalid, "csr_rowwise: bad shape");
  if (nrows > 0 && (!indptr || !y)) return fail(kInvalid, "csr_rowwise: null pointer");
  if (nrows > 0)
    hipLaunchKernelGGL((csr_rowwise_kernel<T>), dim3(grid_for(nrows)), dim3(kBlock), 0, as_stream(s), nrows, indptr,
                       data, p, y, add);
  return check_launch("csr_rowwise");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_matmul_f32(const float* x, float* y, const float* A, int64_t N, int J, int K, int trans, int add,
                           void* s) {
  return matmul_impl<float>(x, y, A, N, J, K, trans, add, s);
}
OPY_API int opy_matmul_f64(const double* x, double* y, const double* A, int64_t N, int J, int K, int trans, int add,
                           void* s) {
  return matmul_impl<double>(x, y, A, N, J, K, trans, add, s);
}
OPY_API int opy_matmul_r_f32(const float* x, float* y, const float* A, int64_t N, int J, int K, int trans, int add,
                             void* s) {
  return matmul_r_impl<float>(x, y, A, N, J, K, trans, add, s);
}
OPY_API int opy_matmul_r_f64(const double* x, double* y, const double* A, int64_t N, int J, int K, int trans,
                             int add, void* s) {
  return matmul_r_impl<double>(x, y, A, N, J, K, trans, add, s);
}
OPY_API int opy_csr_spmv_f32(int64_t nrows, const int64_t* indptr, const int64_t* indices, const float* data,
                             const float* x, float* y, int add, void* s) {
  return csr_impl<float>(nrows, indptr, indices, data, x, y, add, s);
}
OPY_API int opy_csr_spmv_f64(int64_t nrows, const int64_t* indptr, const int64_t* indices, const double* data,
                             const double* x, double* y, int add, void* s) {
  return csr_impl<double>(nrows, indptr, indices, data, x, y, add, s);
}
OPY_API int opy_csr_rowwise_f32(int64_t nrows, const int64_t* indptr, const float* data, double p, float* y, int add,
                                void* s) {
  return csr_rowwise_impl<float>(nrows, indptr, data, p, y, add, s);
}
OPY_API int opy_csr_rowwise_f64(int64_t nrows, const int64_t* indptr, const double* data, double p, double* y,
                                int add, void* s) {
  return csr_rowwise_impl<double>(nrows, indptr, data, p, y, add, s);
}
