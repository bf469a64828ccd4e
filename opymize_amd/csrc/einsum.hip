// Dense per-point matrix products (reference linear/einsum.py:17-128,
// einsum.cu:3-57) and CSR sparse products for operators that only exist as
// a matrix.
#include "opy_common.h"

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
  if (n > 0)
    hipLaunchKernelGGL((matmul_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, A, N, J, K, trans,
                       add);
  return check_launch("matmul");
}

template <typename T>
int matmul_r_impl(const T* x, T* y, const T* A, int64_t N, int J, int K, int trans, int add, void* s) {
  if (N < 0 || J < 1 || K < 1) return fail(kInvalid, "matmul_r: bad shape");
  if (N > 0 && (!x || !y || !A)) return fail(kInvalid, "matmul_r: null pointer");
  const int64_t n = N * J;
  const size_t abytes = (size_t)J * K * sizeof(T);
  const int use_lds = abytes <= 32 * 1024;
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
