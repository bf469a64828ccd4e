// Projection onto polyhedral epigraphs (reference operators/proj.py:376-480,
// proj.cu:222-831).
//
// For region j and function i the constraints are
//     <v[l], y[:-1]> - y[-1] <= b[i, l]   for l = J[j, k] with I[i, l],
// and y = argmin 0.5*|y - x[j,i]|^2 is found by the primal active-set method
// (Nocedal & Wright, Numerical Optimization, 2nd ed., Algorithm 16.3) with at
// most ndim+1 active constraints, started from the feasible point obtained by
// raising x[-1].  Constraints are visited in the order of J[j], which fixes
// the tie-breaking between equally blocking constraints.  Step directions
// are projections of x - y onto the orthogonal complement of the active
// normals (a_l, -1) (Gram-Schmidt); Lagrange multipliers come from the Gram
// system of the active normals.  One thread per (j, i): the work is a few
// hundred flops on (ndim+1)-vectors held in registers, with v, b, I and J
// read through the caches.
#include "opy_common.h"

#include <cmath>

namespace opy {
namespace {

template <typename T, int ND>
struct EpiQP {
  const T* v;         // npoints x ND support points
  const int32_t* jr;  // J[j, :]
  const uint8_t* ir;  // I[i, :]
  const T* br;        // b[i, :]
  int nsub;

  __device__ inline bool on(int k) const { return ir[jr[k]] != 0; }
  __device__ inline T bk(int k) const { return br[jr[k]]; }
  __device__ inline T ax(int k, const T* y) const {  // <(a_k, -1), y>
    const T* a = v + (int64_t)jr[k] * ND;
    T s = -y[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) s += a[d] * y[d];
    return s;
  }
  __device__ inline void normal(int k, T* n) const {
    const T* a = v + (int64_t)jr[k] * ND;
#pragma unroll
    for (int d = 0; d < ND; ++d) n[d] = a[d];
    n[ND] = T(-1);
  }
};

template <typename T, int D>
__device__ inline T dot(const T* a, const T* b) {
  T s = T(0);
#pragma unroll
  for (int d = 0; d < D; ++d) s += a[d] * b[d];
  return s;
}

// solve the m x m (m <= D <= 4) system G mu = r in place (Gaussian elimination, partial pivoting)
template <typename T, int D>
__device__ inline void small_solve(T (&G)[D][D], T (&r)[D], int m) {
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int q = c + 1; q < m; ++q)
      if (fabs(G[q][c]) > fabs(G[p][c])) p = q;
    if (p != c) {
      for (int q = 0; q < m; ++q) {
        const T t = G[c][q];
        G[c][q] = G[p][q];
        G[p][q] = t;
      }
      const T t = r[c];
      r[c] = r[p];
      r[p] = t;
    }
    const T piv = G[c][c];
    if (piv == T(0)) continue;
    for (int q = c + 1; q < m; ++q) {
      const T f = G[q][c] / piv;
      for (int s = c; s < m; ++s) G[q][s] -= f * G[c][s];
      r[q] -= f * r[c];
    }
  }
  for (int c = m - 1; c >= 0; --c) {
    T s = r[c];
    for (int q = c + 1; q < m; ++q) s -= G[c][q] * r[q];
    r[c] = G[c][c] != T(0) ? s / G[c][c] : T(0);
  }
}

template <typename T, int ND>
__device__ void epi_project(const EpiQP<T, ND>& qp, const T* x, T* sol, int maxiter, T tol) {
  constexpr int D = ND + 1;
  int act[D];
  int na = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) sol[d] = x[d];
  // feasible start: raise the last coordinate over every violated constraint;
  // the last one raised is the first active constraint
  for (int k = 0; k < qp.nsub; ++k) {
    if (!qp.on(k)) continue;
    const T a = qp.ax(k, sol);
    const T b = qp.bk(k);
    if (a > b) {
      sol[ND] += a - b;
      act[0] = k;
      na = 1;
    }
  }
  if (na == 0) return;  // feasible points are their own projection
  for (int it = 0; it < maxiter; ++it) {
    int blocking = -1;
    T dir[D];
#pragma unroll
    for (int d = 0; d < D; ++d) dir[d] = x[d] - sol[d];
    // minimiser of 0.5|y - x|^2 on the face of the active set, relative to sol:
    // x - sol with its components along the active normals removed
    T basis[D][D];
    int nb = 0;
    for (int s = 0; s < na; ++s) {
      T n[D];
      qp.normal(act[s], n);
      for (int q = 0; q < nb; ++q) {
        const T c = dot<T, D>(n, basis[q]);
#pragma unroll
        for (int d = 0; d < D; ++d) n[d] -= c * basis[q][d];
      }
      const T nn = sqrt(dot<T, D>(n, n));
      if (nn > T(0)) {
#pragma unroll
        for (int d = 0; d < D; ++d) basis[nb][d] = n[d] / nn;
        ++nb;
      }
    }
    for (int q = 0; q < nb; ++q) {
      const T c = dot<T, D>(dir, basis[q]);
#pragma unroll
      for (int d = 0; d < D; ++d) dir[d] -= c * basis[q][d];
    }
    T l1 = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) l1 += fabs(dir[d]);
    if (l1 > T(0)) {
      // the first inactive constraint hit along dir (only those dir moves towards)
      T step_min = T(1);
      for (int k = 0; k < qp.nsub; ++k) {
        if (!qp.on(k)) continue;
        bool in = false;
        for (int s = 0; s < na; ++s) in |= (act[s] == k);
        if (in) continue;
        const T ad = qp.ax(k, dir);
        if (ad > tol) {
          const T st = (qp.bk(k) - qp.ax(k, sol)) / ad;
          if (st < step_min && st > -tol) {
            step_min = st;
            blocking = k;
          }
        }
      }
      if (step_min != T(0)) {
#pragma unroll
        for (int d = 0; d < D; ++d) sol[d] += step_min * dir[d];
      }
    }
    if (blocking != -1) {
      act[na++] = blocking;
    } else if (na == 1) {
      break;  // projection onto the interior of a single facet
    }
    if (na == D || blocking == -1) {
      // multipliers mu with x - sol = sum_s mu_s n_s (normal equations of the active normals)
      T G[D][D], r[D], ns[D][D], g[D];
#pragma unroll
      for (int d = 0; d < D; ++d) g[d] = x[d] - sol[d];
      for (int s = 0; s < na; ++s) qp.normal(act[s], ns[s]);
      for (int s = 0; s < na; ++s) {
        for (int q = 0; q < na; ++q) G[s][q] = dot<T, D>(ns[s], ns[q]);
        r[s] = dot<T, D>(ns[s], g);
      }
      small_solve<T, D>(G, r, na);
      int m = 0;
      for (int s = 1; s < na; ++s)
        if (r[s] < r[m]) m = s;
      if (r[m] >= T(0)) break;  // KKT point of the full problem
      act[m] = act[--na];       // drop the most negative multiplier
    }
  }
}

template <typename T, int ND>
__global__ void __launch_bounds__(kBlock) epigraph_proj_kernel(T* __restrict__ x, const uint8_t* __restrict__ I,
                                                               const int32_t* __restrict__ J, const T* __restrict__ v,
                                                               const T* __restrict__ b, int nfuns, int nregions,
                                                               int npoints, int nsub, int maxiter, T tol) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nfuns * nregions) return;
  const int j = (int)(t / nfuns), i = (int)(t - (int64_t)j * nfuns);
  constexpr int D = ND + 1;
  EpiQP<T, ND> qp{v, J + (int64_t)j * nsub, I + (int64_t)i * npoints, b + (int64_t)i * npoints, nsub};
  T* xi = x + t * D;
  T xin[D], sol[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xin[d] = xi[d];
  epi_project<T, ND>(qp, xin, sol, maxiter, tol);
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = sol[d];
}

template <typename T>
int proj_impl(T* x, const uint8_t* I, const int32_t* J, const T* v, const T* b, int ndim, int nfuns, int nregions,
              int npoints, int nsub, int maxiter, double tol, void* s) {
  if (ndim < 1 || ndim > 3) return fail(kInvalid, "epigraph_proj: ndim=%d outside [1,3]", ndim);
  if (nfuns < 0 || nregions < 0 || npoints < 1 || nsub < 1) return fail(kInvalid, "epigraph_proj: bad shape");
  if (maxiter < 0) return fail(kInvalid, "epigraph_proj: negative iteration limit");
  const int64_t n = (int64_t)nfuns * nregions;
  if (n == 0) return kOk;
  if (!x || !I || !J || !v || !b) return fail(kInvalid, "epigraph_proj: null pointer");
  dim3 g(grid_for(n)), bl(kBlock);
  hipStream_t st = as_stream(s);
  if (ndim == 1)
    hipLaunchKernelGGL((epigraph_proj_kernel<T, 1>), g, bl, 0, st, x, I, J, v, b, nfuns, nregions, npoints, nsub,
                       maxiter, (T)tol);
  else if (ndim == 2)
    hipLaunchKernelGGL((epigraph_proj_kernel<T, 2>), g, bl, 0, st, x, I, J, v, b, nfuns, nregions, npoints, nsub,
                       maxiter, (T)tol);
  else
    hipLaunchKernelGGL((epigraph_proj_kernel<T, 3>), g, bl, 0, st, x, I, J, v, b, nfuns, nregions, npoints, nsub,
                       maxiter, (T)tol);
  return check_launch("epigraph_proj");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_epigraph_proj_f32(float* x, const uint8_t* I, const int32_t* J, const float* v, const float* b,
                                  int ndim, int nfuns, int nregions, int npoints, int nsub, int maxiter, double tol,
                                  void* s) {
  return proj_impl<float>(x, I, J, v, b, ndim, nfuns, nregions, npoints, nsub, maxiter, tol, s);
}
OPY_API int opy_epigraph_proj_f64(double* x, const uint8_t* I, const int32_t* J, const double* v, const double* b,
                                  int ndim, int nfuns, int nregions, int npoints, int nsub, int maxiter, double tol,
                                  void* s) {
  return proj_impl<double>(x, I, J, v, b, ndim, nfuns, nregions, npoints, nsub, maxiter, tol, s);
}
