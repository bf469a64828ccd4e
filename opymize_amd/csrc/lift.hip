// Fused, batched PDHG iterations of the multi-label lifting model (the
// sublabel-accurate TV lifting of opymize_amd/lifting.py), many instances of
// uneven image size and label count per launch.
//
// Model per instance (image H x W, k = L - 1 pieces, labels g_i = g0 + i*h):
//   primal  U (k planes) in [0, 1], r (k planes) per pixel in the simplex
//   dual    p (2 planes per piece: row and column differences of U_i)
//           z (2 planes per piece: nu_i, s_i)
//   K x = (grad U_i ; a_i, -r_i),  a_i = h U_i + g_i r_i - h sum_{l>i} r_l
//   F* = balls |p_i| <= lam*h, epigraphs s_i >= nu_i^2/4 + f nu_i
// the generic chain of PDHG.solve(steps="precond") on lifting.sublabel_tv:
// GradientOp over the (H, W, k) grid (whose label axis contributes nothing),
// two MatrixMultR products, ClipOp, SimplexProj, L1NormsProj and QuadEpiProj
// (reference pdhg.py:124-153 with linear/diff.py, linear/einsum.py,
// operators/proj.py), with the same diagonal steps: the reference's
// tau_j = 1/sum_i |K_ij|, sigma_i = 1/sum_j |K_ij| (pdhg.py:189-200), made
// constant on the blocks that one projection couples (the simplex, each
// epigraph pair), formed here from the stencil and the labels instead of
// being stored.
//
// Layout: instance b owns the planes of its k_b pieces, each of N_b = H_b*W_b
// values in image order: U and r at offk_b + i*N_b, p and z at
// 2*offk_b + (2i + c)*N_b (c = 0: row difference / nu, c = 1: column
// difference / s), f at off_b.  A wavefront reads 64 consecutive pixels of
// one plane per load.
//
// One iteration is two passes over the running instances' tiles (64 x 4
// pixels, one per thread):
//   primal: U, r <- clip / simplex of (U, r) - tau K^T (p, z); writes the new
//           x in place and x_bar = 2 x+ - x           reads U, r, p, z
//   dual:   p, z <- projections of (p, z) + sigma K x_bar, in place
//                                                      reads x_bar, p, z, f
// Neighbours (p at the previous row / column in the primal pass, x_bar at
// the next in the dual pass) come from the same planes through the caches.
// The FP64 gap of each instance goes through per-tile partials folded in a
// fixed order.
#include "opy_common.h"
#include "quadepi.h"

#include <algorithm>
#include <cmath>

namespace opy {
namespace {

constexpr int kLW = 64;  // tile width: one wavefront row
constexpr int kLH = 4;   // tile height: four wavefronts
constexpr double kLEps = 2.220446049250313e-16;

struct LiftPix {
  int b;
  int row, col;
  bool in;
};

// the g-th running tile -> (instance, pixel of this thread)
__device__ inline LiftPix lift_locate(int g, const int32_t* __restrict__ active_inst,
                                      const int32_t* __restrict__ act_off, int nact,
                                      const opy_lift_inst_t* __restrict__ insts) {
  int lo = 0, hi = nact - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (act_off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  LiftPix p;
  p.b = active_inst[lo];
  const int j = g - act_off[lo];
  const int W = insts[p.b].W;
  const int ntx = (W + kLW - 1) / kLW;
  const int ty = j / ntx;
  p.row = ty * kLH + (threadIdx.x >> 6);
  p.col = (j - ty * ntx) * kLW + (threadIdx.x & 63);
  p.in = p.row < insts[p.b].H && p.col < W;
  return p;
}

template <typename T>
__device__ inline T clamp01(T v) {
  return v < T(0) ? T(0) : (v > T(1) ? T(1) : v);
}

// Michelot's algorithm on K <= KMAX values held in registers (see simplex.hip)
template <typename T, int KMAX>
__device__ inline void simplex_regs(T (&v)[KMAX], int K) {
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
  for (int k = 0; k < KMAX; ++k) {
    const T d = v[k] - theta;
    v[k] = d > T(0) ? d : T(0);
  }
}

// K^T y at one pixel for piece i: the U part (divergence of p plus h nu) needs
// p at the previous row and column
template <typename T>
__device__ inline T kt_u(const T* __restrict__ p, int64_t o2, int64_t N, int W, int row, int col, int H, T h,
                         T nu) {
  const T p0 = p[o2], p1 = p[o2 + N];
  T v = h * nu;
  if (row < H - 1) v -= p0;
  if (row > 0) v += p[o2 - W];
  if (col < W - 1) v -= p1;
  if (col > 0) v += p[o2 + N - 1];
  return v;
}

// Both passes read a pixel's planes in chunks of kChunk: every load of a chunk
// is issued before its arithmetic (planes past K re-read plane K-1, whose
// results are not stored), so that a wavefront keeps many loads in flight.
constexpr int kChunk = 8;

// ---- primal pass --------------------------------------------------------
template <typename T, int KMAX>
__global__ void __launch_bounds__(256) lift_primal_kernel(T* __restrict__ U, T* __restrict__ r, T* __restrict__ Ub,
                                                          T* __restrict__ rb, const T* __restrict__ p,
                                                          const T* __restrict__ z,
                                                          const opy_lift_inst_t* __restrict__ insts,
                                                          const int32_t* __restrict__ active_inst,
                                                          const int32_t* __restrict__ act_off,
                                                          const int32_t* __restrict__ nact_p) {
  const int nact = *nact_p;
  if (nact == 0) return;
  const int total = act_off[nact];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const LiftPix q = lift_locate(g, active_inst, act_off, nact, insts);
    if (!q.in) continue;
    const opy_lift_inst_t I = insts[q.b];
    const int64_t N = (int64_t)I.H * I.W;
    const int64_t pix = (int64_t)q.row * I.W + q.col;
    const int K = I.K;
    const T h = (T)I.h;
    // column sums of |K|: the gradient's entries of U_i at this pixel, and h in a_i
    const bool dn = q.row < I.H - 1, up = q.row > 0, rt = q.col < I.W - 1, lf = q.col > 0;
    const int cnt = (int)dn + (int)up + (int)rt + (int)lf;
    const T tau_u = (T)I.omega / ((T)cnt + h);
    const T tau_r = (T)(I.omega * I.tau_r);
    // divergence weights: p at this pixel enters with -1 unless on the last row / column,
    // p at the previous row / column with +1 if there is one (else a harmless read of this pixel)
    const T wd = dn ? T(1) : T(0), wr = rt ? T(1) : T(0), wu = up ? T(1) : T(0), wl = lf ? T(1) : T(0);
    const int64_t du = up ? I.W : 0, dl = lf ? 1 : 0;
    T rv[KMAX], r0[KMAX];
    T P = T(0);  // sum of nu_l for l < i
#pragma unroll
    for (int c0 = 0; c0 < KMAX; c0 += kChunk) {
      T nu[kChunk], sv[kChunk], uv[kChunk], p0[kChunk], p1[kChunk], pu[kChunk], pl[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int i = min(c0 + j, K - 1);
        const int64_t o = I.offk + i * N + pix;
        const int64_t o2 = 2 * I.offk + 2 * i * N + pix;
        nu[j] = z[o2];
        sv[j] = z[o2 + N];
        uv[j] = U[o];
        r0[c0 + j] = r[o];
        p0[j] = p[o2];
        p1[j] = p[o2 + N];
        pu[j] = p[o2 - du];
        pl[j] = p[o2 + N - dl];
      }
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int i = c0 + j;
        const T ktu = h * nu[j] - wd * p0[j] + wu * pu[j] - wr * p1[j] + wl * pl[j];
        const T un = clamp01(uv[j] - tau_u * ktu);
        const T gi = (T)(I.g0 + i * I.h);
        rv[i] = r0[i] - tau_r * (gi * nu[j] - sv[j] - h * P);
        if (i < K) {
          const int64_t o = I.offk + i * N + pix;
          U[o] = un;
          Ub[o] = T(2) * un - uv[j];
          P += nu[j];
        }
      }
    }
    simplex_regs<T, KMAX>(rv, K);
#pragma unroll
    for (int i = 0; i < KMAX; ++i)
      if (i < K) {
        const int64_t o = I.offk + i * N + pix;
        r[o] = rv[i];
        rb[o] = T(2) * rv[i] - r0[i];
      }
  }
}

// ---- dual pass ----------------------------------------------------------
template <typename T, int KMAX>
__global__ void __launch_bounds__(256) lift_dual_kernel(const T* __restrict__ Ub, const T* __restrict__ rb,
                                                        T* __restrict__ p, T* __restrict__ z,
                                                        const T* __restrict__ f,
                                                        const opy_lift_inst_t* __restrict__ insts,
                                                        const int32_t* __restrict__ active_inst,
                                                        const int32_t* __restrict__ act_off,
                                                        const int32_t* __restrict__ nact_p) {
  const int nact = *nact_p;
  if (nact == 0) return;
  const int total = act_off[nact];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const LiftPix q = lift_locate(g, active_inst, act_off, nact, insts);
    if (!q.in) continue;
    const opy_lift_inst_t I = insts[q.b];
    const int64_t N = (int64_t)I.H * I.W;
    const int64_t pix = (int64_t)q.row * I.W + q.col;
    const int K = I.K;
    const T h = (T)I.h, rad = (T)(I.lam * I.h);
    const T iw = (T)(1.0 / I.omega);
    const bool dn = q.row < I.H - 1, rt = q.col < I.W - 1;
    // gradient rows: two entries of 1 in absolute value, none on the last row / column (p stays)
    const T sd = dn ? T(0.5) * iw : T(0), sr = rt ? T(0.5) * iw : T(0);
    const int64_t dd = dn ? I.W : 0, dr = rt ? 1 : 0;
    const T fv = f[I.off + pix];
    const T s1 = T(-2) * fv, s2 = -fv * fv;  // epi{nu^2/4 + f nu} = epi{0.25 |nu - s1|^2 + s2}
    T rbv[KMAX];
#pragma unroll
    for (int i = 0; i < KMAX; ++i) rbv[i] = rb[I.offk + min(i, K - 1) * N + pix];
    T S = T(0);  // sum of rb_l for l > i
#pragma unroll
    for (int c1 = KMAX; c1 > 0; c1 -= kChunk) {
      T ub[kChunk], ud[kChunk], ur[kChunk], q0[kChunk], q1[kChunk], nu[kChunk], sv[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int i = min(c1 - 1 - j, K - 1);
        const int64_t o = I.offk + i * N + pix;
        const int64_t o2 = 2 * I.offk + 2 * i * N + pix;
        ub[j] = Ub[o];
        ud[j] = Ub[o + dd];
        ur[j] = Ub[o + dr];
        q0[j] = p[o2];
        q1[j] = p[o2 + N];
        nu[j] = z[o2];
        sv[j] = z[o2 + N];
      }
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int i = c1 - 1 - j;
        if (i >= K) continue;
        const int64_t o2 = 2 * I.offk + 2 * i * N + pix;
        T a0 = q0[j] + sd * (ud[j] - ub[j]);
        T a1 = q1[j] + sr * (ur[j] - ub[j]);
        const T n2 = a0 * a0 + a1 * a1;
        if (n2 > rad * rad) {
          const T sc = rad / sqrt(n2);
          a0 *= sc;
          a1 *= sc;
        }
        p[o2] = a0;
        p[o2 + N] = a1;
        // pair rows: |h| + |g_i| + h (k-1-i) for nu, 1 for s; the pair's step is the smaller
        const T gi = (T)(I.g0 + i * I.h);
        const T a = h * ub[j] + gi * rbv[i] - h * S;
        S += rbv[i];
        const T rowsum = h + fabs(gi) + h * (T)(K - 1 - i);
        const T sz = (rowsum > T(1) ? T(1) / rowsum : T(1)) * iw;
        T v1 = nu[j] + sz * a - s1;
        T v2 = sv[j] - sz * rbv[i] - s2;
        quadepi_project1<T>(v1, v2, T(0.5));
        z[o2] = v1 + s1;
        z[o2 + N] = v2 + s2;
      }
    }
  }
}

// ---- FP64 gap partials per tile -------------------------------------------
// [sum |grad U_i|, sum of the data perspectives, sum max(0, -K^T_U y),
//  sum max_l(-K^T_r y), max infeasibility of G (box, simplex), of F (data
//  perspectives at r_i <= 1e-8), of F* (balls, epigraphs), 0]
template <typename T, int KMAX>
__global__ void __launch_bounds__(256) lift_gap_kernel(const T* __restrict__ U, const T* __restrict__ r,
                                                       const T* __restrict__ p, const T* __restrict__ z,
                                                       const T* __restrict__ f,
                                                       const opy_lift_inst_t* __restrict__ insts,
                                                       const int32_t* __restrict__ active_inst,
                                                       const int32_t* __restrict__ act_off,
                                                       const int32_t* __restrict__ nact_p,
                                                       double* __restrict__ partials) {
  __shared__ double sm[256 / 64];
  const int nact = *nact_p;
  if (nact == 0) return;
  const int total = act_off[nact];
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const LiftPix q = lift_locate(g, active_inst, act_off, nact, insts);
    double tv = 0.0, dat = 0.0, du = 0.0, dr = 0.0, infg = 0.0, inff = 0.0, infd = 0.0;
    const opy_lift_inst_t I = insts[q.b];
    if (q.in) {
      const int64_t N = (int64_t)I.H * I.W;
      const int64_t pix = (int64_t)q.row * I.W + q.col;
      const int K = I.K;
      const double h = I.h, g0 = I.g0, rad = I.lam * I.h;
      const bool dn = q.row < I.H - 1, rt = q.col < I.W - 1;
      const double fv = (double)f[I.off + pix];
      double rv[KMAX];
      double rsum = 0.0;
#pragma unroll
      for (int i = 0; i < KMAX; ++i) {
        rv[i] = i < K ? (double)r[I.offk + i * N + pix] : 0.0;
        rsum += rv[i];
      }
      infg = fabs(rsum - 1.0);
      double S = 0.0, P = 0.0, mx = -INFINITY;
      // suffix sums run from the last piece, prefix sums of nu from the first: two loops
#pragma unroll
      for (int i = KMAX - 1; i >= 0; --i) {
        if (i < K) {
          const int64_t o = I.offk + i * N + pix;
          const double u = (double)U[o];
          const double g_0 = dn ? (double)U[o + I.W] - u : 0.0;
          const double g_1 = rt ? (double)U[o + 1] - u : 0.0;
          tv += sqrt(g_0 * g_0 + g_1 * g_1);
          const double a = h * u + (g0 + i * h) * rv[i] - h * S;
          S += rv[i];
          if (rv[i] > 1e-8) dat += a * a / rv[i] - 2.0 * fv * a + rv[i] * fv * fv;
          else inff = fmax(inff, fabs(rv[i]));
          infg = fmax(infg, fmax(-u, u - 1.0));
          infg = fmax(infg, -rv[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < KMAX; ++i) {
        if (i < K) {
          const int64_t o2 = 2 * I.offk + 2 * i * N + pix;
          const double nu = (double)z[o2], s = (double)z[o2 + N];
          const double ku = (double)kt_u<T>(p, o2, N, I.W, q.row, q.col, I.H, (T)h, (T)nu);
          du += fmax(0.0, -ku);
          const double kr = (g0 + i * h) * nu - s - h * P;
          mx = fmax(mx, -kr);
          P += nu;
          const double p0 = (double)p[o2], p1 = (double)p[o2 + N];
          infd = fmax(infd, sqrt(p0 * p0 + p1 * p1) - rad);
          infd = fmax(infd, 0.25 * nu * nu + fv * nu - s);
        }
      }
      dr = mx;
    }
    tv = block_sum(tv, sm);
    dat = block_sum(dat, sm);
    du = block_sum(du, sm);
    dr = block_sum(dr, sm);
    infg = block_max(infg, sm);
    inff = block_max(inff, sm);
    infd = block_max(infd, sm);
    if (threadIdx.x == 0) {
      // tile index within the instance, found again by thread 0 (its pixel is in the tile)
      int lo = 0, hi = nact - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (act_off[mid] <= g) lo = mid;
        else hi = mid - 1;
      }
      double* o = partials + 8 * (int64_t)(I.tile0 + (g - act_off[lo]));
      o[0] = tv;
      o[1] = dat;
      o[2] = du;
      o[3] = dr;
      o[4] = infg;
      o[5] = inff;
      o[6] = infd;
      o[7] = 0.0;
    }
  }
}

// per instance: fold the partials, record objectives, gap, status
__global__ void __launch_bounds__(256) lift_check_kernel(const opy_lift_inst_t* __restrict__ insts,
                                                         opy_tv_check_t* __restrict__ chk,
                                                         const double* __restrict__ partials, double tol_gap,
                                                         double tol_infeas, const int64_t* __restrict__ iter_counter,
                                                         int iters_in_block) {
  __shared__ double sm[256 / 64];
  const int b = blockIdx.x;
  if (chk[b].status != 0) return;
  const opy_lift_inst_t I = insts[b];
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < I.ntiles; t += blockDim.x) {
    const double* q = partials + 8 * (int64_t)(I.tile0 + t);
    for (int j = 0; j < 4; ++j) v[j] += q[j];
    for (int j = 4; j < 7; ++j) v[j] = fmax(v[j], q[j]);
  }
  for (int j = 0; j < 4; ++j) v[j] = block_sum(v[j], sm);
  for (int j = 4; j < 7; ++j) v[j] = block_max(v[j], sm);
  if (threadIdx.x != 0) return;
  opy_tv_check_t& R = chk[b];
  const double P = I.lam * I.h * v[0] + v[1];
  const double D = -(v[2] + v[3]);
  R.objp = P;
  R.objd = D;
  R.gap = P - D;
  R.relgap = (P - D) / fmax(kLEps, fabs(D));
  // as the generic gap: infeasibility of G plus that of F, of G* (none) plus that of F*
  R.infeasp = fmax(0.0, v[4]) + fmax(0.0, v[5]);
  R.infeasd = fmax(0.0, v[6]);
  R.objd_cert = D;
  R.relgap_cert = R.relgap;
  R.iters = (int32_t)(*iter_counter + iters_in_block);
  if (!isfinite(P) || !isfinite(D)) {
    R.status = 2;
    return;
  }
  if (fabs(R.relgap) < tol_gap && fmax(R.infeasp, R.infeasd) < tol_infeas) R.status = 1;
}

// running instances and the prefix of their tile counts; advances the counter
__global__ void __launch_bounds__(256) lift_compact_kernel(const opy_lift_inst_t* __restrict__ insts,
                                                           const opy_tv_check_t* __restrict__ chk, int B,
                                                           int32_t* __restrict__ active_inst,
                                                           int32_t* __restrict__ act_off,
                                                           int32_t* __restrict__ nact_p,
                                                           int64_t* __restrict__ iter_counter, int iters_in_block) {
  __shared__ int cnt[256];
  __shared__ int tiles[256];
  const int per = (B + 255) / 256;
  const int lo = threadIdx.x * per;
  const int hi = min(B, lo + per);
  int n = 0, nt = 0;
  for (int b = lo; b < hi; ++b)
    if (chk[b].status == 0) {
      ++n;
      nt += insts[b].ntiles;
    }
  cnt[threadIdx.x] = n;
  tiles[threadIdx.x] = nt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, t = 0;
    for (int i = 0; i < 256; ++i) {
      const int ci = cnt[i], ti = tiles[i];
      cnt[i] = a;
      tiles[i] = t;
      a += ci;
      t += ti;
    }
    *nact_p = a;
    act_off[a] = t;
    *iter_counter += iters_in_block;
  }
  __syncthreads();
  int pos = cnt[threadIdx.x], off = tiles[threadIdx.x];
  for (int b = lo; b < hi; ++b)
    if (chk[b].status == 0) {
      active_inst[pos] = b;
      act_off[pos] = off;
      ++pos;
      off += insts[b].ntiles;
    }
}

int lift_check_args(const void* a, const void* b, const opy_lift_inst_t* insts, const int32_t* ai,
                    const int32_t* ao, const int32_t* na, int kmax, int grid, const char* what) {
  if (!a || !b || !insts || !ai || !ao || !na) return fail(kInvalid, "%s: null pointer", what);
  if (kmax < 1 || kmax > 32) return fail(kInvalid, "%s: largest piece count %d outside [1, 32]", what, kmax);
  if (grid < 1) return fail(kInvalid, "%s: grid %d < 1", what, grid);
  return kOk;
}

#define OPY_LIFT_KMAX(KERNEL, KM, ...)                                                      \
  do {                                                                                      \
    if ((KM) <= 8) hipLaunchKernelGGL((KERNEL<T, 8>), dim3(grid), dim3(256), 0, st, __VA_ARGS__);   \
    else if ((KM) <= 16) hipLaunchKernelGGL((KERNEL<T, 16>), dim3(grid), dim3(256), 0, st, __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<T, 32>), dim3(grid), dim3(256), 0, st, __VA_ARGS__);   \
  } while (0)

template <typename T>
int primal_impl(T* U, T* r, T* Ub, T* rb, const T* p, const T* z, const opy_lift_inst_t* insts, const int32_t* ai,
                const int32_t* ao, const int32_t* na, int kmax, int grid, void* s) {
  if (int rc = lift_check_args(U, p, insts, ai, ao, na, kmax, grid, "lift_primal")) return rc;
  if (!r || !Ub || !rb || !z) return fail(kInvalid, "lift_primal: null pointer");
  hipStream_t st = as_stream(s);
  OPY_LIFT_KMAX(lift_primal_kernel, kmax, U, r, Ub, rb, p, z, insts, ai, ao, na);
  return check_launch("lift_primal");
}

template <typename T>
int dual_impl(const T* Ub, const T* rb, T* p, T* z, const T* f, const opy_lift_inst_t* insts, const int32_t* ai,
              const int32_t* ao, const int32_t* na, int kmax, int grid, void* s) {
  if (int rc = lift_check_args(Ub, p, insts, ai, ao, na, kmax, grid, "lift_dual")) return rc;
  if (!rb || !z || !f) return fail(kInvalid, "lift_dual: null pointer");
  hipStream_t st = as_stream(s);
  OPY_LIFT_KMAX(lift_dual_kernel, kmax, Ub, rb, p, z, f, insts, ai, ao, na);
  return check_launch("lift_dual");
}

template <typename T>
int gap_impl(const T* U, const T* r, const T* p, const T* z, const T* f, const opy_lift_inst_t* insts,
             const int32_t* ai, const int32_t* ao, const int32_t* na, int kmax, double* partials, int grid, void* s) {
  if (int rc = lift_check_args(U, p, insts, ai, ao, na, kmax, grid, "lift_gap")) return rc;
  if (!r || !z || !f || !partials) return fail(kInvalid, "lift_gap: null pointer");
  hipStream_t st = as_stream(s);
  OPY_LIFT_KMAX(lift_gap_kernel, kmax, U, r, p, z, f, insts, ai, ao, na, partials);
  return check_launch("lift_gap");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_lift_primal_f32(float* U, float* r, float* Ub, float* rb, const float* p, const float* z,
                                const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                                const int32_t* nact, int kmax, int grid, void* s) {
  return primal_impl<float>(U, r, Ub, rb, p, z, insts, active_inst, act_off, nact, kmax, grid, s);
}
OPY_API int opy_lift_primal_f64(double* U, double* r, double* Ub, double* rb, const double* p, const double* z,
                                const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                                const int32_t* nact, int kmax, int grid, void* s) {
  return primal_impl<double>(U, r, Ub, rb, p, z, insts, active_inst, act_off, nact, kmax, grid, s);
}
OPY_API int opy_lift_dual_f32(const float* Ub, const float* rb, float* p, float* z, const float* f,
                              const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                              const int32_t* nact, int kmax, int grid, void* s) {
  return dual_impl<float>(Ub, rb, p, z, f, insts, active_inst, act_off, nact, kmax, grid, s);
}
OPY_API int opy_lift_dual_f64(const double* Ub, const double* rb, double* p, double* z, const double* f,
                              const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                              const int32_t* nact, int kmax, int grid, void* s) {
  return dual_impl<double>(Ub, rb, p, z, f, insts, active_inst, act_off, nact, kmax, grid, s);
}
OPY_API int opy_lift_gap_f32(const float* U, const float* r, const float* p, const float* z, const float* f,
                             const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                             const int32_t* nact, int kmax, double* partials, int grid, void* s) {
  return gap_impl<float>(U, r, p, z, f, insts, active_inst, act_off, nact, kmax, partials, grid, s);
}
OPY_API int opy_lift_gap_f64(const double* U, const double* r, const double* p, const double* z, const double* f,
                             const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                             const int32_t* nact, int kmax, double* partials, int grid, void* s) {
  return gap_impl<double>(U, r, p, z, f, insts, active_inst, act_off, nact, kmax, partials, grid, s);
}
OPY_API int opy_lift_check(const opy_lift_inst_t* insts, opy_tv_check_t* chk, const double* partials, int B,
                           double tol_gap, double tol_infeas, const int64_t* iter_counter, int iters_in_block,
                           void* s) {
  if (!insts || !chk || !partials || !iter_counter || B < 1) return fail(kInvalid, "lift_check: bad arguments");
  hipLaunchKernelGGL(lift_check_kernel, dim3(B), dim3(256), 0, as_stream(s), insts, chk, partials, tol_gap,
                     tol_infeas, iter_counter, iters_in_block);
  return check_launch("lift_check");
}
OPY_API int opy_lift_compact(const opy_lift_inst_t* insts, const opy_tv_check_t* chk, int B, int32_t* active_inst,
                             int32_t* act_off, int32_t* nact, int64_t* iter_counter, int iters_in_block, void* s) {
  if (!insts || !chk || !active_inst || !act_off || !nact || !iter_counter || B < 1)
    return fail(kInvalid, "lift_compact: bad arguments");
  hipLaunchKernelGGL(lift_compact_kernel, dim3(1), dim3(256), 0, as_stream(s), insts, chk, B, active_inst, act_off,
                     nact, iter_counter, iters_in_block);
  return check_launch("lift_compact");
}
