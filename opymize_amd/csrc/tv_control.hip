// Per-instance control of the batched TV solver: adaptive step rule, FP64
// gap partials, termination test, compaction of the running instances and
// cold restarts; see tv_common.h for the layout and the tile map.
#include "tv_common.h"

namespace opy {
namespace {

// ---- per-instance adaptive step rule (reference pdhg.py:155-183) ------------
__global__ void __launch_bounds__(kTB) tv_adapt_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       opy_adapt_state_t* __restrict__ st,
                                                       const opy_tv_check_t* __restrict__ chk,
                                                       const double* __restrict__ partials, int C,
                                                       opy_adapt_params_t p) {
  __shared__ double sm[kTB / 64];
  const int b = blockIdx.x;
  if (chk[b].status != 0) return;
  const opy_tv_inst_t I = insts[b];
  double s[4] = {0, 0, 0, 0};
  for (int t = threadIdx.x; t < I.ntiles; t += blockDim.x)
    for (int j = 0; j < 4; ++j) s[j] += partials[4 * (int64_t)(I.tile0 + t) + j];
  for (int j = 0; j < 4; ++j) s[j] = block_sum(s[j], sm);
  if (threadIdx.x != 0) return;
  opy_adapt_state_t& S = st[b];
  const double n = (double)I.H * I.W * C;
  const double resp = sqrt(s[0]), gnp = sqrt(s[1]), resd = sqrt(s[2]), gnd = sqrt(s[3]);
  const double epsp = sqrt(n) * p.eps_absp + p.eps_relp * gnp;
  const double epsd = sqrt(2.0 * n) * p.eps_absd + p.eps_reld * gnd;
  const double scale = epsp / epsd;
  S.sums[0] = s[0];
  S.sums[1] = s[1];
  S.sums[2] = s[2];
  S.sums[3] = s[3];
  S.resp = resp;
  S.resd = resd;
  S.epsp = epsp;
  S.epsd = epsd;
  if (resp > scale * resd * p.Delta) {
    const double a = S.alpha;
    S.tau /= 1.0 - a;
    S.sigma *= 1.0 - a;
    S.alpha = a * p.eta;
    S.nchanges += 1;
  }
  if (resp < scale * resd / p.Delta) {
    const double a = S.alpha;
    S.tau *= 1.0 - a;
    S.sigma /= 1.0 - a;
    S.alpha = a * p.eta;
    S.nchanges += 1;
  }
}

// ---- accelerated step update (Chambolle & Pock 2011, Algorithm 2) -----------
// tau <- theta tau, sigma <- sigma / theta with the theta each iteration used,
// for `count` consecutive iterations (the same operations as count updates)
__global__ void __launch_bounds__(kTB) tv_accel_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       opy_adapt_state_t* __restrict__ st,
                                                       const opy_tv_check_t* __restrict__ chk, int B, int count) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || chk[b].status != 0) return;
  double tau = st[b].tau, sigma = st[b].sigma;
  for (int c = 0; c < count; ++c) {
    const double th = accel_theta(OPY_TV_STEPS_ACCEL, insts[b], tau);
    tau *= th;
    sigma /= th;
  }
  st[b].tau = tau;
  st[b].sigma = sigma;
}

// ---- restarts of the accelerated schedule ------------------------------------
// After a check, per instance (one thread each): a running TV-L2 instance
// restarts its step schedule at tau0, sigma0 when g = |relgap_cert| has
// fallen below beta times g at the last restart (sufficient decay) or has
// risen above the smallest g since then (no progress), once min_iters
// iterations have run since the last restart.  The decay test needs a
// recorded g; the first check only records it.
__global__ void __launch_bounds__(kTB) tv_restart_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                         opy_adapt_state_t* __restrict__ st,
                                                         const opy_tv_check_t* __restrict__ chk,
                                                         opy_tv_restart_t* __restrict__ rec, int B, double beta,
                                                         int min_iters, double tau0, double sigma0) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || chk[b].status != 0 || insts[b].data != OPY_TV_DATA_L2) return;
  const double g = fabs(chk[b].relgap_cert);
  if (!isfinite(g)) return;
  opy_tv_restart_t R = rec[b];
  const bool fire = chk[b].iters - R.last >= min_iters && ((isfinite(R.ref) && g < beta * R.ref) || g > R.best);
  R.best = fmin(R.best, g);
  if (!isfinite(R.ref)) R.ref = g;
  if (fire) {
    st[b].tau = tau0;
    st[b].sigma = sigma0;
    R.ref = g;
    R.best = INFINITY;
    R.last = chk[b].iters;
    R.count += 1;
  }
  rec[b] = R;
}

// ---- primal-dual gap (reference pdhg.py:23-44) ------------------------------
// per tile partials: [P_data, P_tv, D_lin = sum z f, D_quad = sum z^2, max |y_p|, max |z|] with z = A^T y.
// A thread's rows go in groups of G (all of them for FP32 images of one
// channel) whose loads are all issued before their arithmetic, and the six
// partials leave the block in one LDS round (used by tile parity, so one
// barrier per tile).
template <typename T, int C, int RPT>
constexpr int gap_group() {
  int g = RPT;
  while (g > 1 && (RPT % g != 0 || g * C * (int)sizeof(T) > 64)) --g;
  return g;
}

template <typename T, int C, int RPT>
__global__ void __launch_bounds__(kTB) tv_gap_kernel(const T* __restrict__ x, const T* __restrict__ y,
                                                     const T* __restrict__ f, const opy_tv_inst_t* __restrict__ insts,
                                                     const int32_t* __restrict__ active_inst,
                                                     const int32_t* __restrict__ act_off,
                                                     const int32_t* __restrict__ nact_p, Cfg cfg,
                                                     double* __restrict__ partials) {
  __shared__ double sm[2 * (kTB / 64)][6];
  int parity = 0;
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int tw = cfg.tw;
  const int th = (kTB / tw) * RPT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tx = threadIdx.x % tw, ty = threadIdx.x / tw;
  const bool lhalo = (tx == 0) || (lane == 0);
  const bool rhalo = (tx == tw - 1) || (lane == 63);
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, tw, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const int c = tr.c0 + tx;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xb = x + I.off;
    const T* fb = f + I.off;
    const T* yb = y + 2 * I.off;
    constexpr int G = gap_group<T, C, RPT>();
    // the tile with a pixel to spare on every side inside the image
    const bool inner = tr.r0 >= 1 && tr.r0 + th < H && tr.c0 >= 1 && tr.c0 + tw < W;
    double v[6] = {0, 0, 0, 0, 0, 0};
    T yup0[C];  // y0 of the row above the group
#pragma unroll
    for (int k = 0; k < C; ++k) yup0[k] = T(0);
    if (col && rs > 0 && rs < H) load_c<T, C>(yb + ((int64_t)(rs - 1) * W + c) * 2 * C, yup0);
#pragma unroll
    for (int g0 = 0; g0 < RPT; g0 += G) {
      // ---- loads: x on the group's rows and the next (for the vertical
      // difference), y and f on the group's rows; the halo lanes' right x and
      // left y1
      T xv[G + 1][C], yv[G][2 * C], fv[G][C], xr[G][C], yl1[G][C];
#pragma unroll
      for (int i = 0; i <= G; ++i)
#pragma unroll
        for (int k = 0; k < C; ++k) xv[i][k] = T(0);
#pragma unroll
      for (int i = 0; i < G; ++i)
#pragma unroll
        for (int k = 0; k < C; ++k) fv[i][k] = yv[i][k] = yv[i][C + k] = xr[i][k] = yl1[i][k] = T(0);
#pragma unroll
      for (int i = 0; i <= G; ++i) {
        const int r = rs + g0 + i;
        if (col && r < H) load_c<T, C>(xb + ((int64_t)r * W + c) * C, xv[i]);
      }
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int r = rs + g0 + i;
        if (col && r < H) {
          const int64_t pix = (int64_t)r * W + c;
          load_c<T, 2 * C>(yb + pix * 2 * C, yv[i]);
          load_c<T, C>(fb + pix * C, fv[i]);
          if (rhalo && c + 1 < W) load_c<T, C>(xb + (pix + 1) * C, xr[i]);
          if (lhalo && c > 0) load_c<T, C>(yb + (pix - 1) * 2 * C + C, yl1[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int r = rs + g0 + i;
        const bool ok = col && r < H;
        T xri[C], yli[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
          const T sr = shfl_down1(xv[i][k]);
          const T sl = shfl_up1(yv[i][C + k]);
          xri[k] = rhalo ? xr[i][k] : sr;
          yli[k] = lhalo ? yl1[i][k] : sl;
        }
        // M: boundary masks of the stencils (none on interior tiles)
        auto pixel = [&](auto m_tag) {
          constexpr bool M = decltype(m_tag)::value;
          T gx[2 * C], z[C];
          grad_px<T, C, M>(xv[i], xv[i + 1], xri, r, c, H, W, cfg, gx);
          adjoint_px<T, C, M>(yv[i], yv[i] + C, i == 0 ? yup0 : yv[i - 1], yli, r, c, H, W, cfg, z);
          double gn2 = 0.0, yn2 = 0.0;
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            gn2 += (double)gx[e] * (double)gx[e];
            yn2 += (double)yv[i][e] * (double)yv[i][e];
          }
          v[1] += sqrt(gn2);
          v[4] = fmax(v[4], yn2);  // |y_p|^2: the square root of the maximum is taken per tile
#pragma unroll
          for (int k = 0; k < C; ++k) {
            const double d = (double)xv[i][k] - (double)fv[i][k];
            v[0] += I.data == OPY_TV_DATA_L1 ? fabs(d) : 0.5 * d * d;
            const double zk = (double)z[k];
            v[2] += zk * (double)fv[i][k];
            v[3] += zk * zk;
            v[5] = fmax(v[5], fabs(zk));
          }
        };
        if (ok) {
          if (inner) pixel(std::false_type{});
          else pixel(std::true_type{});
        }
      }
#pragma unroll
      for (int k = 0; k < C; ++k) yup0[k] = yv[G - 1][k];
    }
    // ---- the six partials of the tile: wavefront shuffles, one LDS round
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] += __shfl_down(v[j], off, 64);
#pragma unroll
      for (int j = 4; j < 6; ++j) v[j] = fmax(v[j], __shfl_down(v[j], off, 64));
    }
    double* buf = sm[parity * (kTB / 64) + wave];
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < 6; ++j) buf[j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
      double* p = partials + 6 * (int64_t)(I.tile0 + tr.j);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double a = sm[parity * (kTB / 64)][j];
        for (int w = 1; w < kTB / 64; ++w) {
          const double b = sm[parity * (kTB / 64) + w][j];
          a = j < 4 ? a + b : fmax(a, b);
        }
        p[j] = j == 4 ? sqrt(a) : a;  // max |y_p| = sqrt(max |y_p|^2) (sqrt is monotone and correctly rounded)
      }
    }
    parity ^= 1;
  }
}

// Folds the gap partials of each running instance, records the objective
// values and marks converged / non-finite instances.
__global__ void __launch_bounds__(kTB) tv_check_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       opy_tv_check_t* __restrict__ chk,
                                                       const double* __restrict__ partials, double tol_gap,
                                                       double tol_infeas, int certify,
                                                       const int64_t* __restrict__ iter_counter, int iters_in_block) {
  __shared__ double sm[kTB / 64];
  const int b = blockIdx.x;
  if (chk[b].status != 0) return;
  const opy_tv_inst_t I = insts[b];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < I.ntiles; t += blockDim.x) {
    const double* p = partials + 6 * (int64_t)(I.tile0 + t);
    for (int j = 0; j < 4; ++j) v[j] += p[j];
    v[4] = fmax(v[4], p[4]);
    v[5] = fmax(v[5], p[5]);
  }
  for (int j = 0; j < 4; ++j) v[j] = block_sum(v[j], sm);
  v[4] = block_max(v[4], sm);
  v[5] = block_max(v[5], sm);
  if (threadIdx.x != 0) return;
  opy_tv_check_t& R = chk[b];
  const double P = I.lam * v[0] + I.tvw * v[1];
  double D, infeas_d, Dc;
  const double ball = fmax(0.0, v[4] - I.tvw);
  if (I.data == OPY_TV_DATA_L1) {
    D = v[2];
    infeas_d = ball + fmax(0.0, v[5] - I.lam);
    const double s = v[5] > I.lam ? I.lam / v[5] : 1.0;  // the rescaled dual point is feasible
    Dc = s * v[2];
  } else {
    D = v[2] - 0.5 * v[3] / I.lam;
    infeas_d = ball;
    Dc = D;
  }
  R.objp = P;
  R.objd = D;
  R.gap = P - D;
  R.relgap = (P - D) / fmax(kEps, fabs(D));
  R.infeasp = 0.0;
  R.infeasd = infeas_d;
  R.objd_cert = Dc;
  R.relgap_cert = (P - Dc) / fmax(kEps, fabs(Dc));
  R.iters = (int32_t)(*iter_counter + iters_in_block);
  if (!isfinite(P) || !isfinite(D)) {
    R.status = 2;
    return;
  }
  const bool done = certify ? (fabs(R.relgap_cert) < tol_gap && ball < tol_infeas)
                            : (fabs(R.relgap) < tol_gap && infeas_d < tol_infeas);
  if (done) R.status = 1;
}

// Rebuilds the list of running instances and the prefix of their tile counts;
// advances the iteration counter by the block just run.
__global__ void __launch_bounds__(kTB) tv_compact_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                         const opy_tv_check_t* __restrict__ chk, int B,
                                                         int32_t* __restrict__ active_inst,
                                                         int32_t* __restrict__ act_off, int32_t* __restrict__ nact_p,
                                                         int64_t* __restrict__ iter_counter, int iters_in_block) {
  __shared__ int cnt[kTB];
  __shared__ int tiles[kTB];
  const int per = (B + kTB - 1) / kTB;
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
  if (threadIdx.x == 0) {  // exclusive scan over 256 entries
    int a = 0, t = 0;
    for (int i = 0; i < kTB; ++i) {
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

// Cold restart of flagged instances: zero their x and y (buffer 0) and mark
// them running again.  One block per instance.
template <typename T>
__global__ void __launch_bounds__(kTB) tv_reset_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       const int32_t* __restrict__ flags, int C, T* __restrict__ x,
                                                       T* __restrict__ y, opy_tv_check_t* __restrict__ chk) {
  const int b = blockIdx.x;
  if (!flags[b]) return;
  const opy_tv_inst_t I = insts[b];
  const int64_t n = (int64_t)I.H * I.W * C;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[I.off + i] = T(0);
  for (int64_t i = threadIdx.x; i < 2 * n; i += blockDim.x) y[2 * I.off + i] = T(0);
  if (threadIdx.x == 0) chk[b].status = 0;
}

template <typename T>
int gap_impl(const T* x, const T* y, const T* f, const opy_tv_inst_t* insts, const int32_t* ai, const int32_t* ao,
             const int32_t* na, const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!x || !y || !f || !insts || !ai || !ao || !na || !partials || grid < 0)
    return fail(kInvalid, "tv_gap: null pointer or negative grid");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_gap_kernel<T, CC, RR>), dim3(launch_grid(KF(tv_gap_kernel<T, CC, RR>), grid)), dim3(kTB), 0, \
                     stm, x, y, f, insts, ai, ao, na, c, partials)
  OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  return check_launch("tv_gap");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_tv_accel(const opy_tv_inst_t* insts, opy_adapt_state_t* st, const opy_tv_check_t* chk, int B,
                         void* s) {
  if (!insts || !st || !chk || B < 1) return fail(kInvalid, "tv_accel: bad arguments");
  hipLaunchKernelGGL(tv_accel_kernel, dim3((B + kTB - 1) / kTB), dim3(kTB), 0, as_stream(s), insts, st, chk, B, 1);
  return check_launch("tv_accel");
}
OPY_API int opy_tv_accel_n(const opy_tv_inst_t* insts, opy_adapt_state_t* st, const opy_tv_check_t* chk, int B,
                           int count, void* s) {
  if (!insts || !st || !chk || B < 1 || count < 1) return fail(kInvalid, "tv_accel_n: bad arguments");
  hipLaunchKernelGGL(tv_accel_kernel, dim3((B + kTB - 1) / kTB), dim3(kTB), 0, as_stream(s), insts, st, chk, B,
                     count);
  return check_launch("tv_accel_n");
}
OPY_API int opy_tv_restart(const opy_tv_inst_t* insts, opy_adapt_state_t* st, const opy_tv_check_t* chk,
                           opy_tv_restart_t* rec, int B, double beta, int min_iters, double tau0, double sigma0,
                           void* s) {
  if (!insts || !st || !chk || !rec || B < 1 || !(beta >= 0.0) || min_iters < 0 || !(tau0 > 0.0) ||
      !(sigma0 > 0.0))
    return fail(kInvalid, "tv_restart: bad arguments");
  hipLaunchKernelGGL(tv_restart_kernel, dim3((B + kTB - 1) / kTB), dim3(kTB), 0, as_stream(s), insts, st, chk, rec, B,
                     beta, min_iters, tau0, sigma0);
  return check_launch("tv_restart");
}
OPY_API int opy_tv_adapt(const opy_tv_inst_t* insts, opy_adapt_state_t* st, const opy_tv_check_t* chk,
                         const double* partials, int B, int C, const opy_adapt_params_t* p, void* s) {
  if (!insts || !st || !chk || !partials || !p || B < 1) return fail(kInvalid, "tv_adapt: bad arguments");
  hipLaunchKernelGGL(tv_adapt_kernel, dim3(B), dim3(kTB), 0, as_stream(s), insts, st, chk, partials, C, *p);
  return check_launch("tv_adapt");
}
OPY_API int opy_tv_gap_f32(const float* x, const float* y, const float* f, const opy_tv_inst_t* insts,
                           const int32_t* active_inst, const int32_t* act_off, const int32_t* nact,
                           const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  return gap_impl<float>(x, y, f, insts, active_inst, act_off, nact, cfg, partials, grid, s);
}
OPY_API int opy_tv_gap_f64(const double* x, const double* y, const double* f, const opy_tv_inst_t* insts,
                           const int32_t* active_inst, const int32_t* act_off, const int32_t* nact,
                           const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  return gap_impl<double>(x, y, f, insts, active_inst, act_off, nact, cfg, partials, grid, s);
}
OPY_API int opy_tv_check(const opy_tv_inst_t* insts, opy_tv_check_t* chk, const double* partials, int B,
                         double tol_gap, double tol_infeas, int certify, const int64_t* iter_counter,
                         int iters_in_block, void* s) {
  if (!insts || !chk || !partials || !iter_counter || B < 1) return fail(kInvalid, "tv_check: bad arguments");
  hipLaunchKernelGGL(tv_check_kernel, dim3(B), dim3(kTB), 0, as_stream(s), insts, chk, partials, tol_gap, tol_infeas,
                     certify, iter_counter, iters_in_block);
  return check_launch("tv_check");
}
template <typename T>
static int reset_impl(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, T* x, T* y, opy_tv_check_t* chk,
                      void* s) {
  if (!insts || !flags || !x || !y || !chk || B < 1 || C < 1 || C > 4) return fail(kInvalid, "tv_reset: bad arguments");
  hipLaunchKernelGGL((tv_reset_kernel<T>), dim3(B), dim3(kTB), 0, as_stream(s), insts, flags, C, x, y, chk);
  return check_launch("tv_reset");
}
OPY_API int opy_tv_reset_f32(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, float* x, float* y,
                             opy_tv_check_t* chk, void* s) {
  return reset_impl<float>(insts, flags, B, C, x, y, chk, s);
}
OPY_API int opy_tv_reset_f64(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, double* x, double* y,
                             opy_tv_check_t* chk, void* s) {
  return reset_impl<double>(insts, flags, B, C, x, y, chk, s);
}

OPY_API int opy_tv_compact(const opy_tv_inst_t* insts, const opy_tv_check_t* chk, int B, int32_t* active_inst,
                           int32_t* act_off, int32_t* nact, int64_t* iter_counter, int iters_in_block, void* s) {
  if (!insts || !chk || !active_inst || !act_off || !nact || !iter_counter || B < 1)
    return fail(kInvalid, "tv_compact: bad arguments");
  hipLaunchKernelGGL(tv_compact_kernel, dim3(1), dim3(kTB), 0, as_stream(s), insts, chk, B, active_inst, act_off, nact,
                     iter_counter, iters_in_block);
  return check_launch("tv_compact");
}
