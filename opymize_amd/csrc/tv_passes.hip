// The two-pass fused TV iteration: pass 1 (primal) and pass 2 (dual); see
// tv_common.h for the model, the layout and the tile map.
#include "tv_common.h"

namespace opy {
namespace {

// ---- pass 1 ---------------------------------------------------------------
template <typename T, int C, int RPT, bool PC>
__global__ void __launch_bounds__(kTB) tv_primal_kernel(const T* __restrict__ xin, T* __restrict__ xout,
                                                        const T* __restrict__ y, const T* __restrict__ f,
                                                        const opy_tv_inst_t* __restrict__ insts,
                                                        const opy_adapt_state_t* __restrict__ st,
                                                        const int32_t* __restrict__ active_inst,
                                                        const int32_t* __restrict__ act_off,
                                                        const int32_t* __restrict__ nact_p, Cfg cfg) {
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int tw = cfg.tw;
  const int th = (kTB / tw) * RPT;
  const int lane = threadIdx.x & 63;
  const int tx = threadIdx.x % tw, ty = threadIdx.x / tw;
  const bool lhalo = (tx == 0) || (lane == 0);
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, tw, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const PrimalStep<T, C, PC> prox(st[tr.b].tau, I, cfg);
    const int c = tr.c0 + tx;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xb = xin + I.off;
    const T* fb = f + I.off;
    const T* yb = y + 2 * I.off;
    T* ob = xout + I.off;
    // every load of the segment is issued before any arithmetic, so that a
    // wavefront keeps 3*RPT+1 (+RPT halo) loads in flight
    T yu[C];            // t=0 part of y in the row above the segment
    T yv[RPT][2 * C];   // y of the own rows
    T xv[RPT][C], fv[RPT][C];
    T yl[RPT][C];       // t=1 part of y of the left neighbour (halo lanes)
#pragma unroll
    for (int k = 0; k < C; ++k) yu[k] = T(0);
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int k = 0; k < C; ++k) yv[i][k] = yv[i][C + k] = xv[i][k] = fv[i][k] = yl[i][k] = T(0);
    if (col && rs > 0 && rs < H) load_c<T, C>(yb + ((int64_t)(rs - 1) * W + c) * 2 * C, yu);
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int64_t pix = (int64_t)(rs + i) * W + c;
      if (col && rs + i < H) {
        load_c<T, 2 * C>(yb + pix * 2 * C, yv[i]);
        load_c<T, C>(xb + pix * C, xv[i]);
        load_c<T, C>(fb + pix * C, fv[i]);
      }
    }
    if (lhalo && col && c > 0) {
#pragma unroll
      for (int i = 0; i < RPT; ++i)
        if (rs + i < H) load_c<T, C>(yb + ((int64_t)(rs + i) * W + c - 1) * 2 * C + C, yl[i]);
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rs + i;
      const bool ok = col && r < H;
      const int64_t pix = (int64_t)r * W + c;
      T yl1[C];
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T s = shfl_up1(yv[i][C + k]);
        yl1[k] = lhalo ? yl[i][k] : s;
      }
      if (ok) {
        T d[C], o[C];
        adjoint_px<T, C>(yv[i], yv[i] + C, i == 0 ? yu : yv[i - 1], yl1, r, c, H, W, cfg, d);
#pragma unroll
        for (int k = 0; k < C; ++k) o[k] = prox(xv[i][k], d[k], fv[i][k], r, c, H, W, k, cfg);
        store_c<T, C>(ob + pix * C, o);
      }
    }
  }
}

// ---- pass 2 ---------------------------------------------------------------
template <typename T, int C, int RPT, bool kAdapt, bool PC>
__global__ void __launch_bounds__(kTB) tv_dual_kernel(const T* __restrict__ xo, const T* __restrict__ xn,
                                                      const T* __restrict__ yin, T* __restrict__ yout,
                                                      const opy_tv_inst_t* __restrict__ insts,
                                                      const opy_adapt_state_t* __restrict__ st,
                                                      const int32_t* __restrict__ active_inst,
                                                      const int32_t* __restrict__ act_off,
                                                      const int32_t* __restrict__ nact_p, Cfg cfg,
                                                      double* __restrict__ partials) {
  __shared__ double sm[2 * (kTB / 64)][4];
  int parity = 0;
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int tw = cfg.tw;
  const int th = (kTB / tw) * RPT;
  const int lane = threadIdx.x & 63;
  const int tx = threadIdx.x % tw, ty = threadIdx.x / tw;
  const bool lhalo = (tx == 0) || (lane == 0);
  const bool rhalo = (tx == tw - 1) || (lane == 63);
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, tw, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const double th_d = accel_theta(cfg.steps, I, st[tr.b].tau);
    const T theta = (T)th_d;                      // 1 unless accelerated
    const T sigma = (T)(st[tr.b].sigma / th_d);
    const DualStep<T, C, PC> ds(sigma, cfg);
    const T tau = (T)st[tr.b].tau;
    const T tvw = (T)I.tvw;
    const int c = tr.c0 + tx;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xob = xo + I.off;
    const T* xnb = xn + I.off;
    const T* yib = yin + 2 * I.off;
    T* yob = yout + 2 * I.off;
    // Every load of the segment is issued before any arithmetic.  Row index
    // j of the arrays is image row rs-1+j (row rs-1 only with adaptive steps,
    // whose residual needs y+ of the row above); the left-neighbour arrays
    // start at row rs.
    constexpr int NR = RPT + 2;
    T ax[NR][C], bx[NR][C];             // x_k, x_k+1 -> later x_k - x_k+1, 2x_k+1 - x_k
    T yk[RPT + 1][2 * C];               // y_k
    T axr[RPT + 1][C], bxr[RPT + 1][C];  // right neighbour (right halo lanes), rows rs-1..
    T axl[RPT + 1][C], bxl[RPT + 1][C];  // left neighbour (left halo lanes, adaptive), rows rs..
    T ykl[RPT][2 * C];                  // y_k of the left neighbour (left halo lanes, adaptive)
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) ax[j][k] = bx[j][k] = T(0);
#pragma unroll
    for (int j = 0; j <= RPT; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        yk[j][k] = yk[j][C + k] = T(0);
        axr[j][k] = bxr[j][k] = axl[j][k] = bxl[j][k] = T(0);
        if (j < RPT) ykl[j][k] = ykl[j][C + k] = T(0);
      }
    constexpr int j0 = kAdapt ? 0 : 1;
#pragma unroll
    for (int j = j0; j < NR; ++j) {
      const int r = rs - 1 + j;
      if (col && r >= 0 && r < H) {
        const int64_t p = (int64_t)r * W + c;
        load_c<T, C>(xob + p * C, ax[j]);
        load_c<T, C>(xnb + p * C, bx[j]);
        if (j <= RPT) load_c<T, 2 * C>(yib + p * 2 * C, yk[j]);
      }
    }
    if (rhalo && col && c + 1 < W) {
#pragma unroll
      for (int j = j0; j <= RPT; ++j) {
        const int r = rs - 1 + j;
        if (r >= 0 && r < H) {
          const int64_t p = (int64_t)r * W + c + 1;
          load_c<T, C>(xob + p * C, axr[j]);
          load_c<T, C>(xnb + p * C, bxr[j]);
        }
      }
    }
    if (kAdapt && lhalo && col && c > 0) {
#pragma unroll
      for (int j = 0; j <= RPT; ++j) {
        const int r = rs + j;
        if (r < H) {
          const int64_t p = (int64_t)r * W + c - 1;
          load_c<T, C>(xob + p * C, axl[j]);
          load_c<T, C>(xnb + p * C, bxl[j]);
          if (j < RPT) load_c<T, 2 * C>(yib + p * 2 * C, ykl[j]);
        }
      }
    }
    // (x_k, x_k+1) -> (xd, xbar) = (x_k - x_k+1, 2 x_k+1 - x_k); right neighbours
    // from the next lane unless loaded by a halo lane
#pragma unroll
    for (int j = j0; j < NR; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T a = ax[j][k], b = bx[j][k];
        ax[j][k] = a - b;
        bx[j][k] = overrelax(T(1) + theta, theta, b, a);
      }
#pragma unroll
    for (int j = j0; j <= RPT; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T a = axr[j][k], b = bxr[j][k];
        const T sd = shfl_down1(ax[j][k]), sb = shfl_down1(bx[j][k]);
        axr[j][k] = rhalo ? a - b : sd;
        bxr[j][k] = rhalo ? overrelax(T(1) + theta, theta, b, a) : sb;
        if (kAdapt) {
          const T al = axl[j][k], bl = bxl[j][k];
          axl[j][k] = al - bl;
          bxl[j][k] = overrelax(T(1) + theta, theta, bl, al);
        }
      }
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const T isig = T(1) / sigma, itau = T(1) / tau;
    T ypu[C];  // t=0 part of y+ in the row above (adaptive)
#pragma unroll
    for (int k = 0; k < C; ++k) ypu[k] = T(0);
    if (kAdapt) {
      T yp[2 * C], g[2 * C];
      yplus_px<T, C>(bx[0], bx[1], bxr[0], yk[0], rs - 1, c, H, W, ds, tvw, cfg, yp, g);
#pragma unroll
      for (int k = 0; k < C; ++k) ypu[k] = yp[k];
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rs + i;
      const bool ok = col && r < H;
      const int64_t pix = (int64_t)r * W + c;
      T yp[2 * C], g[2 * C];
      yplus_px<T, C>(bx[i + 1], bx[i + 2], bxr[i + 1], yk[i + 1], r, c, H, W, ds, tvw, cfg, yp, g);
      if (ok) store_c<T, 2 * C>(yob + pix * 2 * C, yp);
      if (kAdapt) {
        // t=1 parts of y_k and y+ of the left neighbour
        T ykl1[C], ypl1[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
          ykl1[k] = shfl_up1(yk[i + 1][C + k]);
          ypl1[k] = shfl_up1(yp[C + k]);
        }
        if (lhalo && ok && c > 0) {
          T ypl[2 * C], gl[2 * C];
          yplus_px<T, C>(bxl[i], bxl[i + 1], bx[i + 1], ykl[i], r, c - 1, H, W, ds, tvw, cfg, ypl, gl);
#pragma unroll
          for (int k = 0; k < C; ++k) {
            ykl1[k] = ykl[i][C + k];
            ypl1[k] = ypl[C + k];
          }
        }
        if (ok) {
          // primal residual (x_k - x_k+1)/tau - A^T(y_k - y+) and its reference term,
          // dual residual (y_k - y+)/sigma - A(x_k - x_k+1) and (y_k - y+)/sigma + A xbar
          T at_old[C], at_new[C], gxd[2 * C];
          adjoint_px<T, C>(yk[i + 1], yk[i + 1] + C, yk[i], ykl1, r, c, H, W, cfg, at_old);
          adjoint_px<T, C>(yp, yp + C, ypu, ypl1, r, c, H, W, cfg, at_new);
          grad_px<T, C>(ax[i + 1], ax[i + 2], axr[i + 1], r, c, H, W, cfg, gxd);
#pragma unroll
          for (int k = 0; k < C; ++k) {
            const T q = ax[i + 1][k] * itau - at_old[k];
            const T rp = q + at_new[k];
            s[0] += (double)rp * (double)rp;
            s[1] += (double)q * (double)q;
          }
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            const T q = (yk[i + 1][e] - yp[e]) * isig;
            const T rd = q - gxd[e];
            const T gd = q + g[e];
            s[2] += (double)rd * (double)rd;
            s[3] += (double)gd * (double)gd;
          }
        }
#pragma unroll
        for (int k = 0; k < C; ++k) ypu[k] = yp[k];
      }
    }
    if (kAdapt) {
      block_sum4(s, sm, parity);
      parity ^= 1;
      if (threadIdx.x == 0) {
        double* p = partials + 4 * (int64_t)(I.tile0 + tr.j);
        p[0] = s[0];
        p[1] = s[1];
        p[2] = s[2];
        p[3] = s[3];
      }
    }
  }
}

template <typename T>
int primal_impl(const T* xin, T* xout, const T* y, const T* f, const opy_tv_inst_t* insts, const opy_adapt_state_t* st,
                const int32_t* ai, const int32_t* ao, const int32_t* na, const opy_tv_cfg_t* cfg, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!xin || !xout || !y || !f || !insts || !st || !ai || !ao || !na || grid < 0)
    return fail(kInvalid, "tv_primal: null pointer or negative grid");
  if (xin == xout) return fail(kInvalid, "tv_primal: x_in and x_out must differ");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_primal_kernel<T, CC, RR, PC>), dim3(launch_grid(KF(tv_primal_kernel<T, CC, RR, PC>), grid)), \
                     dim3(kTB), 0, stm, xin, xout, y, f, insts, st, ai, ao, na, c)
  if (cfg->steps == OPY_TV_STEPS_PRECOND) {
    constexpr bool PC = true;
    OPY_TV_DISPATCH(OPY_L)
  } else {
    constexpr bool PC = false;
    OPY_TV_DISPATCH(OPY_L)
  }
#undef OPY_L
  return check_launch("tv_primal");
}

template <typename T>
int dual_impl(const T* xo, const T* xn, const T* yi, T* yo, const opy_tv_inst_t* insts, const opy_adapt_state_t* st,
              const int32_t* ai, const int32_t* ao, const int32_t* na, const opy_tv_cfg_t* cfg, int adapt,
              double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!xo || !xn || !yi || !yo || !insts || !st || !ai || !ao || !na || grid < 0)
    return fail(kInvalid, "tv_dual: null pointer or negative grid");
  if (yi == yo || xo == xn) return fail(kInvalid, "tv_dual: ping-pong buffers must differ");
  if (adapt && !partials) return fail(kInvalid, "tv_dual: adaptive steps need a partials buffer");
  if (adapt && cfg->steps != OPY_TV_STEPS_SCALAR)
    return fail(kInvalid, "tv_dual: adaptive steps exclude precond and accelerated steps");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_L(CC, RR)                                                                                       \
  hipLaunchKernelGGL((tv_dual_kernel<T, CC, RR, AD, PC>), dim3(launch_grid(KF(tv_dual_kernel<T, CC, RR, AD, PC>), grid)), \
                     dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, partials)
  if (adapt) {
    constexpr bool AD = true, PC = false;
    OPY_TV_DISPATCH(OPY_L)
  } else if (cfg->steps == OPY_TV_STEPS_PRECOND) {
    constexpr bool AD = false, PC = true;
    OPY_TV_DISPATCH(OPY_L)
  } else {
    constexpr bool AD = false, PC = false;
    OPY_TV_DISPATCH(OPY_L)
  }
#undef OPY_L
  return check_launch("tv_dual");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_tv_primal_f32(const float* xin, float* xout, const float* y, const float* f,
                              const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                              const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int grid,
                              void* s) {
  return primal_impl<float>(xin, xout, y, f, insts, st, active_inst, act_off, nact, cfg, grid, s);
}
OPY_API int opy_tv_primal_f64(const double* xin, double* xout, const double* y, const double* f,
                              const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                              const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int grid,
                              void* s) {
  return primal_impl<double>(xin, xout, y, f, insts, st, active_inst, act_off, nact, cfg, grid, s);
}
OPY_API int opy_tv_dual_f32(const float* xo, const float* xn, const float* yi, float* yo, const opy_tv_inst_t* insts,
                            const opy_adapt_state_t* st, const int32_t* active_inst, const int32_t* act_off,
                            const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt, double* partials, int grid,
                            void* s) {
  return dual_impl<float>(xo, xn, yi, yo, insts, st, active_inst, act_off, nact, cfg, adapt, partials, grid, s);
}
OPY_API int opy_tv_dual_f64(const double* xo, const double* xn, const double* yi, double* yo,
                            const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                            const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                            double* partials, int grid, void* s) {
  return dual_impl<double>(xo, xn, yi, yo, insts, st, active_inst, act_off, nact, cfg, adapt, partials, grid, s);
}
