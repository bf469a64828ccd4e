// K fused TV iterations per pass over HBM (temporal blocking); see
// tv_common.h for the model, the layout and the tile map.
#include "tv_common.h"

namespace opy {
namespace {

// LDS bytes of one block: y (2C components) and xbar (C) over the region of
// a tile (the 64 x 4*RPT tile and K pixels on every side, rows rounded up to
// a multiple of the 4 waves) with a guard pixel on every side, and the K
// iterations' scalar steps.
constexpr int multi_lds(int C, int rpt, int k, int size) {
  return (4 * ((4 * rpt + 2 * k + 3) / 4) + 2) * (64 + 2 * k + 2) * 3 * C * size + k * (3 + 4 * C) * size;
}
template <typename T, int C, int RPT, int K>
constexpr int multi_lds_bytes() {
  static_assert(sizeof(FoldedStep<T, C>) == (3 + 4 * C) * sizeof(T), "step table layout");
  return multi_lds(C, RPT, K, (int)sizeof(T));
}
constexpr int kMultiLdsMax = 64 * 1024;

// K iterations x -> x+, y -> y+ of the single-pass kernel (same per-pixel
// arithmetic: primal_px, xbar, dual_px with scalar steps; adjoint_px,
// PrimalStep, yplus_px with preconditioned ones) on one tile per block,
// with one read of x, f, y and one write of x+, y+ per K iterations
// (28 B per pixel at C = 1 FP32 for K iterations instead of 28 K B).
//
// The block loads its tile plus K pixels on each side (the region) and
// iterates there.  Iterate i is exact on a region that shrinks by one pixel
// per iteration: x^i needs y^(i-1) at the pixel, above and to the left, and
// y^i needs xbar^i at the pixel, below and to the right, so x^i is exact on
// rows/columns [i, R - i + 1) of the region and y^i on [i, R - i); after K
// iterations both are exact on the tile.  Every iteration updates the whole
// region anyway (no per-pixel tests): the values outside the exact part are
// never read by an exact one, and the LDS planes have a guard pixel on every
// side (and rows up to a multiple of the 4 waves), so the reads of the
// region's edge pixels stay in bounds (their results are discarded like
// every inexact value).  The image boundary is the stencils' business, as in
// one iteration: their masks are selects, so no value from beyond the image
// (loaded as 0) or from a guard pixel enters an exact one.  A tile whose
// region lies inside the image, with a pixel to spare on every side (most
// tiles), runs the stencils without the masks and loads without bounds
// tests; the data term's kind is a compile-time constant of the loop body.
//
// Threads: the 64 tile columns of every region row belong to one lane each
// (rows split over the 4 waves; aligned vector loads and stores), and the
// 2K halo columns x region rows are spread one pixel per thread (two with
// large tiles), row by row, so that a wavefront's halo loads touch a few
// short row segments instead of 64 rows of one column.  Each thread keeps
// x, f and y of its pixels in registers; LDS holds y and xbar, the values
// its neighbours read.  Two barriers per iteration: after xbar is written
// (phase A) and after y is written (phase B).
//
// Steps: tau, sigma (and theta with accelerated steps) of iteration i are
// derived from the instance's state exactly as the single-step launches and
// opy_tv_accel would advance it, and folded into the stencil weights
// (fold_step), by one thread per tile into a small LDS table (this FP64 work
// was a fifth of the instruction issue when every wave did it); the state
// itself is not written.
template <typename T, int C, int RPT, int K, bool PC>
__global__ void __launch_bounds__(kTB, 3) tv_multi_kernel(const T* __restrict__ xin, T* __restrict__ xout,
                                                       const T* __restrict__ yin, T* __restrict__ yout,
                                                       const T* __restrict__ f, const opy_tv_inst_t* __restrict__ insts,
                                                       const opy_adapt_state_t* __restrict__ st,
                                                       const int32_t* __restrict__ active_inst,
                                                       const int32_t* __restrict__ act_off,
                                                       const int32_t* __restrict__ nact_p, Cfg cfg) {
  constexpr int TH = 4 * RPT;        // tile rows (4 waves x RPT, the tile map of the single-pass kernel)
  constexpr int RH = TH + 2 * K;     // region rows
  constexpr int RW = 64 + 2 * K;     // region columns
  constexpr int NR = (RH + 3) / 4;   // region rows per wave (tile columns); 4 NR - RH rows beyond the region
  constexpr int PW = RW + 2;         // LDS pitch: region row plus a guard pixel each side
  constexpr int NP = (4 * NR + 2) * PW;
  constexpr int NHALO = 2 * K * RH;  // halo-column pixels
  constexpr int NH = (NHALO + kTB - 1) / kTB;
  constexpr int NS = NR + NH;        // pixels per thread: NR in the tile columns, NH halo pixels
  static_assert(multi_lds_bytes<T, C, RPT, K>() <= kMultiLdsMax, "LDS budget");
  __shared__ T sy[2 * C][NP];
  __shared__ T sxb[C][NP];
  __shared__ FoldedStep<T, C> sstep[K];  // scalar steps of the K iterations
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // this thread's pixels j: region row, region column and LDS index.  The
  // main ones (j < NR) are rows wave*NR + j of column K + lane (rows beyond
  // the region up to 4 NR - 1: never loaded or stored), i.e. an LDS index
  // that grows by PW per row; the halo ones are listed (row -1 = none).
  const int mrow = __builtin_amdgcn_readfirstlane(wave * NR);  // uniform over the wavefront
  const int mq = (mrow + 1) * PW + K + lane + 1;
  int hr[NH], hcol[NH], hq[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int t = threadIdx.x + h * kTB;  // halo pixels row by row: 2K per region row
    const int hc = t % (2 * K);
    hr[h] = t < NHALO ? t / (2 * K) : -1;
    hcol[h] = hc < K ? hc : 64 + hc;
    hq[h] = (hr[h] + 1) * PW + hcol[h] + 1;
  }
  auto srow = [&](int j) { return j < NR ? mrow + j : hr[j - NR]; };
  auto scol = [&](int j) { return j < NR ? K + lane : hcol[j - NR]; };
  auto sidx = [&](int j) { return j < NR ? mq + j * PW : hq[j - NR]; };
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, 64, TH);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const int rr0 = tr.r0 - K, cc0 = tr.c0 - K;  // image position of region pixel (0, 0)
    // the region and a pixel around it inside the image: no boundary masks
    const bool inner = rr0 >= 1 && rr0 + RH < H && cc0 >= 1 && cc0 + RW < W;
    const T* xb = xin + I.off;
    const T* fb = f + I.off;
    const T* yb = yin + 2 * I.off;
    T* xob = xout + I.off;
    T* yob = yout + 2 * I.off;
    // ---- loads: every pixel of the thread (0 beyond the image or the region), then y into LDS
    T xv[NS][C], fv[NS][C], yv[NS][2 * C];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int r = rr0 + srow(j), c = cc0 + scol(j);
#pragma unroll
      for (int k = 0; k < C; ++k) xv[j][k] = fv[j][k] = T(0);
#pragma unroll
      for (int e = 0; e < 2 * C; ++e) yv[j][e] = T(0);
      const bool live = srow(j) >= 0 && srow(j) < RH;
      if (live && (inner || (r >= 0 && r < H && c >= 0 && c < W))) {
        const int64_t p = (int64_t)r * W + c;
        load_c<T, 2 * C>(yb + p * 2 * C, yv[j]);
        load_c<T, C>(xb + p * C, xv[j]);
        load_c<T, C>(fb + p * C, fv[j]);
      }
    }
    // the K iterations' scalar steps, formed once per tile (FP64, as K single
    // launches and opy_tv_accel would form them) instead of in every wave,
    // while the loads are in flight
    if (!PC && threadIdx.x == 0) {
      double tau_d = st[tr.b].tau, sig_d = st[tr.b].sigma;
      for (int i = 0; i < K; ++i) {
        const double th_d = accel_theta(cfg.steps, I, tau_d);
        sstep[i] = fold_step<T, C>(tau_d, sig_d / th_d, th_d, I, cfg);
        tau_d *= th_d;
        sig_d /= th_d;
      }
    }
#pragma unroll
    for (int j = 0; j < NS; ++j)
      if (j < NR || srow(j) >= 0)
#pragma unroll
        for (int e = 0; e < 2 * C; ++e) sy[e][sidx(j)] = yv[j][e];
    __syncthreads();
    const T tvw = (T)I.tvw;
    // preconditioned steps do not change between iterations (theta = 1)
    const PrimalStep<T, C, PC> prox_pc(0.0, I, cfg);
    const DualStep<T, C, PC> ds_pc(T(0), cfg);
#pragma unroll 1
    for (int i = 1; i <= K; ++i) {
      const FoldedStep<T, C> fs = sstep[i - 1];  // unused with preconditioned steps
      const T theta = PC ? T(1) : fs.theta;
      // one iteration; L1: data kind, M: boundary masks (both uniform over the tile)
      auto iterate = [&](auto l1_tag, auto m_tag) {
        constexpr bool L1 = decltype(l1_tag)::value, M = decltype(m_tag)::value;
        // ---- phase A: x^i and xbar^i
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          if (j >= NR && srow(j) < 0) continue;
          const int q = sidx(j);
          const int r = rr0 + srow(j), c = cc0 + scol(j);
          T yup0[C], yl1[C], aty[C];
#pragma unroll
          for (int k = 0; k < C; ++k) {
            yup0[k] = sy[k][q - PW];
            yl1[k] = sy[C + k][q - 1];
          }
          if constexpr (PC) adjoint_px<T, C, M>(yv[j], yv[j] + C, yup0, yl1, r, c, H, W, cfg, aty);
#pragma unroll
          for (int k = 0; k < C; ++k) {
            const T p = PC ? prox_pc.template apply<L1, M>(xv[j][k], aty[k], fv[j][k], r, c, H, W, k, cfg)
                           : primal_px<T, C, L1, M>(xv[j][k], fv[j][k], yv[j][k], yv[j][C + k], yup0[k], yl1[k], r,
                                                    c, H, W, k, fs);
            sxb[k][q] = overrelax(T(1) + theta, theta, p, xv[j][k]);
            xv[j][k] = p;
          }
        }
        __syncthreads();
        // ---- phase B: y^i
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          if (j >= NR && srow(j) < 0) continue;
          const int q = sidx(j);
          const int r = rr0 + srow(j), c = cc0 + scol(j);
          T xbc[C], xbd[C], xbr[C], yp[2 * C], gb[2 * C];
#pragma unroll
          for (int k = 0; k < C; ++k) {
            xbc[k] = sxb[k][q];  // its own xbar, written in phase A
            xbd[k] = sxb[k][q + PW];
            xbr[k] = sxb[k][q + 1];
          }
          if constexpr (PC) yplus_px<T, C, PC, M>(xbc, xbd, xbr, yv[j], r, c, H, W, ds_pc, tvw, cfg, yp, gb);
          else dual_px<T, C, M>(xbc, xbd, xbr, yv[j], r, c, H, W, fs, tvw, yp);
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            yv[j][e] = yp[e];
            sy[e][q] = yp[e];
          }
        }
        __syncthreads();
      };
      using TT = std::true_type;
      using FF = std::false_type;
      if (I.data == OPY_TV_DATA_L1) {
        if (inner) iterate(TT{}, FF{});
        else iterate(TT{}, TT{});
      } else {
        if (inner) iterate(FF{}, FF{});
        else iterate(FF{}, TT{});
      }
    }
    // ---- stores: the tile pixels (tile columns only)
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int ri = srow(j);
      const int r = rr0 + ri, c = cc0 + scol(j);
      if (ri >= K && ri < K + TH && r < H && c < W) {
        const int64_t p = (int64_t)r * W + c;
        store_c<T, C>(xob + p * C, xv[j]);
        store_c<T, 2 * C>(yob + p * 2 * C, yv[j]);
      }
    }
  }
}

template <typename T, int C, int RPT, int K, bool PC>
void launch_multi(const T* xi, T* xo, const T* yi, T* yo, const T* f, const opy_tv_inst_t* insts,
                  const opy_adapt_state_t* st, const int32_t* ai, const int32_t* ao, const int32_t* na, const Cfg& c,
                  int grid, hipStream_t stm) {
  if constexpr (multi_lds_bytes<T, C, RPT, K>() <= kMultiLdsMax) {
    hipLaunchKernelGGL((tv_multi_kernel<T, C, RPT, K, PC>),
                       dim3(launch_grid(KF(tv_multi_kernel<T, C, RPT, K, PC>), grid)), dim3(kTB), 0, stm, xi, xo, yi,
                       yo, f, insts, st, ai, ao, na, c);
  }
}

template <typename T>
int multi_impl(const T* xi, T* xo, const T* yi, T* yo, const T* f, const opy_tv_inst_t* insts,
               const opy_adapt_state_t* st, const int32_t* ai, const int32_t* ao, const int32_t* na,
               const opy_tv_cfg_t* cfg, int k, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (cfg->tw != 64) return fail(kInvalid, "tv_multi: needs tw == 64 (got %d)", cfg->tw);
  if (cfg->rpt != 2 && cfg->rpt != 4 && cfg->rpt != 6 && cfg->rpt != 8)
    return fail(kInvalid, "tv_multi: rows per thread %d not in {2,4,6,8}", cfg->rpt);
  if (k != 2 && k != 3 && k != 4) return fail(kInvalid, "tv_multi: k = %d not in {2,3,4}", k);
  if (multi_lds(cfg->C, cfg->rpt, k, (int)sizeof(T)) > kMultiLdsMax)
    return fail(kInvalid, "tv_multi: C=%d, rpt=%d, k=%d needs %d B of LDS (at most %d)", cfg->C, cfg->rpt, k,
                multi_lds(cfg->C, cfg->rpt, k, (int)sizeof(T)), kMultiLdsMax);
  if (!xi || !xo || !yi || !yo || !f || !insts || !st || !ai || !ao || !na || grid < 0)
    return fail(kInvalid, "tv_multi: null pointer or negative grid");
  if (xi == xo || yi == yo) return fail(kInvalid, "tv_multi: ping-pong buffers must differ");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_M(CC, RR, KK) launch_multi<T, CC, RR, KK, PC>(xi, xo, yi, yo, f, insts, st, ai, ao, na, c, grid, stm)
#define OPY_MK(CC, RR)              \
  switch (k) {                      \
    case 2: OPY_M(CC, RR, 2); break; \
    case 3: OPY_M(CC, RR, 3); break; \
    default: OPY_M(CC, RR, 4); break; \
  }
#define OPY_MR(CC)                  \
  switch (cfg->rpt) {               \
    case 2: OPY_MK(CC, 2); break;   \
    case 4: OPY_MK(CC, 4); break;   \
    case 6: OPY_MK(CC, 6); break;   \
    default: OPY_MK(CC, 8); break;  \
  }
#define OPY_MC                     \
  switch (cfg->C) {                \
    case 1: OPY_MR(1); break;      \
    case 2: OPY_MR(2); break;      \
    case 3: OPY_MR(3); break;      \
    default: OPY_MR(4); break;     \
  }
  if (cfg->steps == OPY_TV_STEPS_PRECOND) {
    constexpr bool PC = true;
    OPY_MC
  } else {
    constexpr bool PC = false;
    OPY_MC
  }
#undef OPY_MC
#undef OPY_MR
#undef OPY_MK
#undef OPY_M
  return check_launch("tv_multi");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_tv_multi_f32(const float* xi, float* xo, const float* yi, float* yo, const float* f,
                             const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                             const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int grid,
                             void* s) {
  return multi_impl<float>(xi, xo, yi, yo, f, insts, st, active_inst, act_off, nact, cfg, k, grid, s);
}
OPY_API int opy_tv_multi_f64(const double* xi, double* xo, const double* yi, double* yo, const double* f,
                             const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                             const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int grid,
                             void* s) {
  return multi_impl<double>(xi, xo, yi, yo, f, insts, st, active_inst, act_off, nact, cfg, k, grid, s);
}
