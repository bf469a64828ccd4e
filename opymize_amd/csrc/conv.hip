// 3x3 convolutions with 32 input and 32 output channels for the warm-start
// network: forward, backward-data and backward-weights on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation).
//
// Activations are NHWC float32 in HBM (32 channels per pixel, zero padded
// when a layer has fewer), weights are prepared as BF16 in the layout
// [tap][out channel][in channel] (tap = 3*dy + dx).  A block of 256 threads
// computes a 16x16 tile of output pixels for all 32 output channels as an
// implicit GEMM: M = 16 pixels of one tile row, N = 16 output channels,
// K = 32 input channels per tap.  The input tile with its one-pixel halo is
// staged in LDS as BF16.
//
// MFMA operand layout of the 16x16x32 BF16 instruction (lane l of 64):
//   A (16 x 32): row l%16, k = 8*(l/16) .. 8*(l/16)+7
//   B (32 x 16): col l%16, k = 8*(l/16) .. 8*(l/16)+7
//   D (16 x 16): col l%16, rows 4*(l/16) .. 4*(l/16)+3
#include "opy_common.h"

namespace opy {
namespace {

constexpr int kCh = 32;         // channels
constexpr int kTile = 16;       // output tile side
constexpr int kHalo = kTile + 2;
constexpr int kPixStride = kCh + 8;  // bf16 elements per LDS pixel (padding against bank conflicts)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline __bf16 to_bf16(float v) { return (__bf16)v; }

// The warm start's first and last layers (conv3x3_pipe_kernel) read and
// write the solver's packed buffers directly (image n of a group at
// n * stride elements):
//   first (kFirst): input channel 0 = f - 0.5, channel 1 = loglam[n], the
//   other 30 zero, i.e. the network's input encoding, formed while staging;
//   last (kLast): x = ((f - 0.5) + 0.5) + o0 and y = (o1, o2) per pixel, the
//   decoding of the output, with only the first 16 output channels computed.
enum { kPlain = 0, kFirst = 1, kLast = 2 };
struct ConvIO {
  const float* f;       // packed images
  int64_t stride;       // elements between consecutive images of the group
  const float* loglam;  // kFirst: per image
  float* xo;            // kLast: image n at xo + n * stride
  float* yo;            // kLast: image n at yo + 2 n * stride (pairs)
};

// ---- forward / backward-data -------------------------------------------------
// y[n,h,w,co] = act(b[co] + sum_{tap,ci} W[tap][co][ci] * x[n, h+dy-1, w+dx-1, ci])
// act: 0 none, 1 ReLU.  bias may be NULL.  TX / TY: element type of x / y in
// HBM (float, or __bf16 for activations between layers: the operands are
// rounded to BF16 for the matrix cores anyway, so a BF16 copy of a layer's
// output gives the next layer the same operands at half the traffic).
template <typename TX, typename TY>
__global__ void __launch_bounds__(256) conv3x3_fwd_kernel(const TX* __restrict__ x, TY* __restrict__ y,
                                                          const __bf16* __restrict__ wt, const float* __restrict__ bias,
                                                          int N, int H, int W, int act) {
  __shared__ __align__(16) __bf16 xs[kHalo * kHalo * kPixStride];
  __shared__ __align__(16) __bf16 ws[9 * kCh * kPixStride];
  const int tid = threadIdx.x;
  for (int e = tid; e < 9 * kCh * kCh; e += 256) {
    const int ci = e % kCh, rest = e / kCh;  // rest = tap*32 + co
    ws[rest * kPixStride + ci] = wt[e];
  }
  const int tx = (W + kTile - 1) / kTile, ty = (H + kTile - 1) / kTile;
  const int ntiles = N * tx * ty;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int n = t / (tx * ty);
    const int rem = t - n * tx * ty;
    const int h0 = (rem / tx) * kTile, w0 = (rem % tx) * kTile;
    __syncthreads();  // previous tile done with xs (and ws staged)
    if constexpr (sizeof(TX) == 4) {
      const TX* xb = x + (int64_t)n * H * W * kCh;
      for (int e = tid; e < kHalo * kHalo * (kCh / 4); e += 256) {
        const int q = e % (kCh / 4), p = e / (kCh / 4);
        const int hh = h0 - 1 + p / kHalo, ww = w0 - 1 + p % kHalo;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (hh >= 0 && hh < H && ww >= 0 && ww < W)
          v = *reinterpret_cast<const float4*>(xb + ((int64_t)hh * W + ww) * kCh + 4 * q);
        __bf16* d = xs + p * kPixStride + 4 * q;
        d[0] = to_bf16(v.x);
        d[1] = to_bf16(v.y);
        d[2] = to_bf16(v.z);
        d[3] = to_bf16(v.w);
      }
    } else {
      // 8 BF16 channels (16 B) per access
      const TX* xb = x + (int64_t)n * H * W * kCh;
      for (int e = tid; e < kHalo * kHalo * (kCh / 8); e += 256) {
        const int q = e % (kCh / 8), p = e / (kCh / 8);
        const int hh = h0 - 1 + p / kHalo, ww = w0 - 1 + p % kHalo;
        bf16x8 v = {};
        if (hh >= 0 && hh < H && ww >= 0 && ww < W)
          v = *reinterpret_cast<const bf16x8*>(xb + ((int64_t)hh * W + ww) * kCh + 8 * q);
        *reinterpret_cast<bf16x8*>(xs + p * kPixStride + 8 * q) = v;
      }
    }
    __syncthreads();
    f32x4 acc[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap % 3;
      bf16x8 b[2];
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf16x8*>(ws + (tap * kCh + j * 16 + li) * kPixStride + 8 * lk);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int row = wave * 4 + m;  // tile row of this M block
        const bf16x8 a =
            *reinterpret_cast<const bf16x8*>(xs + ((row + dy) * kHalo + li + dx) * kPixStride + 8 * lk);
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[m][j], 0, 0, 0);
      }
    }
    TY* yb = y + (int64_t)n * H * W * kCh;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int h = h0 + wave * 4 + m;
      if (h >= H) continue;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int co = j * 16 + li;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int w = w0 + 4 * lk + r;
          if (w >= W) continue;
          float v = acc[m][j][r] + bv;
          if (act == 1) v = v > 0.f ? v : 0.f;
          yb[((int64_t)h * W + w) * kCh + co] = (TY)v;
        }
      }
    }
  }
}

// ---- inference layers: BF16 activations, software-pipelined -----------------------
// The arithmetic of conv3x3_fwd_kernel<__bf16, __bf16> (kPlain), and the
// first and last layers reading and writing the solver's buffers (kFirst,
// kLast; see ConvIO), for the warm start's inference: each block walks its
// tiles with the next tile's
// input already on its way into registers while the current one is on the
// matrix cores, and a BF16 output tile is staged in LDS (in the input tile's
// space, which the matrix products are done with) so that every pixel's 64
// bytes leave in four 16-byte stores instead of 32-byte pieces.
template <int MODE>
__global__ void __launch_bounds__(256) conv3x3_pipe_kernel(const __bf16* __restrict__ x, __bf16* __restrict__ y,
                                                           const __bf16* __restrict__ wt,
                                                           const float* __restrict__ bias, int N, int H, int W, int act,
                                                           ConvIO io) {
  constexpr int NJ = MODE == kLast ? 1 : 2;  // blocks of 16 output channels
  constexpr int NHP = kHalo * kHalo;         // pixels of a halo tile
  constexpr int NV = NHP * (kCh / 8);        // 16-byte vectors of a halo tile
  constexpr int NPF = MODE == kFirst ? (NHP + 255) / 256 : (NV + 255) / 256;
  __shared__ __align__(16) __bf16 xs[NHP * kPixStride];
  __shared__ __align__(16) __bf16 ws[9 * kCh * kPixStride];
  static_assert(kTile * kTile <= NHP, "the output tile is staged in the input tile's space");
  const int tid = threadIdx.x;
  for (int e = tid; e < 9 * kCh * kCh; e += 256) {
    const int ci = e % kCh, rest = e / kCh;  // rest = tap*32 + co
    ws[rest * kPixStride + ci] = wt[e];
  }
  const int tx = (W + kTile - 1) / kTile, ty = (H + kTile - 1) / kTile;
  const int ntiles = N * tx * ty;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  auto origin = [&](int t, int& n, int& h0, int& w0) {
    n = t / (tx * ty);
    const int rem = t - n * tx * ty;
    h0 = (rem / tx) * kTile;
    w0 = (rem % tx) * kTile;
  };
  bf16x8 pv[MODE == kFirst ? 1 : NPF];  // the next tile's input vectors
  float pf[MODE == kFirst ? NPF : 1];   // kFirst: the next tile's image values
  auto prefetch = [&](int t) {
    int n, h0, w0;
    origin(t, n, h0, w0);
    if constexpr (MODE == kFirst) {
      const float* fb = io.f + (int64_t)n * io.stride;
#pragma unroll
      for (int i = 0; i < NPF; ++i) {
        const int p = tid + 256 * i;
        const int hh = h0 - 1 + p / kHalo, ww = w0 - 1 + p % kHalo;
        pf[i] = (p < NHP && hh >= 0 && hh < H && ww >= 0 && ww < W) ? fb[(int64_t)hh * W + ww] : 0.f;
      }
    } else {
      const __bf16* xb = x + (int64_t)n * H * W * kCh;
#pragma unroll
      for (int i = 0; i < NPF; ++i) {
        const int e = tid + 256 * i;
        const int q = e % (kCh / 8), p = e / (kCh / 8);
        const int hh = h0 - 1 + p / kHalo, ww = w0 - 1 + p % kHalo;
        pv[i] = bf16x8{};
        if (e < NV && hh >= 0 && hh < H && ww >= 0 && ww < W)
          pv[i] = *reinterpret_cast<const bf16x8*>(xb + ((int64_t)hh * W + ww) * kCh + 8 * q);
      }
    }
  };
  auto to_lds = [&](int t) {
    int n, h0, w0;
    origin(t, n, h0, w0);
    if constexpr (MODE == kFirst) {
      const __bf16 ll = to_bf16(io.loglam[n]);
#pragma unroll
      for (int i = 0; i < NPF; ++i) {
        const int p = tid + 256 * i;
        if (p >= NHP) continue;
        const int hh = h0 - 1 + p / kHalo, ww = w0 - 1 + p % kHalo;
        bf16x8 v = {};
        if (hh >= 0 && hh < H && ww >= 0 && ww < W) {
          v[0] = to_bf16(pf[i] - 0.5f);
          v[1] = ll;
        }
        bf16x8* d = reinterpret_cast<bf16x8*>(xs + p * kPixStride);
        d[0] = v;
        d[1] = bf16x8{};
        d[2] = bf16x8{};
        d[3] = bf16x8{};
      }
    } else {
#pragma unroll
      for (int i = 0; i < NPF; ++i) {
        const int e = tid + 256 * i;
        if (e < NV) *reinterpret_cast<bf16x8*>(xs + (e / (kCh / 8)) * kPixStride + 8 * (e % (kCh / 8))) = pv[i];
      }
    }
  };
  if ((int)blockIdx.x < ntiles) prefetch(blockIdx.x);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    __syncthreads();  // the previous tile's output has left xs (and ws is staged)
    to_lds(t);
    __syncthreads();
    if (t + (int)gridDim.x < ntiles) prefetch(t + gridDim.x);
    f32x4 acc[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap % 3;
      bf16x8 b[2];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        b[j] = *reinterpret_cast<const bf16x8*>(ws + (tap * kCh + j * 16 + li) * kPixStride + 8 * lk);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int row = wave * 4 + m;
        const bf16x8 a =
            *reinterpret_cast<const bf16x8*>(xs + ((row + dy) * kHalo + li + dx) * kPixStride + 8 * lk);
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[m][j], 0, 0, 0);
      }
    }
    int n, h0, w0;
    origin(t, n, h0, w0);
    if constexpr (MODE == kLast) {
      // the 3 output channels of the tile go through LDS (FP32, channel-major,
      // in the input tile's space) so that each thread decodes one pixel: x
      // and the y pair leave as 4- and 8-byte stores of consecutive pixels
      static_assert(3 * kTile * kTile * sizeof(float) <= sizeof(xs), "output staging fits the input tile");
      float* os = reinterpret_cast<float*>(xs);
      __syncthreads();  // every wave is done with the input tile
      if (li < 3) {
        const float bv = bias ? bias[li] : 0.f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float v = acc[m][0][r] + bv;
            if (act == 1) v = v > 0.f ? v : 0.f;
            os[li * kTile * kTile + (wave * 4 + m) * kTile + 4 * lk + r] = v;
          }
        }
      }
      __syncthreads();
      const int h = h0 + tid / kTile, w = w0 + tid % kTile;
      if (h < H && w < W) {
        const int64_t p = (int64_t)h * W + w;
        const float* fb = io.f + (int64_t)n * io.stride;
        float* xo = io.xo + (int64_t)n * io.stride;
        float* yo = io.yo + 2 * (int64_t)n * io.stride;
        xo[p] = ((fb[p] - 0.5f) + 0.5f) + os[tid];
        *reinterpret_cast<float2*>(yo + 2 * p) = make_float2(os[kTile * kTile + tid], os[2 * kTile * kTile + tid]);
      }
    } else {
      __syncthreads();  // every wave is done with the input tile
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int row = wave * 4 + m;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int co = j * 16 + li;
          const float bv = bias ? bias[co] : 0.f;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float v = acc[m][j][r] + bv;
            if (act == 1) v = v > 0.f ? v : 0.f;
            xs[(row * kTile + 4 * lk + r) * kPixStride + co] = to_bf16(v);
          }
        }
      }
      __syncthreads();
      __bf16* yb = y + (int64_t)n * H * W * kCh;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = tid + 256 * i;
        const int px = c / (kCh / 8), part = c % (kCh / 8);
        const int h = h0 + px / kTile, w = w0 + px % kTile;
        if (h < H && w < W)
          *reinterpret_cast<bf16x8*>(yb + ((int64_t)h * W + w) * kCh + 8 * part) =
              *reinterpret_cast<const bf16x8*>(xs + px * kPixStride + 8 * part);
      }
    }
  }
}

// ---- backward-weights ----------------------------------------------------------
// dW[tap][co][ci] = sum_{n,h,w} g[n,h,w,co] * x[n, h+dy-1, w+dx-1, ci]   (partials per block)
// dbias[co] = sum g[..., co]  (partials per block, slot 9*32*32 + co)
// As a GEMM per tap: D (co x ci) += A (co x pixels) * B (pixels x ci).
__global__ void __launch_bounds__(256) conv3x3_wgrad_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                                            float* __restrict__ partials, int N, int H, int W) {
  // channel-major tiles: gsT[co][pixel of the 16x16 tile], xsT[ci][pixel of the 18x18 halo tile]
  constexpr int kGS = kTile * kTile + 8;
  constexpr int kXS = kHalo * kHalo + 8;
  __shared__ __align__(16) __bf16 gsT[kCh * kGS];
  __shared__ __align__(16) __bf16 xsT[kCh * kXS];
  __shared__ float bsum[256];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int tx = (W + kTile - 1) / kTile, ty = (H + kTile - 1) / kTile;
  const int ntiles = N * tx * ty;
  // wave w owns taps w, w+4, w+8 (if < 9); 2 x 2 output blocks (co, ci) per tap
  f32x4 acc[3][2][2];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[a][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.f;  // FP32 bias-gradient partial of channel tid % 32 (the channel this thread stages)
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int n = t / (tx * ty);
    const int rem = t - n * tx * ty;
    const int h0 = (rem / tx) * kTile, w0 = (rem % tx) * kTile;
    __syncthreads();
    const float* xb = x + (int64_t)n * H * W * kCh;
    const float* gb = g + (int64_t)n * H * W * kCh;
    for (int e = tid; e < kHalo * kHalo * kCh; e += 256) {
      const int c = e % kCh, p = e / kCh;
      const int hh = h0 - 1 + p / kHalo, ww = w0 - 1 + p % kHalo;
      float v = 0.f;
      if (hh >= 0 && hh < H && ww >= 0 && ww < W) v = xb[((int64_t)hh * W + ww) * kCh + c];
      xsT[c * kXS + p] = to_bf16(v);
    }
    for (int e = tid; e < kTile * kTile * kCh; e += 256) {
      const int c = e % kCh, p = e / kCh;
      const int hh = h0 + p / kTile, ww = w0 + p % kTile;
      float v = 0.f;
      if (hh < H && ww < W) v = gb[((int64_t)hh * W + ww) * kCh + c];
      gsT[c * kGS + p] = to_bf16(v);
      bacc += v;
    }
    __syncthreads();
    // K loop over the 256 tile pixels in chunks of 32 (two tile rows)
#pragma unroll
    for (int ta = 0; ta < 3; ++ta) {
      const int tap = wave + 4 * ta;
      if (tap >= 9) break;
      const int dy = tap / 3, dx = tap % 3;
#pragma unroll 2
      for (int kc = 0; kc < 8; ++kc) {
        // pixels of this chunk: rows 2kc, 2kc+1; lane group lk covers 8 pixels: row 2kc + lk/2, cols 8*(lk%2)..+7
        const int prow = 2 * kc + (lk >> 1), pcol = 8 * (lk & 1);
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
          a[i] = *reinterpret_cast<const bf16x8*>(gsT + (i * 16 + li) * kGS + prow * kTile + pcol);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          // x at halo coordinates (prow + dy, pcol + dx .. +7): not 16-byte aligned in general
          const __bf16* src = xsT + (j * 16 + li) * kXS + (prow + dy) * kHalo + pcol + dx;
          bf16x8 v;
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = src[q];
          b[j] = v;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[ta][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[ta][i][j], 0, 0, 0);
      }
    }
  }
  // write this block's partials: D rows = co (4*(l/16)+r within the 16-block), cols = ci (l%16)
  float* pb = partials + (int64_t)blockIdx.x * (9 * kCh * kCh + kCh);
#pragma unroll
  for (int ta = 0; ta < 3; ++ta) {
    const int tap = wave + 4 * ta;
    if (tap >= 9) break;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = i * 16 + 4 * lk + r;
          const int ci = j * 16 + li;
          pb[(tap * kCh + co) * kCh + ci] = acc[ta][i][j][r];
        }
  }
  bsum[tid] = bacc;
  __syncthreads();
  if (tid < kCh) {
    float sb = 0.f;
    for (int k = 0; k < 256 / kCh; ++k) sb += bsum[tid + kCh * k];
    pb[9 * kCh * kCh + tid] = sb;
  }
}

// sum of nb partial vectors of length len, in block order (deterministic)
__global__ void sum_partials_kernel(const float* __restrict__ partials, int nb, int len, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += partials[(int64_t)b * len + i];
  out[i] = s;
}

// weights: master W[tap][ci][co] (float) -> bf16 [tap][co][ci] (transpose=0, forward)
// or the backward-data weights Wb[tap][ci][co] = W[8-tap][ci][co] in the [tap][out=ci][in=co] layout
__global__ void prep_weights_kernel(const float* __restrict__ w, __bf16* __restrict__ out, int flip) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 9 * kCh * kCh) return;
  const int tap = e / (kCh * kCh);
  const int r = e % (kCh * kCh);
  const int o = r / kCh, i = r % kCh;  // out[tap][o][i]
  if (!flip) out[e] = to_bf16(w[(tap * kCh + i) * kCh + o]);       // W[tap][ci=i][co=o]
  else out[e] = to_bf16(w[((8 - tap) * kCh + o) * kCh + i]);        // W[8-tap][ci=o][co=i]
}

// g = g * (y > 0): ReLU backward through the stored activation
__global__ void relu_bwd_kernel(float* __restrict__ g, const float* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(y[i] > 0.f)) g[i] = 0.f;
}

// Adam step on n parameters (FP32 master copy)
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// weighted squared error over the first nc channels of NHWC-32 tensors:
// loss = sum_{p, c<nc} wc[c]*(out - tgt)^2 / npix, grad = d loss / d out (zero for c >= nc);
// per-block loss partials in FP64
__global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt,
                                                  float* __restrict__ grad, int64_t npix, int nc, float w0, float w1,
                                                  float w2, double* __restrict__ partials) {
  __shared__ double sm[4];
  const int64_t n = npix * kCh;
  const float inv = 1.0f / (float)npix;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % kCh);
    float gv = 0.f;
    if (c < nc) {
      const float wc = c == 0 ? w0 : (c == 1 ? w1 : w2);
      const float d = out[i] - tgt[i];
      s += (double)wc * (double)d * (double)d;
      gv = 2.f * wc * d * inv;
    }
    grad[i] = gv;
  }
  s = block_sum(s, sm);
  if (threadIdx.x == 0) partials[blockIdx.x] = s / (double)npix;
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_mse3_f32(const float* out, const float* tgt, float* grad, int64_t npix, int nc, double w0, double w1,
                         double w2, double* partials, int grid, void* s) {
  if (!out || !tgt || !grad || !partials || npix < 1 || nc < 1 || nc > 3 || grid < 1)
    return fail(kInvalid, "mse3: bad arguments");
  hipLaunchKernelGGL(mse_kernel, dim3(grid), dim3(256), 0, as_stream(s), out, tgt, grad, npix, nc, (float)w0,
                     (float)w1, (float)w2, partials);
  return check_launch("mse3");
}

OPY_API int opy_conv3x3_fwd_bf16(const float* x, float* y, const void* wt, const float* bias, int N, int H, int W,
                                 int act, int grid, void* s) {
  if (!x || !y || !wt || N < 1 || H < 1 || W < 1 || grid < 1) return fail(kInvalid, "conv3x3_fwd: bad arguments");
  if (x == y) return fail(kInvalid, "conv3x3_fwd: in-place convolution is not supported");
  if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return fail(kInvalid, "conv3x3_fwd: input must be 16-byte aligned");
  hipLaunchKernelGGL((conv3x3_fwd_kernel<float, float>), dim3(grid), dim3(256), 0, as_stream(s), x, y,
                     static_cast<const __bf16*>(wt), bias, N, H, W, act);
  return check_launch("conv3x3_fwd");
}
OPY_API int opy_conv3x3_fwd_io(const void* x, void* y, const void* wt, const float* bias, int N, int H, int W,
                               int act, int x_bf16, int y_bf16, int grid, void* s) {
  if (!x || !y || !wt || N < 1 || H < 1 || W < 1 || grid < 1) return fail(kInvalid, "conv3x3_fwd_io: bad arguments");
  if (x == y) return fail(kInvalid, "conv3x3_fwd_io: in-place convolution is not supported");
  if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return fail(kInvalid, "conv3x3_fwd_io: input must be 16-byte aligned");
  if (y_bf16 && (reinterpret_cast<uintptr_t>(y) & 15) != 0)
    return fail(kInvalid, "conv3x3_fwd_io: BF16 output must be 16-byte aligned");
  const __bf16* w = static_cast<const __bf16*>(wt);
  hipStream_t st = as_stream(s);
  if (x_bf16 && y_bf16)
    hipLaunchKernelGGL((conv3x3_pipe_kernel<kPlain>), dim3(grid), dim3(256), 0, st, static_cast<const __bf16*>(x),
                       static_cast<__bf16*>(y), w, bias, N, H, W, act, ConvIO{});
  else if (x_bf16)
    hipLaunchKernelGGL((conv3x3_fwd_kernel<__bf16, float>), dim3(grid), dim3(256), 0, st,
                       static_cast<const __bf16*>(x), static_cast<float*>(y), w, bias, N, H, W, act);
  else if (y_bf16)
    hipLaunchKernelGGL((conv3x3_fwd_kernel<float, __bf16>), dim3(grid), dim3(256), 0, st,
                       static_cast<const float*>(x), static_cast<__bf16*>(y), w, bias, N, H, W, act);
  else
    hipLaunchKernelGGL((conv3x3_fwd_kernel<float, float>), dim3(grid), dim3(256), 0, st,
                       static_cast<const float*>(x), static_cast<float*>(y), w, bias, N, H, W, act);
  return check_launch("conv3x3_fwd_io");
}

OPY_API int opy_conv3x3_first(const float* f, int64_t stride, const float* loglam, void* y, const void* wt,
                              const float* bias, int N, int H, int W, int grid, void* s) {
  if (!f || !loglam || !y || !wt || N < 1 || H < 1 || W < 1 || grid < 1 || stride < (int64_t)H * W)
    return fail(kInvalid, "conv3x3_first: bad arguments");
  if ((reinterpret_cast<uintptr_t>(y) & 15) != 0) return fail(kInvalid, "conv3x3_first: output must be 16-byte aligned");
  ConvIO io{f, stride, loglam, nullptr, nullptr};
  hipLaunchKernelGGL((conv3x3_pipe_kernel<kFirst>), dim3(grid), dim3(256), 0, as_stream(s), nullptr,
                     static_cast<__bf16*>(y), static_cast<const __bf16*>(wt), bias, N, H, W, 1, io);
  return check_launch("conv3x3_first");
}
OPY_API int opy_conv3x3_last(const void* x, const float* f, int64_t stride, float* xo, float* yo, const void* wt,
                             const float* bias, int N, int H, int W, int grid, void* s) {
  if (!x || !f || !xo || !yo || !wt || N < 1 || H < 1 || W < 1 || grid < 1 || stride < (int64_t)H * W)
    return fail(kInvalid, "conv3x3_last: bad arguments");
  if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return fail(kInvalid, "conv3x3_last: input must be 16-byte aligned");
  if ((reinterpret_cast<uintptr_t>(yo) & 7) != 0) return fail(kInvalid, "conv3x3_last: yo must be 8-byte aligned");
  ConvIO io{f, stride, nullptr, xo, yo};
  hipLaunchKernelGGL((conv3x3_pipe_kernel<kLast>), dim3(grid), dim3(256), 0, as_stream(s),
                     static_cast<const __bf16*>(x), nullptr, static_cast<const __bf16*>(wt), bias, N, H, W, 0, io);
  return check_launch("conv3x3_last");
}

OPY_API int opy_conv3x3_wgrad_bf16(const float* x, const float* g, float* partials, int N, int H, int W, int grid,
                                   void* s) {
  if (!x || !g || !partials || N < 1 || H < 1 || W < 1 || grid < 1) return fail(kInvalid, "conv3x3_wgrad: bad arguments");
  hipLaunchKernelGGL(conv3x3_wgrad_kernel, dim3(grid), dim3(256), 0, as_stream(s), x, g, partials, N, H, W);
  return check_launch("conv3x3_wgrad");
}

OPY_API int opy_sum_partials_f32(const float* partials, int nb, int len, float* out, void* s) {
  if (!partials || !out || nb < 1 || len < 1) return fail(kInvalid, "sum_partials: bad arguments");
  hipLaunchKernelGGL(sum_partials_kernel, dim3((len + 255) / 256), dim3(256), 0, as_stream(s), partials, nb, len, out);
  return check_launch("sum_partials");
}

OPY_API int opy_conv3x3_prep_weights(const float* w, void* out, int flip, void* s) {
  if (!w || !out) return fail(kInvalid, "prep_weights: null pointer");
  hipLaunchKernelGGL(prep_weights_kernel, dim3((9 * 32 * 32 + 255) / 256), dim3(256), 0, as_stream(s), w,
                     static_cast<__bf16*>(out), flip);
  return check_launch("prep_weights");
}

OPY_API int opy_relu_bwd_f32(float* g, const float* y, int64_t n, void* s) {
  if (n > 0 && (!g || !y)) return fail(kInvalid, "relu_bwd: null pointer");
  if (n > 0)
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(s), g, y, n);
  return check_launch("relu_bwd");
}

OPY_API int opy_adam_f32(float* p, const float* g, float* m, float* v, int n, double lr, double b1, double b2,
                         double eps, int step, void* s) {
  if (n < 1 || !p || !g || !m || !v || step < 1) return fail(kInvalid, "adam: bad arguments");
  const float c1 = (float)(1.0 - pow(b1, step)), c2 = (float)(1.0 - pow(b2, step));
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(s), p, g, m, v, n, (float)lr,
                     (float)b1, (float)b2, (float)eps, c1, c2);
  return check_launch("adam");
}
