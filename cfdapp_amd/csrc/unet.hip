// The learned preconditioner's network on the device: weight packing, the
// single-channel first and last layers (VALU), forward and backward passes
// over the MFMA convolutions, the A-norm training loss and Adam.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "reduce.hpp"
#include "unet.hpp"
#include "unet_net.hpp"

namespace cfdhip {

namespace {

__device__ inline float bf2f(bf16 h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ inline bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u)
    return static_cast<bf16>((u >> 16) | ((u & 0xffffu) != 0 ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<bf16>(u >> 16);
}
__device__ inline float roundBf(float f) { return bf2f(f2bf(f)); }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

inline dim3 grid1(int64_t n, int64_t cap = 8192) {
  int64_t b = (n + 255) / 256;
  if (b > cap) b = cap;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : b));
}

// Forward fragments, K walked in chunks of kChunk(cin) input channels:
// element (((ch MSC + m) NB + nb) 64 + lane) 8 + e holds W[tap][ci][co] with
// k = 32 m + 8 (lane / 16) + e = tap * chunk + c, ci = ch * chunk + c,
// co = 16 nb + lane % 16 (zero for k >= 9 chunk).
__device__ inline bf16 packedValue(const float* __restrict__ master, int cin, int cout, int64_t idx,
                                   bool transposed) {
  const int cinK = transposed ? cout : cin;   // K-side channels of this pass
  const int coutN = transposed ? cin : cout;  // N-side channels of this pass
  const int cc = kChunk(cinK);
  const int msc = (9 * cc + 31) / 32;
  const int nbt = coutN / 16;
  const int e = static_cast<int>(idx & 7);
  const int lane = static_cast<int>((idx >> 3) & 63);
  const int64_t blk = idx >> 9;
  const int nb = static_cast<int>(blk % nbt);
  const int64_t qm = blk / nbt;
  const int ch = static_cast<int>(qm / msc), m = static_cast<int>(qm % msc);
  const int k = 32 * m + 8 * (lane >> 4) + e;
  if (k >= 9 * cc) return 0;
  const int tap = k / cc;
  const int c = ch * cc + (k - tap * cc);  // K-side channel
  const int o = 16 * nb + (lane & 15);     // N-side channel
  // transposed conv: cin' = cout, cout' = cin, W'[tap'][c'][o'] = W[8 - tap'][o'][c']
  return transposed ? f2bf(master[(static_cast<int64_t>(8 - tap) * cin + o) * cout + c])
                    : f2bf(master[(static_cast<int64_t>(tap) * cin + c) * cout + o]);
}

__global__ void packKernel(const float* __restrict__ master, bf16* __restrict__ packed, int cin,
                           int cout, int64_t total, int transposed) {
  for (int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<int64_t>(gridDim.x) * blockDim.x)
    packed[idx] = packedValue(master, cin, cout, idx, transposed != 0);
}

// e0a = conv(x, 1 -> 16). x(y, x) = src[y srcW + x] * scale (inside srcH x srcW,
// zero elsewhere); one pixel per thread.
template <class TS>
__global__ __launch_bounds__(256) void firstConvKernel(const TS* __restrict__ src,
                                                       const float* __restrict__ scale, int srcH,
                                                       int srcW, int64_t srcBatch,
                                                       const float* __restrict__ w,
                                                       float* __restrict__ out, int n, int h,
                                                       int wd, int pad) {
  __shared__ float ws[144];
  if (threadIdx.x < 144) ws[threadIdx.x] = roundBf(w[threadIdx.x]);
  __syncthreads();
  const int64_t total = static_cast<int64_t>(n) * h * wd;
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(p % wd);
    const int64_t t = p / wd;
    const int y = static_cast<int>(t % h);
    const int b = static_cast<int>(t / h);
    float v[9];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
      v[tap] = 0.f;
      // the padding reflects at the source's edges (the physical image)
      const float sg = padSign(yy, xx, srcH < h ? srcH : h, srcW < wd ? srcW : wd, pad);
      if (sg != 0.f && yy < srcH && xx < srcW) {
        const int64_t si = b * srcBatch + static_cast<int64_t>(yy) * srcW + xx;
        const float sv = static_cast<float>(src[si]);
        v[tap] = sg * (scale != nullptr ? sv * scale[si] : sv);
      }
    }
    fx4* o = reinterpret_cast<fx4*>(out + p * 16);
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      fx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += v[tap] * ws[tap * 16 + c4 * 4 + k];
      o[c4] = acc;
    }
  }
}

// y = conv(d0, 16 -> 1) on an 8 x 64 output tile: every pixel of the haloed
// tile is read once and reduced against the nine taps (s[tap][pixel] in LDS),
// then each output sums its nine neighbours' tap values.
constexpr int kLTH = 8, kLTW = 64;
constexpr int kLR = kLTH + 2, kLC = kLTW + 2, kLNP = kLR * kLC;

struct LastTileLds {
  float ws[144];
  float s[9 * kLNP];
};

// h x wd: the d0 tensor; ph x pw: the physical image its padding reflects at
__device__ inline void lastTile(const float* __restrict__ d0, LastTileLds& L, int b, int y0, int x0,
                                int h, int wd, int ph, int pw, int pad, float out[2]) {
  constexpr int PER = (kLNP + 255) / 256;
  fx4 v[PER][4];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int px = threadIdx.x + k * 256;
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) v[k][c4] = fx4{0.f, 0.f, 0.f, 0.f};
    if (px < kLNP) {
      const int ty = px / kLC, tx = px - ty * kLC;
      int yy = y0 - 1 + ty, xx = x0 - 1 + tx;
      const float sg = padSign(yy, xx, ph, pw, pad);
      if (sg != 0.f) {
        const fx4* src = reinterpret_cast<const fx4*>(
            d0 + ((static_cast<int64_t>(b) * h + yy) * wd + xx) * 16);
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) v[k][c4] = sg * src[c4];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int px = threadIdx.x + k * 256;
    if (px >= kLNP) continue;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      float dot = 0.f;
#pragma unroll
      for (int c4 = 0; c4 < 4; ++c4)
#pragma unroll
        for (int q = 0; q < 4; ++q) dot += v[k][c4][q] * L.ws[tap * 16 + c4 * 4 + q];
      L.s[tap * kLNP + px] = dot;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int o = threadIdx.x + j * 256;
    const int ty = o / kLTW, tx = o - ty * kLTW;
    float acc = 0.f;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
      acc += L.s[tap * kLNP + (ty + tap / 3) * kLC + tx + tap % 3];
    out[j] = acc;
  }
}

__device__ inline void lastTileCoords(int64_t t, int tilesX, int tilesY, int& b, int& y0, int& x0) {
  x0 = static_cast<int>(t % tilesX) * kLTW;
  t /= tilesX;
  y0 = static_cast<int>(t % tilesY) * kLTH;
  b = static_cast<int>(t / tilesY);
}

// Network output y (FP32, n x ph x pw, compact) of the physical image
// ph x pw inside the h x wd tensors (ph = h, pw = wd for training).
__global__ __launch_bounds__(256) void lastConvKernel(const float* __restrict__ d0,
                                                      const float* __restrict__ w,
                                                      float* __restrict__ y, int n, int h, int wd,
                                                      int ph, int pw, int pad) {
  __shared__ LastTileLds L;
  if (threadIdx.x < 144) L.ws[threadIdx.x] = roundBf(w[threadIdx.x]);
  const int tilesX = (pw + kLTW - 1) / kLTW, tilesY = (ph + kLTH - 1) / kLTH;
  const int64_t tiles = static_cast<int64_t>(tilesX) * tilesY * n;
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    int b, y0, x0;
    lastTileCoords(t, tilesX, tilesY, b, y0, x0);
    __syncthreads();
    float out[2];
    lastTile(d0, L, b, y0, x0, h, wd, ph, pw, pad, out);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int o = threadIdx.x + j * 256;
      const int yy = y0 + o / kLTW, xx = x0 + o % kLTW;
      if (yy < ph && xx < pw) y[(static_cast<int64_t>(b) * ph + yy) * pw + xx] = out[j];
    }
  }
}

// Preconditioner output on the solver grid (ny x nx inside the network's
// h x w): z = scale * y, z[pin] = r[pin] scale[pin]^2, and block partials of
// (z, q).
template <class TZ>
__global__ __launch_bounds__(256) void lastConvApplyKernel(const float* __restrict__ d0,
                                                           const float* __restrict__ w,
                                                           const float* __restrict__ scale,
                                                           const TZ* __restrict__ r,
                                                           const TZ* __restrict__ q,
                                                           TZ* __restrict__ z, int ny, int nx,
                                                           int64_t pin, int h, int wd,
                                                           int pad, double* partial) {
  __shared__ LastTileLds L;
  if (threadIdx.x < 144) L.ws[threadIdx.x] = roundBf(w[threadIdx.x]);
  const int tilesX = (nx + kLTW - 1) / kLTW, tilesY = (ny + kLTH - 1) / kLTH;
  const int64_t tiles = static_cast<int64_t>(tilesX) * tilesY;
  double acc[1] = {0.0};
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    int b, y0, x0;
    lastTileCoords(t, tilesX, tilesY, b, y0, x0);
    __syncthreads();
    float out[2];
    lastTile(d0, L, 0, y0, x0, h, wd, ny, nx, pad, out);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int o = threadIdx.x + j * 256;
      const int yy = y0 + o / kLTW, xx = x0 + o % kLTW;
      if (yy >= ny || xx >= nx) continue;
      const int64_t c = static_cast<int64_t>(yy) * nx + xx;
      TZ v = static_cast<TZ>(out[j] * scale[c]);
      if (c == pin) v = r[c] * static_cast<TZ>(scale[c]) * static_cast<TZ>(scale[c]);
      z[c] = v;
      if (q != nullptr) acc[0] += static_cast<double>(v) * static_cast<double>(q[c]);
    }
  }
  if (partial != nullptr) blockPartials<1>(acc, partial);
}

// gd0 = conv^T(gy, 1 -> 16): gd0[p][c] = sum_tap gy[p - off(tap)] W[tap][c]
// (Neumann padding: gy at the nearest pixel inside, exact for W mirror-symmetric in both axes).
__global__ __launch_bounds__(256) void lastConvBwdKernel(const float* __restrict__ gy,
                                                         const float* __restrict__ w,
                                                         float* __restrict__ gd0, int n, int h,
                                                         int wd, int pad) {
  __shared__ float ws[144];
  if (threadIdx.x < 144) ws[threadIdx.x] = roundBf(w[threadIdx.x]);
  __syncthreads();
  const int64_t total = static_cast<int64_t>(n) * h * wd;
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(p % wd);
    const int64_t t = p / wd;
    const int y = static_cast<int>(t % h);
    const int b = static_cast<int>(t / h);
    float g[9];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      int yy = y - tap / 3 + 1, xx = x - tap % 3 + 1;
      const float sg = padSign(yy, xx, h, wd, pad);
      g[tap] = sg != 0.f ? sg * gy[(static_cast<int64_t>(b) * h + yy) * wd + xx] : 0.f;
    }
    fx4* o = reinterpret_cast<fx4*>(gd0 + p * 16);
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      fx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += g[tap] * ws[tap * 16 + c4 * 4 + k];
      o[c4] = acc;
    }
  }
}

// Weight gradient of a single-channel layer, 144 sums per block (BF16
// operands like the MFMA weight gradients, FP32 sums):
//   first (mode 0): dW[tap][c] = sum_p x[p + off(tap)] g16[p][c]
//   last  (mode 1): dW[tap][c] = sum_p a16[p + off(tap)][c] gy[p]
// x / gy are FP32 scalar images, g16 / a16 FP32 16-channel images.
__global__ __launch_bounds__(256) void singleDwKernel(const float* __restrict__ s1,
                                                      const float* __restrict__ m16, int mode,
                                                      int n, int h, int wd, int pad,
                                                      float* partial) {
  float acc[144];
#pragma unroll
  for (int k = 0; k < 144; ++k) acc[k] = 0.f;
  const int64_t total = static_cast<int64_t>(n) * h * wd;
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(p % wd);
    const int64_t t = p / wd;
    const int y = static_cast<int>(t % h);
    const int b = static_cast<int>(t / h);
    if (mode == 0) {
      float gv[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) gv[c] = roundBf(m16[p * 16 + c]);
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        const float sg = padSign(yy, xx, h, wd, pad);
        if (sg == 0.f) continue;
        const float xv = sg * roundBf(s1[(static_cast<int64_t>(b) * h + yy) * wd + xx]);
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[tap * 16 + c] += xv * gv[c];
      }
    } else {
      const float g = roundBf(s1[p]);
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        const float sg = padSign(yy, xx, h, wd, pad);
        if (sg == 0.f) continue;
        const float* ap = m16 + ((static_cast<int64_t>(b) * h + yy) * wd + xx) * 16;
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[tap * 16 + c] += sg * roundBf(ap[c]) * g;
      }
    }
  }
  __shared__ float lds[4][144];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 144; ++k) {
    float v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) lds[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 144)
    partial[static_cast<int64_t>(blockIdx.x) * 144 + threadIdx.x] =
        lds[0][threadIdx.x] + lds[1][threadIdx.x] + lds[2][threadIdx.x] + lds[3][threadIdx.x];
}

__global__ void singleDwReduceKernel(const float* __restrict__ partial, int parts,
                                     float* __restrict__ grad) {
  const int k = threadIdx.x;
  if (k >= 144) return;
  float s = 0.f;
  for (int b = 0; b < parts; ++b) s += partial[static_cast<int64_t>(b) * 144 + k];
  grad[k] = s;
}

// dW[tap][ci][co] (master layout) = sum over splits, fixed order.
__global__ void dwReduceKernel(const float* __restrict__ partial, int splits, int cin, int cout,
                               float* __restrict__ grad) {
  const int ncb = cin / 16, nob = cout / 16, npair = 9 * ncb;
  const int64_t total = 9LL * cin * cout;
  for (int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; idx < total;
       idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int co = static_cast<int>(idx % cout);
    const int64_t t = idx / cout;
    const int ci = static_cast<int>(t % cin);
    const int tap = static_cast<int>(t / cin);
    const int pair = tap * ncb + ci / 16, ob = co / 16;
    const int e = (ci & 15) * 16 + (co & 15);
    float s = 0.f;
    for (int sp = 0; sp < splits; ++sp)
      s += partial[((static_cast<int64_t>(sp) * npair + pair) * nob + ob) * 256 + e];
    grad[idx] = s;
  }
}

// A-norm loss of the scaled system: d = y - e, A d with diagonal cD (null:
// 1) and couplings cE (to x + 1) and cN (to y + 1) of the sample's system;
// L = (1 / n) sum_s (d, A d)_s / den_s, gy = 2 A d / (den_s n).
__global__ __launch_bounds__(256) void anormLossKernel(
    const float* __restrict__ y, const float* __restrict__ e, const float* __restrict__ cE,
    const float* __restrict__ cN, const float* __restrict__ cD, const int64_t* __restrict__ sysOff,
    const float* __restrict__ den, int n, int h, int wd, float* __restrict__ gy, double* partial) {
  const int64_t hw = static_cast<int64_t>(h) * wd, total = n * hw;
  double acc[1] = {0.0};
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int s = static_cast<int>(p / hw);
    const int64_t c = p - s * hw;
    const int x = static_cast<int>(c % wd), yy = static_cast<int>(c / wd);
    const float* ce = cE + sysOff[s];
    const float* cn = cN + sysOff[s];
    const float d = y[p] - e[p];
    float ad = cD != nullptr ? cD[sysOff[s] + c] * d : d;
    if (x + 1 < wd) ad -= ce[c] * (y[p + 1] - e[p + 1]);
    if (x > 0) ad -= ce[c - 1] * (y[p - 1] - e[p - 1]);
    if (yy + 1 < h) ad -= cn[c] * (y[p + wd] - e[p + wd]);
    if (yy > 0) ad -= cn[c - wd] * (y[p - wd] - e[p - wd]);
    const float inv = 1.f / (den[s] * static_cast<float>(n));
    gy[p] = 2.f * ad * inv;
    acc[0] += static_cast<double>(d) * ad * inv;
  }
  blockPartials<1>(acc, partial);
}

__global__ void sumDoubleKernel(const double* __restrict__ partial, int parts, double* out) {
  double v[1];
  sumPartials<1>(partial, parts, v);
  if (threadIdx.x == 0) out[0] = v[0];
}

// y = B_s x on every image s of the batch: diagonal cD (unit when null),
// couplings cE to x + 1 and cN to y + 1, all at sysOff[s] (the loss operator
// of anormLossKernel).
__global__ __launch_bounds__(256) void opApplyKernel(
    const float* __restrict__ x, const float* __restrict__ cE, const float* __restrict__ cN,
    const float* __restrict__ cD, const int64_t* __restrict__ sysOff, int n, int h, int wd,
    float* __restrict__ y) {
  const int64_t hw = static_cast<int64_t>(h) * wd, total = n * hw;
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int s = static_cast<int>(p / hw);
    const int64_t c = p - s * hw;
    const int xx = static_cast<int>(c % wd), yy = static_cast<int>(c / wd);
    const float* ce = cE + sysOff[s];
    const float* cn = cN + sysOff[s];
    float a = cD != nullptr ? cD[sysOff[s] + c] * x[p] : x[p];
    if (xx + 1 < wd) a -= ce[c] * x[p + 1];
    if (xx > 0) a -= ce[c - 1] * x[p - 1];
    if (yy + 1 < h) a -= cn[c] * x[p + wd];
    if (yy > 0) a -= cn[c - wd] * x[p - wd];
    y[p] = a;
  }
}

// out = a - f[s] b per image s (f null: 1).
__global__ __launch_bounds__(256) void subScaledKernel(const float* __restrict__ a,
                                                       const float* __restrict__ b,
                                                       const float* __restrict__ f, int64_t hw,
                                                       int64_t total, float* __restrict__ out) {
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[p] = a[p] - (f != nullptr ? f[p / hw] : 1.f) * b[p];
}

// x = f[s] x per image s.
__global__ __launch_bounds__(256) void scaleImagesKernel(float* __restrict__ x,
                                                         const float* __restrict__ f, int64_t hw,
                                                         int64_t total) {
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[p] *= f[p / hw];
}

// out[s] = (a_s, b_s) in FP64, one block per image, a fixed summation order.
__global__ __launch_bounds__(256) void imageDotKernel(const float* __restrict__ a,
                                                      const float* __restrict__ b, int64_t hw,
                                                      double* __restrict__ out) {
  const int64_t base = blockIdx.x * hw;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < hw; i += 256)
    acc += static_cast<double>(a[base + i]) * static_cast<double>(b[base + i]);
  __shared__ double lds[256];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) lds[threadIdx.x] += lds[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = lds[0];
}

__global__ void adamKernel(int64_t count, float* __restrict__ w, float* __restrict__ m,
                           float* __restrict__ v, const float* __restrict__ g, float lr, float b1,
                           float b2, float eps, float bc1, float bc2) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

}  // namespace

// ---------------------------------------------------------------------------

Unet::Unet(hipStream_t s) : stream(s) {
  master.ensure(kUnetParams);
  grad.ensure(kUnetParams);
  int64_t fo = 0, bo = 0;
  for (int l = 0; l < kNumLayers; ++l) {
    fwdOff[l] = fo;
    bwdOff[l] = bo;
    if (l != kE0a && l != kOut) {
      fo += packedElems(kLayers[l].cin, kLayers[l].cout);
      bo += packedElems(kLayers[l].cout, kLayers[l].cin);
    }
  }
  fwdPack.ensure(fo);
  bwdPack.ensure(bo);
}

void Unet::setWeights(const float* host) {
  mirrorSymmetric = true;
  for (int l = 0; l < kNumLayers && mirrorSymmetric; ++l) {
    const float* w = host + layerOffset(l);
    const int64_t pair = static_cast<int64_t>(kLayers[l].cin) * kLayers[l].cout;
    for (int tap = 0; tap < 9 && mirrorSymmetric; ++tap) {
      const int ky = tap / 3, kx = tap % 3;
      const int flipY = (2 - ky) * 3 + kx, flipX = ky * 3 + (2 - kx);
      for (int64_t i = 0; i < pair; ++i)
        if (w[tap * pair + i] != w[flipY * pair + i] || w[tap * pair + i] != w[flipX * pair + i]) {
          mirrorSymmetric = false;
          break;
        }
    }
  }
  CFDHIP_CHECK(hipMemcpyAsync(master.get(), host, kUnetParams * sizeof(float),
                              hipMemcpyHostToDevice, stream));
  repack();
  hostMaster.assign(host, host + kUnetParams);
  buildNarrow();
  loaded = true;
}

void Unet::getWeights(float* host) {
  CFDHIP_CHECK(hipMemcpyAsync(host, master.get(), kUnetParams * sizeof(float),
                              hipMemcpyDeviceToHost, stream));
  CFDHIP_CHECK(hipStreamSynchronize(stream));
}

void Unet::repack() {
  for (int l = 0; l < kNumLayers; ++l) {
    if (l == kE0a || l == kOut) continue;
    const int ci = kLayers[l].cin, co = kLayers[l].cout;
    const int64_t nf = packedElems(ci, co), nb = packedElems(co, ci);
    hipLaunchKernelGGL(packKernel, grid1(nf), dim3(256), 0, stream, master.get() + layerOffset(l),
                       fwdPack.get() + fwdOff[l], ci, co, nf, 0);
    hipLaunchKernelGGL(packKernel, grid1(nb), dim3(256), 0, stream, master.get() + layerOffset(l),
                       bwdPack.get() + bwdOff[l], ci, co, nb, 1);
  }
  CFDHIP_CHECK(hipGetLastError());
}

void Unet::ensureShape(int n, int h, int w, bool training) {
  if (h % 8 != 0 || w % 8 != 0 || h < 8 || w < 8 || n < 1)
    throw Error(CFDHIP_ERR_INVALID, "network grid must be a positive multiple of 8");
  if (n == shapeN && h == shapeH && w == shapeW && (training <= shapeTraining)) return;
  const int64_t p0 = static_cast<int64_t>(n) * h * w, p1 = p0 / 4, p2 = p0 / 16, p3 = p0 / 64;
  // forward tensors (elements), then gradient tensors
  const int64_t sizes[kNumTensors] = {
      p0 * 16, p0 * 16, p1 * 16,            // e0a e0 p0
      p1 * 32, p1 * 32, p2 * 32,            // e1a e1 p1
      p2 * 64, p2 * 64, p3 * 64,            // e2a e2 p2
      p3 * 64, p3 * 64,                     // b3a b3
      p2 * 64, p2 * 64, p1 * 32, p1 * 32,   // d2a d2 d1a d1
      p0 * 16, p0 * 16,                     // d0a d0
      // gradients
      p0 * 16, p0 * 16, p1 * 32, p0 * 16,   // gd0 gd0a gd1 ge0s
      p1 * 32, p2 * 64, p1 * 32,            // gd1a gd2 ge1s
      p2 * 64, p3 * 64, p2 * 64,            // gd2a gb3 ge2s
      p3 * 64, p3 * 64,                     // gb3a gp2
      p2 * 64, p2 * 32,                     // ge2a gp1
      p1 * 32, p1 * 16,                     // ge1a gp0
      p0 * 16};                             // ge0a
  const int count = training ? kNumTensors : kNumForward;
  int64_t off = 0;
  for (int t = 0; t < kNumTensors; ++t) {
    tensorOff[t] = off;
    if (t < count) off += (sizes[t] + 127) / 128 * 128;
  }
  act.ensure(off);
  y.ensure(p0);
  gy.ensure(p0);
  shapeN = n;
  shapeH = h;
  shapeW = w;
  shapeTraining = training;
}

int64_t Unet::tensorElems(int t) const {
  require(t >= 0 && t < kNumTensors && (t < kNumForward || shapeTraining),
          "tensor id out of range for the current shape");
  const int64_t p0 = static_cast<int64_t>(shapeN) * shapeH * shapeW;
  static const int kLevel[kNumTensors] = {0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0,
                                          0, 0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 3, 2, 2, 1, 1, 0};
  static const int kChan[kNumTensors] = {16, 16, 16, 32, 32, 32, 64, 64, 64, 64, 64, 64, 64, 32,
                                         32, 16, 16, 16, 16, 32, 16, 32, 64, 32, 64, 64, 64, 64,
                                         64, 64, 32, 32, 16, 16};
  return (p0 >> (2 * kLevel[t])) * kChan[t];
}

void Unet::readTensor(int t, float* host, int64_t count) {
  const int64_t m = tensorElems(t);
  require(count == m, "tensor size mismatch");
  CFDHIP_CHECK(hipMemcpyAsync(host, act.get() + tensorOff[t], m * sizeof(float),
                              hipMemcpyDeviceToHost, stream));
  CFDHIP_CHECK(hipStreamSynchronize(stream));
}

void Unet::setTiming(bool on) {
  timing = on;
  for (int l = 0; l <= kNumLayers; ++l) {
    layerMs[l] = 0.0;
    layerCalls[l] = 0;
  }
}

void Unet::mark(int layer, bool begin) {
  if (!timing) return;
  hipEvent_t e;
  CFDHIP_CHECK(hipEventCreate(&e));
  CFDHIP_CHECK(hipEventRecord(e, stream));
  events.push_back(e);
  eventLayer.push_back(begin ? layer : -1);
}

void Unet::collectTiming() {
  if (!timing || events.empty()) return;
  CFDHIP_CHECK(hipEventSynchronize(events.back()));
  for (size_t i = 0; i + 1 < events.size(); i += 2) {
    float ms = 0.f;
    CFDHIP_CHECK(hipEventElapsedTime(&ms, events[i], events[i + 1]));
    layerMs[eventLayer[i]] += ms;
    layerCalls[eventLayer[i]] += 1;
  }
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  events.clear();
  eventLayer.clear();
}

void Unet::conv(int layer, bool transposed, int gather, int epi, int split, const float* a,
                const float* b, float* out, float* aux, int n, int h, int w) {
  ConvArgs args{};
  args.a = a;
  args.b = b;
  args.out = out;
  args.aux = aux;
  args.n = n;
  args.h = h;
  args.w_ = w;
  args.tilesX = (w + 31) / 32;
  args.tilesY = (h + 7) / 8;
  args.pad = padBits();
  // the physical image on this level: the level-0 extent halved (rounding
  // up) once per level below the level-0 grid
  args.ph = physLevelH > 0 ? physH : h;
  args.pw = physLevelW > 0 ? physW : w;
  for (int hh = physLevelH; hh > h; hh >>= 1) args.ph = (args.ph + 1) >> 1;
  for (int ww = physLevelW; ww > w; ww >>= 1) args.pw = (args.pw + 1) >> 1;
  const int ci = transposed ? kLayers[layer].cout : kLayers[layer].cin;
  const int co = transposed ? kLayers[layer].cin : kLayers[layer].cout;
  args.w = transposed ? static_cast<const void*>(bwdPack.get() + bwdOff[layer])
                      : static_cast<const void*>(fwdPack.get() + fwdOff[layer]);
  args.nbTotal = co / 16;
  int ca = ci;
  if (gather == kGatherConcatUp) ca = ci - (layer == kD2a ? 64 : layer == kD1a ? 32 : 16);
  if (gather == kGatherFirst) {
    args.src = first.src;
    args.scale = first.scale;
    args.w0 = master.get() + layerOffset(kE0a);
    args.srcH = first.srcH;
    args.srcW = first.srcW;
    args.srcDouble = first.srcDouble;
  }
  if (!transposed) mark(layer, true);
  launchConv(stream, ci, ca, co, gather, epi, split, args);
  if (!transposed) mark(layer, false);
  CFDHIP_CHECK(hipGetLastError());
}

void Unet::convDw(int layer, int gx, int gg, const float* xa, const float* xb, const float* ga,
                  const float* gb, int n, int h, int w) {
  DwArgs a{};
  a.xa = xa;
  a.xb = xb;
  a.ga = ga;
  a.gb = gb;
  a.n = n;
  a.h = h;
  a.w_ = w;
  a.tilesX = (w + 31) / 32;
  a.tilesY = (h + 7) / 8;
  a.tiles = static_cast<int64_t>(a.tilesX) * a.tilesY * n;
  a.pad = padBits();
  const int ci = kLayers[layer].cin, co = kLayers[layer].cout;
  const int npair = 9 * (ci / 16);
  const int groups = (npair + 7) / 8;
  int splits = static_cast<int>(std::min<int64_t>(a.tiles, std::max(1, 2048 / groups)));
  const int64_t need = static_cast<int64_t>(splits) * npair * (co / 16) * 256;
  dwPartial.ensure(need);
  a.partial = dwPartial.get();
  int ca = ci;
  if (gx == kGatherConcatUp) ca = ci - (layer == kD2a ? 64 : layer == kD1a ? 32 : 16);
  launchConvDw(stream, ci, ca, co, gx, gg, a, splits);
  hipLaunchKernelGGL(dwReduceKernel, grid1(9LL * ci * co), dim3(256), 0, stream, dwPartial.get(),
                     splits, ci, co, grad.get() + layerOffset(layer));
  CFDHIP_CHECK(hipGetLastError());
}

void Unet::forward(const float* src, const float* scale, int srcH, int srcW, int64_t srcBatch,
                   int n, int h, int w) {
  forwardFrom(src, false, scale, srcH, srcW, srcBatch, n, h, w);
}

void Unet::forwardFrom(const void* src, bool srcDouble, const float* scale, int srcH, int srcW,
                       int64_t srcBatch, int n, int h, int w, bool fuseFirst) {
  if (!loaded) throw Error(CFDHIP_ERR_INVALID, "network weights are not loaded");
  const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
  // a source smaller than the padded grid is the physical image
  physH = std::min(srcH, h);
  physW = std::min(srcW, w);
  physLevelH = h;
  physLevelW = w;
  lastNarrow = false;
  if (fuseFirst && narrowActive()) {
    // inference with unused channels dropped (unet_narrow.hip)
    forwardNarrow(src, srcDouble, scale, srcH, srcW, srcBatch, n, h, w);
    return;
  }
  const int64_t p0 = static_cast<int64_t>(n) * h * w;
  if (fuseFirst) {
    // inference: e0a is computed inside e0b's tile staging and never stored
    require(n == 1 || srcBatch == static_cast<int64_t>(srcH) * srcW,
            "the fused first layer reads a dense batch");
    first = FirstSource{src, scale, srcH, srcW, srcDouble ? 1 : 0};
    conv(kE0b, false, kGatherFirst, kEpiStorePool, 0, nullptr, nullptr, T(kTe0), T(kTp0), n, h, w);
  } else {
    if (srcDouble)
      hipLaunchKernelGGL((firstConvKernel<double>), grid1(p0), dim3(256), 0, stream,
                         static_cast<const double*>(src), scale, srcH, srcW, srcBatch,
                         master.get() + layerOffset(kE0a), T(kTe0a), n, h, w, padBits());
    else
      hipLaunchKernelGGL((firstConvKernel<float>), grid1(p0), dim3(256), 0, stream,
                         static_cast<const float*>(src), scale, srcH, srcW, srcBatch,
                         master.get() + layerOffset(kE0a), T(kTe0a), n, h, w, padBits());
    CFDHIP_CHECK(hipGetLastError());
    conv(kE0b, false, kGatherPlain, kEpiStorePool, 0, T(kTe0a), nullptr, T(kTe0), T(kTp0), n, h,
         w);
  }
  conv(kE1a, false, kGatherPlain, kEpiStore, 0, T(kTp0), nullptr, T(kTe1a), nullptr, n, h1, w1);
  conv(kE1b, false, kGatherPlain, kEpiStorePool, 0, T(kTe1a), nullptr, T(kTe1), T(kTp1), n, h1, w1);
  conv(kE2a, false, kGatherPlain, kEpiStore, 0, T(kTp1), nullptr, T(kTe2a), nullptr, n, h2, w2);
  conv(kE2b, false, kGatherPlain, kEpiStorePool, 0, T(kTe2a), nullptr, T(kTe2), T(kTp2), n, h2, w2);
  conv(kB3a, false, kGatherPlain, kEpiStore, 0, T(kTp2), nullptr, T(kTb3a), nullptr, n, h3, w3);
  if (tail.levels > 0) {
    mark(kNumLayers, true);
    tailApply(T(kTb3a), n, h3, w3, false);
    mark(kNumLayers, false);
  }
  conv(kB3b, false, kGatherPlain, kEpiStore, 0, T(kTb3a), nullptr, T(kTb3), nullptr, n, h3, w3);
  conv(kD2a, false, kGatherConcatUp, kEpiStore, 0, T(kTb3), T(kTe2), T(kTd2a), nullptr, n, h2, w2);
  conv(kD2b, false, kGatherPlain, kEpiStore, 0, T(kTd2a), nullptr, T(kTd2), nullptr, n, h2, w2);
  conv(kD1a, false, kGatherConcatUp, kEpiStore, 0, T(kTd2), T(kTe1), T(kTd1a), nullptr, n, h1, w1);
  conv(kD1b, false, kGatherPlain, kEpiStore, 0, T(kTd1a), nullptr, T(kTd1), nullptr, n, h1, w1);
  conv(kD0a, false, kGatherConcatUp, kEpiStore, 0, T(kTd1), T(kTe0), T(kTd0a), nullptr, n, h, w);
  conv(kD0b, false, kGatherPlain, kEpiStore, 0, T(kTd0a), nullptr, T(kTd0), nullptr, n, h, w);
}

void Unet::outputTrain(int n, int h, int w, int ph, int pw) {
  if (ph <= 0) ph = h;
  if (pw <= 0) pw = w;
  if (lastNarrow) {
    outputNarrow(n, h, w, ph, pw);
    return;
  }
  const int64_t p0 = static_cast<int64_t>(n) * h * w;
  hipLaunchKernelGGL(lastConvKernel, grid1(p0 / 2, 4096), dim3(256), 0, stream, T(kTd0),
                     master.get() + layerOffset(kOut), y.get(), n, h, w, ph, pw, padBits());
  CFDHIP_CHECK(hipGetLastError());
}

template <class TZ>
void Unet::outputApply(const float* scale, const TZ* r, const TZ* q, TZ* z, int ny, int nx,
                       int64_t pin, int h, int w, double* partial, int parts) {
  mark(kOut, true);
  if (lastNarrow)
    outputApplyNarrow<TZ>(scale, r, q, z, ny, nx, pin, h, w, partial, parts);
  else
    hipLaunchKernelGGL((lastConvApplyKernel<TZ>), dim3(static_cast<unsigned>(parts)), dim3(256), 0,
                       stream, T(kTd0), master.get() + layerOffset(kOut), scale, r, q, z, ny, nx,
                       pin, h, w, padBits(), partial);
  mark(kOut, false);
  CFDHIP_CHECK(hipGetLastError());
  collectTiming();
}
template void Unet::outputApply<float>(const float*, const float*, const float*, float*, int, int,
                                       int64_t, int, int, double*, int);
template void Unet::outputApply<double>(const float*, const double*, const double*, double*, int,
                                        int, int64_t, int, int, double*, int);

void Unet::backward(const float* x, const float* gyIn, int n, int h, int w) {
  // training images fill the padded grid: the whole grid is the image
  physH = physLevelH = h;
  physW = physLevelW = w;
  require(padBits() == 0 || mirrorSymmetric,
          "Neumann padding: the backward pass needs kernels mirror-symmetric in both axes "
          "(W[ky][kx] == W[2-ky][kx] == W[ky][2-kx]); load such weights with "
          "cfdhip_unet_set_weights");
  const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
  const int64_t p0 = static_cast<int64_t>(n) * h * w;
  // output layer
  hipLaunchKernelGGL(lastConvBwdKernel, grid1(p0), dim3(256), 0, stream, gyIn,
                     master.get() + layerOffset(kOut), T(kTgd0), n, h, w, padBits());
  singleDw(1, gyIn, T(kTd0), n, h, w, grad.get() + layerOffset(kOut));
  // decoder, level 0
  conv(kD0b, true, kGatherPlain, kEpiStore, 0, T(kTgd0), nullptr, T(kTgd0a), nullptr, n, h, w);
  convDw(kD0b, kGatherPlain, kGatherPlain, T(kTd0a), nullptr, T(kTgd0), nullptr, n, h, w);
  conv(kD0a, true, kGatherPlain, kEpiSplitUp, 32, T(kTgd0a), nullptr, T(kTge0s), T(kTgd1), n, h, w);
  convDw(kD0a, kGatherConcatUp, kGatherPlain, T(kTd1), T(kTe0), T(kTgd0a), nullptr, n, h, w);
  // level 1
  conv(kD1b, true, kGatherPlain, kEpiStore, 0, T(kTgd1), nullptr, T(kTgd1a), nullptr, n, h1, w1);
  convDw(kD1b, kGatherPlain, kGatherPlain, T(kTd1a), nullptr, T(kTgd1), nullptr, n, h1, w1);
  conv(kD1a, true, kGatherPlain, kEpiSplitUp, 64, T(kTgd1a), nullptr, T(kTge1s), T(kTgd2), n, h1, w1);
  convDw(kD1a, kGatherConcatUp, kGatherPlain, T(kTd2), T(kTe1), T(kTgd1a), nullptr, n, h1, w1);
  // level 2
  conv(kD2b, true, kGatherPlain, kEpiStore, 0, T(kTgd2), nullptr, T(kTgd2a), nullptr, n, h2, w2);
  convDw(kD2b, kGatherPlain, kGatherPlain, T(kTd2a), nullptr, T(kTgd2), nullptr, n, h2, w2);
  conv(kD2a, true, kGatherPlain, kEpiSplitUp, 64, T(kTgd2a), nullptr, T(kTge2s), T(kTgb3), n, h2, w2);
  convDw(kD2a, kGatherConcatUp, kGatherPlain, T(kTb3), T(kTe2), T(kTgd2a), nullptr, n, h2, w2);
  // bottom
  conv(kB3b, true, kGatherPlain, kEpiStore, 0, T(kTgb3), nullptr, T(kTgb3a), nullptr, n, h3, w3);
  convDw(kB3b, kGatherPlain, kGatherPlain, T(kTb3a), nullptr, T(kTgb3), nullptr, n, h3, w3);
  tailApply(T(kTgb3a), n, h3, w3, true);
  conv(kB3a, true, kGatherPlain, kEpiStore, 0, T(kTgb3a), nullptr, T(kTgp2), nullptr, n, h3, w3);
  convDw(kB3a, kGatherPlain, kGatherPlain, T(kTp2), nullptr, T(kTgb3a), nullptr, n, h3, w3);
  // encoder, level 2: d e2 = skip + pool gradient
  conv(kE2b, true, kGatherSumUp, kEpiStore, 0, T(kTge2s), T(kTgp2), T(kTge2a), nullptr, n, h2, w2);
  convDw(kE2b, kGatherPlain, kGatherSumUp, T(kTe2a), nullptr, T(kTge2s), T(kTgp2), n, h2, w2);
  conv(kE2a, true, kGatherPlain, kEpiStore, 0, T(kTge2a), nullptr, T(kTgp1), nullptr, n, h2, w2);
  convDw(kE2a, kGatherPlain, kGatherPlain, T(kTp1), nullptr, T(kTge2a), nullptr, n, h2, w2);
  // level 1
  conv(kE1b, true, kGatherSumUp, kEpiStore, 0, T(kTge1s), T(kTgp1), T(kTge1a), nullptr, n, h1, w1);
  convDw(kE1b, kGatherPlain, kGatherSumUp, T(kTe1a), nullptr, T(kTge1s), T(kTgp1), n, h1, w1);
  conv(kE1a, true, kGatherPlain, kEpiStore, 0, T(kTge1a), nullptr, T(kTgp0), nullptr, n, h1, w1);
  convDw(kE1a, kGatherPlain, kGatherPlain, T(kTp0), nullptr, T(kTge1a), nullptr, n, h1, w1);
  // level 0
  conv(kE0b, true, kGatherSumUp, kEpiStore, 0, T(kTge0s), T(kTgp0), T(kTge0a), nullptr, n, h, w);
  convDw(kE0b, kGatherPlain, kGatherSumUp, T(kTe0a), nullptr, T(kTge0s), T(kTgp0), n, h, w);
  singleDw(0, x, T(kTge0a), n, h, w, grad.get() + layerOffset(kE0a));
}

void Unet::singleDw(int mode, const float* s1, const float* m16, int n, int h, int w, float* out) {
  const int64_t p0 = static_cast<int64_t>(n) * h * w;
  const dim3 g = grid1(p0, 1024);
  dwPartial.ensure(static_cast<int64_t>(g.x) * 144);
  hipLaunchKernelGGL(singleDwKernel, g, dim3(256), 0, stream, s1, m16, mode, n, h, w, padBits(),
                     dwPartial.get());
  hipLaunchKernelGGL(singleDwReduceKernel, dim3(1), dim3(256), 0, stream, dwPartial.get(),
                     static_cast<int>(g.x), out);
  CFDHIP_CHECK(hipGetLastError());
}

double Unet::loss(const float* e, const float* cE, const float* cN, const float* cD,
                  const int64_t* sysOffDev,
                  const float* denDev, int n, int h, int w) {
  const int64_t p0 = static_cast<int64_t>(n) * h * w;
  const dim3 g = grid1(p0, 2048);
  lossPartial.ensure(static_cast<int64_t>(g.x) + 8);
  hipLaunchKernelGGL(anormLossKernel, g, dim3(256), 0, stream, y.get(), e, cE, cN, cD, sysOffDev,
                     denDev, n, h, w, gy.get(), lossPartial.get());
  double* out = lossPartial.get() + g.x;
  hipLaunchKernelGGL(sumDoubleKernel, dim3(1), dim3(kFinThreads), 0, stream, lossPartial.get(),
                     static_cast<int>(g.x), out);
  CFDHIP_CHECK(hipGetLastError());
  double v = 0.0;
  CFDHIP_CHECK(hipMemcpyAsync(&v, out, sizeof(double), hipMemcpyDeviceToHost, stream));
  CFDHIP_CHECK(hipStreamSynchronize(stream));
  return v;
}

double Unet::spectral(const float* z, const float* cE, const float* cN, const float* cD,
                      const int64_t* sysOffDev, int n, int h, int w, int steps, float* gradHost) {
  const int64_t hw = static_cast<int64_t>(h) * w, P = n * hw;
  const int k = steps;
  spec.ensure((2LL * (k + 1) + 2) * P);
  specF.ensure(n);
  specDots.ensure(n);
  float* V = spec.get();                          // v_0 .. v_k
  float* BV = V + static_cast<int64_t>(k + 1) * P;  // B v_0 .. B v_k
  float* G = BV + static_cast<int64_t>(k + 1) * P;  // dL/dv_j
  float* Tm = G + P;
  const dim3 g = grid1(P);
  auto op = [&](const float* x, float* out) {
    hipLaunchKernelGGL(opApplyKernel, g, dim3(256), 0, stream, x, cE, cN, cD, sysOffDev, n, h, w,
                       out);
    CFDHIP_CHECK(hipGetLastError());
  };
  std::vector<double> dots(n);
  auto imageDots = [&](const float* a, const float* b) {
    hipLaunchKernelGGL(imageDotKernel, dim3(static_cast<unsigned>(n)), dim3(256), 0, stream, a, b,
                       hw, specDots.get());
    CFDHIP_CHECK(hipGetLastError());
    CFDHIP_CHECK(hipMemcpyAsync(dots.data(), specDots.get(), sizeof(double) * n,
                                hipMemcpyDeviceToHost, stream));
    CFDHIP_CHECK(hipStreamSynchronize(stream));
  };
  std::vector<float> f(n);
  auto setFactors = [&]() {
    CFDHIP_CHECK(hipMemcpyAsync(specF.get(), f.data(), sizeof(float) * n, hipMemcpyHostToDevice,
                                stream));
    CFDHIP_CHECK(hipStreamSynchronize(stream));
  };
  // y = N(x) on the training path (every layer stored, nothing fused)
  auto apply = [&](const float* x) {
    forward(x, nullptr, h, w, hw, n, h, w);
    outputTrain(n, h, w);
  };
  // v_0 = z / |z|_B
  CFDHIP_CHECK(hipMemcpyAsync(V, z, sizeof(float) * P, hipMemcpyDeviceToDevice, stream));
  op(V, BV);
  imageDots(V, BV);
  for (int s = 0; s < n; ++s) {
    if (!(dots[s] > 0.0) || !std::isfinite(dots[s]))
      throw Error(CFDHIP_ERR_INVALID,
                  "spectral loss: an error image has no positive energy in its operator");
    f[s] = static_cast<float>(1.0 / std::sqrt(dots[s]));
  }
  setFactors();
  hipLaunchKernelGGL(scaleImagesKernel, g, dim3(256), 0, stream, V, specF.get(), hw, P);
  hipLaunchKernelGGL(scaleImagesKernel, g, dim3(256), 0, stream, BV, specF.get(), hw, P);
  CFDHIP_CHECK(hipGetLastError());
  // v_j = v_{j-1} - N(B v_{j-1})
  for (int j = 1; j <= k; ++j) {
    apply(BV + (j - 1) * P);
    hipLaunchKernelGGL(subScaledKernel, g, dim3(256), 0, stream, V + (j - 1) * P, y.get(),
                       nullptr, hw, P, V + j * P);
    CFDHIP_CHECK(hipGetLastError());
    op(V + j * P, BV + j * P);
  }
  imageDots(V + k * P, BV + k * P);
  std::vector<double> raw(dots);
  double loss = 0.0;
  for (int s = 0; s < n; ++s) {
    if (!(raw[s] > 0.0) || !std::isfinite(raw[s]))
      throw Error(CFDHIP_ERR_NONFINITE, "spectral loss: the iterated error vanished or overflowed");
    loss += std::log(raw[s]);
  }
  loss /= 2.0 * k * n;
  if (gradHost == nullptr) return loss;
  // L = mean_s log(|v_k|_B^2) / 2k: dL/dv_k = 2 B v_k / (2k n |v_k|_B^2)
  for (int s = 0; s < n; ++s) f[s] = static_cast<float>(1.0 / (k * n * raw[s]));
  setFactors();
  CFDHIP_CHECK(hipMemcpyAsync(G, BV + k * P, sizeof(float) * P, hipMemcpyDeviceToDevice, stream));
  hipLaunchKernelGGL(scaleImagesKernel, g, dim3(256), 0, stream, G, specF.get(), hw, P);
  CFDHIP_CHECK(hipGetLastError());
  std::vector<double> acc(kUnetParams, 0.0);
  std::vector<float> part(kUnetParams);
  for (int j = k; j >= 1; --j) {
    // dL/dW -= d/dW (G, N(B v_{j-1}))
    forward(BV + (j - 1) * P, nullptr, h, w, hw, n, h, w);
    backward(BV + (j - 1) * P, G, n, h, w);
    getGrads(part.data());
    for (int64_t i = 0; i < kUnetParams; ++i) acc[i] -= part[i];
    if (j > 1) {
      // dL/dv_{j-1} = G - B N^T G, with N^T = N (a self-adjoint network)
      apply(G);
      op(y.get(), Tm);
      hipLaunchKernelGGL(subScaledKernel, g, dim3(256), 0, stream, G, Tm, nullptr, hw, P, G);
      CFDHIP_CHECK(hipGetLastError());
    }
  }
  for (int64_t i = 0; i < kUnetParams; ++i) gradHost[i] = static_cast<float>(acc[i]);
  return loss;
}

void Unet::adam(float lr, float b1, float b2, float eps) {
  if (adamM.count < kUnetParams) {
    adamM.ensure(kUnetParams);
    adamV.ensure(kUnetParams);
    CFDHIP_CHECK(hipMemsetAsync(adamM.get(), 0, kUnetParams * sizeof(float), stream));
    CFDHIP_CHECK(hipMemsetAsync(adamV.get(), 0, kUnetParams * sizeof(float), stream));
    adamStep = 0;
  }
  ++adamStep;
  const float bc1 = 1.f - std::pow(b1, static_cast<float>(adamStep));
  const float bc2 = 1.f - std::pow(b2, static_cast<float>(adamStep));
  mirrorSymmetric = false;  // the update is not constrained to mirror-symmetric kernels
  hipLaunchKernelGGL(adamKernel, grid1(kUnetParams), dim3(256), 0, stream,
                     static_cast<int64_t>(kUnetParams), master.get(), adamM.get(), adamV.get(),
                     grad.get(), lr, b1, b2, eps, bc1, bc2);
  CFDHIP_CHECK(hipGetLastError());
  repack();
}

}  // namespace cfdhip
