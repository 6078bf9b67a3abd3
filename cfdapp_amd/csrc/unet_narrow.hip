// Channel-compacted inference of the learned preconditioner.
//
// A network whose weights leave most channels unused (the shipped weights
// are a multigrid V-cycle written as a U-Net: at most two live channels per
// tensor) wastes nearly all of the MFMA layers' HBM traffic on zeros: a
// 16-channel FP32 pixel is a whole 64-byte line, so skipping zero channels
// inside the full-width layout saves no bytes. buildNarrow() therefore finds,
// from the weights and the tail options, the channels of every tensor that
// some output depends on, and when no tensor keeps more than kNarrowMax of
// them, inference stores only those (2 or 4 channels per pixel) and
// runs these kernels:
//  * narrowConvKernel: a 3x3 convolution over a 16 x 64 output tile, input
//    staged once in LDS with its one-pixel halo (plain, the nearest-upsampled
//    concatenation of the decoder, or the fused first layer computed from
//    the scaled solver residual), FP32 activations, BF16-rounded weights,
//    FP32 fused multiply-add accumulation; one thread computes a 2 x 2
//    pixel block, so the encoder's 2 x 2 average pool is thread-local in the
//    epilogue;
//  * narrowLastKernel: the output layer, on the solver grid, with the
//    preconditioner's scaling, pinned row and the (z, q) block partials.
// Padding, physical extents, pooling at odd edges and the coarse tail are the
// MFMA path's (unet_conv.hip, unet_tail.hip); the two paths compute the same
// linear map and differ only in rounding (the MFMA layers see activations
// as a hi + lo BF16 pair, these kernels in FP32).
#include <algorithm>
#include <atomic>
#include <cstring>

#include "reduce.hpp"
#include "unet_net.hpp"

namespace cfdhip {

namespace {

constexpr int kNTH = kNarrowTileH, kNTW = kNarrowTileW;  // output tile
constexpr int kNHR = kNTH + 2, kNHC = kNTW + 2;      // its input with the halo
constexpr int kNHP = kNHR * kNHC;
constexpr int kNXR = kNTH + 4, kNXC = kNTW + 4;      // the fused first layer's scalar input
constexpr int kNBX = kNTW / 2;                       // 2 x 2 blocks per tile row

enum NarrowGather : int { kNPlain = 0, kNConcat = 1, kNFirst = 2 };

struct NarrowArgs {
  const float* a;   // plain: the input (h x w); concat: the coarse half (h/2 x w/2)
  const float* b;   // concat: the skip half (h x w)
  const float* w;   // this layer's weights [tap][slot in][slot out]
  const float* w0;  // first layer: e0a's weights [tap][slot]
  float* out;       // n x h x w x CO
  float* aux;       // pool: n x h/2 x w/2 x CO
  int n, h, w_, ph, pw, pad;
  int tilesX, tilesY;
  // first layer: x = src * scale on srcH x srcW (row stride srcW)
  const void* src;
  const float* scale;
  int srcH, srcW;
  int64_t srcBatch;
  // batched cases of uneven size (cases != null): block entry i of tab is
  // (case, tile of the case at this level); a, b, out, aux are the
  // activation arena and tA, tB, tOut, tAux the tensors read and written
  // (their per-case offsets in NarrowCase::tOff); src and scale are the
  // batch's solver vectors (per-case offset NarrowCase::cells)
  const NarrowCase* cases;
  const int2* tab;
  int64_t tabCount;
  int level, tA, tB, tOut, tAux;
};

// One tile's image: its tensors, extents and padding.
struct NarrowTile {
  const float* a;
  const float* b;
  float* out;
  float* aux;
  const void* src;
  const float* scale;
  int h, w, ph, pw, pad, srcH, srcW, x0, y0;
  int64_t srcRow0;  // src / scale index of the image's first pixel
};

// Tile tb of the launch (after the XCD-aware reordering): the image it
// belongs to and where the tile sits in it.
template <int CA, int CB, int CO, int G>
__device__ inline NarrowTile narrowTile(const NarrowArgs& a, int64_t t) {
  NarrowTile q;
  if (a.cases != nullptr) {
    const int2 e = a.tab[t];
    const NarrowCase& c = a.cases[e.x];
    const int L = a.level;
    q.h = c.h[L];
    q.w = c.w[L];
    q.ph = c.ph[L];
    q.pw = c.pw[L];
    q.pad = c.pad;
    q.x0 = (e.y % c.tilesX[L]) * kNTW;
    q.y0 = (e.y / c.tilesX[L]) * kNTH;
    q.a = a.a != nullptr && a.tA >= 0 ? a.a + c.act + c.tOff[a.tA] : nullptr;
    q.b = a.b != nullptr && a.tB >= 0 ? a.b + c.act + c.tOff[a.tB] : nullptr;
    q.out = a.out + c.act + c.tOff[a.tOut];
    q.aux = a.aux != nullptr && a.tAux >= 0 ? a.aux + c.act + c.tOff[a.tAux] : nullptr;
    q.src = a.src;
    q.scale = a.scale;
    q.srcH = c.ny;
    q.srcW = c.nx;
    q.srcRow0 = c.cells;
  } else {
    q.h = a.h;
    q.w = a.w_;
    q.ph = a.ph;
    q.pw = a.pw;
    q.pad = a.pad;
    // 32-bit decoding (the launch checks tiles < 2^31): a 64-bit division is
    // over a hundred scalar instructions per wave and tile
    const uint32_t tt = static_cast<uint32_t>(t), tx = static_cast<uint32_t>(a.tilesX),
                   ty = static_cast<uint32_t>(a.tilesY);
    const uint32_t rest = tt / tx;
    q.x0 = static_cast<int>(tt - rest * tx) * kNTW;
    const uint32_t b = rest / ty;
    q.y0 = static_cast<int>(rest - b * ty) * kNTH;
    const int64_t hw = static_cast<int64_t>(a.h) * a.w_, hw2 = hw / 4;
    q.a = a.a == nullptr ? nullptr : a.a + b * (G == kNConcat ? hw2 : hw) * CA;
    q.b = a.b == nullptr ? nullptr : a.b + b * hw * CB;
    q.out = a.out + b * hw * CO;
    q.aux = a.aux == nullptr ? nullptr : a.aux + b * hw2 * CO;
    q.src = a.src;
    q.scale = a.scale;
    q.srcH = a.srcH;
    q.srcW = a.srcW;
    q.srcRow0 = b * a.srcBatch;
  }
  return q;
}

template <int CW>
struct Vec;
template <>
struct Vec<2> {
  using T = float2;
  __device__ static inline void load(const float* p, float* v) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    v[0] = t.x;
    v[1] = t.y;
  }
  __device__ static inline void store(float* p, const float* v) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  }
};
template <>
struct Vec<4> {
  using T = float4;
  __device__ static inline void load(const float* p, float* v) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  }
  __device__ static inline void store(float* p, const float* v) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
};

// CA: stored channels of the input (concat: of its coarse half), CB: of the
// skip half (concat only), CO: of the output. G = kNFirst: the input is e0a
// (CA channels) computed in LDS from the scalar source.
// padSign (unet.hpp) without branches: (y, x) moves into the h x w image
// and the factor is returned (0 outside under zero padding), so a thread can
// issue all its loads before it waits for the first.
__device__ inline float padFactor(int& y, int& x, int h, int w, int pad) {
  const bool below = y < 0, above = y >= h, left = x < 0, right = x >= w;
  float s = 1.f;
  if (pad & kPadNeumann) {
    s = (below && (pad & kPadBottom)) ? -s : s;
    s = (above && (pad & kPadTop)) ? -s : s;
    s = (left && (pad & kPadLeft)) ? -s : s;
    s = (right && (pad & kPadRight)) ? -s : s;
  } else {
    s = (below || above || left || right) ? 0.f : 1.f;
  }
  y = min(max(y, 0), h - 1);
  x = min(max(x, 0), w - 1);
  return s;
}

// Workgroup b runs on XCD b % 8 (round-robin dispatch over the 8 XCDs, each
// with its own L2). Workgroup b takes tiles b, b + G, b + 2G, ... of a grid of
// G workgroups; with G and the tile count multiples of 8, tile t then runs on
// XCD t % 8, and mapping it to (t % 8) * (tiles / 8) + t / 8 gives XCD x the
// x-th contiguous eighth of the row-major tile order, so a tile's halo
// neighbours are mostly on its own XCD and the halo is fetched from HBM once
// per L2, not once per neighbour.
__device__ inline int64_t xcdTile(int64_t b, int64_t tiles) {
  if (tiles % 8 != 0 || gridDim.x % 8 != 0) return b;
  return (b % 8) * (tiles / 8) + b / 8;
}

// s * v with the zero of zero padding exact (a non-finite value read from the
// clamped pixel stays out of it)
__device__ inline float signed_(float s, float v) { return s == 0.f ? 0.f : s * v; }

template <int CA, int CB, int CO, int G, bool POOL, class TS>
__global__ __launch_bounds__(256) void narrowConvKernel(NarrowArgs a) {
  constexpr int CI = G == kNConcat ? CA + CB : CA;
  __shared__ float tile[kNHP * CI];
  __shared__ float ws[9 * CI * CO];
  __shared__ float w0s[9 * CA];
  __shared__ float xs[G == kNFirst ? kNXR * kNXC : 1];
  for (int i = threadIdx.x; i < 9 * CI * CO; i += blockDim.x) ws[i] = a.w[i];
  if (G == kNFirst)
    for (int i = threadIdx.x; i < 9 * CA; i += blockDim.x) w0s[i] = a.w0[i];
  const int64_t tiles =
      a.cases != nullptr ? a.tabCount : static_cast<int64_t>(a.tilesX) * a.tilesY * a.n;
  // A block walks its tiles with the next tile's global loads in flight
  // while it computes the current one: fetch() issues a tile's loads into
  // registers (the first layer: x = src * scale on rows y0 - 2 .. y0 + kNTH + 1
  // padded at the physical image; otherwise the input tile with its halo),
  // stage() writes them to LDS.
  constexpr int XPER = (kNXR * kNXC + 255) / 256, PER = (kNHP + 255) / 256;
  constexpr int NL = G == kNFirst ? XPER : PER, NC = G == kNFirst ? 1 : CI;
  float v[NL][NC], sg[NL];
  auto fetch = [&](const NarrowTile& q) {
    if constexpr (G == kNFirst) {
      const TS* src = static_cast<const TS*>(q.src);
      const int sh = q.srcH < q.h ? q.srcH : q.h, sw = q.srcW < q.w ? q.srcW : q.w;
#pragma unroll
      for (int k = 0; k < XPER; ++k) {
        const int i = threadIdx.x + k * 256;
        int yy = q.y0 - 2 + i / kNXC, xx = q.x0 - 2 + i % kNXC;
        sg[k] = padFactor(yy, xx, sh, sw, q.pad);
        v[k][0] = 0.f;
        if (i < kNXR * kNXC) {
          const int64_t si = q.srcRow0 + static_cast<int64_t>(yy) * q.srcW + xx;
          const float sv = static_cast<float>(src[si]);
          v[k][0] = q.scale != nullptr ? sv * q.scale[si] : sv;
        }
      }
    } else {
      const int wd = q.w;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = threadIdx.x + k * 256;
        int yy = q.y0 - 1 + i / kNHC, xx = q.x0 - 1 + i % kNHC;
        sg[k] = padFactor(yy, xx, q.ph, q.pw, q.pad);
#pragma unroll
        for (int c = 0; c < CI; ++c) v[k][c] = 0.f;
        if (i < kNHP) {
          if constexpr (G == kNConcat) {
            const int hw = wd >> 1;
            Vec<CA>::load(q.a + (static_cast<int64_t>(yy >> 1) * hw + (xx >> 1)) * CA, v[k]);
            Vec<CB>::load(q.b + (static_cast<int64_t>(yy) * wd + xx) * CB, v[k] + CA);
          } else {
            Vec<CA>::load(q.a + (static_cast<int64_t>(yy) * wd + xx) * CA, v[k]);
          }
        }
      }
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int i = threadIdx.x + k * 256;
      if constexpr (G == kNFirst) {
        if (i < kNXR * kNXC) xs[i] = signed_(sg[k], v[k][0]);
      } else if (i < kNHP) {
#pragma unroll
        for (int c = 0; c < CI; ++c) tile[i * CI + c] = signed_(sg[k], v[k][c]);
      }
    }
  };
  NarrowTile next;
  if (blockIdx.x < tiles) {
    next = narrowTile<CA, CB, CO, G>(a, xcdTile(blockIdx.x, tiles));
    fetch(next);
  }
  for (int64_t tb = blockIdx.x; tb < tiles; tb += gridDim.x) {
    const NarrowTile q = next;
    const int h = q.h, wd = q.w, x0 = q.x0, y0 = q.y0;
    __syncthreads();  // the weights are in; the previous tile's readers are done
    stage();
    if (tb + gridDim.x < tiles) {
      next = narrowTile<CA, CB, CO, G>(a, xcdTile(tb + gridDim.x, tiles));
      fetch(next);
    }
    if constexpr (G == kNFirst) {
      __syncthreads();
      // e0a on the halo tile; a halo pixel outside the image takes e0a of the
      // nearest pixel inside times the padding's sign (its taps lie inside xs:
      // the tile starts inside the physical image, whose rounded-up tensor
      // extends it by at most 7 pixels)
      for (int i = threadIdx.x; i < kNHP; i += blockDim.x) {
        int yy = y0 - 1 + i / kNHC, xx = x0 - 1 + i % kNHC;
        float acc[CA];
#pragma unroll
        for (int c = 0; c < CA; ++c) acc[c] = 0.f;
        const float sgn = padSign(yy, xx, q.ph, q.pw, q.pad);
        if (sgn != 0.f) {
          const int ry = yy - y0 + 1, rx = xx - x0 + 1;
#pragma unroll
          for (int tap = 0; tap < 9; ++tap) {
            const float xv = xs[(ry + tap / 3) * kNXC + rx + tap % 3];
#pragma unroll
            for (int c = 0; c < CA; ++c) acc[c] = fmaf(xv, w0s[tap * CA + c], acc[c]);
          }
        }
#pragma unroll
        for (int c = 0; c < CA; ++c) tile[i * CI + c] = sgn * acc[c];
      }
    }
    __syncthreads();
    // this thread's 2 x 2 output block from its 4 x 4 input patch, read from
    // LDS once (16 reads, not 36), one patch row at a time: row py feeds
    // output row 0 with kernel row py and output row 1 with kernel row py - 1,
    // so each output still adds its taps in (ky, kx, ci) order
    const int bx = threadIdx.x % kNBX, by = threadIdx.x / kNBX;
    float acc[4][CO];
#pragma unroll
    for (int o = 0; o < 4; ++o)
#pragma unroll
      for (int c = 0; c < CO; ++c) acc[o][c] = 0.f;
    // not unrolled: an unrolled row loop keeps every tap's weights in registers
#pragma unroll 1
    for (int py = 0; py < 4; ++py) {
      float rv[4][CI];
      const float* s = tile + ((2 * by + py) * kNHC + 2 * bx) * CI;
#pragma unroll
      for (int px = 0; px < 4; ++px)
#pragma unroll
        for (int c = 0; c < CI; ++c) rv[px][c] = s[px * CI + c];
#pragma unroll
      for (int oy = 0; oy < 2; ++oy) {
        const int ky = py - oy;
        if (ky < 0 || ky > 2) continue;
        const float* wr = ws + ky * 3 * CI * CO;
#pragma unroll
        for (int kx = 0; kx < 3; ++kx)
#pragma unroll
          for (int ox = 0; ox < 2; ++ox)
#pragma unroll
            for (int ci = 0; ci < CI; ++ci)
#pragma unroll
              for (int co = 0; co < CO; ++co)
                acc[oy * 2 + ox][co] =
                    fmaf(rv[ox + kx][ci], wr[(kx * CI + ci) * CO + co], acc[oy * 2 + ox][co]);
      }
    }
    const int yb = y0 + 2 * by, xb = x0 + 2 * bx;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx)
        if (yb + dy < h && xb + dx < wd)
          Vec<CO>::store(q.out + (static_cast<int64_t>(yb + dy) * wd + xb + dx) * CO,
                         acc[dy * 2 + dx]);
    if constexpr (POOL) {
      // pixels beyond an odd physical edge add nothing (as the MFMA epilogue)
      const bool rowOut = yb + 1 >= q.ph, colOut = xb + 1 >= q.pw;
      const int hh = h >> 1, hw = wd >> 1, py = yb >> 1, px = xb >> 1;
      float s[CO];
#pragma unroll
      for (int c = 0; c < CO; ++c) {
        const float v1 = colOut ? 0.f : acc[1][c];
        const float v2 = rowOut ? 0.f : acc[2][c];
        const float v3 = (rowOut || colOut) ? 0.f : acc[3][c];
        s[c] = ((acc[0][c] + v1) + (v2 + v3)) * 0.25f;
      }
      if (py < hh && px < hw)
        Vec<CO>::store(q.aux + (static_cast<int64_t>(py) * hw + px) * CO, s);
    }
  }
}

// padFactor along one axis: the factor of (y, x) is the product of the
// row's and the column's (both +-1 under Neumann padding, 0 or 1 under zero
// padding), and c moves into [0, n).
__device__ inline float padAxis(int& c, int n, int pad, int flipLo, int flipHi) {
  const bool below = c < 0, above = c >= n;
  float s;
  if (pad & kPadNeumann)
    s = (below && (pad & flipLo)) || (above && (pad & flipHi)) ? -1.f : 1.f;
  else
    s = below || above ? 0.f : 1.f;
  c = min(max(c, 0), n - 1);
  return s;
}

// Output rows per thread in narrowLastKernel (RB): the largest of these that
// still gives every block of the launch a strip.
inline int lastRows(int n, int ph, int pw, int64_t blocks) {
  const int64_t segs = (pw + 255) / 256;
  for (int rb : {8, 2})
    if (static_cast<int64_t>(n) * ((ph + rb - 1) / rb) * segs >= blocks) return rb;
  return 1;
}

// The output layer y = conv(d0, CW -> 1) at the n x ph x pw physical pixels
// (CW: the stored channels of d0).
// APPLY 0: y (compact n x ph x pw). APPLY 1 (n = 1, the solver grid): z =
// scale * y, z[pin] = r[pin] scale[pin]^2 and block partials of (z, q).
// A unit of work is a strip of RB rows by 256 columns of one image, a thread
// one column of it: the RB + 2 rows of d0 it reads are loaded at once (three
// taps each), so each d0 row serves three output rows and every load is in
// flight before the first is used. Units go XCD-contiguously to the blocks
// (xcdTile), so the rows a strip shares with its neighbours meet in one L2.
template <int CW, int APPLY, int RB, class TZ>
__global__ __launch_bounds__(256) void narrowLastKernel(const float* __restrict__ d0,
                                                        const float* __restrict__ w, int n, int h,
                                                        int wd, int ph, int pw, int pad,
                                                        float* __restrict__ y,
                                                        const float* __restrict__ scale,
                                                        const TZ* __restrict__ r,
                                                        const TZ* __restrict__ q,
                                                        TZ* __restrict__ z, int64_t pin,
                                                        double* partial) {
  __shared__ float ws[9 * CW];
  if (threadIdx.x < 9 * CW) ws[threadIdx.x] = w[threadIdx.x];
  __syncthreads();
  double acc[1] = {0.0};
  const int strips = (ph + RB - 1) / RB, segs = (pw + 255) / 256;
  const int64_t units = static_cast<int64_t>(n) * strips * segs;
  for (int64_t ub = blockIdx.x; ub < units; ub += gridDim.x) {
    const int64_t u = xcdTile(ub, units);
    const int seg = static_cast<int>(u % segs);
    const int64_t rest = u / segs;
    const int y0 = static_cast<int>(rest % strips) * RB, b = static_cast<int>(rest / strips);
    const int xx = seg * 256 + static_cast<int>(threadIdx.x);
    if (xx >= pw) continue;
    int cx[3];
    float sx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cx[k] = xx + k - 1;
      sx[k] = padAxis(cx[k], pw, pad, kPadLeft, kPadRight);
    }
    float v[RB + 2][3][CW], sg[RB + 2][3];
#pragma unroll
    for (int i = 0; i < RB + 2; ++i) {
      int ty = y0 - 1 + i;
      const float sy = padAxis(ty, ph, pad, kPadBottom, kPadTop);
      const float* row = d0 + (static_cast<int64_t>(b) * h + ty) * wd * CW;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sg[i][k] = sy * sx[k];
        Vec<CW>::load(row + cx[k] * CW, v[i][k]);
      }
    }
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int yy = y0 + i;
      if (yy >= ph) break;
      float o = 0.f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int ky = tap / 3, kx = tap % 3;
        float dot = 0.f;
#pragma unroll
        for (int c = 0; c < CW; ++c)
          dot += signed_(sg[i + ky][kx], v[i + ky][kx][c]) * ws[tap * CW + c];
        o += dot;
      }
      const int64_t p = (static_cast<int64_t>(b) * ph + yy) * pw + xx;
      if constexpr (APPLY == 0) {
        y[p] = o;
      } else {
        TZ zv = static_cast<TZ>(o * scale[p]);
        if (p == pin) zv = r[p] * static_cast<TZ>(scale[p]) * static_cast<TZ>(scale[p]);
        z[p] = zv;
        if (q != nullptr) acc[0] += static_cast<double>(zv) * static_cast<double>(q[p]);
      }
    }
  }
  if constexpr (APPLY == 1) {
    if (partial != nullptr) blockPartials<1>(acc, partial);
  }
}

// The output layer of a batched application: block entry e = (case, first
// row, rows, partial slot) computes z on those rows of the case's solver
// grid, z = scale conv(d0) and z = r scale^2 at the pinned cell (as
// narrowLastKernel with APPLY 1), and the block's partial of (z, q) into
// partial[slot] (q null: none).
template <int CW>
__global__ __launch_bounds__(256) void narrowLastBatchKernel(
    const float* __restrict__ act, const NarrowCase* __restrict__ cases,
    const int4* __restrict__ blocks, const float* __restrict__ w, const float* __restrict__ scale,
    const double* __restrict__ r, const double* __restrict__ q, double* __restrict__ z,
    double* partial) {
  __shared__ float ws[9 * CW];
  __shared__ double red[256 / 64];
  if (threadIdx.x < 9 * CW) ws[threadIdx.x] = w[threadIdx.x];
  __syncthreads();
  const int4 e = blocks[blockIdx.x];
  const NarrowCase& c = cases[e.x];
  const float* d0 = act + c.act + c.tOff[kTd0];
  const int wd = c.w[0], ph = c.ny, pw = c.nx, pad = c.pad;
  double acc = 0.0;
  const int rowEnd = min(e.y + e.z, ph);
  for (int yy = e.y; yy < rowEnd; ++yy)
    for (int xx = threadIdx.x; xx < pw; xx += blockDim.x) {
      const int64_t lp = static_cast<int64_t>(yy) * pw + xx, p = c.cells + lp;
      float v[9][CW], sg[9];
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        int ty = yy + tap / 3 - 1, tx = xx + tap % 3 - 1;
        sg[tap] = padFactor(ty, tx, ph, pw, pad);
        Vec<CW>::load(d0 + (static_cast<int64_t>(ty) * wd + tx) * CW, v[tap]);
      }
      float o = 0.f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        float dot = 0.f;
#pragma unroll
        for (int k = 0; k < CW; ++k) dot += signed_(sg[tap], v[tap][k]) * ws[tap * CW + k];
        o += dot;
      }
      double zv = static_cast<double>(o * scale[p]);
      if (lp == c.pin) zv = r[p] * static_cast<double>(scale[p]) * static_cast<double>(scale[p]);
      z[p] = zv;
      if (q != nullptr) acc += zv * q[p];
    }
  if (q == nullptr || partial == nullptr) return;
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t += red[i];
    partial[e.w] = t;
  }
}

// BF16 round to nearest even, as the device's f2bf (NaN stays NaN).
float roundBfHost(float f) {
  uint32_t u;
  std::memcpy(&u, &f, sizeof(u));
  if ((u & 0x7f800000u) == 0x7f800000u) {
    u = ((u >> 16) | ((u & 0xffffu) != 0 ? 0x40u : 0u)) << 16;
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
  }
  float r;
  std::memcpy(&r, &u, sizeof(r));
  return r;
}

// channels of each forward tensor in the full network
constexpr int kTensorChannels[kNumForward] = {16, 16, 16, 32, 32, 32, 64, 64, 64,
                                              64, 64, 64, 64, 32, 32, 16, 16};

// Resident workgroups per CU the convolution launches use (at most what the
// kernel's occupancy allows; 0: one workgroup per tile): each walks several
// tiles, the next tile's loads in flight behind the current tile's
// arithmetic. Measured at 2048^2 (learned FCG, 16 iterations): 8 (the
// occupancy, 7 for the first layer) 5.68 ms per solve, one per tile 5.73,
// 4 6.06, 2 6.73.
constexpr int kConvBlocksPerCU = 8;

// A convolution's launch: one workgroup per tile up to kConvBlocksPerCU per
// CU, the grid a multiple of 8 (xcdTile). The resident count is taken from
// the occupancy query once, outside any stream capture; a launch recorded
// into a graph before that uses one workgroup per tile (the grid changes no
// result: each tile's arithmetic is the same).
template <int CA, int CB, int CO, int G, bool POOL, class TS>
void launchConv(hipStream_t s, int64_t tiles, const NarrowArgs& a) {
  static std::atomic<int64_t> resident{0};
  if (kConvBlocksPerCU > 0 && resident.load(std::memory_order_relaxed) == 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    CFDHIP_CHECK(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) {
      int dev = 0, cus = 0, per = 0;
      CFDHIP_CHECK(hipGetDevice(&dev));
      CFDHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      CFDHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per, narrowConvKernel<CA, CB, CO, G, POOL, TS>, 256, 0));
      resident.store(static_cast<int64_t>(std::max(1, std::min(per, kConvBlocksPerCU))) *
                         std::max(1, cus),
                     std::memory_order_relaxed);
    }
  }
  const int64_t r = resident.load(std::memory_order_relaxed);
  int64_t g = r > 0 ? std::min<int64_t>(tiles, r) : tiles;
  if (g < tiles && g >= 8) g -= g % 8;
  hipLaunchKernelGGL((narrowConvKernel<CA, CB, CO, G, POOL, TS>),
                     dim3(static_cast<unsigned>(std::max<int64_t>(1, g))), dim3(256), 0, s, a);
}

template <int CA, int CB, int CO>
void launchCombo(hipStream_t s, int gather, bool pool, bool srcDouble, int64_t tiles,
                 const NarrowArgs& a) {
  if (gather == kNFirst) {
    if (srcDouble)
      launchConv<CA, 0, CO, kNFirst, true, double>(s, tiles, a);
    else
      launchConv<CA, 0, CO, kNFirst, true, float>(s, tiles, a);
  } else if (gather == kNConcat) {
    launchConv<CA, CB, CO, kNConcat, false, float>(s, tiles, a);
  } else if (pool) {
    launchConv<CA, 0, CO, kNPlain, true, float>(s, tiles, a);
  } else {
    launchConv<CA, 0, CO, kNPlain, false, float>(s, tiles, a);
  }
}

// ca, cb, co: stored widths (2 or 4; cb only for the concatenation)
void launchNarrowConv(hipStream_t s, int ca, int cb, int co, int gather, bool pool,
                      bool srcDouble, const NarrowArgs& a) {
  const int64_t tiles =
      a.cases != nullptr ? a.tabCount : static_cast<int64_t>(a.tilesX) * a.tilesY * a.n;
  require(tiles < (int64_t(1) << 31), "narrow convolution: too many tiles for one launch");
  const int key = (ca == 4 ? 4 : 0) | (cb == 4 ? 2 : 0) | (co == 4 ? 1 : 0);
  switch (key) {
    case 0: launchCombo<2, 2, 2>(s, gather, pool, srcDouble, tiles, a); break;
    case 1: launchCombo<2, 2, 4>(s, gather, pool, srcDouble, tiles, a); break;
    case 2: launchCombo<2, 4, 2>(s, gather, pool, srcDouble, tiles, a); break;
    case 3: launchCombo<2, 4, 4>(s, gather, pool, srcDouble, tiles, a); break;
    case 4: launchCombo<4, 2, 2>(s, gather, pool, srcDouble, tiles, a); break;
    case 5: launchCombo<4, 2, 4>(s, gather, pool, srcDouble, tiles, a); break;
    case 6: launchCombo<4, 4, 2>(s, gather, pool, srcDouble, tiles, a); break;
    default: launchCombo<4, 4, 4>(s, gather, pool, srcDouble, tiles, a); break;
  }
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace

void Unet::setDense(bool dense) { dense_ = dense; }

void Unet::levelExtent(int h, int w, int& ph, int& pw) const {
  ph = physLevelH > 0 ? physH : h;
  pw = physLevelW > 0 ? physW : w;
  for (int hh = physLevelH; hh > h; hh >>= 1) ph = (ph + 1) >> 1;
  for (int ww = physLevelW; ww > w; ww >>= 1) pw = (pw + 1) >> 1;
}

void Unet::buildNarrow() {
  narrow = NarrowPlan{};
  if (static_cast<int64_t>(hostMaster.size()) != kUnetParams) return;
  auto W = [&](int l, int tap, int ci, int co) {
    return hostMaster[layerOffset(l) +
                      (static_cast<int64_t>(tap) * kLayers[l].cin + ci) * kLayers[l].cout + co];
  };
  std::vector<std::vector<char>> need(kNumForward);
  for (int t = 0; t < kNumForward; ++t) need[t].assign(kTensorChannels[t], 0);
  // layer l produces the channels `wanted` of its output from input A (and,
  // concatenated after A's ca channels, input B): mark what it reads
  auto back = [&](int l, const std::vector<char>& wanted, int A, int B, int ca) {
    for (int co = 0; co < kLayers[l].cout; ++co) {
      if (!wanted[co]) continue;
      for (int tap = 0; tap < 9; ++tap)
        for (int ci = 0; ci < kLayers[l].cin; ++ci) {
          if (W(l, tap, ci, co) == 0.f) continue;
          if (B < 0 || ci < ca) need[A][ci] = 1;
          else need[B][ci - ca] = 1;
        }
    }
  };
  const std::vector<char> yWanted(1, 1);
  back(kOut, yWanted, kTd0, -1, 0);
  back(kD0b, need[kTd0], kTd0a, -1, 0);
  back(kD0a, need[kTd0a], kTd1, kTe0, 32);
  back(kD1b, need[kTd1], kTd1a, -1, 0);
  back(kD1a, need[kTd1a], kTd2, kTe1, 64);
  back(kD2b, need[kTd2], kTd2a, -1, 0);
  back(kD2a, need[kTd2a], kTb3, kTe2, 64);
  back(kB3b, need[kTb3], kTb3a, -1, 0);
  // the tail overwrites its destination channel from its source channel
  const bool tailLive = tail.levels > 0 && need[kTb3a][tail.dst];
  std::vector<char> convWanted = need[kTb3a];
  if (tailLive) {
    need[kTb3a][tail.src] = 1;
    convWanted = need[kTb3a];
    convWanted[tail.dst] = 0;
  }
  back(kB3a, convWanted, kTp2, -1, 0);
  // a pooled tensor keeps its source's channels
  auto pooled = [&](int e, int p) {
    for (int c = 0; c < kTensorChannels[e]; ++c) need[e][c] = need[e][c] || need[p][c];
    need[p] = need[e];
  };
  pooled(kTe2, kTp2);
  back(kE2b, need[kTe2], kTe2a, -1, 0);
  back(kE2a, need[kTe2a], kTp1, -1, 0);
  pooled(kTe1, kTp1);
  back(kE1b, need[kTe1], kTe1a, -1, 0);
  back(kE1a, need[kTe1a], kTp0, -1, 0);
  pooled(kTe0, kTp0);
  back(kE0b, need[kTe0], kTe0a, -1, 0);
  for (int t = 0; t < kNumForward; ++t) {
    int k = 0;
    for (int c = 0; c < kTensorChannels[t]; ++c)
      if (need[t][c]) {
        if (k == kNarrowMax) return;  // too many live channels: the MFMA layers
        narrow.chan[t][k++] = c;
      }
    if (k == 0) narrow.chan[t][k++] = 0;  // nothing reads it: one placeholder slot
    narrow.width[t] = k;
    narrow.sw[t] = k <= 2 ? 2 : 4;
  }
  if (tailLive) {
    for (int k = 0; k < narrow.width[kTb3a]; ++k) {
      if (narrow.chan[kTb3a][k] == tail.src) narrow.tailSrc = k;
      if (narrow.chan[kTb3a][k] == tail.dst) narrow.tailDst = k;
    }
  }
  // compacted weights, BF16-rounded, [tap][slot in][slot out] over the stored
  // widths; slots beyond a tensor's live width get zeros
  auto slotChannel = [&](int t, int k) { return k < narrow.width[t] ? narrow.chan[t][k] : -1; };
  std::vector<float>& wc = narrow.w;
  auto layer = [&](int l, int out, int A, int B, int ca) {
    narrow.wOff[l] = static_cast<int64_t>(wc.size());
    const int sa = narrow.sw[A], sb = B < 0 ? 0 : narrow.sw[B];
    const int co = out < 0 ? 1 : narrow.sw[out];  // out < 0: the one-channel output layer
    for (int tap = 0; tap < 9; ++tap)
      for (int i = 0; i < sa + sb; ++i)
        for (int o = 0; o < co; ++o) {
          int src = -1;
          if (i < sa) {
            src = slotChannel(A, i);
          } else if (slotChannel(B, i - sa) >= 0) {
            src = ca + slotChannel(B, i - sa);
          }
          const int dst = out < 0 ? 0 : slotChannel(out, o);
          wc.push_back(src < 0 || dst < 0 ? 0.f : roundBfHost(W(l, tap, src, dst)));
        }
  };
  // e0a takes the scalar input: [tap][slot]
  narrow.wOff[kE0a] = 0;
  for (int tap = 0; tap < 9; ++tap)
    for (int o = 0; o < narrow.sw[kTe0a]; ++o) {
      const int dst = slotChannel(kTe0a, o);
      wc.push_back(dst < 0 ? 0.f : roundBfHost(W(kE0a, tap, 0, dst)));
    }
  layer(kE0b, kTe0, kTe0a, -1, 0);
  layer(kE1a, kTe1a, kTp0, -1, 0);
  layer(kE1b, kTe1, kTe1a, -1, 0);
  layer(kE2a, kTe2a, kTp1, -1, 0);
  layer(kE2b, kTe2, kTe2a, -1, 0);
  layer(kB3a, kTb3a, kTp2, -1, 0);
  layer(kB3b, kTb3, kTb3a, -1, 0);
  layer(kD2a, kTd2a, kTb3, kTe2, 64);
  layer(kD2b, kTd2, kTd2a, -1, 0);
  layer(kD1a, kTd1a, kTd2, kTe1, 64);
  layer(kD1b, kTd1, kTd1a, -1, 0);
  layer(kD0a, kTd0a, kTd1, kTe0, 32);
  layer(kD0b, kTd0, kTd0a, -1, 0);
  layer(kOut, -1, kTd0, -1, 0);  // [tap][slot]
  nW.ensure(static_cast<int64_t>(wc.size()));
  CFDHIP_CHECK(hipMemcpyAsync(nW.get(), wc.data(), wc.size() * sizeof(float),
                              hipMemcpyHostToDevice, stream));
  CFDHIP_CHECK(hipStreamSynchronize(stream));
  narrowN = 0;  // the stored widths may have changed: size the tensors again
  narrow.usable = true;
}

void Unet::ensureNarrow(int n, int h, int w) {
  if (n == narrowN && h == narrowH && w == narrowW) return;
  const int64_t p0 = static_cast<int64_t>(n) * h * w;
  const int64_t px[4] = {p0, static_cast<int64_t>(n) * (h / 2) * (w / 2),
                         static_cast<int64_t>(n) * (h / 4) * (w / 4),
                         static_cast<int64_t>(n) * (h / 8) * (w / 8)};
  // level of each forward tensor (e0a is fused into e0b and not stored)
  constexpr int kLevel[kNumForward] = {-1, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0};
  int64_t off = 0;
  for (int t = 0; t < kNumForward; ++t) {
    nOff[t] = off;
    if (kLevel[t] >= 0) off += (px[kLevel[t]] * narrow.sw[t] + 63) / 64 * 64;
  }
  nact.ensure(off);
  narrowBytes = off * static_cast<int64_t>(sizeof(float));
  narrowN = n;
  narrowH = h;
  narrowW = w;
}

void Unet::forwardNarrow(const void* src, bool srcDouble, const float* scale, int srcH, int srcW,
                         int64_t srcBatch, int n, int h, int w) {
  require(h - std::min(srcH, h) < 8 && w - std::min(srcW, w) < 8,
          "the network grid rounds the image up to a multiple of 8");
  ensureNarrow(n, h, w);
  const float* wts = nW.get();
  const int* sw = narrow.sw;
  auto run = [&](int layer, int gather, bool pool, int A, int B, int out, int aux, int hh,
                 int ww) {
    NarrowArgs a{};
    a.a = A < 0 ? nullptr : NT(A);
    a.b = B < 0 ? nullptr : NT(B);
    a.w = wts + narrow.wOff[layer];
    a.w0 = wts + narrow.wOff[kE0a];
    a.out = NT(out);
    a.aux = aux < 0 ? nullptr : NT(aux);
    a.n = n;
    a.h = hh;
    a.w_ = ww;
    levelExtent(hh, ww, a.ph, a.pw);
    a.pad = padBits();
    a.tilesX = (ww + kNTW - 1) / kNTW;
    a.tilesY = (hh + kNTH - 1) / kNTH;
    a.src = src;
    a.scale = scale;
    a.srcH = srcH;
    a.srcW = srcW;
    a.srcBatch = srcBatch;
    const int ca = gather == kNFirst ? sw[kTe0a] : sw[A];
    mark(layer, true);
    launchNarrowConv(stream, ca, B < 0 ? 2 : sw[B], sw[out], gather, pool, srcDouble, a);
    mark(layer, false);
  };
  const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
  run(kE0b, kNFirst, true, -1, -1, kTe0, kTp0, h, w);
  run(kE1a, kNPlain, false, kTp0, -1, kTe1a, -1, h1, w1);
  run(kE1b, kNPlain, true, kTe1a, -1, kTe1, kTp1, h1, w1);
  run(kE2a, kNPlain, false, kTp1, -1, kTe2a, -1, h2, w2);
  run(kE2b, kNPlain, true, kTe2a, -1, kTe2, kTp2, h2, w2);
  run(kB3a, kNPlain, false, kTp2, -1, kTb3a, -1, h3, w3);
  if (narrow.tailSrc >= 0 && narrow.tailDst >= 0) {
    mark(kNumLayers, true);
    tailApply(NT(kTb3a), n, h3, w3, false, sw[kTb3a], narrow.tailSrc, narrow.tailDst);
    mark(kNumLayers, false);
  }
  run(kB3b, kNPlain, false, kTb3a, -1, kTb3, -1, h3, w3);
  run(kD2a, kNConcat, false, kTb3, kTe2, kTd2a, -1, h2, w2);
  run(kD2b, kNPlain, false, kTd2a, -1, kTd2, -1, h2, w2);
  run(kD1a, kNConcat, false, kTd2, kTe1, kTd1a, -1, h1, w1);
  run(kD1b, kNPlain, false, kTd1a, -1, kTd1, -1, h1, w1);
  run(kD0a, kNConcat, false, kTd1, kTe0, kTd0a, -1, h, w);
  run(kD0b, kNPlain, false, kTd0a, -1, kTd0, -1, h, w);
  lastNarrow = true;
}

// The output layer's launch: RB from lastRows, CW from the stored channels of d0.
template <int APPLY, class TZ>
void launchLast(hipStream_t stream, dim3 grid, int rb, int cw, const float* d0, const float* w,
                int n, int h, int wd, int ph, int pw, int pad, float* y, const float* scale,
                const TZ* r, const TZ* q, TZ* z, int64_t pin, double* partial) {
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, stream, d0, w, n, h, wd, ph, pw, pad, y, scale,
                       r, q, z, pin, partial);
  };
  if (cw == 2) {
    if (rb == 8) go(narrowLastKernel<2, APPLY, 8, TZ>);
    else if (rb == 2) go(narrowLastKernel<2, APPLY, 2, TZ>);
    else go(narrowLastKernel<2, APPLY, 1, TZ>);
  } else {
    if (rb == 8) go(narrowLastKernel<4, APPLY, 8, TZ>);
    else if (rb == 2) go(narrowLastKernel<4, APPLY, 2, TZ>);
    else go(narrowLastKernel<4, APPLY, 1, TZ>);
  }
  CFDHIP_CHECK(hipGetLastError());
}

void Unet::outputNarrow(int n, int h, int w, int ph, int pw) {
  // at most 8192 blocks, enough for 8 per CU; each block walks its units
  const int rb = lastRows(n, ph, pw, 2048);
  const int64_t units =
      static_cast<int64_t>(n) * ((ph + rb - 1) / rb) * ((pw + 255) / 256);
  const dim3 grid(static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(units, 8192))));
  launchLast<0, float>(stream, grid, rb, narrow.sw[kTd0], NT(kTd0), nW.get() + narrow.wOff[kOut],
                       n, h, w, ph, pw, padBits(), y.get(), nullptr, nullptr, nullptr, nullptr,
                       int64_t(-1), nullptr);
}

template <class TZ>
void Unet::outputApplyNarrow(const float* scale, const TZ* r, const TZ* q, TZ* z, int ny, int nx,
                             int64_t pin, int h, int w, double* partial, int parts) {
  // one block per partial slot (the caller sums parts partials)
  launchLast<1, TZ>(stream, dim3(static_cast<unsigned>(parts)), lastRows(1, ny, nx, parts),
                    narrow.sw[kTd0], NT(kTd0), nW.get() + narrow.wOff[kOut], 1, h, w, ny, nx,
                    padBits(), nullptr, scale, r, q, z, pin, partial);
}
template void Unet::outputApplyNarrow<float>(const float*, const float*, const float*, float*, int,
                                             int, int64_t, int, int, double*, int);
template void Unet::outputApplyNarrow<double>(const float*, const double*, const double*, double*,
                                              int, int, int64_t, int, int, double*, int);

int64_t Unet::narrowBatchLayout(std::vector<NarrowCase>& cases) const {
  require(narrow.usable && !dense_,
          "batched inference runs the narrow kernels: the weights must leave at most 4 live "
          "channels per tensor");
  // level of each forward tensor (e0a is fused into e0b and not stored)
  constexpr int kLevel[kNumForward] = {-1, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0};
  int64_t off = 0;
  for (NarrowCase& c : cases) {
    c.h[0] = (c.ny + 7) / 8 * 8;
    c.w[0] = (c.nx + 7) / 8 * 8;
    c.ph[0] = c.ny;
    c.pw[0] = c.nx;
    for (int l = 1; l < 4; ++l) {
      c.h[l] = c.h[l - 1] / 2;
      c.w[l] = c.w[l - 1] / 2;
      c.ph[l] = (c.ph[l - 1] + 1) / 2;
      c.pw[l] = (c.pw[l - 1] + 1) / 2;
    }
    for (int l = 0; l < 4; ++l) c.tilesX[l] = (c.w[l] + kNTW - 1) / kNTW;
    c.act = off;
    int64_t o = 0;
    for (int t = 0; t < kNumForward; ++t) {
      c.tOff[t] = o;
      if (kLevel[t] >= 0)
        o += (static_cast<int64_t>(c.h[kLevel[t]]) * c.w[kLevel[t]] * narrow.sw[t] + 15) / 16 * 16;
    }
    off += o;
  }
  return off;
}

void Unet::forwardNarrowBatch(const NarrowBatch& nb, const double* r, const float* scale,
                              const double* q, double* z, double* partial) {
  const float* wts = nW.get();
  const int* sw = narrow.sw;
  auto run = [&](int layer, int gather, bool pool, int A, int B, int out, int aux, int L) {
    NarrowArgs a{};
    a.a = nb.act;
    a.b = B < 0 ? nullptr : nb.act;
    a.out = nb.act;
    a.aux = aux < 0 ? nullptr : nb.act;
    a.w = wts + narrow.wOff[layer];
    a.w0 = wts + narrow.wOff[kE0a];
    a.src = r;
    a.scale = scale;
    a.cases = nb.cases;
    a.tab = nb.tab[L];
    a.tabCount = nb.tabCount[L];
    a.level = L;
    a.tA = A;
    a.tB = B;
    a.tOut = out;
    a.tAux = aux;
    const int ca = gather == kNFirst ? sw[kTe0a] : sw[A];
    mark(layer, true);
    launchNarrowConv(stream, ca, B < 0 ? 2 : sw[B], sw[out], gather, pool, true, a);
    mark(layer, false);
  };
  run(kE0b, kNFirst, true, -1, -1, kTe0, kTp0, 0);
  run(kE1a, kNPlain, false, kTp0, -1, kTe1a, -1, 1);
  run(kE1b, kNPlain, true, kTe1a, -1, kTe1, kTp1, 1);
  run(kE2a, kNPlain, false, kTp1, -1, kTe2a, -1, 2);
  run(kE2b, kNPlain, true, kTe2a, -1, kTe2, kTp2, 2);
  run(kB3a, kNPlain, false, kTp2, -1, kTb3a, -1, 3);
  if (narrow.tailSrc >= 0 && narrow.tailDst >= 0) {
    mark(kNumLayers, true);
    tailBatch(nb, sw[kTb3a], narrow.tailSrc, narrow.tailDst);
    mark(kNumLayers, false);
  }
  run(kB3b, kNPlain, false, kTb3a, -1, kTb3, -1, 3);
  run(kD2a, kNConcat, false, kTb3, kTe2, kTd2a, -1, 2);
  run(kD2b, kNPlain, false, kTd2a, -1, kTd2, -1, 2);
  run(kD1a, kNConcat, false, kTd2, kTe1, kTd1a, -1, 1);
  run(kD1b, kNPlain, false, kTd1a, -1, kTd1, -1, 1);
  run(kD0a, kNConcat, false, kTd1, kTe0, kTd0a, -1, 0);
  run(kD0b, kNPlain, false, kTd0a, -1, kTd0, -1, 0);
  const float* wo = wts + narrow.wOff[kOut];
  const dim3 grid(static_cast<unsigned>(std::max<int64_t>(1, nb.lastCount)));
  mark(kOut, true);
  if (sw[kTd0] == 2)
    hipLaunchKernelGGL((narrowLastBatchKernel<2>), grid, dim3(256), 0, stream, nb.act, nb.cases,
                       nb.last, wo, scale, r, q, z, partial);
  else
    hipLaunchKernelGGL((narrowLastBatchKernel<4>), grid, dim3(256), 0, stream, nb.act, nb.cases,
                       nb.last, wo, scale, r, q, z, partial);
  mark(kOut, false);
  CFDHIP_CHECK(hipGetLastError());
  collectTiming();
}

}  // namespace cfdhip
