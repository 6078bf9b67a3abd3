// MFMA convolutions of the learned preconditioner (gfx950).
//
// Forward and transposed 3x3 convolutions are one implicit GEMM: M = pixels
// of an 8 x 32 output tile, N = output channels, K = 9 taps x input
// channels, on v_mfma_f32_16x16x32_bf16 with FP32 accumulators.
//  * Activations are FP32 in HBM. The input tile (one-pixel halo, gathered
//    through upsampling, concatenation or the pool-gradient sum) is staged in
//    LDS one chunk of input channels at a time, split into hi = bf16(x) and
//    lo = bf16(x - hi); every weight fragment meets both (two MFMAs).
//  * The layer's BF16 weight fragments for the workgroup's output channels
//    are staged in LDS once, beside the tile.
//  * Each wave owns two tile rows: four 16-pixel M-blocks, each a 2 x 8 patch
//    ordered so that a lane's four accumulator rows are one 2 x 2 pixel block;
//    the 2x2 average pool and the nearest-upsample gradient (2x2 sum) are then
//    lane-local in the epilogue.
//
// The weight gradient is a GEMM over pixels, D[ci][co] += X[p + tap][ci] G[p][co]
// per (tap, 16-channel block), with BF16 operands (gradients need no more),
// X and G tiles in LDS, partial sums per pixel split reduced in a fixed order.
#include <algorithm>

#include "unet.hpp"

namespace cfdhip {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf8v;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4v;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float fx4 __attribute__((ext_vector_type(4)));

constexpr int kTH = 8, kTW = 32;
constexpr int kInRows = kTH + 2, kInCols = kTW + 2;
constexpr int kInPix = kInRows * kInCols;
// The upsampled half of a concatenated input is staged at coarse resolution:
// fine halo rows y0 - 1 .. y0 + kTH map to coarse rows y0/2 - 1 .. y0/2 +
// kTH/2 (tile row t -> coarse row (t + 1) >> 1), likewise for columns.
constexpr int kCoRows = kTH / 2 + 2, kCoCols = kTW / 2 + 2;

__device__ inline float bf2f(bf16 h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Round to nearest even; NaN stays NaN.
__device__ inline bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u)
    return static_cast<bf16>((u >> 16) | ((u & 0xffffu) != 0 ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<bf16>(u >> 16);
}

// 4 channels starting at c of the gathered input at pixel (y, x) inside the image.
template <int C, int CA, int G>
__device__ inline fx4 gather4(const float* __restrict__ A, const float* __restrict__ B, int n,
                              int h, int w, int y, int x, int c) {
  if constexpr (G == kGatherPlain) {
    return *reinterpret_cast<const fx4*>(A + ((static_cast<int64_t>(n) * h + y) * w + x) * C + c);
  } else if constexpr (G == kGatherConcatUp) {
    if (c < CA) {
      const int hh = h >> 1, hw = w >> 1;
      return *reinterpret_cast<const fx4*>(
          A + ((static_cast<int64_t>(n) * hh + (y >> 1)) * hw + (x >> 1)) * CA + c);
    }
    return *reinterpret_cast<const fx4*>(
        B + ((static_cast<int64_t>(n) * h + y) * w + x) * (C - CA) + (c - CA));
  } else {
    const int hh = h >> 1, hw = w >> 1;
    const fx4 va =
        *reinterpret_cast<const fx4*>(A + ((static_cast<int64_t>(n) * h + y) * w + x) * C + c);
    const fx4 vb = *reinterpret_cast<const fx4*>(
        B + ((static_cast<int64_t>(n) * hh + (y >> 1)) * hw + (x >> 1)) * C + c);
    return va + 0.25f * vb;
  }
}

// Stages ROWS x COLS pixels at origin (y0, x0) (zero outside the image),
// channels [c0, c0 + CC) of the gathered input: hi[(ty COLS + tx) cs + c] =
// bf16(v), lo[...] = bf16(v - hi) (lo == nullptr: hi only). Every load of the
// thread's items is issued before the first conversion, so a thread waits on
// one memory round trip, not one per item.
template <int C, int CA, int G, int CC, int ROWS, int COLS>
__device__ inline void stageSplit(bf16* hi, bf16* lo, int cs, const float* A, const float* B,
                                  int n, int h, int w, int y0, int x0, int c0, int pad) {
  constexpr int C4 = CC / 4;
  constexpr int ITEMS = ROWS * COLS * C4;
  constexpr int PER = (ITEMS + 255) / 256;
  static_assert(PER <= 32, "one sign bit per item");
  fx4 v[PER];
  uint32_t neg = 0;  // items read through a Dirichlet side: negated at the store
  // a window inside the image (uniform over the workgroup) skips the padding
  const bool inside = y0 >= 0 && x0 >= 0 && y0 + ROWS <= h && x0 + COLS <= w;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int it = threadIdx.x + k * 256;
    v[k] = fx4{0.f, 0.f, 0.f, 0.f};
    if (it < ITEMS) {
      const int c4 = it % C4;
      const int px = it / C4;
      const int ty = px / COLS, tx = px - ty * COLS;
      int y = y0 + ty, x = x0 + tx;
      const float sg = inside ? 1.f : padSign(y, x, h, w, pad);
      if (sg != 0.f) v[k] = gather4<C, CA, G>(A, B, n, h, w, y, x, c0 + c4 * 4);
      if (sg < 0.f) neg |= 1u << k;
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int it = threadIdx.x + k * 256;
    if (it >= ITEMS) continue;
    const int c4 = it % C4;
    const int px = it / C4;
    bf16 hb[4], lb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      hb[q] = f2bf(v[k][q]);
      lb[q] = f2bf(v[k][q] - bf2f(hb[q]));
    }
    // bf16(-x) = -bf16(x) and the split of -x is the negated split: flip the signs
    const uint32_t flip = (neg >> k & 1u) ? 0x80008000u : 0u;
    u32x2 vh;
    vh[0] = (static_cast<uint32_t>(hb[0]) | (static_cast<uint32_t>(hb[1]) << 16)) ^ flip;
    vh[1] = (static_cast<uint32_t>(hb[2]) | (static_cast<uint32_t>(hb[3]) << 16)) ^ flip;
    const int o = px * cs + c4 * 4;
    *reinterpret_cast<u32x2*>(hi + o) = vh;
    if (lo != nullptr) {
      u32x2 vl;
      vl[0] = (static_cast<uint32_t>(lb[0]) | (static_cast<uint32_t>(lb[1]) << 16)) ^ flip;
      vl[1] = (static_cast<uint32_t>(lb[2]) | (static_cast<uint32_t>(lb[3]) << 16)) ^ flip;
      *reinterpret_cast<u32x2*>(lo + o) = vl;
    }
  }
}

// The two halves of stageSplit for software pipelining: loadChunk issues a
// chunk's loads into registers, storeChunk converts and writes them to LDS.
// Pixels read through a Dirichlet side are negated afterwards in LDS by
// flipSides, on the windows that touch such a side only, so that no sign
// state is held in registers across the multiply.
template <int C, int CA, int G, int CC, int ROWS, int COLS>
struct ChunkStage {
  static constexpr int C4 = CC / 4;
  static constexpr int ITEMS = ROWS * COLS * C4;
  static constexpr int PER = (ITEMS + 255) / 256;
  fx4 v[PER];
  // h x w: the tensor (addressing); ph x pw: the physical image (padding)
  __device__ inline void load(const float* A, const float* B, int n, int h, int w, int ph, int pw,
                              int y0, int x0, int c0, int pad) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int it = threadIdx.x + k * 256;
      v[k] = fx4{0.f, 0.f, 0.f, 0.f};
      if (it < ITEMS) {
        const int c4 = it % C4;
        const int px = it / C4;
        const int ty = px / COLS, tx = px - ty * COLS;
        int y = y0 + ty, x = x0 + tx;
        if (padSign(y, x, ph, pw, pad) != 0.f)
          v[k] = gather4<C, CA, G>(A, B, n, h, w, y, x, c0 + c4 * 4);
      }
    }
  }
  __device__ inline void store(bf16* hi, bf16* lo, int cs) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int it = threadIdx.x + k * 256;
      if (it >= ITEMS) continue;
      const int c4 = it % C4;
      const int px = it / C4;
      bf16 hb[4], lb[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        hb[q] = f2bf(v[k][q]);
        lb[q] = f2bf(v[k][q] - bf2f(hb[q]));
      }
      u32x2 vh, vl;
      vh[0] = static_cast<uint32_t>(hb[0]) | (static_cast<uint32_t>(hb[1]) << 16);
      vh[1] = static_cast<uint32_t>(hb[2]) | (static_cast<uint32_t>(hb[3]) << 16);
      vl[0] = static_cast<uint32_t>(lb[0]) | (static_cast<uint32_t>(lb[1]) << 16);
      vl[1] = static_cast<uint32_t>(lb[2]) | (static_cast<uint32_t>(lb[3]) << 16);
      const int o = px * cs + c4 * 4;
      *reinterpret_cast<u32x2*>(hi + o) = vh;
      *reinterpret_cast<u32x2*>(lo + o) = vl;
    }
  }
  // Whether the window (y0, x0) of an h x w image reads through a Dirichlet
  // side (uniform over the workgroup).
  __device__ static inline bool hasSigns(int h, int w, int y0, int x0, int pad) {
    return (pad & kPadNeumann) && (pad >> 1) != 0 &&
           !(y0 >= 0 && x0 >= 0 && y0 + ROWS <= h && x0 + COLS <= w);
  }
  // Negates the staged items of this thread that were read through a
  // Dirichlet side: bf16(-x) = -bf16(x) and the split of -x is the negated
  // split, so the sign bits of both halves flip. Same item-to-thread map as
  // store(), hence no barrier between the two.
  __device__ static void flipSides(bf16* hi, bf16* lo, int cs, int h, int w, int y0, int x0,
                                   int pad) {
#pragma unroll 1
    for (int it = threadIdx.x; it < ITEMS; it += 256) {
      const int c4 = it % C4;
      const int px = it / C4;
      const int ty = px / COLS, tx = px - ty * COLS;
      int y = y0 + ty, x = x0 + tx;
      if (padSign(y, x, h, w, pad) < 0.f) {
        const int o = px * cs + c4 * 4;
        u32x2* ph = reinterpret_cast<u32x2*>(hi + o);
        u32x2* pl = reinterpret_cast<u32x2*>(lo + o);
        u32x2 th = *ph, tl = *pl;
        th[0] ^= 0x80008000u;
        th[1] ^= 0x80008000u;
        tl[0] ^= 0x80008000u;
        tl[1] ^= 0x80008000u;
        *ph = th;
        *pl = tl;
      }
    }
  }
};

__device__ inline void splitStore(fx4 v, bf16* hi, bf16* lo, int o) {
  bf16 hb[4], lb[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    hb[k] = f2bf(v[k]);
    lb[k] = f2bf(v[k] - bf2f(hb[k]));
  }
  u32x2 vh, vl;
  vh[0] = static_cast<uint32_t>(hb[0]) | (static_cast<uint32_t>(hb[1]) << 16);
  vh[1] = static_cast<uint32_t>(hb[2]) | (static_cast<uint32_t>(hb[3]) << 16);
  vl[0] = static_cast<uint32_t>(lb[0]) | (static_cast<uint32_t>(lb[1]) << 16);
  vl[1] = static_cast<uint32_t>(lb[2]) | (static_cast<uint32_t>(lb[3]) << 16);
  *reinterpret_cast<u32x2*>(hi + o) = vh;
  *reinterpret_cast<u32x2*>(lo + o) = vl;
}

// The first layer fused into the staging of the second: the 16 channels of
// e0a = conv(x, 1 -> 16) on the (kInRows x kInCols) tile with origin (y0, x0),
// from x staged once with its own halo. Same tap order and FP32 arithmetic as
// the stand-alone first-layer kernel, so the values are identical.
template <class TS>
__device__ inline void stageFirst(bf16* hi, bf16* lo, int cs, const ConvArgs& a, int n, int y0,
                                  int x0) {
  constexpr int XR = kInRows + 2, XC = kInCols + 2;
  __shared__ float xs[XR * XC];
  __shared__ float w0s[144];
  if (threadIdx.x < 144) w0s[threadIdx.x] = bf2f(f2bf(a.w0[threadIdx.x]));
  const TS* src = static_cast<const TS*>(a.src);
  for (int i = threadIdx.x; i < XR * XC; i += blockDim.x) {
    const int ty = i / XC, tx = i - ty * XC;
    int y = y0 - 1 + ty, x = x0 - 1 + tx;
    float v = 0.f;
    const float sg = padSign(y, x, a.ph, a.pw, a.pad);
    if (sg != 0.f && y < a.srcH && x < a.srcW) {
      const int64_t si = static_cast<int64_t>(n) * a.srcH * a.srcW +
                         static_cast<int64_t>(y) * a.srcW + x;
      const float sv = static_cast<float>(src[si]);
      v = sg * (a.scale != nullptr ? sv * a.scale[si] : sv);
    }
    xs[i] = v;
  }
  __syncthreads();
  // a thread always handles the same 4 channels (blockDim is a multiple of 4):
  // their 36 weights stay in registers
  const int c4 = threadIdx.x & 3;
  float wr[36];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int k = 0; k < 4; ++k) wr[tap * 4 + k] = w0s[tap * 16 + c4 * 4 + k];
  for (int it = threadIdx.x; it < kInPix * 4; it += blockDim.x) {
    const int px = it >> 2;
    // a halo pixel outside the image takes e0a of the nearest pixel inside
    // (times the padding's sign), whose taps lie inside the staged window
    int y = y0 + px / kInCols, x = x0 + px % kInCols;
    fx4 acc = {0.f, 0.f, 0.f, 0.f};
    const float sg = padSign(y, x, a.ph, a.pw, a.pad);
    if (sg != 0.f) {
      const int ty = y - y0, tx = x - x0;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const float v = xs[(ty + tap / 3) * XC + tx + tap % 3];
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += v * wr[tap * 4 + k];
      }
      acc = acc * sg;
    }
    splitStore(acc, hi, lo, px * cs + c4 * 4);
  }
}

__device__ inline f4v mfma(u32x4 a, u32x4 b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a),
                                                 __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}

template <int A, int B>
constexpr int cmax() {
  return A > B ? A : B;
}

// SIDES: the padding may have Dirichlet sides (odd reflection), so staged
// windows on such a side get the sign fix-up; without it that code is not
// compiled in (it costs registers on the concatenating layers).
template <int CIN, int CA, int COUT, int NBG, int G, int E, int SPLIT, bool SIDES>
__global__ __launch_bounds__(256) void convKernel(ConvArgs a) {
  constexpr int CC = kChunk(CIN);
  constexpr int NCH = CIN / CC;
  constexpr int MSC = (9 * CC + 31) / 32;
  constexpr int CS = CC + 8;  // LDS channel stride (BF16 elements), 16-byte rows
  constexpr int TILE_E = kInPix * CS;
  constexpr int W_FRAGS = NCH * MSC * NBG * 64;
  constexpr int SS = NBG * 16 + 4;  // FP32 output staging stride
  constexpr int REGION = cmax<2 * TILE_E * 2, kTH * kTW * SS * 4>();
  __shared__ u32x4 wl[W_FRAGS];
  __shared__ __attribute__((aligned(16))) unsigned char region[REGION];
  bf16* const hiT = reinterpret_cast<bf16*>(region);
  bf16* const loT = hiT + TILE_E;
  float* const stage = reinterpret_cast<float*>(region);

  // Persistent over tiles: block x takes tiles x, x + gridDim.x, ... (the
  // same count for every thread of the block); the weights are staged once and
  // the next tile's first chunk is loaded while this tile is multiplied.
  const int nTiles = a.tilesX * a.tilesY * a.n;
  int tile = blockIdx.x;
  if (tile >= nTiles) return;
  auto origin = [&](int t, int& tn, int& ty0, int& tx0) {
    tx0 = (t % a.tilesX) * kTW;
    t /= a.tilesX;
    ty0 = (t % a.tilesY) * kTH;
    tn = t / a.tilesY;
  };
  int n, y0, x0;
  origin(tile, n, y0, x0);
  const int nb0 = blockIdx.y * NBG;

  // the first chunk's loads go out before the weights are staged; chunk
  // ch + 1's loads are in flight while chunk ch is multiplied. For a
  // concatenation, chunks below CA come from the coarse tensor at coarse
  // resolution (a third of the pixels), the rest from the fine tensor.
  constexpr bool kConcat = G == kGatherConcatUp;
  constexpr int NCO = kConcat ? CA / CC : 0;  // coarse chunks (CC divides CA)
  constexpr int CF = kConcat ? CIN - CA : CIN;  // channels of the tensor stageBuf reads
  ChunkStage<CF, kConcat ? CF : CA, (G == kGatherFirst || kConcat) ? kGatherPlain : G, CC, kInRows,
             kInCols>
      stageBuf;
  ChunkStage<CA, CA, kGatherPlain, CC, kCoRows, kCoCols> coarseBuf;
  auto loadChunk = [&](int ch, int tn, int ty0, int tx0) {
    if constexpr (kConcat) {
      if (ch < NCO)
        coarseBuf.load(a.a, nullptr, tn, a.h >> 1, a.w_ >> 1, (a.ph + 1) >> 1, (a.pw + 1) >> 1,
                       (ty0 >> 1) - 1, (tx0 >> 1) - 1, ch * CC, a.pad);
      else
        stageBuf.load(a.b, nullptr, tn, a.h, a.w_, a.ph, a.pw, ty0 - 1, tx0 - 1, ch * CC - CA,
                      a.pad);
    } else if constexpr (G != kGatherFirst) {
      stageBuf.load(a.a, a.b, tn, a.h, a.w_, a.ph, a.pw, ty0 - 1, tx0 - 1, ch * CC, a.pad);
    }
  };
  loadChunk(0, n, y0, x0);

  // this group's weight fragments: packed [(chunk step) * nbTotal + nb] * 64 + lane
  const u32x4* __restrict__ wp = reinterpret_cast<const u32x4*>(a.w);
  {
    constexpr int PER = (W_FRAGS + 255) / 256;
    constexpr int BATCH = PER < 8 ? PER : 8;
#pragma unroll
    for (int k0 = 0; k0 < PER; k0 += BATCH) {
      u32x4 t[BATCH];
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int i = threadIdx.x + (k0 + k) * 256;
        if (k0 + k < PER && i < W_FRAGS) {
          const int ln = i & 63, f = i >> 6;
          const int nb = f % NBG, qm = f / NBG;
          t[k] = wp[(static_cast<int64_t>(qm) * a.nbTotal + nb0 + nb) * 64 + ln];
        }
      }
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int i = threadIdx.x + (k0 + k) * 256;
        if (k0 + k < PER && i < W_FRAGS) wl[i] = t[k];
      }
    }
  }

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, g = lane >> 4;
  // A-operand pixel of this lane (row li of each M-block): element li & 3 of
  // the 2 x 2 block li >> 2; M-block q adds 8 q columns.
  const int ayl = 2 * wv + ((li & 3) >> 1);
  const int axl = 2 * (li >> 2) + (li & 1);
  for (;;) {
  f4v acc[4][NBG];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int nb = 0; nb < NBG; ++nb) acc[q][nb] = f4v{0.f, 0.f, 0.f, 0.f};
  const int nextTile = tile + static_cast<int>(gridDim.x);
  int nn = 0, ny0 = 0, nx0 = 0;
  if (nextTile < nTiles) origin(nextTile, nn, ny0, nx0);

  for (int ch = 0; ch < NCH; ++ch) {
    if (ch > 0) __syncthreads();  // the previous chunk is consumed
    if constexpr (G == kGatherFirst) {
      if (a.srcDouble)
        stageFirst<double>(hiT, loT, CS, a, n, y0 - 1, x0 - 1);
      else
        stageFirst<float>(hiT, loT, CS, a, n, y0 - 1, x0 - 1);
    } else if (kConcat && ch < NCO) {
      coarseBuf.store(hiT, loT, CS);
      const int ch2 = (a.ph + 1) >> 1, cw2 = (a.pw + 1) >> 1, cy = (y0 >> 1) - 1,
                cx = (x0 >> 1) - 1;
      if (SIDES && decltype(coarseBuf)::hasSigns(ch2, cw2, cy, cx, a.pad))
        decltype(coarseBuf)::flipSides(hiT, loT, CS, ch2, cw2, cy, cx, a.pad);
    } else {
      stageBuf.store(hiT, loT, CS);
      if (SIDES && decltype(stageBuf)::hasSigns(a.ph, a.pw, y0 - 1, x0 - 1, a.pad))
        decltype(stageBuf)::flipSides(hiT, loT, CS, a.ph, a.pw, y0 - 1, x0 - 1, a.pad);
    }
    __syncthreads();
    if constexpr (G != kGatherFirst) {
      if (ch + 1 < NCH)
        loadChunk(ch + 1, n, y0, x0);
      else if (nextTile < nTiles)
        loadChunk(0, nn, ny0, nx0);  // the next tile's first chunk
    }
    const bool coarse = kConcat && ch < NCO;
    // Dirichlet top / right sides of odd fine length, and whether this tile
    // reaches the fine ghost row / column beyond them (uniform per tile)
    const bool oddTop = (a.pad & kPadTop) && (a.ph & 1);
    const bool oddRight = (a.pad & kPadRight) && (a.pw & 1);
    const bool oddEdgeTile = (oddTop && a.ph >= y0 - 1 && a.ph <= y0 + kTH) ||
                             (oddRight && a.pw >= x0 - 1 && a.pw <= x0 + kTW);
    (void)oddEdgeTile;
#pragma unroll 2
    for (int m = 0; m < MSC; ++m) {
      const int kk = 32 * m + 8 * g;
      u32x4 ah[4], al[4];
      if (kk < 9 * CC) {
        const int tap = kk / CC;
        const int c0 = kk - tap * CC;
        const int ky = tap / 3, kx = tap - 3 * ky;
        // fine: tile pixel (ayl + ky, axl + kx + 8 q); coarse: its (t + 1) >> 1
        const int base = coarse ? (((ayl + ky + 1) >> 1) * kCoCols + ((axl + kx + 1) >> 1)) * CS + c0
                                : ((ayl + ky) * kInCols + axl + kx) * CS + c0;
        const int qs = coarse ? 4 * CS : 8 * CS;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ah[q] = *reinterpret_cast<const u32x4*>(hiT + base + q * qs);
          al[q] = *reinterpret_cast<const u32x4*>(loT + base + q * qs);
        }
        if constexpr (SIDES && kConcat) {
          // the upsampled half one pixel beyond an odd fine edge maps to the
          // last coarse pixel inside, which the coarse staging keeps as it is;
          // across a Dirichlet side that fine ghost is its negation
          if (coarse && oddEdgeTile) {
            const int fy = y0 - 1 + ayl + ky;
            const bool flipY = oddTop && fy == a.ph;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int fx = x0 - 1 + axl + kx + 8 * q;
              if (flipY != (oddRight && fx == a.pw)) {
                ah[q] ^= u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
                al[q] ^= u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
              }
            }
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) ah[q] = al[q] = u32x4{0u, 0u, 0u, 0u};
      }
      const int wbase = ((ch * MSC + m) * NBG) * 64 + lane;
#pragma unroll
      for (int nb = 0; nb < NBG; ++nb) {
        const u32x4 bw = wl[wbase + nb * 64];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc[q][nb] = mfma(ah[q], bw, acc[q][nb]);
          acc[q][nb] = mfma(al[q], bw, acc[q][nb]);
        }
      }
    }
  }
  __syncthreads();  // the tile is dead; its LDS becomes the output staging

  // Lane holds D[4 g + r][li]: pixel (2 wv + (r >> 1), 8 q + 2 g + (r & 1)),
  // channel 16 (nb0 + nb) + li.
  const bool coarseGroup = (E == kEpiSplitUp) && (nb0 * 16 < SPLIT);
  if (E == kEpiStorePool || coarseGroup) {
    const int hh = a.h >> 1, hw = a.w_ >> 1;
    const int py = (y0 >> 1) + wv;
    constexpr int AUXC = E == kEpiStorePool ? COUT : SPLIT;
    // a 2 x 2 block that straddles an odd physical edge: the pixels beyond it
    // add nothing (the pool is then a quarter of the transposed cropped
    // upsampling, as inside, which keeps a symmetric network symmetric)
    const bool rowOut = y0 + 2 * wv + 1 >= a.ph;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int px = (x0 >> 1) + 4 * q + g;
      const bool colOut = x0 + 8 * q + 2 * g + 1 >= a.pw;
#pragma unroll
      for (int nb = 0; nb < NBG; ++nb) {
        f4v v = acc[q][nb];
        if (E == kEpiStorePool) {
          if (rowOut) v[2] = v[3] = 0.f;
          if (colOut) v[1] = v[3] = 0.f;
        }
        float s = (v[0] + v[1]) + (v[2] + v[3]);
        if (E == kEpiStorePool) s *= 0.25f;
        if (py < hh && px < hw)
          a.aux[((static_cast<int64_t>(n) * hh + py) * hw + px) * AUXC + 16 * (nb0 + nb) + li] = s;
      }
    }
  }
  if (!coarseGroup) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int nb = 0; nb < NBG; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int yl = 2 * wv + (r >> 1), xl = 8 * q + 2 * g + (r & 1);
          stage[(yl * kTW + xl) * SS + 16 * nb + li] = acc[q][nb][r];
        }
    __syncthreads();
    constexpr int OUTC = E == kEpiSplitUp ? COUT - SPLIT : COUT;
    const int cOff = E == kEpiSplitUp ? nb0 * 16 - SPLIT : nb0 * 16;
    constexpr int G4 = NBG * 4;
    for (int it = threadIdx.x; it < kTH * kTW * G4; it += blockDim.x) {
      const int c4 = it % G4;
      const int px = it / G4;
      const int yl = px / kTW, xl = px - yl * kTW;
      const int y = y0 + yl, x = x0 + xl;
      if (y < a.h && x < a.w_)
        *reinterpret_cast<fx4*>(a.out + ((static_cast<int64_t>(n) * a.h + y) * a.w_ + x) * OUTC +
                                cOff + c4 * 4) =
            *reinterpret_cast<const fx4*>(stage + px * SS + c4 * 4);
    }
  }
  tile = nextTile;
  if (tile >= nTiles) break;
  n = nn;
  y0 = ny0;
  x0 = nx0;
  __syncthreads();  // the staged output is read; the region takes the next tile
  }
}

// Eight BF16 values p[0], p[stride], ..., p[7 stride] packed into a fragment.
__device__ inline u32x4 pack8(const bf16* p, int stride) {
  u32x4 r;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    r[k] = static_cast<uint32_t>(p[(2 * k) * stride]) |
           (static_cast<uint32_t>(p[(2 * k + 1) * stride]) << 16);
  return r;
}

template <int CIN, int CA, int COUT, int GX, int GG, int PPW>
__global__ __launch_bounds__(256) void dwKernel(DwArgs a) {
  constexpr int CS = CIN + 8, GS = COUT + 8;
  constexpr int NCB = CIN / 16, NOB = COUT / 16, NPAIR = 9 * NCB;
  __shared__ __attribute__((aligned(16))) bf16 xs[kInPix * CS];
  __shared__ __attribute__((aligned(16))) bf16 gs[kTH * kTW * GS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, g = lane >> 4;
  const int pair0 = (blockIdx.y * 4 + wv) * PPW;
  f4v acc[PPW][NOB];
#pragma unroll
  for (int p = 0; p < PPW; ++p)
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob) acc[p][ob] = f4v{0.f, 0.f, 0.f, 0.f};
  int tapY[PPW], tapX[PPW], cbk[PPW];
#pragma unroll
  for (int p = 0; p < PPW; ++p) {
    const int pair = pair0 + p < NPAIR ? pair0 + p : 0;
    const int tap = pair / NCB;
    cbk[p] = pair - tap * NCB;
    tapY[p] = tap / 3;
    tapX[p] = tap - 3 * tapY[p];
  }
  for (int64_t t = blockIdx.x; t < a.tiles; t += gridDim.x) {
    int64_t tile = t;
    const int tx = static_cast<int>(tile % a.tilesX);
    tile /= a.tilesX;
    const int ty = static_cast<int>(tile % a.tilesY);
    const int n = static_cast<int>(tile / a.tilesY);
    const int y0 = ty * kTH, x0 = tx * kTW;
    __syncthreads();
    if constexpr (CIN <= 64) {
      stageSplit<CIN, CA, GX, CIN, kInRows, kInCols>(xs, nullptr, CS, a.xa, a.xb, n, a.h, a.w_,
                                                     y0 - 1, x0 - 1, 0, a.pad);
    } else {  // two halves keep the register batch small
      stageSplit<CIN, CA, GX, CIN / 2, kInRows, kInCols>(xs, nullptr, CS, a.xa, a.xb, n, a.h,
                                                         a.w_, y0 - 1, x0 - 1, 0, a.pad);
      stageSplit<CIN, CA, GX, CIN / 2, kInRows, kInCols>(xs + CIN / 2, nullptr, CS, a.xa, a.xb, n,
                                                         a.h, a.w_, y0 - 1, x0 - 1, CIN / 2,
                                                         a.pad);
    }
    // the output-gradient tile has no halo: zero beyond the image
    stageSplit<COUT, COUT, GG, COUT, kTH, kTW>(gs, nullptr, GS, a.ga, a.gb, n, a.h, a.w_, y0, x0,
                                               0, 0);
    __syncthreads();
    for (int s = 0; s < kTH; ++s) {
      u32x4 bfr[NOB];
#pragma unroll
      for (int ob = 0; ob < NOB; ++ob) bfr[ob] = pack8(gs + (s * kTW + 8 * g) * GS + 16 * ob + li, GS);
#pragma unroll
      for (int p = 0; p < PPW; ++p) {
        if (pair0 + p >= NPAIR) continue;  // uniform across the wave
        const u32x4 afr =
            pack8(xs + ((s + tapY[p]) * kInCols + 8 * g + tapX[p]) * CS + 16 * cbk[p] + li, CS);
#pragma unroll
        for (int ob = 0; ob < NOB; ++ob) acc[p][ob] = mfma(afr, bfr[ob], acc[p][ob]);
      }
    }
  }
#pragma unroll
  for (int p = 0; p < PPW; ++p) {
    const int pair = pair0 + p;
    if (pair >= NPAIR) continue;
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        a.partial[((static_cast<int64_t>(blockIdx.x) * NPAIR + pair) * NOB + ob) * 256 +
                  (4 * g + r) * 16 + li] = acc[p][ob][r];
  }
}

// Resident blocks of a kernel on the whole device (occupancy x compute units).
template <class K>
int residentBlocks(K kernel) {
  int dev = 0, cus = 0, perCu = 0;
  CFDHIP_CHECK(hipGetDevice(&dev));
  CFDHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  CFDHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, 256, 0));
  return std::max(1, cus) * std::max(1, perCu);
}

template <int CIN, int CA, int COUT, int NBG, int G, int E, int SPLIT, bool SIDES>
void convWith(hipStream_t s, const ConvArgs& a) {
  static const int resident = residentBlocks(convKernel<CIN, CA, COUT, NBG, G, E, SPLIT, SIDES>);
  constexpr int groups = COUT / (16 * NBG);
  const int64_t tiles = static_cast<int64_t>(a.tilesX) * a.tilesY * a.n;
  require(tiles < (int64_t{1} << 31), "the network batch has too many tiles");
  const int64_t perGroup = std::max(1, resident / groups);
  const dim3 grid(static_cast<unsigned>(std::min(tiles, perGroup)), static_cast<unsigned>(groups));
  hipLaunchKernelGGL((convKernel<CIN, CA, COUT, NBG, G, E, SPLIT, SIDES>), grid, dim3(256), 0, s,
                     a);
}

template <int CIN, int CA, int COUT, int NBG, int G, int E, int SPLIT>
void conv(hipStream_t s, const ConvArgs& a) {
  if ((a.pad & kPadNeumann) && (a.pad >> 1) != 0)
    convWith<CIN, CA, COUT, NBG, G, E, SPLIT, true>(s, a);
  else
    convWith<CIN, CA, COUT, NBG, G, E, SPLIT, false>(s, a);
}

constexpr int kPpw = 2;

template <int CIN, int CA, int COUT, int GX, int GG>
int dw(hipStream_t s, const DwArgs& a, int splits) {
  constexpr int NPAIR = 9 * (CIN / 16);
  constexpr int groups = (NPAIR + 4 * kPpw - 1) / (4 * kPpw);
  const dim3 grid(static_cast<unsigned>(splits), static_cast<unsigned>(groups));
  hipLaunchKernelGGL((dwKernel<CIN, CA, COUT, GX, GG, kPpw>), grid, dim3(256), 0, s, a);
  return NPAIR;
}

}  // namespace

// Every (cin, ca, cout, gather, epilogue, split) the network uses. The
// N-group width (NBG 16-channel blocks per workgroup) keeps the staged weights
// and the accumulators small and never straddles the split of a transposed
// concat layer.
void launchConv(hipStream_t s, int cin, int ca, int cout, int gather, int epi, int split,
                const ConvArgs& a) {
#define CFDHIP_CONV(CI, CAV, CO, NBG, G, E, SP)                                            \
  if (cin == CI && (G != kGatherConcatUp || ca == CAV) && cout == CO && gather == G &&     \
      epi == E && (E != kEpiSplitUp || split == SP)) {                                     \
    conv<CI, CAV, CO, NBG, G, E, SP>(s, a);                                                \
    return;                                                                                \
  }
  // forward
  CFDHIP_CONV(16, 16, 16, 1, kGatherFirst, kEpiStorePool, 0)
  CFDHIP_CONV(16, 16, 16, 1, kGatherPlain, kEpiStorePool, 0)
  CFDHIP_CONV(16, 16, 32, 2, kGatherPlain, kEpiStore, 0)
  CFDHIP_CONV(32, 32, 32, 2, kGatherPlain, kEpiStorePool, 0)
  CFDHIP_CONV(32, 32, 64, 2, kGatherPlain, kEpiStore, 0)
  CFDHIP_CONV(64, 64, 64, 2, kGatherPlain, kEpiStorePool, 0)
  CFDHIP_CONV(64, 64, 64, 2, kGatherPlain, kEpiStore, 0)
  CFDHIP_CONV(128, 64, 64, 2, kGatherConcatUp, kEpiStore, 0)
  CFDHIP_CONV(96, 64, 32, 2, kGatherConcatUp, kEpiStore, 0)
  CFDHIP_CONV(32, 32, 32, 2, kGatherPlain, kEpiStore, 0)
  CFDHIP_CONV(48, 32, 16, 1, kGatherConcatUp, kEpiStore, 0)
  CFDHIP_CONV(16, 16, 16, 1, kGatherPlain, kEpiStore, 0)
  // transposed (gradient of the input)
  CFDHIP_CONV(16, 16, 48, 1, kGatherPlain, kEpiSplitUp, 32)
  CFDHIP_CONV(32, 32, 96, 2, kGatherPlain, kEpiSplitUp, 64)
  CFDHIP_CONV(64, 64, 128, 2, kGatherPlain, kEpiSplitUp, 64)
  CFDHIP_CONV(64, 64, 64, 2, kGatherSumUp, kEpiStore, 0)
  CFDHIP_CONV(64, 64, 32, 2, kGatherPlain, kEpiStore, 0)
  CFDHIP_CONV(32, 32, 32, 2, kGatherSumUp, kEpiStore, 0)
  CFDHIP_CONV(32, 32, 16, 1, kGatherPlain, kEpiStore, 0)
  CFDHIP_CONV(16, 16, 16, 1, kGatherSumUp, kEpiStore, 0)
#undef CFDHIP_CONV
  throw Error(CFDHIP_ERR_UNSUPPORTED, "no convolution kernel for this layer shape");
}

int launchConvDw(hipStream_t s, int cin, int ca, int cout, int gx, int gg, const DwArgs& a,
                 int splits) {
#define CFDHIP_DW(CI, CAV, CO, GX, GG)                                                      \
  if (cin == CI && (GX != kGatherConcatUp || ca == CAV) && cout == CO && gx == GX && gg == GG) \
    return dw<CI, CAV, CO, GX, GG>(s, a, splits);
  CFDHIP_DW(16, 16, 16, kGatherPlain, kGatherPlain)     // d0b
  CFDHIP_DW(48, 32, 16, kGatherConcatUp, kGatherPlain)  // d0a
  CFDHIP_DW(32, 32, 32, kGatherPlain, kGatherPlain)     // d1b
  CFDHIP_DW(96, 64, 32, kGatherConcatUp, kGatherPlain)  // d1a
  CFDHIP_DW(64, 64, 64, kGatherPlain, kGatherPlain)     // d2b, b3b, b3a
  CFDHIP_DW(128, 64, 64, kGatherConcatUp, kGatherPlain) // d2a
  CFDHIP_DW(64, 64, 64, kGatherPlain, kGatherSumUp)     // e2b
  CFDHIP_DW(32, 32, 64, kGatherPlain, kGatherPlain)     // e2a
  CFDHIP_DW(32, 32, 32, kGatherPlain, kGatherSumUp)     // e1b
  CFDHIP_DW(16, 16, 32, kGatherPlain, kGatherPlain)     // e1a
  CFDHIP_DW(16, 16, 16, kGatherPlain, kGatherSumUp)     // e0b
#undef CFDHIP_DW
  throw Error(CFDHIP_ERR_UNSUPPORTED, "no weight-gradient kernel for this layer shape");
}

}  // namespace cfdhip
