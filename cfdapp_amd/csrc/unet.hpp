// Learned p' preconditioner: a linear 4-level convolutional U-Net.
//
// All convolutions are 3x3 with zero padding and no bias; there are no
// activations, so the network is a fixed linear map r -> z and flexible CG
// sees a fixed preconditioner up to BF16 rounding.
//
//   e0a = conv(x, 1->16)            e0  = conv(e0a, 16->16)     (H x W)
//   e1a = conv(pool e0, 16->32)     e1  = conv(e1a, 32->32)     (H/2)
//   e2a = conv(pool e1, 32->64)     e2  = conv(e2a, 64->64)     (H/4)
//   b3a = conv(pool e2, 64->64)     b3  = conv(b3a, 64->64)     (H/8)
//   d2a = conv([up b3, e2], 128->64) d2 = conv(d2a, 64->64)
//   d1a = conv([up d2, e1], 96->32)  d1 = conv(d1a, 32->32)
//   d0a = conv([up d1, e0], 48->16)  d0 = conv(d0a, 16->16)
//   y   = conv(d0, 16->1)
//
// pool = 2x2 average, up = nearest x2, [a, b] = channel concatenation.
// Activations: FP32, NHWC, batch-major. Master weights: FP32,
// W[ky][kx][cin][cout] per layer, layers concatenated in the order above.
// The MFMA layers run from BF16 copies packed in the operand order of
// v_mfma_f32_16x16x32_bf16 (one 16-byte fragment per lane); activations enter
// the MFMA as a hi + lo pair of BF16 values (two products per weight), which
// keeps about 16 bits of each activation: the preconditioner's output must be
// accurate in the A-norm, and BF16-rounded activations are not (their rounding
// noise at the fine levels outweighs the smooth part of z on large grids).
#pragma once

#include <cstdint>

#include "context.hpp"

namespace cfdhip {

using bf16 = uint16_t;

enum UnetLayer : int {
  kE0a = 0, kE0b, kE1a, kE1b, kE2a, kE2b, kB3a, kB3b,
  kD2a, kD2b, kD1a, kD1b, kD0a, kD0b, kOut, kNumLayers
};

struct LayerShape {
  int cin, cout;
};

constexpr LayerShape kLayers[kNumLayers] = {
    {1, 16},  {16, 16}, {16, 32}, {32, 32}, {32, 64}, {64, 64}, {64, 64}, {64, 64},
    {128, 64}, {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}, {16, 1}};

constexpr int64_t layerWeights(int l) { return 9LL * kLayers[l].cin * kLayers[l].cout; }
constexpr int64_t layerOffset(int l) {
  int64_t o = 0;
  for (int i = 0; i < l; ++i) o += layerWeights(i);
  return o;
}
constexpr int64_t kUnetParams = layerOffset(kNumLayers);

// The K dimension (9 taps x cin) is walked in chunks of kChunk(cin) input
// channels (one LDS tile each); within a chunk k = tap * chunk + c in steps of
// 32. Packed MFMA fragments: chunks x steps x (cout / 16) N-blocks x 64 lanes
// x 8 BF16.
constexpr int kChunk(int cin) { return cin == 48 ? 16 : cin > 32 ? 32 : cin; }
constexpr int64_t packedSteps(int cin) {
  return (cin / kChunk(cin)) * ((9LL * kChunk(cin) + 31) / 32);
}
constexpr int64_t packedElems(int cin, int cout) { return packedSteps(cin) * (cout / 16) * 512; }

// Gathers applied while a tile of the conv input is staged in LDS.
enum Gather : int {
  kGatherPlain = 0,   // channels of a at (y, x)
  kGatherConcatUp,    // [a at (y/2, x/2) (ca channels), b at (y, x)]
  kGatherSumUp,       // a at (y, x) + 0.25 b at (y/2, x/2)  (gradient of pool + skip)
  kGatherFirst        // conv(x, 1 -> 16) of the scalar input, computed while staging
};
// What the conv does with its FP32 accumulators.
enum Epilogue : int {
  kEpiStore = 0,   // BF16 to out
  kEpiStorePool,   // BF16 to out, 2x2 average to aux (H/2 x W/2)
  kEpiSplitUp      // channels < split: 2x2 sum to aux (coarse); the rest to out
};

struct ConvArgs {
  const float* a;
  const float* b;
  const void* w;  // packed fragments of the layer
  float* out;
  float* aux;
  int n, h, w_;   // batch, image size of the output (= input) grid
  int tilesX, tilesY;
  int nbTotal;    // N-blocks of the layer (cout / 16)
  // kGatherFirst: x = src * scale on srcH x srcW (row stride srcW, zero
  // outside), first-layer master weights w0[tap][16]
  const void* src;
  const float* scale;
  const float* w0;
  int srcH, srcW, srcDouble;
  // padding (padSign): 0 zero; bit 0 Neumann (replicate) padding, a tap
  // outside the image reads the nearest pixel inside it; bits 1-4 (with bit
  // 0) Dirichlet sides (left, right, bottom, top) where it reads the negated
  // nearest pixel (odd reflection: zero at the face)
  int pad;
  // the image's physical extent on this level (ph <= h, pw <= w): a grid
  // whose sides are not multiples of 8 sits in the top-left corner of the
  // padded h x w tensors, and the padding reflects at its edges, not at the
  // tensors' (pixels beyond it are computed but read by no live pixel)
  int ph, pw;
};

struct DwArgs {
  const float* xa;
  const float* xb;
  const float* ga;
  const float* gb;
  float* partial;  // [split][pair][ob][16 x 16]
  int n, h, w_;
  int tilesX, tilesY;
  int64_t tiles;
  int pad;         // padding of the staged input tile (as ConvArgs::pad)
};

// Pad bits (ConvArgs::pad): Neumann padding, and the Dirichlet sides.
enum PadBits : int {
  kPadNeumann = 1, kPadLeft = 2, kPadRight = 4, kPadBottom = 8, kPadTop = 16
};

// Pixel (y, x) of a staged window and the factor its value takes. Inside
// the h x w image: 1. Zero padding: 0 outside. Neumann padding: (y, x) moves
// to the nearest pixel inside, the factor is -1 per crossed Dirichlet side
// (y < 0 bottom, y >= h top, x < 0 left, x >= w right), else 1. A conv with
// this padding and a kernel that is mirror-symmetric in both axes
// (W[ky][kx] == W[2-ky][kx] == W[ky][2-kx] per channel pair) has as its
// adjoint the conv with the transposed kernel and the same padding, so the
// backward pass keeps the same staging; the host holds training to such
// kernels when the padding is not zero.
__device__ inline float padSign(int& y, int& x, int h, int w, int pad) {
  if (y >= 0 && y < h && x >= 0 && x < w) return 1.f;
  if (!(pad & kPadNeumann)) return 0.f;
  float s = 1.f;
  if (y < 0) { y = 0; if (pad & kPadBottom) s = -s; }
  else if (y >= h) { y = h - 1; if (pad & kPadTop) s = -s; }
  if (x < 0) { x = 0; if (pad & kPadLeft) s = -s; }
  else if (x >= w) { x = w - 1; if (pad & kPadRight) s = -s; }
  return s;
}

// The coarse tail below the network's deepest level: a symmetric V(1,1)
// multigrid cycle on one channel, Neumann padding throughout, levels H/16 ..
// down to the coarsest grid with even sides (at most `levels` of them):
//   r_{t+1} = 4 pool(K r_t - omega K A r_t)   (restriction of the residual
//             after one Jacobi sweep from zero),
//   e_t     = (I - omega A)(omega r_t + K up(e_{t+1})) + omega r_t,
//   coarsest: the same without the coarse term, e = (I - omega A) omega r + omega r.
// Applied to channel `src` of the H/8 x W/8 bottom tensor (b3a), the tail
// writes K up(e_0) (the prolonged coarse correction) into channel `dst`;
// its adjoint (the backward pass) adds the same operator applied to
// channel dst into channel src and clears dst. A and K are symmetric 3 x 3
// stencils (centre, edge, corner).
struct TailParams {
  int levels;         // 0: no tail
  float omega;        // damped-Jacobi weight
  float a[3];         // level operator A: centre, edge, corner
  float k[3];         // transfer kernel K: centre, edge, corner
  int src, dst;       // channels of the bottom tensor
};

// One convolution (forward or transposed) on the device.
void launchConv(hipStream_t s, int cin, int ca, int cout, int gather, int epi, int split,
                const ConvArgs& a);
// Weight gradient of an MFMA layer: partial sums over pixel splits.
int launchConvDw(hipStream_t s, int cin, int ca, int cout, int gx, int gg, const DwArgs& a,
                 int splits);

}  // namespace cfdhip
