// Coarse tail of the learned preconditioner (TailParams in unet.hpp): a
// one-channel symmetric V(1,1) cycle below the network's bottom level, with
// Neumann padding, so that the preconditioner reaches the coarsest grid.
//
// Launches per application: the first restriction of the input channel of
// the bottom tensor (all coarse pixels in parallel, read in place); the
// passes of the large levels, each over the whole chip; the small levels in
// one workgroup per image (a workgroup barrier between dependent passes;
// the levels stay in L2); and a scatter of the prolonged correction
// K up(e_0) into the output channel. The operator is symmetric (restriction
// 4 pool(K .) is the transpose of prolongation K up(.), the same Jacobi sweep
// before and after the coarse correction), so the backward pass runs the
// same launches from the output channel's gradient into the input channel's.
#include "unet_net.hpp"

namespace cfdhip {

namespace {

constexpr int kTailThreads = 1024;
// a tail level with more pixels than this runs outside the one-workgroup cycle
constexpr int kTailCyclePixels = 4096;

dim3 tailGrid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : (b > 8192 ? 8192 : b)));
}

// One image's tail workspace: per level t the residual R_t, a work image
// S_t and the correction E_t. Level t is the (t + 1)-th halving of the bottom
// grid, rounding up: a side of odd length keeps a last coarse pixel that
// covers one fine row (column) only; the restriction then sums the fine
// pixels it covers and the prolongation drops the copies beyond the fine
// edge (restriction = prolongation transposed, as on even sides).
struct TailShape {
  int h3, w3, levels;
  int hs[16], ws[16];
  int64_t perImage;
  int64_t off[3 * 16];  // R_t, S_t, E_t offsets
};

__host__ __device__ inline int tailH(const TailShape& s, int t) { return s.hs[t]; }
__host__ __device__ inline int tailW(const TailShape& s, int t) { return s.ws[t]; }

TailShape tailShape(int h3, int w3, int maxLevels) {
  TailShape s{};
  s.h3 = h3;
  s.w3 = w3;
  int hh = h3, ww = w3, L = 0;
  while (L < maxLevels && L < 16 && hh >= 2 && ww >= 2) {
    hh = (hh + 1) / 2;
    ww = (ww + 1) / 2;
    s.hs[L] = hh;
    s.ws[L] = ww;
    ++L;
  }
  s.levels = L;
  int64_t o = 0;
  for (int t = 0; t < L; ++t) {
    const int64_t m = static_cast<int64_t>(s.hs[t]) * s.ws[t];
    for (int q = 0; q < 3; ++q) {
      s.off[3 * t + q] = o;
      o += m;
    }
  }
  s.perImage = o;
  return s;
}

// The shape copied from the kernel arguments into LDS by the whole block.
// Read with a run-time level index, a kernel-argument struct is fetched
// through the scalar cache from the argument buffer, a cache line at a
// time and again for every level a pass visits (measured: about 1.6 us per
// pass of the one-workgroup cycle even on a 1 x 1 level); staged once, every
// later read is an LDS read.
__device__ inline void stageShape(const TailShape& s, TailShape& sh) {
  static_assert(sizeof(TailShape) % 4 == 0, "the shape is copied in 4-byte words");
  constexpr int kWords = static_cast<int>(sizeof(TailShape) / 4);
  const int* src = reinterpret_cast<const int*>(&s);
  int* dst = reinterpret_cast<int*>(&sh);
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// Pixel (y, x) of an h x w image (pixel stride cs) with the network's
// padding (padSign).
__device__ inline float at(const float* img, int h, int w, int y, int x, int pad, int cs = 1) {
  const float s = padSign(y, x, h, w, pad);
  return s * img[(static_cast<int64_t>(y) * w + x) * cs];
}

// The symmetric 3 x 3 stencil (centre, edge, corner) over a 3 x 3 patch.
__device__ inline float stencilPatch(const float (&q)[3][3], float c, float e, float k) {
  const float edges = (q[0][1] + q[2][1]) + (q[1][0] + q[1][2]);
  const float corners = (q[0][0] + q[0][2]) + (q[2][0] + q[2][2]);
  return c * q[1][1] + e * edges + k * corners;
}

// Symmetric 3 x 3 stencil at (y, x); pixels off the image through the
// padding, interior pixels by direct loads.
__device__ inline float stencil3(const float* img, int h, int w, int y, int x, float c, float e,
                                 float k, int pad) {
  float q[3][3];
  if (y > 0 && y + 1 < h && x > 0 && x + 1 < w) {
    const float* r = img + static_cast<int64_t>(y - 1) * w + (x - 1);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) q[a][b] = r[a * w + b];
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) q[a][b] = at(img, h, w, y - 1 + a, x - 1 + b, pad);
  }
  return stencilPatch(q, c, e, k);
}

// 4 pool(K v) at coarse pixel (y, x) of v (fine h x w, pixel stride cs):
// the four stencils of the 2 x 2 block read from one 4 x 4 patch.
__device__ inline float restrictK(const float* v, int h, int w, int y, int x, const TailParams& p,
                                  int pad, int cs = 1) {
  float t[4][4];
  const int y0 = 2 * y - 1, x0 = 2 * x - 1;
  if (y0 >= 0 && y0 + 3 < h && x0 >= 0 && x0 + 3 < w) {
    const float* r = v + (static_cast<int64_t>(y0) * w + x0) * cs;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) t[a][b] = r[static_cast<int64_t>(a * w + b) * cs];
  } else {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) t[a][b] = at(v, h, w, y0 + a, x0 + b, pad, cs);
  }
  float s = 0.f;
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
      if (2 * y + dy >= h || 2 * x + dx >= w) continue;  // beyond an odd fine edge
      float q[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) q[a][b] = t[dy + a][dx + b];
      s += stencilPatch(q, p.k[0], p.k[1], p.k[2]);
    }
  return s;  // 4 x the 2 x 2 average
}

// K up(e) at fine pixel (y, x) of the h x w grid (e on the rounded-up
// halves ceil(h / 2) x ceil(w / 2)).
__device__ inline float prolongK(const float* e, int h, int w, int y, int x, const TailParams& p,
                                 int pad) {
  const int wc = (w + 1) >> 1;
  float q[3][3];
  if (y > 0 && y + 1 < h && x > 0 && x + 1 < w) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        q[a][b] = e[static_cast<int64_t>((y - 1 + a) >> 1) * wc + ((x - 1 + b) >> 1)];
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        int yy = y - 1 + a, xx = x - 1 + b;  // the upsampled image with the padding
        const float sg = padSign(yy, xx, h, w, pad);
        q[a][b] = sg * e[static_cast<int64_t>(yy >> 1) * wc + (xx >> 1)];
      }
  }
  return stencilPatch(q, p.k[0], p.k[1], p.k[2]);
}

// R_0 = 4 pool(K x) of the input channel, all coarse pixels of all images in
// parallel, read in place from the C-channel bottom tensor.
__global__ __launch_bounds__(256) void tailRestrictKernel(const float* __restrict__ t, int C, int ch,
                                                          TailShape s, TailParams p, int pad,
                                                          int64_t pix, float* __restrict__ work) {
  const int h0 = tailH(s, 0), w0 = tailW(s, 0);
  const int64_t hw0 = static_cast<int64_t>(h0) * w0, hw = static_cast<int64_t>(s.h3) * s.w3;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < pix;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = q / hw0, i = q - b * hw0;
    const float* img = t + b * hw * C + ch;
    work[b * s.perImage + s.off[0] + i] =
        restrictK(img, s.h3, s.w3, static_cast<int>(i / w0), static_cast<int>(i % w0), p, pad, C);
  }
}

// The four passes of a tail level t, at pixel i of its output image:
//   kPre:      S_t = R_t - omega A R_t (the pre-sweep from zero);
//   kRestrict: R_{t+1} = 4 pool(K S_t) (i on level t+1);
//   kProlong:  S_t = omega R_t + K up(E_{t+1});
//   kPost:     E_t = S_t - omega A S_t + omega R_t (the post-sweep).
enum TailPass { kPre = 0, kRestrict = 1, kProlong = 2, kPost = 3 };

// base + s.off[k] - shift is level array k (shift: where base's level
// arrays start in the workspace layout).
__device__ inline void tailPass(float* base, const TailShape& s, const TailParams& p, int pad,
                                int t, int pass, int i, int64_t shift = 0) {
  const float om = p.omega;
  float* R = base + (s.off[3 * t] - shift);
  float* S = base + (s.off[3 * t + 1] - shift);
  float* E = base + (s.off[3 * t + 2] - shift);
  const int h = tailH(s, t), w = tailW(s, t);
  if (pass == kRestrict) {
    const int wc = tailW(s, t + 1);
    base[s.off[3 * (t + 1)] - shift + i] = restrictK(S, h, w, i / wc, i % wc, p, pad);
    return;
  }
  const int y = i / w, x = i % w;
  if (pass == kPre) {
    S[i] = R[i] - om * stencil3(R, h, w, y, x, p.a[0], p.a[1], p.a[2], pad);
  } else if (pass == kProlong) {
    S[i] = om * R[i] + prolongK(base + (s.off[3 * (t + 1) + 2] - shift), h, w, y, x, p, pad);
  } else {
    E[i] = S[i] - om * stencil3(S, h, w, y, x, p.a[0], p.a[1], p.a[2], pad) + om * R[i];
  }
}

__host__ __device__ inline int tailPassPixels(const TailShape& s, int t, int pass) {
  const int u = pass == kRestrict ? t + 1 : t;
  return tailH(s, u) * tailW(s, u);
}

// One pass of a large level over all its pixels of all images.
__global__ __launch_bounds__(256) void tailLevelKernel(TailShape sArg, TailParams p, int pad, int t,
                                                       int pass, int64_t pix,
                                                       float* __restrict__ work) {
  __shared__ TailShape s;
  stageShape(sArg, s);
  const int64_t m = tailPassPixels(s, t, pass);
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < pix;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = q / m;
    tailPass(work + b * s.perImage, s, p, pad, t, pass, static_cast<int>(q - b * m));
  }
}

// Levels t0 .. L-1 of one image: down from R_t0, the coarsest level, up to
// E_t0; base holds the image's levels at the offsets of s.
// (y, x) moved into the h x w image through the Neumann padding and the
// factor the pixel there takes (-1 per crossed Dirichlet side): padSign
// without branches, so a thread issues all of a stencil's reads at once.
__device__ inline float reflect(int& y, int& x, int h, int w, int pad) {
  float s = 1.f;
  s = (y < 0 && (pad & kPadBottom)) ? -s : s;
  s = (y >= h && (pad & kPadTop)) ? -s : s;
  s = (x < 0 && (pad & kPadLeft)) ? -s : s;
  s = (x >= w && (pad & kPadRight)) ? -s : s;
  y = min(max(y, 0), h - 1);
  x = min(max(x, 0), w - 1);
  return s;
}

// The 3 x 3 patch of img (h x w) around (y, x), padded.
__device__ inline void patch3(const float* img, int h, int w, int y, int x, int pad,
                              float (&q)[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      int yy = y - 1 + a, xx = x - 1 + b;
      const float sg = reflect(yy, xx, h, w, pad);
      q[a][b] = sg * img[yy * w + xx];
    }
}

// Every pixel of an h x w level once: wave v takes rows v, v + waves, ...,
// lane l columns l, l + 64, ... (no division per pixel).
template <class F>
__device__ inline void forPixels(int h, int w, F&& f) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  for (int y = wave; y < h; y += waves)
    for (int x = lane; x < w; x += 64) f(y, x);
}

// Levels t0 .. L-1 of one image: down from R_t0, the coarsest level, up to
// E_t0; base holds the image's levels at the offsets of s (minus shift).
// The same arithmetic, in the same order, as tailPass.
__device__ inline void cycleLevels(float* base, const TailShape& s, const TailParams& p, int pad,
                                   int t0, int64_t shift) {
  const int L = s.levels;
  const float om = p.omega;
  const float a0 = p.a[0], a1 = p.a[1], a2 = p.a[2];
  const float k0 = p.k[0], k1 = p.k[1], k2 = p.k[2];
  auto level = [&](int t, int q) { return base + (s.off[3 * t + q] - shift); };
  for (int t = t0; t + 1 < L; ++t) {
    const int h = tailH(s, t), w = tailW(s, t);
    const float* R = level(t, 0);
    float* S = level(t, 1);
    // S_t = R_t - omega A R_t
    forPixels(h, w, [&](int y, int x) {
      float q[3][3];
      patch3(R, h, w, y, x, pad, q);
      S[y * w + x] = R[y * w + x] - om * stencilPatch(q, a0, a1, a2);
    });
    __syncthreads();
    // R_{t+1} = 4 pool(K S_t): the K stencils of the 2 x 2 fine block from
    // one 4 x 4 patch; fine pixels beyond an odd edge add nothing
    const int hc = tailH(s, t + 1), wc = tailW(s, t + 1);
    float* Rc = level(t + 1, 0);
    forPixels(hc, wc, [&](int y, int x) {
      float v[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          int yy = 2 * y - 1 + a, xx = 2 * x - 1 + b;
          const float sg = reflect(yy, xx, h, w, pad);
          v[a][b] = sg * S[yy * w + xx];
        }
      float acc = 0.f;
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          if (2 * y + dy >= h || 2 * x + dx >= w) continue;
          float q[3][3];
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) q[a][b] = v[dy + a][dx + b];
          acc += stencilPatch(q, k0, k1, k2);
        }
      Rc[y * wc + x] = acc;
    });
    __syncthreads();
  }
  // coarsest level: the two sweeps without a coarse correction,
  // e = omega r + (I - omega A) omega r
  {
    const int h = tailH(s, L - 1), w = tailW(s, L - 1);
    const float* R = level(L - 1, 0);
    float* E = level(L - 1, 2);
    forPixels(h, w, [&](int y, int x) {
      float q[3][3];
      patch3(R, h, w, y, x, pad, q);
      const float ar = stencilPatch(q, a0, a1, a2);
      E[y * w + x] = 2.f * om * R[y * w + x] - om * om * ar;
    });
  }
  __syncthreads();
  for (int t = L - 2; t >= t0; --t) {
    const int h = tailH(s, t), w = tailW(s, t);
    const int wc = tailW(s, t + 1);
    const float* R = level(t, 0);
    float* S = level(t, 1);
    float* E = level(t, 2);
    const float* Ec = level(t + 1, 2);
    // S_t = omega R_t + K up(E_{t+1})
    forPixels(h, w, [&](int y, int x) {
      float q[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          int yy = y - 1 + a, xx = x - 1 + b;
          const float sg = reflect(yy, xx, h, w, pad);
          q[a][b] = sg * Ec[(yy >> 1) * wc + (xx >> 1)];
        }
      S[y * w + x] = om * R[y * w + x] + stencilPatch(q, k0, k1, k2);
    });
    __syncthreads();
    // E_t = S_t - omega A S_t + omega R_t
    forPixels(h, w, [&](int y, int x) {
      float q[3][3];
      patch3(S, h, w, y, x, pad, q);
      E[y * w + x] = S[y * w + x] - om * stencilPatch(q, a0, a1, a2) + om * R[y * w + x];
    });
    __syncthreads();
  }
}

// One image per workgroup, its levels in the global workspace.
__global__ __launch_bounds__(kTailThreads) void tailCycleKernel(TailShape sArg, TailParams p,
                                                               int pad, int t0,
                                                               float* __restrict__ work) {
  __shared__ TailShape s;
  stageShape(sArg, s);
  cycleLevels(work + static_cast<int64_t>(blockIdx.x) * s.perImage, s, p, pad, t0, 0);
}

// The same with levels t0 .. L-1 staged in LDS (R_t0 in, E_t0 out): every
// pass then waits on LDS instead of L2. Levels t0.. hold at most
// 3 x 4/3 x kTailCyclePixels floats.
constexpr int kTailLdsFloats = 16384;
__global__ __launch_bounds__(kTailThreads) void tailCycleLdsKernel(TailShape sArg, TailParams p,
                                                                  int pad, int t0,
                                                                  float* __restrict__ work) {
  __shared__ float lds[kTailLdsFloats];  // with the shape: 64.5 KiB (gfx950: 160 KiB per block)
  __shared__ TailShape s;
  stageShape(sArg, s);
  float* g = work + static_cast<int64_t>(blockIdx.x) * s.perImage;
  const int64_t o0 = s.off[3 * t0];  // LDS holds the workspace from o0 on
  const int m0 = tailH(s, t0) * tailW(s, t0);
  for (int i = threadIdx.x; i < m0; i += blockDim.x) lds[i] = g[o0 + i];  // R_t0 at offset 0
  __syncthreads();
  cycleLevels(lds, s, p, pad, t0, o0);
  const int64_t e0 = s.off[3 * t0 + 2];
  for (int i = threadIdx.x; i < m0; i += blockDim.x) g[e0 + i] = lds[e0 - o0 + i];
}

// The whole tail of one case of a batch per workgroup: R_0 = 4 pool(K x) of
// the source channel of the case's bottom tensor into LDS, the cycle over
// every level in LDS, and the destination channel = K up(E_0) -- the
// launches of tailApply for one image, in one kernel (its levels fit LDS).
__global__ __launch_bounds__(kTailThreads) void tailBatchKernel(const TailShape* __restrict__ shapes,
                                                               const NarrowCase* __restrict__ cases,
                                                               float* __restrict__ act, int C,
                                                               int srcCh, int dstCh, TailParams p) {
  __shared__ float lds[kTailLdsFloats];
  __shared__ TailShape s;
  stageShape(shapes[blockIdx.x], s);
  const NarrowCase& c = cases[blockIdx.x];
  float* b3a = act + c.act + c.tOff[kTb3a];
  const int pad = c.pad;
  if (s.levels > 0) {
    const int h0 = tailH(s, 0), w0 = tailW(s, 0);
    forPixels(h0, w0, [&](int y, int x) {
      lds[s.off[0] + y * w0 + x] = restrictK(b3a + srcCh, s.h3, s.w3, y, x, p, pad, C);
    });
    __syncthreads();
    cycleLevels(lds, s, p, pad, 0, 0);
  }
  const float* e0 = lds + s.off[2];
  forPixels(s.h3, s.w3, [&](int y, int x) {
    b3a[(static_cast<int64_t>(y) * s.w3 + x) * C + dstCh] =
        s.levels > 0 ? prolongK(e0, s.h3, s.w3, y, x, p, pad) : 0.f;
  });
}

// out channel = K up(E_0) (forward), or in channel += K up(E_0) and the
// consumed channel cleared (backward).
__global__ __launch_bounds__(256) void tailScatterKernel(float* __restrict__ t, int C, int outCh,
                                                         int clearCh, int accumulate, TailShape s,
                                                         TailParams p, int pad,
                                                         const float* __restrict__ work,
                                                         int64_t pix) {
  const int64_t hw = static_cast<int64_t>(s.h3) * s.w3;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < pix;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = q / hw, i = q - b * hw;
    float v = 0.f;
    if (s.levels > 0) {
      const float* e0 = work + b * s.perImage + s.off[2];
      v = prolongK(e0, s.h3, s.w3, static_cast<int>(i / s.w3), static_cast<int>(i % s.w3), p, pad);
    }
    float* px = t + q * C;
    if (accumulate) {
      px[outCh] += v;
      px[clearCh] = 0.f;
    } else {
      px[outCh] = v;
    }
  }
}

}  // namespace

void Unet::setSides(int dirichletSides) {
  require(dirichletSides >= 0 && dirichletSides < 16, "Dirichlet sides must be a 4-bit mask");
  sides = dirichletSides;
}

void Unet::setOptions(int neumannPadding, int dirichletSides, const TailParams& t) {
  require(neumannPadding == 0 || neumannPadding == 1, "padding must be 0 (zero) or 1 (Neumann)");
  setSides(dirichletSides);
  require(t.levels >= 0 && t.levels <= 16, "tail levels must be in 0..16");
  require(t.levels == 0 || (t.src >= 0 && t.src < 64 && t.dst >= 0 && t.dst < 64 && t.src != t.dst),
          "tail channels must be two distinct channels of the bottom tensor");
  require(t.levels == 0 || neumannPadding == 1, "the coarse tail needs Neumann padding");
  neumann = neumannPadding;
  tail = t;
  buildNarrow();  // the tail's channels take part in which channels are live
}

int64_t Unet::tailShapeBytes() { return static_cast<int64_t>(sizeof(TailShape)); }

void Unet::narrowBatchTail(const std::vector<NarrowCase>& cases, void* hostStaging, void* device) {
  TailShape* h = static_cast<TailShape*>(hostStaging);
  for (size_t c = 0; c < cases.size(); ++c) {
    h[c] = tailShape(cases[c].h[3], cases[c].w[3], tail.levels);
    require(h[c].perImage <= kTailLdsFloats,
            "a case of the batch is too large for the batched coarse tail (sides up to 1024)");
  }
  CFDHIP_CHECK(hipMemcpyAsync(device, hostStaging, sizeof(TailShape) * cases.size(),
                              hipMemcpyHostToDevice, stream));
}

void Unet::tailBatch(const NarrowBatch& nb, int C, int srcCh, int dstCh) {
  if (tail.levels == 0 || nb.nCases == 0) return;
  hipLaunchKernelGGL(tailBatchKernel, dim3(static_cast<unsigned>(nb.nCases)), dim3(kTailThreads),
                     0, stream, static_cast<const TailShape*>(nb.tailShapes), nb.cases, nb.act, C,
                     srcCh, dstCh, tail);
  CFDHIP_CHECK(hipGetLastError());
}

void Unet::tailApply(float* tensor, int n, int h3, int w3, bool backward, int C, int srcCh,
                     int dstCh) {
  if (tail.levels == 0) return;
  const TailShape s = tailShape(h3, w3, tail.levels);
  const int64_t hw = static_cast<int64_t>(h3) * w3, pix = n * hw;
  if (srcCh < 0) srcCh = tail.src;
  if (dstCh < 0) dstCh = tail.dst;
  const int inCh = backward ? dstCh : srcCh;
  const int outCh = backward ? srcCh : dstCh;
  if (s.levels > 0) {
    tailWork.ensure(static_cast<int64_t>(n) * s.perImage);
    const int64_t pix0 = static_cast<int64_t>(n) * tailH(s, 0) * tailW(s, 0);
    hipLaunchKernelGGL(tailRestrictKernel, tailGrid(pix0), dim3(256), 0, stream, tensor, C, inCh, s,
                       tail, padBits(), pix0, tailWork.get());
    // levels above kTailCyclePixels run pass by pass over the whole chip; the
    // rest in the one-workgroup-per-image cycle
    int t0 = 0;
    while (t0 + 1 < s.levels && tailH(s, t0) * tailW(s, t0) > kTailCyclePixels) ++t0;
    auto level = [&](int t, int pass) {
      const int64_t m = static_cast<int64_t>(n) * tailPassPixels(s, t, pass);
      hipLaunchKernelGGL(tailLevelKernel, tailGrid(m), dim3(256), 0, stream, s, tail, padBits(), t,
                         pass, m, tailWork.get());
    };
    for (int t = 0; t < t0; ++t) {
      level(t, kPre);
      level(t, kRestrict);
    }
    if (s.perImage - s.off[3 * t0] <= kTailLdsFloats)
      hipLaunchKernelGGL(tailCycleLdsKernel, dim3(static_cast<unsigned>(n)), dim3(kTailThreads), 0,
                         stream, s, tail, padBits(), t0, tailWork.get());
    else
      hipLaunchKernelGGL(tailCycleKernel, dim3(static_cast<unsigned>(n)), dim3(kTailThreads), 0,
                         stream, s, tail, padBits(), t0, tailWork.get());
    for (int t = t0 - 1; t >= 0; --t) {
      level(t, kProlong);
      level(t, kPost);
    }
  }
  hipLaunchKernelGGL(tailScatterKernel, tailGrid(pix), dim3(256), 0, stream, tensor, C, outCh, inCh,
                     backward ? 1 : 0, s, tail, padBits(), tailWork.get(), pix);
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace cfdhip
