// Coarse tail of the learned preconditioner (TailParams in unet.hpp): a
// one-channel symmetric V(1,1) cycle below the network's bottom level, with
// Neumann padding, so that the preconditioner reaches the coarsest grid.
//
// Three launches per application: a gather of the input channel of the
// bottom tensor into a compact image (all pixels in parallel), the cycle
// itself (one workgroup per image walks every tail level, a workgroup
// barrier between dependent passes; the levels are at most H/16 x W/16 and
// stay in L2), and a scatter of the prolonged correction K up(e_0) into the
// output channel. The operator is symmetric (restriction 4 pool(K .) is the
// transpose of prolongation K up(.), the same Jacobi sweep before and after
// the coarse correction), so the backward pass runs the same three launches
// from the output channel's gradient into the input channel's.
#include "unet_net.hpp"

namespace cfdhip {

namespace {

constexpr int kTailThreads = 1024;

dim3 tailGrid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : (b > 8192 ? 8192 : b)));
}

// One image's tail workspace: the compact bottom channel, then per level t
// the residual R_t, a work image S_t and the correction E_t.
struct TailShape {
  int h3, w3, levels;
  int64_t perImage;
  int64_t off[3 * 16];  // R_t, S_t, E_t offsets (after the compact image)
};

__host__ __device__ inline int tailH(const TailShape& s, int t) { return s.h3 >> (t + 1); }
__host__ __device__ inline int tailW(const TailShape& s, int t) { return s.w3 >> (t + 1); }

TailShape tailShape(int h3, int w3, int maxLevels) {
  TailShape s{};
  s.h3 = h3;
  s.w3 = w3;
  int hh = h3, ww = w3, L = 0;
  while (L < maxLevels && L < 16 && hh % 2 == 0 && ww % 2 == 0 && hh >= 2 && ww >= 2) {
    hh /= 2;
    ww /= 2;
    ++L;
  }
  s.levels = L;
  int64_t o = static_cast<int64_t>(h3) * w3;
  for (int t = 0; t < L; ++t) {
    const int64_t m = static_cast<int64_t>(h3 >> (t + 1)) * (w3 >> (t + 1));
    for (int q = 0; q < 3; ++q) {
      s.off[3 * t + q] = o;
      o += m;
    }
  }
  s.perImage = o;
  return s;
}

// Pixel (y, x) of an h x w image with the network's padding (padSign).
__device__ inline float at(const float* img, int h, int w, int y, int x, int pad) {
  const float s = padSign(y, x, h, w, pad);
  return s * img[static_cast<int64_t>(y) * w + x];
}

// Symmetric 3 x 3 stencil (centre, edge, corner) at (y, x).
__device__ inline float stencil3(const float* img, int h, int w, int y, int x, float c, float e,
                                 float k, int pad) {
  const float edges = (at(img, h, w, y - 1, x, pad) + at(img, h, w, y + 1, x, pad)) +
                      (at(img, h, w, y, x - 1, pad) + at(img, h, w, y, x + 1, pad));
  const float corners = (at(img, h, w, y - 1, x - 1, pad) + at(img, h, w, y - 1, x + 1, pad)) +
                        (at(img, h, w, y + 1, x - 1, pad) + at(img, h, w, y + 1, x + 1, pad));
  return c * img[static_cast<int64_t>(y) * w + x] + e * edges + k * corners;
}

// 4 pool(K v) at coarse pixel (y, x) of v (fine h x w).
__device__ inline float restrictK(const float* v, int h, int w, int y, int x, const TailParams& p,
                                  int pad) {
  float s = 0.f;
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
      s += stencil3(v, h, w, 2 * y + dy, 2 * x + dx, p.k[0], p.k[1], p.k[2], pad);
  return s;  // 4 x the 2 x 2 average
}

// K up(e) at fine pixel (y, x) of the h x w grid (e on h/2 x w/2).
__device__ inline float prolongK(const float* e, int h, int w, int y, int x, const TailParams& p,
                                 int pad) {
  const int wc = w >> 1;
  auto up = [&](int yy, int xx) {  // the upsampled image with the padding
    const float s = padSign(yy, xx, h, w, pad);
    return s * e[static_cast<int64_t>(yy >> 1) * wc + (xx >> 1)];
  };
  const float edges = (up(y - 1, x) + up(y + 1, x)) + (up(y, x - 1) + up(y, x + 1));
  const float corners = (up(y - 1, x - 1) + up(y - 1, x + 1)) + (up(y + 1, x - 1) + up(y + 1, x + 1));
  return p.k[0] * up(y, x) + p.k[1] * edges + p.k[2] * corners;
}

__global__ __launch_bounds__(256) void tailGatherKernel(const float* __restrict__ t, int C, int ch,
                                                        int64_t pix, int64_t perImage, int64_t hw,
                                                        float* __restrict__ work) {
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < pix;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = p / hw, q = p - b * hw;
    work[b * perImage + q] = t[p * C + ch];
  }
}

__global__ __launch_bounds__(kTailThreads) void tailCycleKernel(TailShape s, TailParams p, int pad,
                                                               float* __restrict__ work) {
  float* base = work + static_cast<int64_t>(blockIdx.x) * s.perImage;
  const int L = s.levels;
  const float om = p.omega;
  auto R = [&](int t) { return base + s.off[3 * t]; };
  auto S = [&](int t) { return base + s.off[3 * t + 1]; };
  auto E = [&](int t) { return base + s.off[3 * t + 2]; };
  // level 0 residual: restriction of the bottom channel
  {
    const int h = tailH(s, 0), w = tailW(s, 0);
    for (int i = threadIdx.x; i < h * w; i += blockDim.x)
      R(0)[i] = restrictK(base, s.h3, s.w3, i / w, i % w, p, pad);
  }
  __syncthreads();
  // down: one Jacobi sweep from zero, restriction of the remaining residual
  for (int t = 0; t + 1 < L; ++t) {
    const int h = tailH(s, t), w = tailW(s, t);
    for (int i = threadIdx.x; i < h * w; i += blockDim.x) {
      const int y = i / w, x = i % w;
      S(t)[i] = R(t)[i] - om * stencil3(R(t), h, w, y, x, p.a[0], p.a[1], p.a[2], pad);
    }
    __syncthreads();
    const int hc = tailH(s, t + 1), wc = tailW(s, t + 1);
    for (int i = threadIdx.x; i < hc * wc; i += blockDim.x)
      R(t + 1)[i] = restrictK(S(t), h, w, i / wc, i % wc, p, pad);
    __syncthreads();
  }
  // coarsest level: the two sweeps without a coarse correction,
  // e = omega r + (I - omega A) omega r
  {
    const int h = tailH(s, L - 1), w = tailW(s, L - 1);
    for (int i = threadIdx.x; i < h * w; i += blockDim.x) {
      const float ar = stencil3(R(L - 1), h, w, i / w, i % w, p.a[0], p.a[1], p.a[2], pad);
      E(L - 1)[i] = 2.f * om * R(L - 1)[i] - om * om * ar;
    }
  }
  __syncthreads();
  // up: coarse correction, then the post-sweep
  for (int t = L - 2; t >= 0; --t) {
    const int h = tailH(s, t), w = tailW(s, t);
    for (int i = threadIdx.x; i < h * w; i += blockDim.x)
      S(t)[i] = om * R(t)[i] + prolongK(E(t + 1), h, w, i / w, i % w, p, pad);
    __syncthreads();
    for (int i = threadIdx.x; i < h * w; i += blockDim.x) {
      const int y = i / w, x = i % w;
      const float v = S(t)[i];
      E(t)[i] = v - om * stencil3(S(t), h, w, y, x, p.a[0], p.a[1], p.a[2], pad) + om * R(t)[i];
    }
    __syncthreads();
  }
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
}

void Unet::tailApply(float* tensor, int n, int h3, int w3, bool backward) {
  if (tail.levels == 0) return;
  const TailShape s = tailShape(h3, w3, tail.levels);
  const int64_t hw = static_cast<int64_t>(h3) * w3, pix = n * hw;
  const int inCh = backward ? tail.dst : tail.src;
  const int outCh = backward ? tail.src : tail.dst;
  if (s.levels > 0) {
    tailWork.ensure(static_cast<int64_t>(n) * s.perImage);
    hipLaunchKernelGGL(tailGatherKernel, tailGrid(pix), dim3(256), 0, stream, tensor, 64, inCh, pix,
                       s.perImage, hw, tailWork.get());
    hipLaunchKernelGGL(tailCycleKernel, dim3(static_cast<unsigned>(n)), dim3(kTailThreads), 0,
                       stream, s, tail, padBits(), tailWork.get());
  }
  hipLaunchKernelGGL(tailScatterKernel, tailGrid(pix), dim3(256), 0, stream, tensor, 64, outCh, inCh,
                     backward ? 1 : 0, s, tail, padBits(), tailWork.get(), pix);
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace cfdhip
