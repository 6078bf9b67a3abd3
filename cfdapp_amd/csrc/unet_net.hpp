// Device state of the learned preconditioner's network.
#pragma once

#include <vector>

#include "unet.hpp"

namespace cfdhip {

// Activation and gradient tensors (FP32, NHWC), in allocation order: the
// forward tensors first, so an inference-only shape allocates only those.
enum UnetTensor : int {
  kTe0a = 0, kTe0, kTp0, kTe1a, kTe1, kTp1, kTe2a, kTe2, kTp2, kTb3a, kTb3,
  kTd2a, kTd2, kTd1a, kTd1, kTd0a, kTd0,
  kNumForward,
  kTgd0 = kNumForward, kTgd0a, kTgd1, kTge0s, kTgd1a, kTgd2, kTge1s, kTgd2a, kTgb3, kTge2s,
  kTgb3a, kTgp2, kTge2a, kTgp1, kTge1a, kTgp0, kTge0a,
  kNumTensors
};

// The network with the channels no output depends on dropped, built from
// the weights and the tail options: per forward tensor the channels that
// some later weight reads (slots; a tensor stores only those, 2 or 4 per pixel),
// per layer the weights between the kept channels (BF16-rounded, as every
// layer applies them). When every tensor keeps at most kNarrowMax channels,
// inference runs on the narrow kernels of unet_narrow.hip instead of the
// MFMA layers; the result is the same linear map.
constexpr int kNarrowMax = 4;
// the narrow convolutions' output tile (rows x columns) and the rows of one
// output-layer block of a batched application
constexpr int kNarrowTileH = 16, kNarrowTileW = 64, kNarrowLastRows = 8;
struct NarrowPlan {
  bool usable = false;
  int sw[kNumForward] = {};                 // stored channels per tensor (2 or 4)
  int width[kNumForward] = {};              // kept channels per tensor (<= sw)
  int chan[kNumForward][kNarrowMax] = {};   // the original channel of each slot
  int tailSrc = -1, tailDst = -1;           // slots of the tail's channels in b3a
  std::vector<float> w;                     // per layer [tap][slot in][slot out]
  int64_t wOff[kNumLayers] = {};
};

// One case of a batched application of the narrow network (many 2D cases
// of uneven size in one launch per layer): per network level its tensor size
// (level 0 rounded up to multiples of 8), physical image (the rounded-up
// halves of the solver grid) and conv tiles per tensor row; its padding bits;
// the offsets of its solver vectors (cells) and of its activations in the
// batch's arena (act + tOff[tensor], floats).
struct NarrowCase {
  int h[4], w[4], ph[4], pw[4], tilesX[4];
  int pad, ny, nx;
  int64_t cells, pin, act;
  int64_t tOff[kNumForward];
};

// The device tables of a batched application: the case table, per level the
// (case, tile) entries of the conv launches, the (case, first row, rows,
// partial slot) entries of the output layer, and the coarse tail's shape of
// each case (an opaque device array written by Unet::narrowBatchTables).
struct NarrowBatch {
  const NarrowCase* cases = nullptr;
  int64_t nCases = 0;
  const int2* tab[4] = {};
  int64_t tabCount[4] = {};
  const int4* last = nullptr;
  int64_t lastCount = 0;
  const void* tailShapes = nullptr;
  float* act = nullptr;
};

// The coarse tail of the C options (channels 1 -> 2 of the bottom tensor).
TailParams tailFrom(const cfdhip_unet_options& o);

class Unet {
 public:
  explicit Unet(hipStream_t s);

  void setWeights(const float* host);
  // Padding of every convolution (0 zero, 1 Neumann), the Dirichlet sides
  // of Neumann padding (bits: 1 left, 2 right, 4 bottom, 8 top; odd
  // reflection there) and the coarse tail below the bottom level (levels 0:
  // none; needs Neumann padding).
  void setOptions(int neumannPadding, int dirichletSides, const TailParams& t);
  void setSides(int dirichletSides);
  // dense: run inference through the MFMA layers over every channel even
  // where the weights leave channels unused
  void setDense(bool dense);
  // whether inference takes the narrow kernels
  bool narrowActive() const { return narrow.usable && !dense_; }
  const NarrowPlan& narrowPlan() const { return narrow; }
  int padding() const { return neumann; }
  // ConvArgs::pad of the current options
  int padBits() const { return neumann ? (kPadNeumann | (sides << 1)) : 0; }
  void getWeights(float* host);
  void getGrads(float* host) {
    CFDHIP_CHECK(hipMemcpyAsync(host, grad.get(), kUnetParams * sizeof(float),
                                hipMemcpyDeviceToHost, stream));
    CFDHIP_CHECK(hipStreamSynchronize(stream));
  }
  bool ready() const { return loaded; }

  // Activations for n images of h x w (multiples of 8); training adds the
  // gradient tensors.
  void ensureShape(int n, int h, int w, bool training);
  // Encoder/decoder up to d0 on x = src * scale (src: srcH x srcW per image,
  // zero-extended to h x w).
  void forward(const float* src, const float* scale, int srcH, int srcW, int64_t srcBatch, int n,
               int h, int w);
  // The same from an FP32 or FP64 source.
  // fuseFirst: compute the first layer inside the second's tile staging
  // (inference only: the first layer's output is not stored).
  void forwardFrom(const void* src, bool srcDouble, const float* scale, int srcH, int srcW,
                   int64_t srcBatch, int n, int h, int w, bool fuseFirst = false);
  // y = conv(d0) into the FP32 output, compact n x ph x pw for the physical
  // image ph x pw of the last forward pass (0: the whole h x w grid).
  void outputTrain(int n, int h, int w, int ph = 0, int pw = 0);
  // z = scale * conv(d0) on the ny x nx solver grid, z[pin] = r[pin] scale^2
  // (the inverse of the decoupled pinned row), and block partials of (z, q)
  // (parts blocks; none when partial is null).
  template <class TZ>
  void outputApply(const float* scale, const TZ* r, const TZ* q, TZ* z, int ny, int nx,
                   int64_t pin, int h, int w, double* partial, int parts);
  // Weight gradients for dL/dy = gyIn (x: the FP32 network input).
  void backward(const float* x, const float* gyIn, int n, int h, int w);
  // A-norm loss of the training output against e; fills gy, returns the loss.
  double loss(const float* e, const float* cE, const float* cN, const float* cD,
              const int64_t* sysOffDev,
              const float* denDev, int n, int h, int w);
  void adam(float lr, float b1, float b2, float eps);
  // Multi-step spectral loss of the stationary iteration x <- x + N(b - B x)
  // on n error images z (h x w, multiples of 8) with the operators of the
  // A-norm loss: v_0 = z / |z|_B, v_j = v_{j-1} - N(B v_{j-1}) (j = 1..steps),
  // L = mean_s log(|v_steps|_B^2) / (2 steps). gradHost (param count floats)
  // receives dL/dW by reverse accumulation through the steps with N^T = N,
  // or nothing when null.
  double spectral(const float* z, const float* cE, const float* cN, const float* cD,
                  const int64_t* sysOffDev, int n, int h, int w, int steps, float* gradHost);

  float* output() { return y.get(); }
  // Tensor t of the current shape, copied to the host.
  int64_t tensorElems(int t) const;
  void readTensor(int t, float* host, int64_t count);
  float* outputGrad() { return gy.get(); }
  hipStream_t stream;

  // Per-layer kernel timing (HIP events on the stream around each layer's
  // launch; the host waits at the end of every forward pass while enabled).
  void setTiming(bool on);
  // Batched inference over cases of uneven size (narrow kernels only: the
  // weights must leave at most kNarrowMax live channels per tensor). The
  // layout fills each case's level sizes, tiles and activation offsets
  // (cases[c].ny, nx, pad, cells, pin set by the caller) and returns the
  // arena size in floats; the tables are then built on the host (tabs,
  // last) with kNarrowTile* and the tail shapes uploaded by narrowBatchTail.
  int64_t narrowBatchLayout(std::vector<NarrowCase>& cases) const;
  // the per-case coarse tail shapes, as a device array of nCases entries
  // (bytes of one entry: tailShapeBytes())
  static int64_t tailShapeBytes();
  void narrowBatchTail(const std::vector<NarrowCase>& cases, void* hostStaging, void* device);
  // z = S N(S r) on every case (r, q, z, scale: the batch's FP64 vectors and
  // FP32 scaling at NarrowCase::cells), z = r scale^2 at pinned cells, and the
  // block partials of (z, q) in each output-layer entry's slot (q null: none)
  void forwardNarrowBatch(const NarrowBatch& nb, const double* r, const float* scale,
                          const double* q, double* z, double* partial);
  // index kNumLayers: the coarse tail
  double layerMs[kNumLayers + 1] = {};
  int64_t layerCalls[kNumLayers + 1] = {};

 private:
  void repack();
  void buildNarrow();
  void ensureNarrow(int n, int h, int w);
  void forwardNarrow(const void* src, bool srcDouble, const float* scale, int srcH, int srcW,
                     int64_t srcBatch, int n, int h, int w);
  void outputNarrow(int n, int h, int w, int ph, int pw);
  template <class TZ>
  void outputApplyNarrow(const float* scale, const TZ* r, const TZ* q, TZ* z, int ny, int nx,
                         int64_t pin, int h, int w, double* partial, int parts);
  float* NT(int t) { return nact.get() + nOff[t]; }
  // the physical extent of the level whose tensors are h x w
  void levelExtent(int h, int w, int& ph, int& pw) const;
  NarrowPlan narrow;
  bool dense_ = false;
  bool lastNarrow = false;  // the last forward pass ran on the narrow kernels
  std::vector<float> hostMaster;
  DeviceArray<float> nact, nW;
  int64_t nOff[kNumForward] = {};
  int narrowN = 0, narrowH = 0, narrowW = 0;
  int64_t narrowBytes = 0;
  float* T(int t) { return act.get() + tensorOff[t]; }
  void conv(int layer, bool transposed, int gather, int epi, int split, const float* a,
            const float* b, float* out, float* aux, int n, int h, int w);
  void convDw(int layer, int gx, int gg, const float* xa, const float* xb, const float* ga,
              const float* gb, int n, int h, int w);
  void singleDw(int mode, const float* s1, const float* m16, int n, int h, int w, float* out);
  // The coarse tail on the bottom tensor (h3 x w3 per image): forward writes
  // channel tail.dst from tail.src; backward is its adjoint on the gradient.
  // C: channels of the tensor; srcCh / dstCh (-1: the options' channels)
  // The coarse tail of every case of a batched application, one workgroup
  // per case (cases whose bottom level is at most 128 x 128: the whole tail in
  // LDS); C, srcCh, dstCh: the stored channels of b3a and the tail's slots.
  void tailBatch(const NarrowBatch& nb, int C, int srcCh, int dstCh);
  void tailApply(float* tensor, int n, int h3, int w3, bool backward, int C = 64, int srcCh = -1,
                 int dstCh = -1);
  int neumann = 0;
  int sides = 0;
  TailParams tail{};
  // every kernel of the loaded weights is mirror-symmetric in both axes (the Neumann-padded
  // backward pass is the exact adjoint only for such kernels)
  bool mirrorSymmetric = false;
  DeviceArray<float> tailWork;

  DeviceArray<float> master, grad, adamM, adamV, y, gy, dwPartial;
  DeviceArray<double> lossPartial;
  DeviceArray<float> spec, specF;  // spectral loss: iterates, their B images, per-image factors
  DeviceArray<double> specDots;
  DeviceArray<bf16> fwdPack, bwdPack;
  DeviceArray<float> act;
  int64_t fwdOff[kNumLayers] = {};
  int64_t bwdOff[kNumLayers] = {};
  int64_t tensorOff[kNumTensors] = {};
  struct FirstSource {
    const void* src = nullptr;
    const float* scale = nullptr;
    int srcH = 0, srcW = 0, srcDouble = 0;
  } first;
  // physical image of the last forward pass (its level-0 extent and the
  // level-0 tensor size it sits in); ConvArgs::ph / pw of every conv
  int physH = 0, physW = 0, physLevelH = 0, physLevelW = 0;
  int shapeN = 0, shapeH = 0, shapeW = 0;
  bool shapeTraining = false;
  bool loaded = false;
  bool timing = false;
  std::vector<hipEvent_t> events;
  std::vector<int> eventLayer;
  void mark(int layer, bool begin);
  void collectTiming();
  int64_t adamStep = 0;
};

}  // namespace cfdhip
