// extern "C" entry points of the learned preconditioner: weights on a solver
// context, harvesting of training pairs, and a network handle that trains on
// caller-owned device buffers. None of these has a counterpart in the
// reference; they serve the flexible-CG p' solve (CFDHIP_PREC_LEARNED).
#include <cstring>
#include <new>

#include "unet_net.hpp"

using namespace cfdhip;

struct cfdhip_ctx {
  Context c;
};

struct cfdhip_unet {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<Unet> net;
  std::string lastError;
  DeviceArray<float> x;
  ~cfdhip_unet() {
    net.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace cfdhip {
int64_t harvestPairs(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                     int64_t stride, int64_t maxPairs, int64_t h, int64_t w, float* rOut,
                     float* eOut, float* ceOut, float* cnOut, float* denOut,
                     cfdhip_solve_result* result);
int64_t harvestNoisePairs(Context& ctx, uint64_t seed, int64_t count, int64_t h, int64_t w,
                          float* rOut, float* eOut, float* denOut);
}

namespace {

template <class F>
int guardedCtx(cfdhip_ctx* ctx, F&& fn) {
  if (ctx == nullptr) return CFDHIP_ERR_INVALID;
  try {
    CFDHIP_CHECK(hipSetDevice(ctx->c.device));
    fn(ctx->c);
    ctx->c.lastError.clear();
    return CFDHIP_OK;
  } catch (const Error& e) {
    ctx->c.lastError = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.lastError = e.what();
    return CFDHIP_ERR_RUNTIME;
  }
}

template <class F>
int guardedNet(cfdhip_unet* h, F&& fn) {
  if (h == nullptr) return CFDHIP_ERR_INVALID;
  try {
    CFDHIP_CHECK(hipSetDevice(h->device));
    fn(*h->net);
    CFDHIP_CHECK(hipStreamSynchronize(h->stream));
    h->lastError.clear();
    return CFDHIP_OK;
  } catch (const Error& e) {
    h->lastError = e.what();
    return e.code;
  } catch (const std::exception& e) {
    h->lastError = e.what();
    return CFDHIP_ERR_RUNTIME;
  }
}


void checkShape(int64_t batch, int64_t h, int64_t w) {
  require(batch > 0 && h >= 8 && w >= 8 && h % 8 == 0 && w % 8 == 0,
          "the network grid must be a positive multiple of 8 in each direction");
  require(batch * h * w < (int64_t{1} << 31), "the network batch is too large");
}

}  // namespace

TailParams cfdhip::tailFrom(const cfdhip_unet_options& o) {
  TailParams t{};
  t.levels = o.tail_levels;
  t.omega = o.tail_omega;
  for (int q = 0; q < 3; ++q) {
    t.a[q] = o.tail_a[q];
    t.k[q] = o.tail_k[q];
  }
  t.src = 1;
  t.dst = 2;
  return t;
}

extern "C" {

int64_t cfdhip_unet_param_count(void) { return kUnetParams; }

int cfdhip_set_learned_weights(cfdhip_ctx* ctx, const float* weights, int64_t count) {
  return guardedCtx(ctx, [&](Context& c) {
    require(weights != nullptr && count == kUnetParams,
            "learned weights: expected cfdhip_unet_param_count() floats");
    loadLearnedWeights(c, weights);
  });
}

int cfdhip_set_learned_options(cfdhip_ctx* ctx, const cfdhip_unet_options* o) {
  return guardedCtx(ctx, [&](Context& c) {
    require(o != nullptr, "options are required");
    if (!c.net) c.net = std::make_unique<Unet>(c.stream);
    c.net->setDense(o->dense != 0);
    c.net->setOptions(o->padding, 0, tailFrom(*o));
  });
}

int cfdhip_learned_channels(cfdhip_ctx* ctx, int32_t* stored, int32_t count) {
  return guardedCtx(ctx, [&](Context& c) {
    require(stored != nullptr && count == kNumForward, "stored must hold 17 entries");
    require(c.net && c.net->ready(), "no learned weights loaded on this context");
    constexpr int kFull[kNumForward] = {16, 16, 16, 32, 32, 32, 64, 64, 64,
                                        64, 64, 64, 64, 32, 32, 16, 16};
    const bool narrow = c.net->narrowActive();
    for (int t = 0; t < kNumForward; ++t)
      stored[t] = narrow ? (t == kTe0a ? 0 : c.net->narrowPlan().sw[t]) : kFull[t];
  });
}

int cfdhip_pressure_harvest(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, const double* b,
                            int64_t stride, int64_t max_pairs, int64_t h, int64_t w, float* r_out,
                            float* e_out, float* ce_out, float* cn_out, float* den_out,
                            int64_t* pairs, cfdhip_solve_result* result) {
  return guardedCtx(ctx, [&](Context& c) {
    require(s != nullptr, "solver settings are required");
    c.requireMesh();
    require(c.pressureAssembled, "no p' system assembled");
    require(max_pairs == 0 || (r_out && e_out && den_out), "output buffers are required");
    if (b)
      CFDHIP_CHECK(hipMemcpyAsync(c.f(CFDHIP_FIELD_PC_RHS), b, c.view.n * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
    const int64_t k = harvestPairs(c, *s, c.f(CFDHIP_FIELD_PC_RHS), stride, max_pairs, h, w, r_out,
                                   e_out, ce_out, cn_out, den_out, result);
    if (pairs) *pairs = k;
  });
}

int cfdhip_pressure_harvest_noise(cfdhip_ctx* ctx, uint64_t seed, int64_t count, int64_t h,
                                  int64_t w, float* r_out, float* e_out, float* den_out,
                                  int64_t* pairs) {
  return guardedCtx(ctx, [&](Context& c) {
    c.requireMesh();
    require(c.pressureAssembled, "no p' system assembled");
    require(count == 0 || (r_out && e_out && den_out), "output buffers are required");
    const int64_t k = harvestNoisePairs(c, seed, count, h, w, r_out, e_out, den_out);
    if (pairs) *pairs = k;
  });
}

int cfdhip_unet_create(cfdhip_unet** out) {
  if (out == nullptr) return CFDHIP_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CFDHIP_ERR_NO_DEVICE;
  auto* h = new (std::nothrow) cfdhip_unet;
  if (h == nullptr) return CFDHIP_ERR_RUNTIME;
  try {
    CFDHIP_CHECK(hipGetDevice(&h->device));
    CFDHIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->net = std::make_unique<Unet>(h->stream);
  } catch (...) {
    delete h;
    return CFDHIP_ERR_RUNTIME;
  }
  *out = h;
  return CFDHIP_OK;
}

void cfdhip_unet_destroy(cfdhip_unet* h) {
  if (h == nullptr) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
}

const char* cfdhip_unet_last_error(const cfdhip_unet* h) {
  return h == nullptr ? "null network handle" : h->lastError.c_str();
}

int cfdhip_unet_set_weights(cfdhip_unet* h, const float* weights, int64_t count) {
  return guardedNet(h, [&](Unet& net) {
    require(weights != nullptr && count == kUnetParams,
            "expected cfdhip_unet_param_count() floats");
    net.setWeights(weights);
  });
}

int cfdhip_unet_set_options(cfdhip_unet* h, const cfdhip_unet_options* o) {
  return guardedNet(h, [&](Unet& net) {
    require(o != nullptr, "options are required");
    net.setDense(o->dense != 0);
    net.setOptions(o->padding, o->dirichlet_sides, tailFrom(*o));
  });
}

int cfdhip_unet_get_weights(cfdhip_unet* h, float* weights, int64_t count) {
  return guardedNet(h, [&](Unet& net) {
    require(weights != nullptr && count == kUnetParams,
            "expected cfdhip_unet_param_count() floats");
    net.getWeights(weights);
  });
}

int cfdhip_unet_forward(cfdhip_unet* h, const float* x, float* y, int64_t batch, int64_t hh,
                        int64_t w) {
  return guardedNet(h, [&](Unet& net) {
    checkShape(batch, hh, w);
    require(x != nullptr && y != nullptr, "device buffers are required");
    const int n = static_cast<int>(batch), H = static_cast<int>(hh), W = static_cast<int>(w);
    net.ensureShape(n, H, W, false);
    net.forward(x, nullptr, H, W, hh * w, n, H, W);
    net.outputTrain(n, H, W);
    CFDHIP_CHECK(hipMemcpyAsync(y, net.output(), sizeof(float) * batch * hh * w,
                                hipMemcpyDeviceToDevice, net.stream));
  });
}

int cfdhip_unet_apply(cfdhip_unet* h, const float* x, float* y, int64_t batch, int64_t hh,
                      int64_t w) {
  return guardedNet(h, [&](Unet& net) {
    // any image size: the network runs on the grid rounded up to multiples
    // of 8 and reflects its padding at the image's edges
    require(hh >= 1 && w >= 1, "the image must have at least one pixel");
    const int64_t h8 = (hh + 7) / 8 * 8, w8 = (w + 7) / 8 * 8;
    checkShape(batch, h8, w8);
    require(x != nullptr && y != nullptr, "device buffers are required");
    const int n = static_cast<int>(batch), H = static_cast<int>(h8), W = static_cast<int>(w8);
    net.ensureShape(n, H, W, false);
    net.forwardFrom(x, false, nullptr, static_cast<int>(hh), static_cast<int>(w), hh * w, n, H, W,
                    true);
    net.outputTrain(n, H, W, static_cast<int>(hh), static_cast<int>(w));
    CFDHIP_CHECK(hipMemcpyAsync(y, net.output(), sizeof(float) * batch * hh * w,
                                hipMemcpyDeviceToDevice, net.stream));
  });
}

int cfdhip_unet_backward(cfdhip_unet* h, const float* x, const float* gy, int64_t batch, int64_t hh,
                         int64_t w, float* grad) {
  return guardedNet(h, [&](Unet& net) {
    checkShape(batch, hh, w);
    require(x != nullptr && gy != nullptr && grad != nullptr, "buffers are required");
    const int n = static_cast<int>(batch), H = static_cast<int>(hh), W = static_cast<int>(w);
    net.ensureShape(n, H, W, true);
    net.forward(x, nullptr, H, W, hh * w, n, H, W);
    net.backward(x, gy, n, H, W);
    net.getGrads(grad);
  });
}

int cfdhip_unet_tensor(cfdhip_unet* h, int tensor, float* out, int64_t count) {
  return guardedNet(h, [&](Unet& net) {
    require(out != nullptr, "output buffer is required");
    net.readTensor(tensor, out, count);
  });
}

int64_t cfdhip_unet_tensor_size(cfdhip_unet* h, int tensor) {
  if (h == nullptr) return -1;
  try {
    return h->net->tensorElems(tensor);
  } catch (...) {
    return -1;
  }
}

int cfdhip_unet_loss_grad(cfdhip_unet* h, const float* x, const float* e, const float* ce,
                          const float* cn, const float* cd, const int64_t* sys_offset,
                          const float* den, int64_t batch, int64_t hh, int64_t w, double* loss,
                          float* grad) {
  return guardedNet(h, [&](Unet& net) {
    checkShape(batch, hh, w);
    require(x && e && ce && cn && sys_offset && den && grad, "buffers are required");
    const int n = static_cast<int>(batch), H = static_cast<int>(hh), W = static_cast<int>(w);
    net.ensureShape(n, H, W, true);
    net.forward(x, nullptr, H, W, hh * w, n, H, W);
    net.outputTrain(n, H, W);
    const double l = net.loss(e, ce, cn, cd, sys_offset, den, n, H, W);
    if (loss) *loss = l;
    net.backward(x, net.outputGrad(), n, H, W);
    net.getGrads(grad);
  });
}

int cfdhip_unet_spectral_loss_grad(cfdhip_unet* h, const float* z, const float* ce,
                                   const float* cn, const float* cd, const int64_t* sys_offset,
                                   int64_t batch, int64_t hh, int64_t w, int32_t steps,
                                   double* loss, float* grad) {
  return guardedNet(h, [&](Unet& net) {
    checkShape(batch, hh, w);
    require(z && ce && cn && sys_offset && loss, "buffers are required");
    require(steps >= 1 && steps <= 32, "steps must be in 1..32");
    const int n = static_cast<int>(batch), H = static_cast<int>(hh), W = static_cast<int>(w);
    net.ensureShape(n, H, W, grad != nullptr);
    *loss = net.spectral(z, ce, cn, cd, sys_offset, n, H, W, steps, grad);
  });
}

int cfdhip_unet_train_step(cfdhip_unet* h, const float* x, const float* e, const float* ce,
                           const float* cn, const float* cd, const int64_t* sys_offset,
                           const float* den, int64_t batch, int64_t hh, int64_t w, float lr,
                           int update, double* loss) {
  return guardedNet(h, [&](Unet& net) {
    checkShape(batch, hh, w);
    require(x && e && ce && cn && sys_offset && den, "device buffers are required");
    require(lr >= 0.f, "the learning rate must be non-negative");
    const int n = static_cast<int>(batch), H = static_cast<int>(hh), W = static_cast<int>(w);
    net.ensureShape(n, H, W, update != 0);
    net.forward(x, nullptr, H, W, hh * w, n, H, W);
    net.outputTrain(n, H, W);
    const double l = net.loss(e, ce, cn, cd, sys_offset, den, n, H, W);
    if (loss) *loss = l;
    if (update) {
      net.backward(x, net.outputGrad(), n, H, W);
      net.adam(lr, 0.9f, 0.999f, 1e-8f);
    }
  });
}

}  // extern "C"
