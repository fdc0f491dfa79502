// dr_unet.cpp -- the learned denoiser: a 5-level U-Net on the HIP kernels
// of dr_denoise.hip, with training (forward, backward, Adam) on the device.
//
// Network (NHWC BF16 activations; every conv 3x3 + bias, ReLU except the
// last; pool = 2x2 average; "up" = nearest 2x upsampling, concatenated
// after the skip tensor):
//   x0[16] -> c0a[32] -> e0[32]          -> pool -> c1a[32] -> e1[32]
//   -> pool -> c2a[48] -> e2[48] -> pool -> c3a[64] -> e3[64]
//   -> pool -> c4a[128] -> b4[128]
//   (e3, up b4) -> c5a[64] -> d3[64];  (e2, up d3) -> c6a[48] -> d2[48]
//   (e1, up d2) -> c7a[32] -> d1[32];  (e0, up d1) -> c8a[32] -> d0[32]
//   d0 -> c9 -> y[16] (FP32; channels 0-2 = demodulated log radiance)
// 19 convolutions, 633,600 weights + 976 biases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dr_denoise.h"
#include "dr_internal.h"

namespace {

enum Tensor { X0, T0A, E0, P0, T1A, E1, P1, T2A, E2, P2, T3A, E3, P3, T4A, B4, T5A, D3, T6A, D2, T7A, D1, T8A, D0, Y, NT };
struct TensorDef {
    int c, level;
};
const TensorDef kTensors[NT] = {{16, 0}, {32, 0},  {32, 0},  {32, 1}, {32, 1}, {32, 1}, {32, 2}, {48, 2},
                                {48, 2}, {48, 3},  {64, 3},  {64, 3}, {64, 4}, {128, 4}, {128, 4}, {64, 3},
                                {64, 3}, {48, 2},  {48, 2},  {32, 1}, {32, 1}, {32, 0}, {32, 0}, {16, 0}};
struct Op {
    bool conv;
    int a, b, out;
};
const Op kOps[] = {
    {true, X0, -1, T0A},  {true, T0A, -1, E0}, {false, E0, -1, P0},  {true, P0, -1, T1A},  {true, T1A, -1, E1},
    {false, E1, -1, P1},  {true, P1, -1, T2A}, {true, T2A, -1, E2},  {false, E2, -1, P2},  {true, P2, -1, T3A},
    {true, T3A, -1, E3},  {false, E3, -1, P3}, {true, P3, -1, T4A},  {true, T4A, -1, B4},  {true, E3, B4, T5A},
    {true, T5A, -1, D3},  {true, E2, D3, T6A}, {true, T6A, -1, D2},  {true, E1, D2, T7A},  {true, T7A, -1, D1},
    {true, E0, D1, T8A},  {true, T8A, -1, D0}, {true, D0, -1, Y},
};
constexpr int kNumOps = int(sizeof(kOps) / sizeof(kOps[0]));

struct Layer {
    int op, ca, cb, cout;
    size_t w, b, wb;  // FP32 offsets of W and b in the parameter vector; BF16 offset of W
};

// Activation buffers for one batch shape; gradients only for training.
struct Workspace {
    int n = 0, H = 0, W = 0;
    bool grads = false;
    void *block = nullptr;
    size_t bytes = 0;
    uint16_t *act[NT] = {};
    float *y = nullptr;
    float *grad[NT] = {};
    size_t gradFloats = 0;
    float *gradBase = nullptr;
    uint16_t *dpre = nullptr;
    float *din = nullptr;
    float *target = nullptr;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

struct dr_unet {
    dr_device *device = nullptr;
    hipStream_t stream = nullptr;
    std::vector<Layer> layers;
    size_t paramCount = 0, weightCount = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wf = nullptr, *wt = nullptr;
    float *loss = nullptr;  // per-step loss sums
    uint32_t lossCap = 0;
    uint32_t step = 0;
    // Test hook (environment DENRIM_HIP_UNET_PLAIN_CONV=1 when the network
    // is created): every convolution on the per-layer LDS kernel.
    bool plainConv = false;
    Workspace train, infer;
    // Device-resident training set.
    uint16_t *dsFeat = nullptr;
    float *dsTarget = nullptr;
    uint32_t dsCap = 0, dsCount = 0, dsH = 0, dsW = 0;
    bool batchReady = false;
};

namespace {

dr_status ufail(dr_status s, const std::string &m) { return drFail(s, m); }

#define DR_UHIP(call, what)                                                                                   \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return ufail(e_ == hipErrorOutOfMemory ? DR_ERR_OUT_OF_MEMORY : DR_ERR_COMMAND_BUFFER_FAILED,     \
                         std::string(what) + ": " + hipGetErrorString(e_));                                   \
    } while (0)

void buildLayers(dr_unet *u) {
    size_t off = 0, woff = 0;
    for (int i = 0; i < kNumOps; ++i) {
        if (!kOps[i].conv) continue;
        Layer L;
        L.op = i;
        L.ca = kTensors[kOps[i].a].c;
        L.cb = kOps[i].b >= 0 ? kTensors[kOps[i].b].c : 0;
        L.cout = kTensors[kOps[i].out].c;
        L.w = off;
        off += size_t(9) * L.cout * (L.ca + L.cb);
        L.b = off;
        off += L.cout;
        L.wb = woff;
        woff += size_t(9) * L.cout * (L.ca + L.cb);
        u->layers.push_back(L);
    }
    u->paramCount = off;
    u->weightCount = woff;
}

void freeWorkspace(Workspace &w) {
    DR_DEVICE_LOCK;
    if (w.block) (void)hipFree(w.block);
    w = Workspace{};
}

// Sizes (and on growth reallocates) the buffers for batch n x H x W.
dr_status ensureWorkspace(dr_unet *u, Workspace &w, int n, int H, int W, bool grads) {
    if (w.block && w.n == n && w.H == H && w.W == W && w.grads == grads) return DR_OK;
    std::vector<size_t> actOff(NT), gradOff(NT);
    size_t bytes = 0, gradFloats = 0, dpre = 0, din = 0;
    auto pix = [&](int level) { return size_t(n) * (H >> level) * (W >> level); };
    for (int t = 0; t < NT; ++t) {
        actOff[t] = bytes;
        const size_t e = pix(kTensors[t].level) * kTensors[t].c;
        bytes += align256(e * (t == Y ? 4 : 2));
        if (grads && t != X0 && t != Y) {
            gradOff[t] = gradFloats;
            gradFloats += (e + 63) & ~size_t(63);
        }
    }
    for (const Layer &L : u->layers) {
        const size_t p = pix(kTensors[kOps[L.op].out].level);
        dpre = std::max(dpre, p * L.cout);
        din = std::max(din, p * (L.ca + L.cb));
    }
    const size_t gradAt = bytes;
    if (grads) {
        bytes += align256(gradFloats * 4);
    }
    const size_t dpreAt = bytes;
    if (grads) bytes += align256(dpre * 2);
    const size_t dinAt = bytes;
    if (grads) bytes += align256(din * 4);
    const size_t targetAt = bytes;
    bytes += align256(pix(0) * 16);
    if (w.block && bytes <= w.bytes) {
        // reuse the block
    } else {
        DR_DEVICE_LOCK;
        (void)hipStreamSynchronize(u->stream);
        if (w.block) (void)hipFree(w.block);
        w.block = nullptr;
        DR_UHIP(hipMalloc(&w.block, bytes), "denoiser workspace");
        w.bytes = bytes;
    }
    char *base = static_cast<char *>(w.block);
    w.n = n, w.H = H, w.W = W, w.grads = grads;
    for (int t = 0; t < NT; ++t) {
        w.act[t] = reinterpret_cast<uint16_t *>(base + actOff[t]);
        w.grad[t] = grads && t != X0 && t != Y ? reinterpret_cast<float *>(base + gradAt) + gradOff[t] : nullptr;
    }
    w.y = reinterpret_cast<float *>(base + actOff[Y]);
    w.gradBase = grads ? reinterpret_cast<float *>(base + gradAt) : nullptr;
    w.gradFloats = grads ? gradFloats : 0;
    w.dpre = grads ? reinterpret_cast<uint16_t *>(base + dpreAt) : nullptr;
    w.din = grads ? reinterpret_cast<float *>(base + dinAt) : nullptr;
    w.target = reinterpret_cast<float *>(base + targetAt);
    return DR_OK;
}

dr_status refreshBf16(dr_unet *u) {
    for (const Layer &L : u->layers)
        DR_UHIP(dr::launchWeightsBf16(u->params + L.w, u->wf + L.wb, u->wt + L.wb, L.cout, L.ca + L.cb, u->stream),
                "weights to bf16");
    return DR_OK;
}

dr::ConvArgs convArgs(const dr_unet *u, const Workspace &w, const Layer &L) {
    const Op &op = kOps[L.op];
    const int lv = kTensors[op.out].level;
    dr::ConvArgs c;
    c.a = w.act[op.a];
    c.ca = L.ca;
    c.b = op.b >= 0 ? w.act[op.b] : nullptr;
    c.cb = L.cb;
    c.wgt = u->wf + L.wb;
    c.bias = u->params + L.b;
    c.out = op.out == Y ? static_cast<void *>(w.y) : static_cast<void *>(w.act[op.out]);
    c.cout = L.cout;
    c.n = w.n, c.H = w.H >> lv, c.W = w.W >> lv;
    c.relu = op.out != Y;
    c.f32out = op.out == Y;
    c.plain = u->plainConv ? 1 : 0;
    return c;
}

// unpack: inference only; the output layer then writes the denoised beauty
// (launchConvUnpack) instead of leaving its output for a separate pass.
dr_status forward(dr_unet *u, Workspace &w, hipStream_t st, const dr::ConvArgs *unpack = nullptr) {
    size_t li = 0;
    for (int i = 0; i < kNumOps; ++i) {
        const Op &op = kOps[i];
        if (op.conv) {
            dr::ConvArgs c = convArgs(u, w, u->layers[li++]);
            // A pool of this output right after it is written by the
            // convolution's epilogue (launchConv falls back to launchPool).
            if (i + 1 < kNumOps && !kOps[i + 1].conv && kOps[i + 1].a == op.out) {
                c.poolOut = w.act[kOps[i + 1].out];
                ++i;
            }
            if (unpack && op.out == Y) {
                c.unpackAlbedo = unpack->unpackAlbedo;
                c.unpackBeauty = unpack->unpackBeauty;
                c.unpackOut = unpack->unpackOut;
                c.imgW = unpack->imgW;
                c.imgH = unpack->imgH;
                DR_UHIP(dr::launchConvUnpack(c, st), "output layer + unpack");
            } else {
                DR_UHIP(dr::launchConv(c, st), "conv forward");
            }
        } else {
            const int lv = kTensors[op.a].level;
            DR_UHIP(dr::launchPool(w.act[op.a], w.act[op.out], w.n, w.H >> lv, w.W >> lv, kTensors[op.a].c, st),
                    "pool");
        }
    }
    return DR_OK;
}

// Loss into u->loss[slot], gradients into u->grads (zeroed first).
dr_status forwardBackward(dr_unet *u, uint32_t slot) {
    Workspace &w = u->train;
    hipStream_t st = u->stream;
    dr_status rc = forward(u, w, st);
    if (rc != DR_OK) return rc;
    const size_t pixels = size_t(w.n) * w.H * w.W;
    DR_UHIP(dr::launchFill(u->grads, u->paramCount, 0.f, st), "zero grads");
    DR_UHIP(dr::launchFill(w.gradBase, w.gradFloats, 0.f, st), "zero activation grads");
    DR_UHIP(dr::launchLossL1(w.y, w.target, w.dpre, u->loss + slot, pixels, 1.f / float(pixels * 3), st), "loss");
    size_t li = u->layers.size();
    for (int i = kNumOps - 1; i >= 0; --i) {
        const Op &op = kOps[i];
        const int lv = kTensors[op.out].level;
        const int H = w.H >> lv, W = w.W >> lv;
        if (!op.conv) {
            const int la = kTensors[op.a].level;
            DR_UHIP(dr::launchPoolBwd(w.grad[op.a], w.grad[op.out], w.n, w.H >> la, w.W >> la, kTensors[op.a].c, st),
                    "pool backward");
            continue;
        }
        const Layer &L = u->layers[--li];
        const size_t outCount = size_t(w.n) * H * W * L.cout;
        if (op.out != Y)
            DR_UHIP(dr::launchReluBwd(w.dpre, w.grad[op.out], w.act[op.out], outCount, st), "relu backward");
        dr::WgradArgs g;
        g.a = w.act[op.a], g.ca = L.ca;
        g.b = op.b >= 0 ? w.act[op.b] : nullptr, g.cb = L.cb;
        g.dpre = w.dpre, g.cout = L.cout;
        g.dw = u->grads + L.w, g.db = u->grads + L.b;
        g.n = w.n, g.H = H, g.W = W;
        DR_UHIP(dr::launchWgrad(g, st), "weight gradient");
        if (op.a == X0) continue;  // no gradient into the features
        dr::ConvArgs c;
        c.a = w.dpre, c.ca = L.cout;
        c.wgt = u->wt + L.wb;
        c.out = w.din;
        c.cout = L.ca + L.cb;
        c.n = w.n, c.H = H, c.W = W;
        c.relu = 0, c.f32out = 1;
        c.plain = u->plainConv ? 1 : 0;
        DR_UHIP(dr::launchConv(c, st), "data gradient");
        DR_UHIP(dr::launchGradSlice(w.grad[op.a], w.din, L.ca + L.cb, 0, L.ca, w.n, H, W, false, st), "grad slice");
        if (op.b >= 0)
            DR_UHIP(dr::launchGradSlice(w.grad[op.b], w.din, L.ca + L.cb, L.ca, L.cb, w.n, H, W, true, st),
                    "grad upsum");
    }
    return DR_OK;
}

dr_status adam(dr_unet *u, float lr) {
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = float(1.0 / (1.0 - std::pow(double(b1), double(u->step))));
    const float c2 = float(1.0 / (1.0 - std::pow(double(b2), double(u->step))));
    DR_UHIP(dr::launchAdam(u->params, u->grads, u->m, u->v, u->paramCount, lr, b1, b2, 1e-8f, c1, c2, u->stream),
            "adam");
    return refreshBf16(u);
}

dr_status ensureLoss(dr_unet *u, uint32_t n) {
    if (n <= u->lossCap) return DR_OK;
    DR_DEVICE_LOCK;
    (void)hipStreamSynchronize(u->stream);
    if (u->loss) (void)hipFree(u->loss);
    u->loss = nullptr;
    DR_UHIP(hipMalloc(&u->loss, size_t(n) * 4), "loss buffer");
    u->lossCap = n;
    return DR_OK;
}

}  // namespace

// Learned denoising of a session's beauty into session->denoised (enqueued
// on the session stream; the workspace is the session's own).
namespace {
// Inference activations of one imgW x imgH image (padded to multiples of 16).
size_t inferLayout(int imgW, int imgH, size_t *off) {
    const int W = (imgW + 15) & ~15, H = (imgH + 15) & ~15;
    size_t bytes = 0;
    for (int t = 0; t < NT; ++t) {
        if (off) off[t] = bytes;
        bytes += align256(size_t(H >> kTensors[t].level) * (W >> kTensors[t].level) * kTensors[t].c *
                          (t == Y ? 4 : 2));
    }
    return bytes;
}
}  // namespace

size_t drUnetWorkBytes(uint32_t width, uint32_t height) { return inferLayout(int(width), int(height), nullptr); }

dr_status drUnetDenoiseSession(dr_unet *u, dr_session *s, void **work, size_t *workBytes) {
    const int imgW = int(s->settings.width), imgH = int(s->settings.height);
    const int W = (imgW + 15) & ~15, H = (imgH + 15) & ~15;
    // The session's inference workspace: activations of one image.
    Workspace w;
    std::vector<size_t> off(NT);
    const size_t bytes = inferLayout(imgW, imgH, off.data());
    if (*workBytes < bytes) {
        DR_DEVICE_LOCK;
        (void)hipStreamSynchronize(s->stream);
        if (*work) (void)hipFree(*work);
        *work = nullptr;
        *workBytes = 0;
        DR_UHIP(hipMalloc(work, bytes), "denoiser workspace");
        *workBytes = bytes;
    }
    char *base = static_cast<char *>(*work);
    w.n = 1, w.H = H, w.W = W;
    for (int t = 0; t < NT; ++t) w.act[t] = reinterpret_cast<uint16_t *>(base + off[t]);
    w.y = reinterpret_cast<float *>(base + off[Y]);
    // Parameters may have been updated on the network's stream.
    DR_UHIP(hipStreamSynchronize(u->stream), "denoiser stream");
    const float invSpp = 1.f / float(std::max(1u, s->sampleCount));
    // The features are packed by their own kernel: computed in the first
    // layer's tile fetch (halo pixels twice, two threads per pixel) that
    // layer ran 279 us instead of 88 + 34.
    DR_UHIP(dr::launchPackFeatures(s->out.accumulation, s->out.albedo, s->out.normal, s->out.depth, imgW, imgH, invSpp,
                                   w.act[X0], H, W, s->stream),
            "pack features");
    dr::ConvArgs up;
    up.unpackAlbedo = s->out.albedo;
    up.unpackBeauty = s->out.accumulation;
    up.unpackOut = s->denoised;
    up.imgW = imgW;
    up.imgH = imgH;
    return forward(u, w, s->stream, &up);
}

extern "C" {

dr_status dr_hip_unet_create(dr_device *device, uint32_t seed, dr_unet **out) {
    if (!device || !out) return ufail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    DR_UHIP(hipSetDevice(device->ordinal), "hipSetDevice");
    auto *u = new dr_unet;
    u->device = device;
    if (const char *v = std::getenv("DENRIM_HIP_UNET_PLAIN_CONV")) u->plainConv = std::atoi(v) != 0;
    buildLayers(u);
    DR_DEVICE_LOCK;
    hipError_t e = hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking);
    const size_t P = u->paramCount;
    if (e == hipSuccess) e = hipMalloc(&u->params, P * 4);
    if (e == hipSuccess) e = hipMalloc(&u->grads, P * 4);
    if (e == hipSuccess) e = hipMalloc(&u->m, P * 4);
    if (e == hipSuccess) e = hipMalloc(&u->v, P * 4);
    if (e == hipSuccess) e = hipMalloc(&u->wf, u->weightCount * 2);
    if (e == hipSuccess) e = hipMalloc(&u->wt, u->weightCount * 2);
    if (e != hipSuccess) {
        dr_hip_unet_destroy(u);
        return ufail(DR_ERR_OUT_OF_MEMORY, std::string("denoiser parameters: ") + hipGetErrorString(e));
    }
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases.
    std::vector<float> h(P, 0.f);
    std::mt19937 rng(seed);
    for (const Layer &L : u->layers) {
        const float bound = std::sqrt(6.f / float(9 * (L.ca + L.cb)));
        std::uniform_real_distribution<float> dist(-bound, bound);
        for (size_t i = 0; i < size_t(9) * L.cout * (L.ca + L.cb); ++i) h[L.w + i] = dist(rng);
    }
    dr_status rc = dr_hip_unet_set_params(u, h.data(), P);
    if (rc != DR_OK) {
        dr_hip_unet_destroy(u);
        return rc;
    }
    *out = u;
    return DR_OK;
}

dr_status dr_hip_unet_destroy(dr_unet *u) {
    if (!u) return DR_OK;
    DR_DEVICE_LOCK;
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    for (void *p : {static_cast<void *>(u->params), static_cast<void *>(u->grads), static_cast<void *>(u->m),
                    static_cast<void *>(u->v), static_cast<void *>(u->wf), static_cast<void *>(u->wt),
                    static_cast<void *>(u->loss), static_cast<void *>(u->dsFeat), static_cast<void *>(u->dsTarget)})
        if (p) (void)hipFree(p);
    freeWorkspace(u->train);
    freeWorkspace(u->infer);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
    return DR_OK;
}

dr_status dr_hip_unet_param_count(const dr_unet *u, uint64_t *count) {
    if (!u || !count) return ufail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    *count = u->paramCount;
    return DR_OK;
}

dr_status dr_hip_unet_get_params(dr_unet *u, float *p, size_t count) {
    if (!u || !p || count < u->paramCount) return ufail(DR_ERR_INVALID_ARGUMENT, "parameter buffer too small");
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    DR_UHIP(drCopy(p, u->params, u->paramCount * 4, hipMemcpyDeviceToHost, u->stream), "read parameters");
    return DR_OK;
}

dr_status dr_hip_unet_set_params(dr_unet *u, const float *p, size_t count) {
    if (!u || !p || count != u->paramCount) return ufail(DR_ERR_INVALID_ARGUMENT, "parameter count mismatch");
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    DR_UHIP(drCopy(u->params, p, u->paramCount * 4, hipMemcpyHostToDevice, u->stream), "write parameters");
    DR_UHIP(dr::launchFill(u->m, u->paramCount, 0.f, u->stream), "reset adam");
    DR_UHIP(dr::launchFill(u->v, u->paramCount, 0.f, u->stream), "reset adam");
    u->step = 0;
    dr_status rc = refreshBf16(u);
    if (rc != DR_OK) return rc;
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    return DR_OK;
}

dr_status dr_hip_unet_set_batch(dr_unet *u, const float *features, const float *targets, uint32_t n, uint32_t h,
                                uint32_t w) {
    if (!u || !features || !targets) return ufail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0 || h == 0 || w == 0 || h % 16 || w % 16)
        return ufail(DR_ERR_INVALID_ARGUMENT, "batch height and width must be positive multiples of 16");
    dr_status rc = ensureWorkspace(u, u->train, int(n), int(h), int(w), true);
    if (rc != DR_OK) return rc;
    const size_t px = size_t(n) * h * w;
    DR_DEVICE_LOCK;
    float *tmp = nullptr;
    DR_UHIP(hipMalloc(&tmp, px * 16 * 4), "staging");
    hipError_t e = drCopy(tmp, features, px * 16 * 4, hipMemcpyHostToDevice, u->stream);
    if (e == hipSuccess) e = dr::launchToBf16(tmp, u->train.act[X0], px * 16, u->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(u->stream);
    (void)hipFree(tmp);
    DR_UHIP(e, "features upload");
    DR_UHIP(drCopy(u->train.target, targets, px * 4 * 4, hipMemcpyHostToDevice, u->stream), "targets upload");
    u->batchReady = true;
    return DR_OK;
}

dr_status dr_hip_unet_forward_backward(dr_unet *u, float *loss) {
    if (!u) return ufail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!u->batchReady) return ufail(DR_ERR_INVALID_ARGUMENT, "no training batch set");
    dr_status rc = ensureLoss(u, 1);
    if (rc != DR_OK) return rc;
    DR_UHIP(dr::launchFill(u->loss, 1, 0.f, u->stream), "zero loss");
    rc = forwardBackward(u, 0);
    if (rc != DR_OK) return rc;
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    if (loss) DR_UHIP(drCopy(loss, u->loss, 4, hipMemcpyDeviceToHost, u->stream), "loss readback");
    return DR_OK;
}

dr_status dr_hip_unet_get_grads(dr_unet *u, float *g, size_t count) {
    if (!u || !g || count < u->paramCount) return ufail(DR_ERR_INVALID_ARGUMENT, "gradient buffer too small");
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    DR_UHIP(drCopy(g, u->grads, u->paramCount * 4, hipMemcpyDeviceToHost, u->stream), "read gradients");
    return DR_OK;
}

dr_status dr_hip_unet_get_output(dr_unet *u, float *out, size_t floats) {
    if (!u || !out || !u->batchReady) return ufail(DR_ERR_INVALID_ARGUMENT, "no forward pass");
    const size_t n = size_t(u->train.n) * u->train.H * u->train.W * 16;
    if (floats < n) return ufail(DR_ERR_INVALID_ARGUMENT, "output buffer too small");
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    DR_UHIP(drCopy(out, u->train.y, n * 4, hipMemcpyDeviceToHost, u->stream), "read output");
    return DR_OK;
}

dr_status dr_hip_unet_train_step(dr_unet *u, float lr, float *loss) {
    dr_status rc = dr_hip_unet_forward_backward(u, loss);
    if (rc != DR_OK) return rc;
    rc = adam(u, lr);
    if (rc != DR_OK) return rc;
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    return DR_OK;
}

dr_status dr_hip_unet_dataset_reserve(dr_unet *u, uint32_t capacity, uint32_t h, uint32_t w) {
    if (!u || capacity == 0 || h == 0 || w == 0 || h % 16 || w % 16)
        return ufail(DR_ERR_INVALID_ARGUMENT, "dataset size must be positive, h and w multiples of 16");
    DR_DEVICE_LOCK;
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    if (u->dsFeat) (void)hipFree(u->dsFeat);
    if (u->dsTarget) (void)hipFree(u->dsTarget);
    u->dsFeat = nullptr, u->dsTarget = nullptr;
    const size_t px = size_t(capacity) * h * w;
    DR_UHIP(hipMalloc(&u->dsFeat, px * 16 * 2), "dataset features");
    DR_UHIP(hipMalloc(&u->dsTarget, px * 4 * 4), "dataset targets");
    u->dsCap = capacity, u->dsCount = 0, u->dsH = h, u->dsW = w;
    return DR_OK;
}

dr_status dr_hip_unet_dataset_add(dr_unet *u, dr_session *noisy, dr_session *target, uint32_t *index) {
    if (!u || !noisy || !target) return ufail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (u->dsCount >= u->dsCap) return ufail(DR_ERR_INVALID_ARGUMENT, "dataset is full");
    const uint32_t W = noisy->settings.width, H = noisy->settings.height;
    if (W != target->settings.width || H != target->settings.height || W != u->dsW || H != u->dsH)
        return ufail(DR_ERR_INVALID_ARGUMENT, "session size differs from the dataset's example size");
    DR_UHIP(hipStreamSynchronize(noisy->stream), "noisy session");
    DR_UHIP(hipStreamSynchronize(target->stream), "target session");
    const size_t px = size_t(H) * W;
    const uint32_t k = u->dsCount;
    const float invSpp = 1.f / float(std::max(1u, noisy->sampleCount));
    DR_UHIP(dr::launchPackFeatures(noisy->out.accumulation, noisy->out.albedo, noisy->out.normal, noisy->out.depth,
                                   int(W), int(H), invSpp, u->dsFeat + k * px * 16, int(H), int(W), u->stream),
            "pack features");
    // The target is demodulated by the noisy render's albedo (the network's).
    DR_UHIP(dr::launchPackTarget(target->out.accumulation, noisy->out.albedo, int(W), int(H), u->dsTarget + k * px * 4,
                                 int(H), int(W), u->stream),
            "pack target");
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    u->dsCount++;
    if (index) *index = k;
    return DR_OK;
}

dr_status dr_hip_unet_train_dataset(dr_unet *u, uint32_t steps, uint32_t batch, float lr, float lrFinal, uint32_t seed,
                                    float *losses) {
    if (!u || steps == 0 || batch == 0) return ufail(DR_ERR_INVALID_ARGUMENT, "steps and batch must be positive");
    if (u->dsCount == 0) return ufail(DR_ERR_INVALID_ARGUMENT, "dataset is empty");
    dr_status rc = ensureWorkspace(u, u->train, int(batch), int(u->dsH), int(u->dsW), true);
    if (rc != DR_OK) return rc;
    rc = ensureLoss(u, steps);
    if (rc != DR_OK) return rc;
    DR_UHIP(dr::launchFill(u->loss, steps, 0.f, u->stream), "zero losses");
    std::mt19937 rng(seed);
    std::uniform_int_distribution<uint32_t> pick(0, u->dsCount - 1);
    const size_t px = size_t(u->dsH) * u->dsW;
    for (uint32_t s = 0; s < steps; ++s) {
        for (uint32_t b = 0; b < batch; ++b) {
            const uint32_t k = pick(rng);
            DR_UHIP(hipMemcpyAsync(u->train.act[X0] + b * px * 16, u->dsFeat + k * px * 16, px * 32,
                                   hipMemcpyDeviceToDevice, u->stream),
                    "batch features");
            DR_UHIP(hipMemcpyAsync(u->train.target + b * px * 4, u->dsTarget + k * px * 4, px * 16,
                                   hipMemcpyDeviceToDevice, u->stream),
                    "batch targets");
        }
        // Cosine decay from lr to lrFinal.
        const float f = steps > 1 ? float(s) / float(steps - 1) : 1.f;
        const float rate = lrFinal + 0.5f * (lr - lrFinal) * (1.f + std::cos(3.14159265f * f));
        rc = forwardBackward(u, s);
        if (rc != DR_OK) return rc;
        rc = adam(u, rate);
        if (rc != DR_OK) return rc;
    }
    DR_UHIP(hipStreamSynchronize(u->stream), "stream synchronize");
    u->batchReady = true;
    if (losses) DR_UHIP(drCopy(losses, u->loss, size_t(steps) * 4, hipMemcpyDeviceToHost, u->stream), "loss readback");
    return DR_OK;
}

}  // extern "C"
