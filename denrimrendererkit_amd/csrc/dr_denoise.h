// dr_denoise.h -- denoiser kernels: the reference's a-trous bilateral filter
// and the building blocks of the learned U-Net (BF16 MFMA convolutions with
// FP32 accumulation, their backward passes, pooling, loss and Adam).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dr {

// GPUDenoiseConstants as SimpleSpatialDenoiser resolves them
// (Denoising/Denoiser.swift:112-164).
struct AtrousParams {
    uint32_t width, height, radius, step;
    float normalSigma, depthSigma, albedoSigma, colorSigma;
};
hipError_t launchAtrous(const float4 *src, const float4 *depth, const float4 *normal, const float4 *albedo, float4 *dst,
                        const AtrousParams &p, hipStream_t stream);

// 3x3 convolution, stride 1, zero padding, NHWC BF16 tensors. The input is
// the channel concatenation of `a` (ca channels, H x W) and `b` (cb
// channels, H/2 x W/2, read nearest-upsampled); weights BF16
// [9][cout][ca + cb] (tap = 3 * (dy + 1) + (dx + 1)). Output = bias +
// sum, optionally ReLU, BF16 or FP32 NHWC [n][H][W][cout]. All channel
// counts are multiples of 16; H and W even when b is used.
struct ConvArgs {
    const uint16_t *a = nullptr;
    int ca = 0;
    const uint16_t *b = nullptr;
    int cb = 0;
    const uint16_t *wgt = nullptr;
    const float *bias = nullptr;
    void *out = nullptr;
    int cout = 0;
    int n = 0, H = 0, W = 0;
    int relu = 0;
    int f32out = 0;
    // 1: always the per-layer LDS kernel (no persistent pipelined kernel, no
    // output-channel split); the other paths give the same bits, and a test
    // checks that against this one.
    int plain = 0;
    // Output layer of an inference (f32out, n = 1): when unpackOut is set,
    // the denoised beauty of the imgW x imgH image is written to unpackOut
    // as launchUnpack computes it from this layer's output and the planes
    // below; `out` is written only where the layer cannot write the beauty
    // itself (launchConvUnpack then runs the unpack kernel over it).
    const float4 *unpackAlbedo = nullptr;
    const float4 *unpackBeauty = nullptr;
    float4 *unpackOut = nullptr;
    int imgW = 0, imgH = 0;
    // When set (BF16 outputs, H and W even), the 2x2 average pool of the
    // output (launchPool's values, n x H/2 x W/2 x cout) is written here too.
    uint16_t *poolOut = nullptr;
};
hipError_t launchConv(const ConvArgs &c, hipStream_t stream);
// launchConv of an output layer with unpackOut set: the denoised beauty
// directly from the pipelined kernel's epilogue, else the convolution
// followed by launchUnpack (the same values either way).
hipError_t launchConvUnpack(const ConvArgs &c, hipStream_t stream);

// Weight and bias gradients of a ConvArgs convolution: dw[9][cout][cin] +=
// sum over pixels of in(p + tap) * dpre(p), db[cout] += sum dpre (FP32
// atomics), dpre = BF16 gradient at the convolution's output (before ReLU).
struct WgradArgs {
    const uint16_t *a = nullptr;
    int ca = 0;
    const uint16_t *b = nullptr;
    int cb = 0;
    const uint16_t *dpre = nullptr;
    int cout = 0;
    float *dw = nullptr;
    float *db = nullptr;
    int n = 0, H = 0, W = 0;
};
hipError_t launchWgrad(const WgradArgs &g, hipStream_t stream);

hipError_t launchPool(const uint16_t *in, uint16_t *out, int n, int H, int W, int C, hipStream_t stream);
hipError_t launchPoolBwd(float *din, const float *dout, int n, int H, int W, int C, hipStream_t stream);
// dst += src[..., off:off + C] (src has srcC channels, n x H x W); with
// upsum, dst is n x H/2 x W/2 and receives the sum of each 2x2 block.
hipError_t launchGradSlice(float *dst, const float *src, int srcC, int off, int C, int n, int H, int W, bool upsum,
                           hipStream_t stream);
hipError_t launchReluBwd(uint16_t *dpre, const float *dout, const uint16_t *out, size_t count, hipStream_t stream);
// Mean L1 over channels 0-2: pred FP32 [pixels][16], target FP32
// [pixels][4]; dpre BF16 [pixels][16] = sign(pred - target) * scale in
// channels 0-2, else 0; lossSum += scale * sum |pred - target|.
hipError_t launchLossL1(const float *pred, const float *target, uint16_t *dpre, float *lossSum, size_t pixels,
                        float scale, hipStream_t stream);
hipError_t launchAdam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                      float c1, float c2, hipStream_t stream);
// BF16 copies of FP32 weights [9][cout][cin]: wf in the same layout, wt
// flipped and transposed [8 - tap][cin][cout] for the data gradient.
hipError_t launchWeightsBf16(const float *w, uint16_t *wf, uint16_t *wt, int cout, int cin, hipStream_t stream);
hipError_t launchFill(float *p, size_t n, float v, hipStream_t stream);
hipError_t launchToBf16(const float *src, uint16_t *dst, size_t n, hipStream_t stream);

// Network input of one image (H x W, multiples of 16, edge-clamped over
// the imgW x imgH render): per pixel 16 BF16 channels
//   0-2 log(1 + max(beauty, 0) / (albedo + 0.01)), 3-5 albedo, 6-8 encoded
//   normal, 9 log(1 + max(depth, 0)), 10 1/spp, 11 beauty alpha, 12-15 zero.
hipError_t launchPackFeatures(const float4 *beauty, const float4 *albedo, const float4 *normal, const float4 *depth,
                              int imgW, int imgH, float invSpp, uint16_t *x, int H, int W, hipStream_t stream);
// Target of one image: channels 0-2 log(1 + max(beauty, 0) / (albedo + 0.01)).
hipError_t launchPackTarget(const float4 *beauty, const float4 *albedo, int imgW, int imgH, float *t, int H, int W,
                            hipStream_t stream);
// Denoised beauty: (exp(max(y, 0)) - 1) * (albedo + 0.01), alpha of the input beauty.
hipError_t launchUnpack(const float *y, const float4 *albedo, const float4 *beauty, int imgW, int imgH, int W,
                        float4 *out, hipStream_t stream);

}  // namespace dr
