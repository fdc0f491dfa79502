// dr_denoise.hip -- denoiser kernels (declarations and contracts in dr_denoise.h).
//
// The a-trous filter restates simpleSpatialDenoiseKernel
// (Resources/Shaders/Denoise.metal:39-115). The U-Net kernels are this
// backend's own: 3x3 convolutions as implicit GEMMs on the gfx950 matrix
// cores (BF16 operands, FP32 accumulation). The forward and data-gradient
// convolution runs v_mfma_f32_16x16x32_bf16 on 16-pixel row blocks; the
// weight gradient runs v_mfma_f32_16x16x16_bf16 over 8x8 pixel tiles of the
// output gradient and the shifted input staged in LDS, transposed so that
// each MFMA operand is one 8-byte LDS read, reduced over pixels in registers.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dr_denoise.h"

namespace dr {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> BF16 (NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// ---------------------------------------------------------------- a-trous

// B3-spline taps of the 5x5 kernel (Denoise.metal:29-38).
__device__ __forceinline__ float atrousTap(int o) {
    o = o < 0 ? -o : o;
    return o == 0 ? 0.375f : (o == 1 ? 0.25f : 0.0625f);
}

__device__ __forceinline__ float3 decodeNormal(float4 e) {
    float3 n = make_float3(e.x * 2.f - 1.f, e.y * 2.f - 1.f, e.z * 2.f - 1.f);
    float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    return make_float3(n.x / len, n.y / len, n.z / len);
}

__device__ __forceinline__ float dist2(float3 a, float3 b) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

__global__ __launch_bounds__(256) void k_atrous(const float4 *__restrict__ src, const float4 *__restrict__ depth,
                                                const float4 *__restrict__ normal, const float4 *__restrict__ albedo,
                                                float4 *__restrict__ dst, AtrousParams p) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= int(p.width) || y >= int(p.height)) return;
    const int W = int(p.width), H = int(p.height);
    const size_t idx = size_t(y) * W + x;
    const float4 c = src[idx];
    const float cd = depth[idx].x;
    if (c.w <= 0.0001f || !isfinite(cd) || cd <= 0.f) {  // Denoise.metal:56-60
        dst[idx] = c;
        return;
    }
    const float3 cn = decodeNormal(normal[idx]);
    const float4 ca4 = albedo[idx];
    const float3 ca = make_float3(ca4.x, ca4.y, ca4.z), cc = make_float3(c.x, c.y, c.z);
    const int r = int(p.radius), step = int(p.step > 0 ? p.step : 1);
    const float ns2 = p.normalSigma * p.normalSigma;
    const float ds = fmaxf(p.depthSigma * fmaxf(cd, 1.f), 0.0001f), ds2 = ds * ds;
    const float as2 = p.albedoSigma * p.albedoSigma, cs2 = p.colorSigma * p.colorSigma;
    float sr = 0.f, sg = 0.f, sb = 0.f, tot = 0.f;
    for (int oy = -r; oy <= r; ++oy) {
        for (int ox = -r; ox <= r; ++ox) {
            const int sx = min(max(x + ox * step, 0), W - 1), sy = min(max(y + oy * step, 0), H - 1);
            const size_t j = size_t(sy) * W + sx;
            const float4 s = src[j];
            const float sd = depth[j].x;
            if (s.w <= 0.0001f || !isfinite(sd) || sd <= 0.f) continue;
            const float spatial = atrousTap(ox) * atrousTap(oy);
            const float3 sn = decodeNormal(normal[j]);
            const float nd = 1.f - fminf(fmaxf(cn.x * sn.x + cn.y * sn.y + cn.z * sn.z, 0.f), 1.f);
            const float wn = expf(-(nd * nd) / (2.f * ns2));
            const float dd = sd - cd;
            const float wd = expf(-(dd * dd) / (2.f * ds2));
            const float4 sa4 = albedo[j];
            const float wa = expf(-dist2(ca, make_float3(sa4.x, sa4.y, sa4.z)) / (2.f * as2));
            const float wc = expf(-dist2(cc, make_float3(s.x, s.y, s.z)) / (2.f * cs2));
            const float w = spatial * wn * wd * wa * wc;
            sr += s.x * w;
            sg += s.y * w;
            sb += s.z * w;
            tot += w;
        }
    }
    float dr_ = c.x, dg = c.y, db = c.z;
    if (tot > 0.f) {
        dr_ = sr / tot;
        dg = sg / tot;
        db = sb / tot;
    }
    // mix(centre, denoised, 0.85) with mix(x, y, a) = x + (y - x) * a.
    dst[idx] = make_float4(c.x + (dr_ - c.x) * 0.85f, c.y + (dg - c.y) * 0.85f, c.z + (db - c.z) * 0.85f, c.w);
}

// ------------------------------------------------------------ convolution

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// One workgroup = a 16-pixel-wide tile of 4 * MB rows; wave w computes rows
// w * MB ... w * MB + MB - 1 (one 16-pixel MFMA row block each) for all cout
// channels (NB blocks of 16). K runs over (tap, 32-channel chunk): the
// lane's A operand is 8 channels of one pixel (16 bytes), its B operand 8
// weights of one output channel, and each weight fragment is reused by the
// MB row blocks. Channel chunks past cin (cin % 32 == 16) read zeros.
template <int NB, int MB>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs c) {
    constexpr int TH = 4 * MB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tilesX = (c.W + 15) >> 4, tilesY = (c.H + TH - 1) / TH;
    int t = blockIdx.x;
    const int img = t / (tilesX * tilesY);
    t -= img * tilesX * tilesY;
    const int ty = t / tilesX, tx = t - ty * tilesX;
    const int px = tx * 16 + (lane & 15);
    const int row0 = ty * TH + wave * MB;
    const int kg = (lane >> 4) * 8;
    const int cin = c.ca + c.cb;
    const int Hb = c.H >> 1, Wb = c.W >> 1;
    f4 acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
    const bf8 zero8 = {};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const int sx = px + dx;
        const bool xin = sx >= 0 && sx < c.W;
        const uint16_t *wt = c.wgt + size_t(tap) * c.cout * cin + size_t(lane & 15) * cin;
        for (int kc = 0; kc < cin; kc += 32) {
            const int k0 = kc + kg;
            const bool kin = k0 < cin;
            bf8 bv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                bv[nb] = kin ? *reinterpret_cast<const bf8 *>(wt + size_t(nb) * 16 * cin + k0) : zero8;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const int sy = row0 + mb + dy;
                bf8 av = zero8;
                if (kin && xin && sy >= 0 && sy < c.H) {
                    const uint16_t *src =
                        k0 < c.ca ? c.a + ((size_t(img) * c.H + sy) * c.W + sx) * c.ca + k0
                                  : c.b + ((size_t(img) * Hb + (sy >> 1)) * Wb + (sx >> 1)) * c.cb + (k0 - c.ca);
                    av = *reinterpret_cast<const bf8 *>(src);
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[nb], acc[mb][nb], 0, 0, 0);
            }
        }
    }
    // D: lane holds rows (pixels) 4 * (lane / 16) + j, column (channel) lane % 16.
    const int co0 = lane & 15;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int oy = row0 + mb;
        if (oy >= c.H) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ox = tx * 16 + (lane >> 4) * 4 + j;
            if (ox >= c.W) continue;
            const size_t o = ((size_t(img) * c.H + oy) * c.W + ox) * c.cout;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = nb * 16 + co0;
                float v = acc[mb][nb][j] + (c.bias ? c.bias[co] : 0.f);
                if (c.relu) v = fmaxf(v, 0.f);
                if (c.f32out) static_cast<float *>(c.out)[o + co] = v;
                else static_cast<uint16_t *>(c.out)[o + co] = f2bf(v);
            }
        }
    }
}

// The same convolution with the input tile staged in LDS: per 32-channel
// chunk the workgroup loads its (4 * MB + 2) x 18 pixels (tile plus 1-pixel
// halo) once, and the nine taps read their A operands from LDS instead of
// re-reading global memory nine times. Pixels are 80 B apart in LDS (32
// channels + 8 of padding), so the 16 lanes of a channel group read 16
// disjoint 4-bank spans. Accumulation runs chunk-major (tap-major inside),
// which only reorders the FP32 sums of inputs wider than 32 channels.
// Whether the output tile of k_conv3x3_lds (4 * MB rows x 16 pixels x
// 16 * NB channels of `bytes` each, 8 elements of padding per pixel) fits
// the LDS space of its input tile.
template <int NB, int MB>
constexpr bool outTileFits(int bytes) {
    return 4 * MB * 16 * (NB * 16 + 8) * bytes <= (4 * MB + 2) * 18 * 40 * 2;
}

// 2x2 average pool of a TH x 16-pixel output tile held in LDS (BF16,
// `ost` elements per pixel, channels [coBase, coBase + chans) of c.cout) into
// c.poolOut, with k_pool's arithmetic ((a + b) + (c + d)) * 0.25.
__device__ __forceinline__ void poolTile(const ConvArgs &c, const uint16_t *tile, int ost, int TH, int chans, int img,
                                         int ty, int tx, int coBase) {
    const int Ho = c.H >> 1, Wo = c.W >> 1, vc = chans / 8;
    for (int v = threadIdx.x; v < (TH / 2) * 8 * vc; v += blockDim.x) {
        const int pp = v / vc, cv = v - pp * vc;
        const int py = pp >> 3, px = pp & 7;
        const int oy = ty * (TH / 2) + py, ox = tx * 8 + px;
        if (oy >= Ho || ox >= Wo) continue;
        auto at = [&](int r, int x) { return *reinterpret_cast<const uint4 *>(&tile[(r * 16 + x) * ost + cv * 8]); };
        const uint4 a = at(2 * py, 2 * px), b = at(2 * py, 2 * px + 1), cc = at(2 * py + 1, 2 * px),
                    d = at(2 * py + 1, 2 * px + 1);
        const uint32_t wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w};
        const uint32_t wc[4] = {cc.x, cc.y, cc.z, cc.w}, wd[4] = {d.x, d.y, d.z, d.w};
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t lohi[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int sh = h * 16;
                const float val = ((bf2f(uint16_t(wa[k] >> sh)) + bf2f(uint16_t(wb[k] >> sh))) +
                                   (bf2f(uint16_t(wc[k] >> sh)) + bf2f(uint16_t(wd[k] >> sh)))) *
                                  0.25f;
                lohi[h] = f2bf(val);
            }
            o[k] = lohi[0] | (lohi[1] << 16);
        }
        *reinterpret_cast<uint4 *>(c.poolOut + ((size_t(img) * Ho + oy) * Wo + ox) * c.cout + coBase + cv * 8) =
            make_uint4(o[0], o[1], o[2], o[3]);
    }
}

template <int NB, int MB>
__global__ __launch_bounds__(256) void k_conv3x3_lds(ConvArgs c) {
    constexpr int TH = 4 * MB, LR = TH + 2, LC = 18, PIX = 40;
    // Up to 64 output channels the chunk's weights (9 taps x cout x 32
    // channels, zero past cin, rows padded to 40 elements like the pixels)
    // are staged in LDS with the input tile: one round trip per chunk
    // instead of one per tap, and one copy per workgroup instead of per wave.
    constexpr bool WLDS = NB <= 4;
    constexpr int WROW = 40;
    __shared__ __attribute__((aligned(16))) uint16_t tile[LR * LC * PIX];
    __shared__ __attribute__((aligned(16))) uint16_t wlds[WLDS ? 9 * NB * 16 * WROW : 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tilesX = (c.W + 15) >> 4, tilesY = (c.H + TH - 1) / TH;
    int t = blockIdx.x;
    const int img = t / (tilesX * tilesY);
    t -= img * tilesX * tilesY;
    const int ty = t / tilesX, tx = t - ty * tilesX;
    const int row0 = ty * TH + wave * MB;
    const int kg = (lane >> 4) * 8;
    // Output channels [coBase, coBase + 16 * NB) of c.cout: layers with few
    // tiles spread their channel groups over blockIdx.y.
    const int coBase = blockIdx.y * NB * 16;
    const int cin = c.ca + c.cb;
    const int Hb = c.H >> 1, Wb = c.W >> 1;
    f4 acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
    const bf8 zero8 = {};
    for (int kc = 0; kc < cin; kc += 32) {
        __syncthreads();  // the previous chunk's reads are done
        for (int v = threadIdx.x; v < LR * LC * 4; v += 256) {
            const int pix = v >> 2, q = v & 3;
            const int ly = pix / LC, lx = pix - ly * LC;
            const int sy = ty * TH - 1 + ly, sx = tx * 16 - 1 + lx;
            const int k0 = kc + q * 8;
            bf8 val = zero8;
            if (k0 < cin && sy >= 0 && sy < c.H && sx >= 0 && sx < c.W) {
                const uint16_t *src =
                    k0 < c.ca ? c.a + ((size_t(img) * c.H + sy) * c.W + sx) * c.ca + k0
                              : c.b + ((size_t(img) * Hb + (sy >> 1)) * Wb + (sx >> 1)) * c.cb + (k0 - c.ca);
                val = *reinterpret_cast<const bf8 *>(src);
            }
            *reinterpret_cast<bf8 *>(&tile[pix * PIX + q * 8]) = val;
        }
        if constexpr (WLDS) {
            for (int v = threadIdx.x; v < 9 * NB * 16 * 4; v += 256) {
                const int row = v >> 2, q = v & 3;  // row = tap * 16 * NB + output channel - coBase
                const int tap = row / (NB * 16), col = row - tap * (NB * 16);
                const int k = kc + q * 8;
                bf8 val = zero8;
                if (k < cin)
                    val = *reinterpret_cast<const bf8 *>(c.wgt + (size_t(tap) * c.cout + coBase + col) * cin + k);
                *reinterpret_cast<bf8 *>(&wlds[row * WROW + q * 8]) = val;
            }
        }
        __syncthreads();
        const int k0 = kc + kg;
        const bool kin = k0 < cin;
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
            bf8 bv[NB];
            if constexpr (WLDS) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    bv[nb] = *reinterpret_cast<const bf8 *>(&wlds[(tap * NB * 16 + nb * 16 + (lane & 15)) * WROW + kg]);
            } else {
                const uint16_t *wt = c.wgt + size_t(tap) * c.cout * cin + size_t(coBase + (lane & 15)) * cin;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    bv[nb] = kin ? *reinterpret_cast<const bf8 *>(wt + size_t(nb) * 16 * cin + k0) : zero8;
            }
            const int lx = (lane & 15) + dx + 1;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const int ly = wave * MB + mb + dy + 1;
                const bf8 av = *reinterpret_cast<const bf8 *>(&tile[(ly * LC + lx) * PIX + kg]);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[nb], acc[mb][nb], 0, 0, 0);
            }
        }
    }
    // D: lane holds rows (pixels) 4 * (lane / 16) + j, column (channel) lane % 16.
    const int co0 = lane & 15;
    float bias[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bias[nb] = c.bias ? c.bias[coBase + nb * 16 + co0] : 0.f;
    if (c.f32out ? outTileFits<NB, MB>(4) : outTileFits<NB, MB>(2)) {
        // Through LDS: the tile's outputs (bias, ReLU, BF16 or FP32) go to
        // the input tile's space, pixel-major, then each thread stores whole
        // 16-byte vectors of one pixel's channels.
        constexpr int OST = NB * 16 + 8;  // BF16 elements per pixel (8 of padding)
        __syncthreads();                  // every wave is done with the input tile
        const bool f32 = c.f32out != 0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = (wave * MB + mb) * 16 + (lane >> 4) * 4 + j;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    const int co = nb * 16 + co0;
                    float v = acc[mb][nb][j] + bias[nb];
                    if (c.relu) v = fmaxf(v, 0.f);
                    if (f32) reinterpret_cast<float *>(tile)[p * OST + co] = v;
                    else tile[p * OST + co] = f2bf(v);
                }
            }
        __syncthreads();
        const int esize = f32 ? 4 : 2;
        const int vecs = NB * 16 * esize / 16;  // 16-byte vectors per pixel
        for (int v = threadIdx.x; v < TH * 16 * vecs; v += 256) {
            const int p = v / vecs, q = v - p * vecs;
            const int oy = ty * TH + (p >> 4), ox = tx * 16 + (p & 15);
            if (oy >= c.H || ox >= c.W) continue;
            const size_t o = ((size_t(img) * c.H + oy) * c.W + ox) * c.cout + coBase;
            const uint4 val = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(tile) +
                                                               size_t(p) * OST * esize + q * 16);
            *reinterpret_cast<uint4 *>(static_cast<char *>(c.out) + o * esize + q * 16) = val;
        }
        if (c.poolOut && !f32) poolTile(c, tile, OST, TH, NB * 16, img, ty, tx, coBase);
        return;
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int oy = row0 + mb;
        if (oy >= c.H) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ox = tx * 16 + (lane >> 4) * 4 + j;
            if (ox >= c.W) continue;
            const size_t o = ((size_t(img) * c.H + oy) * c.W + ox) * c.cout;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = coBase + nb * 16 + co0;
                float v = acc[mb][nb][j] + bias[nb];
                if (c.relu) v = fmaxf(v, 0.f);
                if (c.f32out) static_cast<float *>(c.out)[o + co] = v;
                else static_cast<uint16_t *>(c.out)[o + co] = f2bf(v);
            }
        }
    }
}

// The LDS convolution as a persistent, pipelined kernel for full-resolution
// layers (16 * NB output channels, NB <= 2, 16-row tiles): a workgroup
// stages every channel chunk's weights in LDS once, then walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ...; while the matrix cores work on
// one (tile, chunk) item from LDS, the next item's input tile is already
// being fetched into registers. Accumulation order and epilogue are those of
// k_conv3x3_lds, so the outputs are the same bits.
// UNPACK: the epilogue writes the denoised beauty as k_unpack does.
template <int NB, int CHUNKS, bool UNPACK = false>
__global__ __launch_bounds__(256) void k_conv3x3_pipe(ConvArgs c) {
    constexpr int MB = 4, TH = 16, LR = TH + 2, LC = 18, PIX = 40, WROW = 40;
    constexpr int PIECES = LR * LC * 4;             // 16-byte pieces of one input tile chunk
    constexpr int PER = (PIECES + 255) / 256;       // pieces per thread
    __shared__ __attribute__((aligned(16))) uint16_t tile[LR * LC * PIX];
    __shared__ __attribute__((aligned(16))) uint16_t wlds[CHUNKS * 9 * NB * 16 * WROW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tilesX = (c.W + 15) >> 4, tilesY = (c.H + TH - 1) / TH;
    const int tiles = c.n * tilesX * tilesY;
    const int kg = (lane >> 4) * 8;
    const int cin = c.ca + c.cb;
    const int Hb = c.H >> 1, Wb = c.W >> 1;
    const bf8 zero8 = {};
    for (int v = threadIdx.x; v < CHUNKS * 9 * NB * 16 * 4; v += 256) {
        const int row = v >> 2, q = v & 3;  // row = (chunk * 9 + tap) * 16 * NB + output channel
        const int chunk = row / (9 * NB * 16), r = row - chunk * (9 * NB * 16);
        const int tap = r / (NB * 16), col = r - tap * (NB * 16);
        const int k = chunk * 32 + q * 8;
        bf8 val = zero8;
        if (k < cin) val = *reinterpret_cast<const bf8 *>(c.wgt + (size_t(tap) * c.cout + col) * cin + k);
        *reinterpret_cast<bf8 *>(&wlds[row * WROW + q * 8]) = val;
    }
    // Fetches item (t, chunk)'s input tile into registers.
    bf8 pre[PER];
    auto fetch = [&](int t, int chunk) {
        const int img = t / (tilesX * tilesY);
        const int tt = t - img * tilesX * tilesY;
        const int ty = tt / tilesX, tx = tt - ty * tilesX;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int v = threadIdx.x + i * 256;
            const int pix = v >> 2, q = v & 3;
            const int ly = pix / LC, lx = pix - ly * LC;
            const int sy = ty * TH - 1 + ly, sx = tx * 16 - 1 + lx;
            const int k0 = chunk * 32 + q * 8;
            bf8 val = zero8;
            if (v < PIECES && k0 < cin && sy >= 0 && sy < c.H && sx >= 0 && sx < c.W) {
                const uint16_t *src =
                    k0 < c.ca ? c.a + ((size_t(img) * c.H + sy) * c.W + sx) * c.ca + k0
                              : c.b + ((size_t(img) * Hb + (sy >> 1)) * Wb + (sx >> 1)) * c.cb + (k0 - c.ca);
                val = *reinterpret_cast<const bf8 *>(src);
            }
            pre[i] = val;
        }
    };
    int t = blockIdx.x, chunk = 0;
    if (t < tiles) fetch(t, 0);
    const int co0 = lane & 15;
    float bias[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bias[nb] = c.bias ? c.bias[nb * 16 + co0] : 0.f;
    f4 acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
    while (t < tiles) {
        __syncthreads();  // the previous item's reads of the tile (and the epilogue's) are done
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int v = threadIdx.x + i * 256;
            if (v < PIECES) *reinterpret_cast<bf8 *>(&tile[(v >> 2) * PIX + (v & 3) * 8]) = pre[i];
        }
        __syncthreads();
        // Next item: the following chunk of this tile, else the next tile.
        const int nt = chunk + 1 < CHUNKS ? t : t + int(gridDim.x);
        const int nc = chunk + 1 < CHUNKS ? chunk + 1 : 0;
        if (nt < tiles) fetch(nt, nc);
        const uint16_t *w = wlds + chunk * (9 * NB * 16 * WROW);
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
            bf8 bv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                bv[nb] = *reinterpret_cast<const bf8 *>(&w[(tap * NB * 16 + nb * 16 + co0) * WROW + kg]);
            const int lx = co0 + dx + 1;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const int ly = wave * MB + mb + dy + 1;
                const bf8 av = *reinterpret_cast<const bf8 *>(&tile[(ly * LC + lx) * PIX + kg]);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[nb], acc[mb][nb], 0, 0, 0);
            }
        }
        if (chunk + 1 == CHUNKS) {
            // Epilogue of tile t through the tile's LDS, as in k_conv3x3_lds.
            const int img = t / (tilesX * tilesY);
            const int tt = t - img * tilesX * tilesY;
            const int ty = tt / tilesX, tx = tt - ty * tilesX;
            constexpr int OST = NB * 16 + 8;
            const bool f32 = c.f32out != 0;
            __syncthreads();
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int p = (wave * MB + mb) * 16 + (lane >> 4) * 4 + j;
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) {
                        const int co = nb * 16 + co0;
                        float v = acc[mb][nb][j] + bias[nb];
                        if (c.relu) v = fmaxf(v, 0.f);
                        if (f32) reinterpret_cast<float *>(tile)[p * OST + co] = v;
                        else tile[p * OST + co] = f2bf(v);
                    }
                }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
            __syncthreads();
            if (UNPACK) {
                // The denoised beauty of this tile's image pixels (k_unpack's
                // arithmetic on the same FP32 outputs).
                const float *yt = reinterpret_cast<const float *>(tile);
                for (int p = threadIdx.x; p < TH * 16; p += 256) {
                    const int oy = ty * TH + (p >> 4), ox = tx * 16 + (p & 15);
                    if (oy >= c.imgH || ox >= c.imgW) continue;
                    const size_t i = size_t(oy) * c.imgW + ox;
                    const float *yp = yt + p * OST;
                    const float4 a = c.unpackAlbedo[i];
                    c.unpackOut[i] = make_float4(expm1f(fminf(fmaxf(yp[0], 0.f), 30.f)) * (a.x + 0.01f),
                                                 expm1f(fminf(fmaxf(yp[1], 0.f), 30.f)) * (a.y + 0.01f),
                                                 expm1f(fminf(fmaxf(yp[2], 0.f), 30.f)) * (a.z + 0.01f),
                                                 c.unpackBeauty[i].w);
                }
                t = nt;
                chunk = nc;
                continue;
            }
            const int esize = f32 ? 4 : 2;
            const int vecs = NB * 16 * esize / 16;
            for (int v = threadIdx.x; v < TH * 16 * vecs; v += 256) {
                const int p = v / vecs, q = v - p * vecs;
                const int oy = ty * TH + (p >> 4), ox = tx * 16 + (p & 15);
                if (oy >= c.H || ox >= c.W) continue;
                const size_t o = ((size_t(img) * c.H + oy) * c.W + ox) * c.cout;
                const uint4 val = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(tile) +
                                                                   size_t(p) * OST * esize + q * 16);
                *reinterpret_cast<uint4 *>(static_cast<char *>(c.out) + o * esize + q * 16) = val;
            }
            if (c.poolOut && !f32) poolTile(c, tile, OST, TH, NB * 16, img, ty, tx, 0);
        }
        t = nt;
        chunk = nc;
    }
}

constexpr int kWgRow = 64 + 4;  // LDS row: 64 pixels + 4 elements of padding
constexpr int kWgMaxQ = 16;     // weight tiles per wave (cout/16 * cin/16 <= 64)

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs g) {
    extern __shared__ uint16_t lds[];
    const int cin = g.ca + g.cb;
    uint16_t *sD = lds;                    // [cout][kWgRow]: output gradient, pixel-minor
    uint16_t *sI = lds + g.cout * kWgRow;  // [cin][kWgRow]: shifted input, pixel-minor
    const int tap = blockIdx.y, dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kq = (lane >> 4) * 4;
    const int nbk = cin / 16, tiles = (g.cout / 16) * nbk;
    const int tilesX = (g.W + 7) >> 3, tilesY = (g.H + 7) >> 3, ntile = g.n * tilesX * tilesY;
    const int Hb = g.H >> 1, Wb = g.W >> 1;
    const bool bias = g.db && tap == 4;
    f4 acc[kWgMaxQ];
#pragma unroll
    for (int q = 0; q < kWgMaxQ; ++q) acc[q] = f4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f;
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int img = t / (tilesX * tilesY), r = t - img * (tilesX * tilesY);
        const int ty = r / tilesX, tx = r - ty * tilesX;
        const int vo = g.cout / 8;
        for (int e = threadIdx.x; e < 64 * vo; e += 256) {
            const int p = e / vo, cv = e - p * vo;
            const int y = ty * 8 + (p >> 3), x = tx * 8 + (p & 7);
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (y < g.H && x < g.W)
                v = *reinterpret_cast<const uint4 *>(g.dpre + ((size_t(img) * g.H + y) * g.W + x) * g.cout + cv * 8);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) sD[(cv * 8 + k) * kWgRow + p] = uint16_t(w4[k >> 1] >> ((k & 1) * 16));
        }
        const int vi = cin / 8;
        for (int e = threadIdx.x; e < 64 * vi; e += 256) {
            const int p = e / vi, cv = e - p * vi, c8 = cv * 8;
            const int y = ty * 8 + (p >> 3) + dy, x = tx * 8 + (p & 7) + dx;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (y >= 0 && y < g.H && x >= 0 && x < g.W) {
                if (c8 < g.ca)
                    v = *reinterpret_cast<const uint4 *>(g.a + ((size_t(img) * g.H + y) * g.W + x) * g.ca + c8);
                else
                    v = *reinterpret_cast<const uint4 *>(g.b + ((size_t(img) * Hb + (y >> 1)) * Wb + (x >> 1)) * g.cb +
                                                         (c8 - g.ca));
            }
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) sI[(c8 + k) * kWgRow + p] = uint16_t(w4[k >> 1] >> ((k & 1) * 16));
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kWgMaxQ; ++q) {
            const int tt = wave + 4 * q;
            if (tt < tiles) {
                const int mb = tt / nbk, nb = tt - mb * nbk;
                const uint16_t *ra = sD + (mb * 16 + (lane & 15)) * kWgRow + kq;
                const uint16_t *rb = sI + (nb * 16 + (lane & 15)) * kWgRow + kq;
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const s4 av = *reinterpret_cast<const s4 *>(ra + ks * 16);
                    const s4 bv = *reinterpret_cast<const s4 *>(rb + ks * 16);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[q], 0, 0, 0);
                }
            }
        }
        if (bias && int(threadIdx.x) < g.cout)
            for (int p = 0; p < 64; ++p) bacc += bf2f(sD[threadIdx.x * kWgRow + p]);
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kWgMaxQ; ++q) {
        const int tt = wave + 4 * q;
        if (tt < tiles) {
            const int mb = tt / nbk, nb = tt - mb * nbk;
            const int ci = nb * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = mb * 16 + kq + j;
                unsafeAtomicAdd(&g.dw[(size_t(tap) * g.cout + co) * cin + ci], acc[q][j]);
            }
        }
    }
    if (bias && int(threadIdx.x) < g.cout) unsafeAtomicAdd(&g.db[threadIdx.x], bacc);
}

// ---------------------------------------------------------- elementwise

__global__ void k_pool(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int n, int H, int W, int C) {
    const int Ho = H >> 1, Wo = W >> 1, vc = C / 8;
    const size_t total = size_t(n) * Ho * Wo * vc;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int cv = int(i % vc);
    size_t rest = i / vc;
    const int x = int(rest % Wo);
    rest /= Wo;
    const int y = int(rest % Ho), img = int(rest / Ho);
    auto at = [&](int yy, int xx) {
        return *reinterpret_cast<const uint4 *>(in + ((size_t(img) * H + yy) * W + xx) * C + cv * 8);
    };
    const uint4 a = at(2 * y, 2 * x), b = at(2 * y, 2 * x + 1), c = at(2 * y + 1, 2 * x), d = at(2 * y + 1, 2 * x + 1);
    const uint32_t wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w};
    const uint32_t wc[4] = {c.x, c.y, c.z, c.w}, wd[4] = {d.x, d.y, d.z, d.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t lohi[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sh = h * 16;
            const float v = ((bf2f(uint16_t(wa[k] >> sh)) + bf2f(uint16_t(wb[k] >> sh))) +
                             (bf2f(uint16_t(wc[k] >> sh)) + bf2f(uint16_t(wd[k] >> sh)))) *
                            0.25f;
            lohi[h] = f2bf(v);
        }
        o[k] = lohi[0] | (lohi[1] << 16);
    }
    *reinterpret_cast<uint4 *>(out + ((size_t(img) * Ho + y) * Wo + x) * C + cv * 8) = make_uint4(o[0], o[1], o[2], o[3]);
}

__global__ void k_pool_bwd(float *__restrict__ din, const float *__restrict__ dout, int n, int H, int W, int C) {
    const size_t total = size_t(n) * H * W * C;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = int(i % C);
    size_t rest = i / C;
    const int x = int(rest % W);
    rest /= W;
    const int y = int(rest % H), img = int(rest / H);
    din[i] += dout[((size_t(img) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * C + c] * 0.25f;
}

__global__ void k_grad_slice(float *__restrict__ dst, const float *__restrict__ src, int srcC, int off, int C, int n,
                             int H, int W, int upsum) {
    const int Ho = upsum ? H >> 1 : H, Wo = upsum ? W >> 1 : W;
    const size_t total = size_t(n) * Ho * Wo * C;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = int(i % C);
    size_t rest = i / C;
    const int x = int(rest % Wo);
    rest /= Wo;
    const int y = int(rest % Ho), img = int(rest / Ho);
    if (!upsum) {
        dst[i] += src[((size_t(img) * H + y) * W + x) * srcC + off + c];
        return;
    }
    auto at = [&](int yy, int xx) { return src[((size_t(img) * H + yy) * W + xx) * srcC + off + c]; };
    dst[i] += (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1)) + (at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
}

__global__ void k_relu_bwd(uint16_t *__restrict__ dpre, const float *__restrict__ dout, const uint16_t *__restrict__ out,
                           size_t count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    dpre[i] = f2bf(bf2f(out[i]) > 0.f ? dout[i] : 0.f);
}

__global__ __launch_bounds__(256) void k_loss_l1(const float *__restrict__ pred, const float *__restrict__ target,
                                                 uint16_t *__restrict__ dpre, float *__restrict__ lossSum,
                                                 size_t pixels, float scale) {
    __shared__ float part[4];
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < pixels) {
        uint16_t gv[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) gv[c] = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float d = pred[p * 16 + c] - target[p * 4 + c];
            l += fabsf(d);
            gv[c] = f2bf(d > 0.f ? scale : (d < 0.f ? -scale : 0.f));
        }
        uint4 *o = reinterpret_cast<uint4 *>(dpre + p * 16);
        o[0] = make_uint4(gv[0] | (uint32_t(gv[1]) << 16), gv[2] | (uint32_t(gv[3]) << 16), 0u, 0u);
        o[1] = make_uint4(0u, 0u, 0u, 0u);
    }
    for (int s = 32; s > 0; s >>= 1) l += __shfl_xor(l, s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(lossSum, ((part[0] + part[1]) + (part[2] + part[3])) * scale);
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_weights_bf16(const float *__restrict__ w, uint16_t *__restrict__ wf, uint16_t *__restrict__ wt,
                               int cout, int cin) {
    const size_t total = size_t(9) * cout * cin;
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int tap = int(e / (size_t(cout) * cin));
    const int r = int(e - size_t(tap) * cout * cin), co = r / cin, ci = r - co * cin;
    const uint16_t h = f2bf(w[e]);
    wf[e] = h;
    wt[(size_t(8 - tap) * cin + ci) * cout + co] = h;
}

__global__ void k_fill(float *p, size_t n, float v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_to_bf16(const float *__restrict__ s, uint16_t *__restrict__ d, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = f2bf(s[i]);
}

__global__ void k_pack_features(const float4 *__restrict__ beauty, const float4 *__restrict__ albedo,
                                const float4 *__restrict__ normal, const float4 *__restrict__ depth, int imgW,
                                int imgH, float invSpp, uint16_t *__restrict__ xout, int H, int W) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const size_t i = size_t(min(y, imgH - 1)) * imgW + min(x, imgW - 1);
    const float4 b = beauty[i], a = albedo[i], nr = normal[i];
    const float d = depth[i].x;
    float f[16];
    f[0] = log1pf(fmaxf(b.x, 0.f) / (a.x + 0.01f));
    f[1] = log1pf(fmaxf(b.y, 0.f) / (a.y + 0.01f));
    f[2] = log1pf(fmaxf(b.z, 0.f) / (a.z + 0.01f));
    f[3] = a.x, f[4] = a.y, f[5] = a.z;
    f[6] = nr.x, f[7] = nr.y, f[8] = nr.z;
    f[9] = log1pf(fmaxf(d, 0.f));
    f[10] = invSpp;
    f[11] = b.w;
    f[12] = f[13] = f[14] = f[15] = 0.f;
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = uint32_t(f2bf(f[2 * k])) | (uint32_t(f2bf(f[2 * k + 1])) << 16);
    uint4 *o = reinterpret_cast<uint4 *>(xout + (size_t(y) * W + x) * 16);
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

__global__ void k_pack_target(const float4 *__restrict__ beauty, const float4 *__restrict__ albedo, int imgW, int imgH,
                              float *__restrict__ t, int H, int W) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const size_t i = size_t(min(y, imgH - 1)) * imgW + min(x, imgW - 1);
    const float4 b = beauty[i], a = albedo[i];
    float4 v;
    v.x = log1pf(fmaxf(b.x, 0.f) / (a.x + 0.01f));
    v.y = log1pf(fmaxf(b.y, 0.f) / (a.y + 0.01f));
    v.z = log1pf(fmaxf(b.z, 0.f) / (a.z + 0.01f));
    v.w = 0.f;
    reinterpret_cast<float4 *>(t)[size_t(y) * W + x] = v;
}

__global__ void k_unpack(const float *__restrict__ yv, const float4 *__restrict__ albedo,
                         const float4 *__restrict__ beauty, int imgW, int imgH, int W, float4 *__restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= imgW || y >= imgH) return;
    const size_t i = size_t(y) * imgW + x;
    const float *p = yv + (size_t(y) * W + x) * 16;
    const float4 a = albedo[i];
    // Clamped to e^30 - 1: far beyond any radiance the network is trained on.
    out[i] = make_float4(expm1f(fminf(fmaxf(p[0], 0.f), 30.f)) * (a.x + 0.01f),
                         expm1f(fminf(fmaxf(p[1], 0.f), 30.f)) * (a.y + 0.01f),
                         expm1f(fminf(fmaxf(p[2], 0.f), 30.f)) * (a.z + 0.01f), beauty[i].w);
}

inline unsigned blocks(size_t n, unsigned b = 256) { return unsigned((n + b - 1) / b); }

}  // namespace

hipError_t launchAtrous(const float4 *src, const float4 *depth, const float4 *normal, const float4 *albedo, float4 *dst,
                        const AtrousParams &p, hipStream_t stream) {
    dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
    hipLaunchKernelGGL(k_atrous, grid, dim3(256), 0, stream, src, depth, normal, albedo, dst, p);
    return hipGetLastError();
}

#ifndef DR_CONV_LDS
#define DR_CONV_LDS 1
#endif
// Multiprocessors of the current device (cached per device ordinal).
int deviceCUs() {
    static int cache[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cache[dev]) {
        int n = 0;
        cache[dev] = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0 ? n : 256;
    }
    return cache[dev];
}

template <int NB, int CHUNKS, bool UNPACK = false>
hipError_t convPipeLaunch(const ConvArgs &c, hipStream_t stream, unsigned tiles) {
    // One resident wave of workgroups (registers and LDS decide how many per CU).
    static int perCU = 0;
    if (!perCU &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_conv3x3_pipe<NB, CHUNKS, UNPACK>, 256, 0) !=
             hipSuccess || perCU < 1))
        perCU = 1;
    const unsigned grid = std::min<unsigned>(tiles, unsigned(deviceCUs()) * unsigned(perCU));
    hipLaunchKernelGGL((k_conv3x3_pipe<NB, CHUNKS, UNPACK>), dim3(grid), dim3(256), 0, stream, c);
    return hipGetLastError();
}

template <int NB, int MB>
hipError_t convLaunch(const ConvArgs &c, hipStream_t stream, unsigned groups = 1) {
    const unsigned grid = unsigned(c.n * ((c.W + 15) / 16) * ((c.H + 4 * MB - 1) / (4 * MB)));
    if (grid && DR_CONV_LDS) hipLaunchKernelGGL((k_conv3x3_lds<NB, MB>), dim3(grid, groups), dim3(256), 0, stream, c);
    else if (grid) hipLaunchKernelGGL((k_conv3x3<NB, MB>), dim3(grid), dim3(256), 0, stream, c);
    return hipGetLastError();
}

hipError_t launchConvOnly(const ConvArgs &c, hipStream_t stream);

// Whether the kernel launchConvOnly picks for c writes its output tile
// through LDS (k_conv3x3_pipe, or k_conv3x3_lds with a fitting BF16 tile),
// where poolTile can read it.
template <int NB, int MB>
bool poolsInLds() {
    return DR_CONV_LDS && outTileFits<NB, MB>(2);
}
bool convPoolsItself(const ConvArgs &c) {
    if (!DR_CONV_LDS || c.f32out) return false;
    const unsigned tiles16 = unsigned(c.n * ((c.W + 15) / 16) * ((c.H + 15) / 16));
    const unsigned groups = unsigned(c.cout / 16);
    const int cin = c.ca + c.cb;
    if (!c.plain && groups > 1 && tiles16 < 256u)
        return tiles16 * groups >= 256u ? poolsInLds<1, 4>() : poolsInLds<1, 1>();
    if (!c.plain && tiles16 >= 4096u && cin <= 64 && (groups == 1 || groups == 2)) return true;
    switch (c.cout / 16) {
    case 1: return poolsInLds<1, 4>();
    case 2: return poolsInLds<2, 4>();
    case 3: return poolsInLds<3, 4>();
    case 4: return poolsInLds<4, 4>();
    case 5: return poolsInLds<5, 2>();
    case 7: return poolsInLds<7, 2>();
    case 8: return poolsInLds<8, 2>();
    case 12: return poolsInLds<12, 1>();
    default: return false;
    }
}

hipError_t launchConv(const ConvArgs &c, hipStream_t stream) {
    if (c.poolOut && (c.f32out || c.H % 2 || c.W % 2)) return hipErrorInvalidValue;
    if (!c.poolOut || convPoolsItself(c)) return launchConvOnly(c, stream);
    ConvArgs plainOut = c;
    plainOut.poolOut = nullptr;
    hipError_t e = launchConvOnly(plainOut, stream);
    if (e != hipSuccess) return e;
    return launchPool(static_cast<const uint16_t *>(c.out), c.poolOut, c.n, c.H, c.W, c.cout, stream);
}

hipError_t launchConvOnly(const ConvArgs &c, hipStream_t stream) {
    if (c.ca % 16 || c.cb % 16 || c.cout % 16 || (c.cb && (c.H % 2 || c.W % 2))) return hipErrorInvalidValue;
    // Layers with fewer 16x16 tiles than CUs (the deep levels) give each
    // workgroup one 16-channel group of the output, with 4-row tiles when
    // even that leaves fewer workgroups than CUs.
    const unsigned tiles16 = unsigned(c.n * ((c.W + 15) / 16) * ((c.H + 15) / 16));
    const unsigned groups = unsigned(c.cout / 16);
    if (DR_CONV_LDS && !c.plain && groups > 1 && tiles16 < 256u)
        return tiles16 * groups >= 256u ? convLaunch<1, 4>(c, stream, groups) : convLaunch<1, 1>(c, stream, groups);
    // Full-resolution layers (many tiles, at most 32 output channels and 64
    // input channels): the persistent pipelined kernel, when its output tile
    // fits the tile's LDS (BF16 outputs, or FP32 with 16 channels).
    const int cin = c.ca + c.cb;
    if (DR_CONV_LDS && !c.plain && tiles16 >= 4096u && cin <= 64 && (!c.f32out || groups == 1)) {
        if (groups == 1) return cin > 32 ? convPipeLaunch<1, 2>(c, stream, tiles16) : convPipeLaunch<1, 1>(c, stream, tiles16);
        if (groups == 2) return cin > 32 ? convPipeLaunch<2, 2>(c, stream, tiles16) : convPipeLaunch<2, 1>(c, stream, tiles16);
    }
    // Row blocks per wave: 4 for narrow outputs, fewer as the accumulators grow.
    switch (c.cout / 16) {
    case 1: return convLaunch<1, 4>(c, stream);
    case 2: return convLaunch<2, 4>(c, stream);
    case 3: return convLaunch<3, 4>(c, stream);
    case 4: return convLaunch<4, 4>(c, stream);
    case 5: return convLaunch<5, 2>(c, stream);
    case 7: return convLaunch<7, 2>(c, stream);
    case 8: return convLaunch<8, 2>(c, stream);
    case 12: return convLaunch<12, 1>(c, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launchConvUnpack(const ConvArgs &c, hipStream_t stream) {
    if (!c.unpackOut) return launchConv(c, stream);
    if (!c.f32out || c.relu || c.n != 1 || c.cout != 16 || c.cb || c.imgW > c.W || c.imgH > c.H)
        return hipErrorInvalidValue;
    // The same choice launchConv makes for this layer: the pipelined kernel
    // takes full-resolution 16-channel outputs.
    const unsigned tiles16 = unsigned(((c.W + 15) / 16) * ((c.H + 15) / 16));
    if (DR_CONV_LDS && !c.plain && tiles16 >= 4096u && c.ca <= 64)
        return c.ca > 32 ? convPipeLaunch<1, 2, true>(c, stream, tiles16) : convPipeLaunch<1, 1, true>(c, stream, tiles16);
    hipError_t e = launchConv(c, stream);
    if (e != hipSuccess) return e;
    return launchUnpack(static_cast<const float *>(c.out), c.unpackAlbedo, c.unpackBeauty, c.imgW, c.imgH, c.W,
                        c.unpackOut, stream);
}

hipError_t launchWgrad(const WgradArgs &g, hipStream_t stream) {
    const int cin = g.ca + g.cb;
    if (g.ca % 16 || g.cb % 16 || g.cout % 16 || (g.cout / 16) * (cin / 16) > 4 * kWgMaxQ ||
        (g.cb && (g.H % 2 || g.W % 2)))
        return hipErrorInvalidValue;
    const int ntile = g.n * ((g.W + 7) / 8) * ((g.H + 7) / 8);
    if (ntile == 0) return hipSuccess;
    const unsigned gx = unsigned(ntile < 96 ? ntile : 96);
    const size_t lds = size_t(g.cout + cin) * kWgRow * sizeof(uint16_t);
    hipLaunchKernelGGL(k_wgrad, dim3(gx, 9), dim3(256), lds, stream, g);
    return hipGetLastError();
}

hipError_t launchPool(const uint16_t *in, uint16_t *out, int n, int H, int W, int C, hipStream_t stream) {
    const size_t total = size_t(n) * (H / 2) * (W / 2) * (C / 8);
    if (total) hipLaunchKernelGGL(k_pool, dim3(blocks(total)), dim3(256), 0, stream, in, out, n, H, W, C);
    return hipGetLastError();
}

hipError_t launchPoolBwd(float *din, const float *dout, int n, int H, int W, int C, hipStream_t stream) {
    const size_t total = size_t(n) * H * W * C;
    if (total) hipLaunchKernelGGL(k_pool_bwd, dim3(blocks(total)), dim3(256), 0, stream, din, dout, n, H, W, C);
    return hipGetLastError();
}

hipError_t launchGradSlice(float *dst, const float *src, int srcC, int off, int C, int n, int H, int W, bool upsum,
                           hipStream_t stream) {
    const size_t total = size_t(n) * (upsum ? H / 2 : H) * (upsum ? W / 2 : W) * C;
    if (total)
        hipLaunchKernelGGL(k_grad_slice, dim3(blocks(total)), dim3(256), 0, stream, dst, src, srcC, off, C, n, H, W,
                           upsum ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchReluBwd(uint16_t *dpre, const float *dout, const uint16_t *out, size_t count, hipStream_t stream) {
    if (count) hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(count)), dim3(256), 0, stream, dpre, dout, out, count);
    return hipGetLastError();
}

hipError_t launchLossL1(const float *pred, const float *target, uint16_t *dpre, float *lossSum, size_t pixels,
                        float scale, hipStream_t stream) {
    if (pixels)
        hipLaunchKernelGGL(k_loss_l1, dim3(blocks(pixels)), dim3(256), 0, stream, pred, target, dpre, lossSum, pixels,
                           scale);
    return hipGetLastError();
}

hipError_t launchAdam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                      float c1, float c2, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launchWeightsBf16(const float *w, uint16_t *wf, uint16_t *wt, int cout, int cin, hipStream_t stream) {
    const size_t total = size_t(9) * cout * cin;
    hipLaunchKernelGGL(k_weights_bf16, dim3(blocks(total)), dim3(256), 0, stream, w, wf, wt, cout, cin);
    return hipGetLastError();
}

hipError_t launchFill(float *p, size_t n, float v, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(k_fill, dim3(blocks(n)), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launchToBf16(const float *src, uint16_t *dst, size_t n, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

hipError_t launchPackFeatures(const float4 *beauty, const float4 *albedo, const float4 *normal, const float4 *depth,
                              int imgW, int imgH, float invSpp, uint16_t *x, int H, int W, hipStream_t stream) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(k_pack_features, grid, dim3(256), 0, stream, beauty, albedo, normal, depth, imgW, imgH, invSpp,
                       x, H, W);
    return hipGetLastError();
}

hipError_t launchPackTarget(const float4 *beauty, const float4 *albedo, int imgW, int imgH, float *t, int H, int W,
                            hipStream_t stream) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(k_pack_target, grid, dim3(256), 0, stream, beauty, albedo, imgW, imgH, t, H, W);
    return hipGetLastError();
}

hipError_t launchUnpack(const float *y, const float4 *albedo, const float4 *beauty, int imgW, int imgH, int W,
                        float4 *out, hipStream_t stream) {
    dim3 grid((imgW + 15) / 16, (imgH + 15) / 16);
    hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, stream, y, albedo, beauty, imgW, imgH, W, out);
    return hipGetLastError();
}

}  // namespace dr
