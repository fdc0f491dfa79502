// dr_api.cpp -- C ABI of the HIP backend (include/denrim_hip.h).
//
// Owns the device, the device-resident scene (caller's records plus the
// traversal BVH) and sessions (path-state buffers, output planes, sample
// count). Session semantics follow RenderSession (Renderer/RenderSession.swift):
// samples accumulate as a running mean, resetAccumulation only zeroes the
// sample count (the next sample has weight 1), tiles advance the count only
// when they complete a sample.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/denrim_hip.h"
#include "dr_bvh.h"
#include "dr_kernels.h"
#include "dr_scene.h"
#include "dr_denoise.h"
#include "dr_internal.h"

namespace {

thread_local std::string g_error;

dr_status fail(dr_status s, const std::string &msg) {
    g_error = msg;
    return s;
}

}  // namespace

dr_status drFail(dr_status s, const std::string &msg) { return fail(s, msg); }

std::recursive_mutex &drDeviceMutex() {
    static std::recursive_mutex m;
    return m;
}

namespace {

#define DR_HIP(call, what)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) {                                                                              \
            dr_status s_ = e_ == hipErrorOutOfMemory ? DR_ERR_OUT_OF_MEMORY : DR_ERR_COMMAND_BUFFER_FAILED; \
            return fail(s_, std::string(what) + ": " + hipGetErrorString(e_));                               \
        }                                                                                                    \
    } while (0)

template <class T> hipError_t devAlloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <class T> hipError_t upload(T **p, const void *src, size_t count, hipStream_t stream) {
    hipError_t e = devAlloc(p, count);
    if (e != hipSuccess || count == 0) return e;
    return drCopy(*p, src, count * sizeof(T), hipMemcpyHostToDevice, stream);
}

void devFree(void *p) {
    if (p) (void)hipFree(p);
}

DrCamera toCam(const dr_gpu_camera &c) {
    DrCamera d;
    d.origin = make_float4(c.origin[0], c.origin[1], c.origin[2], c.origin[3]);
    d.lowerLeft = make_float4(c.lower_left[0], c.lower_left[1], c.lower_left[2], c.lower_left[3]);
    d.horizontal = make_float4(c.horizontal[0], c.horizontal[1], c.horizontal[2], c.horizontal[3]);
    d.vertical = make_float4(c.vertical[0], c.vertical[1], c.vertical[2], c.vertical[3]);
    d.lens = make_float4(c.lens[0], c.lens[1], c.lens[2], c.lens[3]);
    return d;
}

}  // namespace

// Path-state block of a session: 8 float4 arrays of N slots, 4 float4
// arrays of 3N (shadow records), u32 queues (2 x N ping-pong, 3N shadow,
// N walk, N resume), counters, claim cursors, stats, the sample base word and the
// per-lane traversal stack spill.
struct PathLayout {
    size_t rayO, rayD, thr, rad, meta, hit, hitN, neeT, shO, shD, shC, shR, q0, q1, sq, sss, res, ctr, claims, stats, base,
        ovf;
    size_t bytes;
};
static PathLayout pathLayout(size_t N, size_t stride, uint32_t stackMax) {
    PathLayout l{};
    size_t bytes = 0;
    auto take = [&](size_t n) {
        size_t off = bytes;
        bytes += (n + 255) & ~size_t(255);
        return off;
    };
    l.rayO = take(N * 16), l.rayD = take(N * 16), l.thr = take(N * 16), l.rad = take(N * 16), l.meta = take(N * 16);
    l.hit = take(N * 16), l.hitN = take(N * 16), l.neeT = take(N * 16);
    l.shO = take(3 * N * 16), l.shD = take(3 * N * 16), l.shC = take(3 * N * 16), l.shR = take(3 * N * 16);
    l.q0 = take(N * 4), l.q1 = take(N * 4), l.sq = take(3 * N * 4), l.sss = take(N * 4), l.res = take(N * 4);
    l.ctr = take(DR_CTR_WORDS * 4), l.claims = take(DR_CLAIM_WORDS * 4), l.stats = take(DR_CNT_SLOTS * 8);
    l.base = take(4);
    l.ovf = take(stride * (stackMax > DR_LDS_STACK ? stackMax - DR_LDS_STACK : 0u) * 4);
    l.bytes = bytes;
    return l;
}

// Samples of one pixel in flight together when the caller leaves it open:
// about 64 M path slots per batch (348 B each, 23 GB of path state), at
// most 64 samples. The paths still alive after a few bounces are too few to
// fill the GPU on their own, so a batch of one sample spends most of its
// time in thin late-bounce launches whose length is set by their slowest
// chain; with k samples in flight each launch carries k times the paths
// (1080p layered terrain, 64 spp: 562 Mrays/s at k = 1, 1118 at 8, 1217 at
// 16 with the tracers of the time; later 1598 at 16, 1689 at 32, 1727 at
// 64).
uint32_t drAutoSamplesPerLaunch(uint64_t pixels) {
    return uint32_t(std::min<uint64_t>(64, std::max<uint64_t>(1, (64u << 20) / std::max<uint64_t>(pixels, 1))));
}

static void dropGraphs(dr_session *s) {
    DR_DEVICE_LOCK;
    for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.second);
    s->graphs.clear();
}

int drDeviceOrdinal(const dr_device *device) { return device->ordinal; }

extern "C" {

const char *dr_last_error(void) { return g_error.c_str(); }

dr_status dr_hip_device_create(int ordinal, dr_device **out) {
    if (!out) return fail(DR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= ordinal || ordinal < 0)
        return fail(DR_ERR_NO_DEVICE, "no HIP device with ordinal " + std::to_string(ordinal));
    DR_HIP(hipSetDevice(ordinal), "hipSetDevice");
    hipDeviceProp_t prop;
    DR_HIP(hipGetDeviceProperties(&prop, ordinal), "hipGetDeviceProperties");
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0)
        return fail(DR_ERR_MISSING_SHADER_FUNCTION, "kernels are built for gfx950, device is " + arch);
    uint32_t *scratch = nullptr;
    DR_HIP(dr::intersectScratchAddress(&scratch), "traversal scratch address");
    auto *d = new dr_device;
    d->ordinal = ordinal;
    d->cuCount = prop.multiProcessorCount;
    d->arch = arch;
    d->intersectScratch = scratch;
    *out = d;
    return DR_OK;
}

dr_status dr_hip_device_destroy(dr_device *device) {
    delete device;
    return DR_OK;
}

static void freeScene(dr_scene *s) {
    DR_DEVICE_LOCK;
    devFree(s->nodes);
    devFree(s->triPos);
    devFree(s->triangles);
    devFree(s->materials);
    devFree(s->matKey);
    devFree(s->lights);
    devFree(s->env);
    devFree(s->envGuide);
    devFree(s->envAlias);
    devFree(s->texDesc);
    devFree(s->texels);
    devFree(s->volumes);
    devFree(s->denseSamples);
    devFree(s->bricks);
    devFree(s->dist16);
    devFree(s->dist32);
    devFree(s->corner16);
    devFree(s->sampleMaterial);
    devFree(s->grids);
    devFree(s->gridIndices);
}

namespace {

// Reference-layout sparse-volume records (PathTrace.metal:92-148).
struct RefBrick {
    float worldBoundsMin[4], worldBoundsMax[4], localBoundsMin[4], localBoundsMax[4];
    float sampleBoundsMin[4], sampleBoundsMax[4];
    uint32_t gridOriginAndVolume[4], dimensionsAndSampleOffset[4];
};
static_assert(sizeof(RefBrick) == 128, "GPUVolumeBrick layout");
struct RefBrickSample {
    float distance;
    uint32_t materialA, materialB;
    float materialBlend;
};
static_assert(sizeof(RefBrickSample) == 16, "GPUVolumeBrickSample layout");

bool halfExact(float v, uint16_t &bits) {
    _Float16 h = (_Float16)v;
    if ((float)h != v && !(v != v)) return false;
    if (v != v) return false;  // NaN payloads are not carried
    bits = __builtin_bit_cast(uint16_t, h);
    return true;
}

}  // namespace

dr_status dr_hip_scene_upload(dr_device *device, const dr_scene_desc *d, dr_scene **out) {
    if (!device || !d || !out) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    // RenderSession.init checks (Renderer/RenderSession.swift:322-327).
    if ((d->triangle_count == 0 || !d->triangles) && d->volume_count == 0)
        return fail(DR_ERR_INVALID_SCENE, "Scene contains no renderable geometry.");
    if (d->material_count == 0 || !d->materials) return fail(DR_ERR_INVALID_SCENE, "Scene contains no materials.");
    if (d->triangle_count >= DR_MAX_TRIANGLES)
        return fail(DR_ERR_INVALID_SCENE, "Scene has more triangles than the traversal can index (2^27).");
    if (d->volume_count &&
        (!d->volumes || (d->volume_sample_count && !d->volume_samples) || (d->volume_brick_count && !d->volume_bricks) ||
         (d->volume_brick_sample_count && !d->volume_brick_samples) ||
         (d->volume_brick_grid_count && !d->volume_brick_grids) ||
         (d->volume_brick_grid_index_count && !d->volume_brick_grid_indices)))
        return fail(DR_ERR_INVALID_ARGUMENT, "volume arrays are NULL");
    if (d->volume_count && d->material_count > 65536u)
        return fail(DR_ERR_INVALID_SCENE, "scenes with sparse volumes are limited to 65536 materials");
    if (d->volume_count >= (1u << 20)) return fail(DR_ERR_INVALID_SCENE, "too many volumes");
    if (d->light_count && !d->lights) return fail(DR_ERR_INVALID_ARGUMENT, "lights is NULL");
    if (d->environment_sample_count && !d->environment_samples)
        return fail(DR_ERR_INVALID_ARGUMENT, "environment_samples is NULL");
    if (d->texture_count && (!d->texture_descriptors || !d->texture_pixels))
        return fail(DR_ERR_INVALID_ARGUMENT, "texture arrays are NULL");
    const auto *tris = static_cast<const DrTriangle *>(d->triangles);
    const auto *texd = static_cast<const uint4 *>(d->texture_descriptors);
    for (uint32_t i = 0; i < d->texture_count; ++i) {
        uint64_t end = uint64_t(texd[i].x) + uint64_t(texd[i].y) * texd[i].z;
        if (end > d->texture_pixel_count) return fail(DR_ERR_INVALID_SCENE, "texture descriptor exceeds texel array");
    }
    const auto *lights = static_cast<const DrLight *>(d->lights);
    for (uint32_t i = 0; i < d->light_count; ++i)
        if (lights[i].triangleIndex >= d->triangle_count)
            return fail(DR_ERR_INVALID_SCENE, "light record references an unknown triangle");
    DR_HIP(hipSetDevice(device->ordinal), "hipSetDevice");

    auto *s = new dr_scene;
    s->device = device;
    std::vector<DrNode4> nodes;
    std::vector<float4> triPos;
    if (d->triangle_count)
        dr::buildBvh(tris, d->triangle_count, d->max_leaf_triangles ? d->max_leaf_triangles : 4u, nodes, triPos,
                     &s->stats);
    if (nodes.size() >= (1u << 25)) {  // node byte offsets are 32-bit in traversal
        delete s;
        return fail(DR_ERR_INVALID_SCENE, "Scene's BVH has more nodes than the traversal can address (2^25).");
    }
    s->stackMax = 7u * s->stats.maxDepth + 2u;

    // Sparse volumes: compact brick headers, FP16 distances when every
    // distance is FP16-exact (else FP32), per-sample material pairs with the
    // material indices clamped as materialForHit does (:1096-1137).
    std::vector<DrBrick> bricks(d->volume_brick_count);
    std::vector<uint16_t> dist16;
    std::vector<float> dist32;
    std::vector<uint2> sampleMat;
    uint32_t sampleCount = 0;
    bool distHalf = true;
    if (d->volume_count) {
        const auto *rb = static_cast<const RefBrick *>(d->volume_bricks);
        for (uint32_t i = 0; i < d->volume_brick_count; ++i) {
            const RefBrick &b = rb[i];
            DrBrick &o = bricks[i];
            o.smin = make_float4(b.sampleBoundsMin[0], b.sampleBoundsMin[1], b.sampleBoundsMin[2], b.sampleBoundsMin[3]);
            o.smax = make_float4(b.sampleBoundsMax[0], b.sampleBoundsMax[1], b.sampleBoundsMax[2], b.sampleBoundsMax[3]);
            uint32_t dx = b.dimensionsAndSampleOffset[0], dy = b.dimensionsAndSampleOffset[1],
                     dz = b.dimensionsAndSampleOffset[2];
            // Dimensions above 255 cannot be packed; they also never pass the
            // reference's range check against a 32-bit sample count at
            // useful sizes, so they are stored as invalid (0).
            uint32_t packed = (dx < 256u && dy < 256u && dz < 256u) ? (dx | dy << 8 | dz << 16) : 0u;
            float pf, of;
            std::memcpy(&pf, &packed, 4);
            std::memcpy(&of, &b.dimensionsAndSampleOffset[3], 4);
            o.cmin = make_float4(b.localBoundsMin[0], b.localBoundsMin[1], b.localBoundsMin[2], pf);
            o.cmax = make_float4(b.localBoundsMax[0], b.localBoundsMax[1], b.localBoundsMax[2], of);
        }
        // The FP16 corner records are indexed by sample, so the valid bricks'
        // sample ranges must not overlap; if the caller's do, each valid brick
        // gets a private copy of its samples (distances and materials alike,
        // so every lookup reads the same values) and invalid bricks keep
        // failing the range check through zero dimensions.
        const auto *rs = static_cast<const RefBrickSample *>(d->volume_brick_samples);
        std::vector<uint32_t> src;  // new sample index -> caller's; empty: identity
        {
            const uint32_t n0 = d->volume_brick_sample_count;
            std::vector<std::pair<uint32_t, uint32_t>> ranges;  // (offset, brick)
            auto cellsOf = [](const DrBrick &o) {
                uint32_t packed;
                std::memcpy(&packed, &o.cmin.w, 4);
                return uint64_t(packed & 0xffu) * ((packed >> 8) & 0xffu) * ((packed >> 16) & 0xffu);
            };
            auto validOf = [&](const DrBrick &o) {
                uint32_t packed, offset;
                std::memcpy(&packed, &o.cmin.w, 4);
                std::memcpy(&offset, &o.cmax.w, 4);
                return (packed & 0xffu) >= 2u && ((packed >> 8) & 0xffu) >= 2u && ((packed >> 16) & 0xffu) >= 2u &&
                       offset < n0 && cellsOf(o) <= uint64_t(n0 - offset);
            };
            for (uint32_t i = 0; i < d->volume_brick_count; ++i)
                if (validOf(bricks[i])) {
                    uint32_t offset;
                    std::memcpy(&offset, &bricks[i].cmax.w, 4);
                    ranges.emplace_back(offset, i);
                }
            std::sort(ranges.begin(), ranges.end());
            bool overlap = false;
            for (size_t k = 1; k < ranges.size() && !overlap; ++k)
                overlap = uint64_t(ranges[k].first) < uint64_t(ranges[k - 1].first) + cellsOf(bricks[ranges[k - 1].second]);
            if (overlap) {
                std::vector<bool> valid(d->volume_brick_count);
                for (auto &r : ranges) valid[r.second] = true;
                for (uint32_t i = 0; i < d->volume_brick_count; ++i) {
                    DrBrick &o = bricks[i];
                    if (!valid[i]) {
                        const uint32_t zero = 0u;
                        std::memcpy(&o.cmin.w, &zero, 4);
                        continue;
                    }
                    uint32_t offset;
                    std::memcpy(&offset, &o.cmax.w, 4);
                    const uint64_t cells = cellsOf(o);
                    if (src.size() + cells > 0xffffffffull) {
                        freeScene(s);
                        delete s;
                        return fail(DR_ERR_INVALID_SCENE, "brick samples exceed 2^32 once overlapping bricks are split");
                    }
                    const uint32_t base = uint32_t(src.size());
                    for (uint64_t k = 0; k < cells; ++k) src.push_back(offset + uint32_t(k));
                    std::memcpy(&o.cmax.w, &base, 4);
                }
            }
        }
        const uint32_t n = src.empty() ? d->volume_brick_sample_count : uint32_t(src.size());
        auto at = [&](uint32_t i) -> const RefBrickSample & { return rs[src.empty() ? i : src[i]]; };
        sampleCount = n;
        dist16.resize(n);
        for (uint32_t i = 0; i < n && distHalf; ++i) distHalf = halfExact(at(i).distance, dist16[i]);
        if (!distHalf) {
            dist16.clear();
            dist32.resize(n);
            for (uint32_t i = 0; i < n; ++i) dist32[i] = at(i).distance;
        }
        sampleMat.resize(n);
        const uint32_t last = d->material_count - 1u;
        for (uint32_t i = 0; i < n; ++i) {
            const RefBrickSample &r = at(i);
            uint32_t a = std::min(r.materialA, last), b = std::min(r.materialB, last);
            uint32_t blend;
            std::memcpy(&blend, &r.materialBlend, 4);
            sampleMat[i] = make_uint2(a | b << 16, blend);
        }
    }

    bool opaqueOnly = true;
    const auto *mats = static_cast<const DrMaterial *>(d->materials);
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const DrMaterial &m = mats[i];
        // A shadow ray passes a surface with opacity <= 0.001 or
        // transmission > 0.001 (PathTrace.metal:4249-4258); base-colour
        // textures can lower opacity too.
        if (m.s[M_BASE].w <= 0.001f || std::min(std::max(m.s[M_EMISSION].w, 0.f), 1.f) > 0.001f || m.s[M_PARAMS].z != 0.f)
            opaqueOnly = false;
    }

    // Shading class of each material: the features that decide k_shade's
    // branches (pass-through, emission, transmission, volume scattering,
    // subsurface, clearcoat, sheen, thin film, textures), each distinct set
    // numbered in order of first use, modulo 14. k_shade orders a block's
    // paths by the class of what they hit.
    std::vector<uint32_t> matKey(d->material_count);
    {
        std::vector<uint32_t> seen;
        for (uint32_t i = 0; i < d->material_count; ++i) {
            const DrMaterial &m = mats[i];
            auto cl = [](float x) { return std::min(std::max(x, 0.f), 1.f); };
            const float transmission = cl(m.s[M_EMISSION].w);
            const uint32_t bits =
                (m.s[M_BASE].w <= 0.001f ? 1u : 0u) |
                (std::max(std::max(m.s[M_EMISSION].x, m.s[M_EMISSION].y), m.s[M_EMISSION].z) > 0.f ? 2u : 0u) |
                (transmission > 0.001f ? 4u : 0u) | (cl(m.s[M_VOL_SCATTER].w) > 0.f ? 8u : 0u) |
                (cl(m.s[M_SSS_COLOR].w) > 0.f ? 16u : 0u) | (cl(m.s[M_PARAMS2].z) > 0.f ? 32u : 0u) |
                (cl(m.s[M_SHEEN].w) > 0.f ? 64u : 0u) |
                (m.s[M_THINFILM].x > 0.f && m.s[M_THINFILM].y > 0.f ? 128u : 0u) |
                (m.s[M_PARAMS].z != 0.f || m.s[M_PARAMS].w != 0.f ? 256u : 0u);
            size_t k = 0;
            while (k < seen.size() && seen[k] != bits) ++k;
            if (k == seen.size()) seen.push_back(bits);
            matKey[i] = uint32_t(k % 14u);
        }
    }

    // Guide table of the HDRI CDF (G a power of two, at least one slot per
    // texel up to 2^22 slots): the bracket of a lookup is then a texel or
    // two, so an NEE sample's inversion costs two dependent loads instead of
    // a binary search (k_shade 1-2% faster than with 8 texels per slot).
    std::vector<uint32_t> guide;
    uint32_t guideSize = 0;
    if (d->environment_sample_count > 1) {
        const uint32_t n = d->environment_sample_count;
        guideSize = 1;
        while (guideSize < n && guideSize < (1u << 22)) guideSize <<= 1;
        const auto *env = static_cast<const float2 *>(d->environment_samples);
        guide.resize(size_t(guideSize) + 1);
        uint32_t i = 0;
        for (uint32_t k = 0; k <= guideSize; ++k) {
            float threshold = float(k) / float(guideSize);
            while (i < n - 1 && !(threshold <= env[i].x)) ++i;
            guide[k] = i;
        }
        guide[guideSize] = n - 1;
    }

    // Alias table of the same texel distribution (Walker's method, Vose's
    // construction in double): worklists of under- and over-full texels in
    // index order, processed from the back; leftovers keep probability 1.
    std::vector<uint2> alias;
    if (d->environment_sample_count > 1) {
        const uint32_t n = d->environment_sample_count;
        const auto *env = static_cast<const float2 *>(d->environment_samples);
        alias.resize(n);
        std::vector<double> scaled(n);
        std::vector<float> prob(n, 1.0f);
        std::vector<uint32_t> small, large;
        double prev = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const double c = double(env[i].x);
            scaled[i] = std::max(c - prev, 0.0) * double(n);
            prev = std::max(prev, c);
            alias[i].y = i;
            (scaled[i] < 1.0 ? small : large).push_back(i);
        }
        while (!small.empty() && !large.empty()) {
            const uint32_t l = small.back(), g = large.back();
            small.pop_back();
            large.pop_back();
            prob[l] = float(scaled[l]);
            alias[l].y = g;
            scaled[g] = (scaled[g] + scaled[l]) - 1.0;
            (scaled[g] < 1.0 ? small : large).push_back(g);
        }
        for (uint32_t i = 0; i < n; ++i) std::memcpy(&alias[i].x, &prob[i], 4);
    }

    DR_DEVICE_LOCK;
    hipStream_t up = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking);
    auto chain = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    chain(upload(&s->nodes, nodes.data(), nodes.size(), up));
    chain(upload(&s->envAlias, alias.data(), alias.size(), up));
    chain(upload(&s->triPos, triPos.data(), triPos.size(), up));
    chain(upload(&s->triangles, d->triangles, d->triangle_count, up));
    chain(upload(&s->materials, d->materials, d->material_count, up));
    chain(upload(&s->matKey, matKey.data(), matKey.size(), up));
    chain(upload(&s->lights, d->lights, d->light_count, up));
    chain(upload(&s->env, d->environment_samples, d->environment_sample_count, up));
    chain(upload(&s->envGuide, guide.data(), guide.size(), up));
    chain(upload(&s->texDesc, d->texture_descriptors, d->texture_count, up));
    chain(upload(&s->texels, d->texture_pixels, d->texture_pixel_count, up));
    chain(upload(&s->volumes, d->volumes, d->volume_count, up));
    chain(upload(&s->denseSamples, d->volume_samples, d->volume_count ? d->volume_sample_count : 0u, up));
    chain(upload(&s->bricks, bricks.data(), bricks.size(), up));
    chain(upload(&s->dist16, dist16.data(), dist16.size(), up));
    chain(upload(&s->dist32, dist32.data(), dist32.size(), up));
    chain(devAlloc(&s->corner16, distHalf ? size_t(sampleCount) : 0u));
    if (e == hipSuccess && distHalf && sampleCount)
        chain(dr::launchBrickCorners(s->bricks, d->volume_brick_count, s->dist16, sampleCount, s->corner16, up));
    if (e == hipSuccess) chain(hipStreamSynchronize(up));
    chain(upload(&s->sampleMaterial, sampleMat.data(), sampleMat.size(), up));
    chain(upload(&s->grids, d->volume_brick_grids, d->volume_brick_grid_count, up));
    chain(upload(&s->gridIndices, d->volume_brick_grid_indices, d->volume_brick_grid_index_count, up));
    if (up) (void)hipStreamDestroy(up);
    if (e != hipSuccess) {
        freeScene(s);
        delete s;
        return fail(e == hipErrorOutOfMemory ? DR_ERR_OUT_OF_MEMORY : DR_ERR_COMMAND_BUFFER_FAILED,
                    std::string("scene upload: ") + hipGetErrorString(e));
    }
    DrSceneDev &g = s->dev;
    g.nodes = s->nodes;
    g.triPos = s->triPos;
    g.triangles = s->triangles;
    g.materials = s->materials;
    g.matKey = s->matKey;
    g.lights = s->lights;
    g.env = s->env;
    g.envGuide = s->envGuide;
    g.envAlias = s->envAlias;
    g.envGuideSize = guideSize;
    g.texDesc = s->texDesc;
    g.texels = s->texels;
    g.triangleCount = d->triangle_count;
    g.materialCount = d->material_count;
    g.lightCount = d->light_count;
    g.envCount = d->environment_sample_count;
    g.texCount = d->texture_count;
    g.envTexIndexPlusOne = d->environment_texture_index_plus_one;
    g.envIntensity = d->environment_intensity;
    g.envRotationY = d->environment_rotation_y;
    g.envMaxRadiance = d->environment_max_radiance;
    g.opaqueOnly = opaqueOnly ? 1u : 0u;
    // Volume hits blend two materials (mix of opacity and transmission), so
    // the occlusion test of volume scenes asks for a margin on both sides of
    // the 0.001 thresholds, which a blend's rounding cannot cross.
    bool opaqueMargin = opaqueOnly;
    for (uint32_t i = 0; i < d->material_count && opaqueMargin; ++i) {
        const DrMaterial &m = mats[i];
        opaqueMargin = m.s[M_BASE].w >= 0.002f && std::min(std::max(m.s[M_EMISSION].w, 0.f), 1.f) <= 0.0005f;
    }
    bool denseFields = false;
    if (d->volume_count && d->volume_sample_count) {
        const auto *vs = static_cast<const DrVolumeSample *>(d->volume_samples);
        for (uint32_t i = 0; i < d->volume_sample_count && !denseFields; ++i)
            denseFields = vs[i].materialFieldFlags.x != 0u;
    }
    DrSdfDev &v = g.sdf;
    v.volumes = s->volumes;
    v.dense = s->denseSamples;
    v.denseCount = d->volume_count ? d->volume_sample_count : 0u;
    v.bricks = s->bricks;
    v.dist16 = s->dist16;
    v.dist32 = s->dist32;
    v.corner16 = s->corner16;
    v.material = s->sampleMaterial;
    v.grids = s->grids;
    v.gridIndices = s->gridIndices;
    v.volumeCount = d->volume_count;
    v.brickCount = d->volume_brick_count;
    v.sampleCount = sampleCount;
    v.gridCount = d->volume_brick_grid_count;
    v.gridIndexCount = d->volume_brick_grid_index_count;
    v.distHalf = distHalf ? 1u : 0u;
    v.opaqueShadows = opaqueMargin && !denseFields ? 1u : 0u;
    *out = s;
    return DR_OK;
}

}  // extern "C"

namespace {
// True when p is device (or managed) memory of `ordinal`.
bool onDevice(const void *p, int ordinal) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unknown host pointer: clear the sticky error
        return false;
    }
    return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == ordinal;
}
}  // namespace

extern "C" dr_status dr_hip_scene_intersect_device(dr_scene *scene, const float *rays, uint32_t count,
                                                   uint32_t intersection, float *t, uint32_t *ids, void *stream) {
    if (!scene || (count && (!rays || !t || !ids))) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (intersection > DR_INTERSECTION_WATERTIGHT) return fail(DR_ERR_INVALID_ARGUMENT, "unknown intersection test");
    if (count == 0) return DR_OK;
    const dr_device *dev = scene->device;
    DR_HIP(hipSetDevice(dev->ordinal), "hipSetDevice");
    if (!onDevice(rays, dev->ordinal) || !onDevice(t, dev->ordinal) || !onDevice(ids, dev->ordinal))
        return fail(DR_ERR_INVALID_ARGUMENT, "rays, t and ids must be device memory of the scene's device");
    // Each launch holds as many lanes as the scratch spills for at the
    // scene's stack depth; a depth whose spill for one block of lanes
    // exceeds the scratch is refused.
    const uint32_t spill = scene->stackMax > DR_LDS_STACK ? scene->stackMax - DR_LDS_STACK : 0u;
    uint32_t lanes = 0;
    if (spill) {
        lanes = DR_INTERSECT_SCRATCH_WORDS / spill / DR_TRACE_BLOCK * DR_TRACE_BLOCK;
        if (lanes == 0)
            return fail(DR_ERR_INVALID_ARGUMENT,
                        "scene BVH too deep for the traversal scratch: " + std::to_string(size_t(spill) * 4 * DR_TRACE_BLOCK) +
                            " B per block of rays, " + std::to_string(size_t(DR_INTERSECT_SCRATCH_WORDS) * 4) + " B available");
    }
    DrSceneDev S = scene->dev;
    S.watertight = intersection == DR_INTERSECTION_WATERTIGHT ? 1u : 0u;
    DR_HIP(dr::launchIntersect(S, reinterpret_cast<const float4 *>(rays), count, t, ids, dev->intersectScratch,
                               static_cast<hipStream_t>(stream), lanes),
           "intersect launch");
    return DR_OK;
}

// Scene lifetime: sessions (and requests queued on a render server) hold a
// reference; a released scene is freed when the last holder lets go. The
// server's scheduler thread takes and drops references, hence the lock.
static std::mutex gSceneRefs;

void drSceneRetain(dr_scene *scene) {
    std::lock_guard<std::mutex> lk(gSceneRefs);
    scene->sessions++;
}

void drSceneDrop(dr_scene *scene) {
    bool last;
    {
        std::lock_guard<std::mutex> lk(gSceneRefs);
        last = --scene->sessions == 0 && scene->released;
    }
    if (last) {
        freeScene(scene);
        delete scene;
    }
}

extern "C" {

dr_status dr_hip_scene_release(dr_scene *scene) {
    if (!scene) return DR_OK;
    {
        std::lock_guard<std::mutex> lk(gSceneRefs);
        if (scene->sessions > 0) {
            scene->released = true;  // freed with its last holder
            return DR_OK;
        }
    }
    freeScene(scene);
    delete scene;
    return DR_OK;
}

dr_status dr_hip_scene_intersect(dr_scene *scene, const float *rays, uint32_t count, uint32_t intersection, float *t,
                                 uint32_t *ids) {
    if (!scene || (count && (!rays || !t || !ids))) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (intersection > DR_INTERSECTION_WATERTIGHT) return fail(DR_ERR_INVALID_ARGUMENT, "unknown intersection test");
    if (count == 0) return DR_OK;
    DR_HIP(hipSetDevice(scene->device->ordinal), "hipSetDevice");
    DrSceneDev S = scene->dev;
    S.watertight = intersection == DR_INTERSECTION_WATERTIGHT ? 1u : 0u;
    const size_t lanes = (size_t(count) + DR_TRACE_BLOCK - 1) / DR_TRACE_BLOCK * DR_TRACE_BLOCK;
    const size_t bytes = size_t(count) * 32 + size_t(count) * 8 + lanes * (scene->stackMax > DR_LDS_STACK ? scene->stackMax - DR_LDS_STACK : 0u) * 4;
    DR_DEVICE_LOCK;
    char *buf = nullptr;
    DR_HIP(hipMalloc(&buf, bytes), "intersect buffers");
    float4 *dRays = reinterpret_cast<float4 *>(buf);
    float *dT = reinterpret_cast<float *>(buf + size_t(count) * 32);
    uint32_t *dIds = reinterpret_cast<uint32_t *>(buf + size_t(count) * 36);
    uint32_t *dOvf = reinterpret_cast<uint32_t *>(buf + size_t(count) * 40);
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = drCopy(dRays, rays, size_t(count) * 32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = dr::launchIntersect(S, dRays, count, dT, dIds, dOvf, st);
    if (e == hipSuccess) e = drCopy(t, dT, size_t(count) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = drCopy(ids, dIds, size_t(count) * 4, hipMemcpyDeviceToHost, st);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    DR_HIP(e, "intersect");
    return DR_OK;
}

dr_status dr_hip_scene_info(const dr_scene *s, uint32_t *nodes, uint32_t *leaves, uint32_t *depth, double *secs) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (nodes) *nodes = s->stats.nodeCount;
    if (leaves) *leaves = s->stats.leafCount;
    if (depth) *depth = s->stats.maxDepth;
    if (secs) *secs = s->stats.buildSeconds;
    return DR_OK;
}

static void freeSession(dr_session *s) {
    DR_DEVICE_LOCK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    dropGraphs(s);
    devFree(s->pathBlock);
    devFree(s->outBlock);
    devFree(s->unetWork);
    if (s->evBegin) (void)hipEventDestroy(s->evBegin);
    if (s->evEnd) (void)hipEventDestroy(s->evEnd);
    for (auto ev : s->stageEvents) (void)hipEventDestroy(ev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// Shape of a session for settings: samples per batch, path slots, lanes of
// the persistent kernels (at most 16 x 128 per CU, no more than the slots
// need), stack-spill stride, and the bytes of both blocks.
// Workgroups per CU of the persistent tracer launches.
#ifndef DR_TRACE_BLOCKS_PER_CU
#define DR_TRACE_BLOCKS_PER_CU 12u
#endif
struct SessionShape {
    uint32_t k, nSlots, traceGrid, stride;
    PathLayout lay;
    size_t outBytes;
    int planes;
};
static dr_status sessionShape(const dr_scene *scene, const dr_settings *st, SessionShape &sh) {
    if (st->width == 0 || st->height == 0)  // RenderSession.swift:304-306
        return fail(DR_ERR_INVALID_SCENE, "Render dimensions must be positive.");
    if (st->quality > 2) return fail(DR_ERR_INVALID_ARGUMENT, "quality must be 0 (preview), 1 (interactive) or 2 (final)");
    if (st->max_bounces > DR_MAX_BOUNCES) return fail(DR_ERR_INVALID_ARGUMENT, "max_bounces exceeds 64");
    const uint64_t pixels = uint64_t(st->width) * st->height;
    sh.k = st->samples_per_launch ? st->samples_per_launch : drAutoSamplesPerLaunch(pixels);
    if (pixels * sh.k > 0xFFFFFFFFull / 3) return fail(DR_ERR_INVALID_ARGUMENT, "image too large");
    sh.nSlots = uint32_t(pixels * sh.k);
    const uint32_t full = uint32_t(std::max(1, scene->device->cuCount)) * DR_TRACE_BLOCKS_PER_CU;
    sh.traceGrid = std::max(1u, std::min(full, (sh.nSlots + DR_TRACE_BLOCK - 1) / DR_TRACE_BLOCK));
    // Traversal-stack spill lanes: the tracers' persistent lanes, and one per
    // pixel for the per-pixel kernels (preview, first-hit AOVs at finalize).
    sh.stride = uint32_t(std::max<uint64_t>(pixels, uint64_t(sh.traceGrid) * DR_TRACE_BLOCK));
    sh.lay = pathLayout(sh.nSlots, sh.stride, scene->stackMax);
    if (st->denoiser == DR_DENOISER_APPLE_SVGF)  // RenderSession.swift:575-576
        return fail(DR_ERR_INVALID_SCENE, "Apple SVGF denoising is not available on this platform.");
    if (st->denoiser > DR_DENOISER_LEARNED) return fail(DR_ERR_INVALID_ARGUMENT, "unknown denoiser");
    if (st->intersection > DR_INTERSECTION_WATERTIGHT) return fail(DR_ERR_INVALID_ARGUMENT, "unknown intersection test");
    if (st->environment_sampling > DR_ENV_SAMPLING_ALIAS)
        return fail(DR_ERR_INVALID_ARGUMENT, "unknown environment sampling");
    sh.planes = st->denoiser != DR_DENOISER_NONE ? 10 : 8;  // + denoised beauty and its ping-pong temporary
    sh.outBytes = size_t(pixels) * 16 * sh.planes;
    return DR_OK;
}

// Points the session's arrays into its blocks for `sh` and clears the
// outputs and counters (on the session stream): a fresh session's state.
static dr_status bindSession(dr_session *s, dr_scene *scene, const dr_settings *st, const dr_gpu_camera *camera,
                             const dr_gpu_camera *prev, const SessionShape &sh) {
    s->scene = scene;
    s->settings = *st;
    s->sceneDev = scene->dev;
    s->sceneDev.sdf.renderQuality = st->quality;  // constants.renderQuality (RenderSettings.swift:106-115)
    s->sceneDev.watertight = st->intersection == DR_INTERSECTION_WATERTIGHT ? 1u : 0u;
    s->sceneDev.envAliasSampling = st->environment_sampling == DR_ENV_SAMPLING_ALIAS ? 1u : 0u;
    s->camera = toCam(*camera);
    s->previousCamera = prev ? toCam(*prev) : s->camera;
    s->samplesPerLaunch = sh.k;
    s->nSlots = sh.nSlots;
    s->traceGrid = sh.traceGrid;
    s->sampleCount = 0;
    dropGraphs(s);
    const PathLayout &l = sh.lay;
    char *b = static_cast<char *>(s->pathBlock);
    DrPaths &P = s->paths;
    P.rayO = reinterpret_cast<float4 *>(b + l.rayO);
    P.rayD = reinterpret_cast<float4 *>(b + l.rayD);
    P.thr = reinterpret_cast<float4 *>(b + l.thr);
    P.rad = reinterpret_cast<float4 *>(b + l.rad);
    P.meta = reinterpret_cast<uint4 *>(b + l.meta);
    P.hit = reinterpret_cast<float4 *>(b + l.hit);
    P.hitN = reinterpret_cast<float4 *>(b + l.hitN);
    P.neeT = reinterpret_cast<float4 *>(b + l.neeT);
    P.shO = reinterpret_cast<float4 *>(b + l.shO);
    P.shD = reinterpret_cast<float4 *>(b + l.shD);
    P.shC = reinterpret_cast<float4 *>(b + l.shC);
    P.shR = reinterpret_cast<float4 *>(b + l.shR);
    P.queue[0] = reinterpret_cast<uint32_t *>(b + l.q0);
    P.queue[1] = reinterpret_cast<uint32_t *>(b + l.q1);
    P.shadowQueue = reinterpret_cast<uint32_t *>(b + l.sq);
    P.sssQueue = reinterpret_cast<uint32_t *>(b + l.sss);
    P.resumeQueue = reinterpret_cast<uint32_t *>(b + l.res);
    P.counters = reinterpret_cast<uint32_t *>(b + l.ctr);
    P.claims = reinterpret_cast<uint32_t *>(b + l.claims);
    P.stats = reinterpret_cast<unsigned long long *>(b + l.stats);
    s->sampleBase = reinterpret_cast<uint32_t *>(b + l.base);
    P.overflow = reinterpret_cast<uint32_t *>(b + l.ovf);
    P.overflowStride = sh.stride;
    const size_t plane = sh.outBytes / size_t(sh.planes);
    char *o = static_cast<char *>(s->outBlock);
    DrOutputs &O = s->out;
    O.accumulation = reinterpret_cast<float4 *>(o + 0 * plane);
    O.depth = reinterpret_cast<float4 *>(o + 1 * plane);
    O.normal = reinterpret_cast<float4 *>(o + 2 * plane);
    O.albedo = reinterpret_cast<float4 *>(o + 3 * plane);
    O.materialID = reinterpret_cast<float4 *>(o + 4 * plane);
    O.objectID = reinterpret_cast<float4 *>(o + 5 * plane);
    O.motion = reinterpret_cast<float4 *>(o + 6 * plane);
    O.triangleID = reinterpret_cast<float4 *>(o + 7 * plane);
    s->denoised = sh.planes > 8 ? reinterpret_cast<float4 *>(o + 8 * plane) : nullptr;
    s->denoiseTemp = sh.planes > 8 ? reinterpret_cast<float4 *>(o + 9 * plane) : nullptr;
    s->denoisedCurrent = false;
    DR_HIP(hipMemsetAsync(s->outBlock, 0, sh.outBytes, s->stream), "clear outputs");
    DR_HIP(hipMemsetAsync(P.stats, 0, DR_CNT_SLOTS * 8, s->stream), "clear stats");
    for (int i = 0; i < DR_STAGE_COUNT; ++i) {
        s->stageMs[i] = 0.0;
        s->stageLaunches[i] = 0;
    }
    return DR_OK;
}

dr_status dr_hip_session_create(dr_scene *scene, const dr_settings *st, const dr_gpu_camera *camera,
                                const dr_gpu_camera *prev, dr_session **out) {
    if (!scene || !st || !camera || !out) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    SessionShape sh;
    dr_status rc = sessionShape(scene, st, sh);
    if (rc != DR_OK) return rc;
    DR_HIP(hipSetDevice(scene->device->ordinal), "hipSetDevice");
    DR_DEVICE_LOCK;
    auto *s = new dr_session;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->evBegin);
    if (e == hipSuccess) e = hipEventCreate(&s->evEnd);
    if (e == hipSuccess) e = hipMalloc(&s->pathBlock, sh.lay.bytes);
    if (e == hipSuccess) e = hipMemsetAsync(s->pathBlock, 0, sh.lay.bytes, s->stream);
    if (e == hipSuccess) e = hipMalloc(&s->outBlock, sh.outBytes);
    if (e != hipSuccess) {
        freeSession(s);
        delete s;
        return fail(e == hipErrorOutOfMemory ? DR_ERR_OUT_OF_MEMORY : DR_ERR_COMMAND_BUFFER_FAILED,
                    std::string("session buffers: ") + hipGetErrorString(e));
    }
    s->pathCapacity = sh.lay.bytes;
    s->outCapacity = sh.outBytes;
    rc = bindSession(s, scene, st, camera, prev, sh);
    if (rc != DR_OK) {
        freeSession(s);
        delete s;
        return rc;
    }
    drSceneRetain(scene);
    *out = s;
    return DR_OK;
}

dr_status dr_hip_session_retarget(dr_session *s, dr_scene *scene, const dr_settings *st, const dr_gpu_camera *camera,
                                  const dr_gpu_camera *prev) {
    if (!s || !scene || !st || !camera) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (scene->device != s->scene->device) return fail(DR_ERR_INVALID_ARGUMENT, "scene is on another device");
    SessionShape sh;
    dr_status rc = sessionShape(scene, st, sh);
    if (rc != DR_OK) return rc;
    if (sh.lay.bytes > s->pathCapacity || sh.outBytes > s->outCapacity)
        return fail(DR_ERR_OUT_OF_MEMORY, "settings need more memory than this session holds");
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    dr_scene *old = s->scene;
    rc = bindSession(s, scene, st, camera, prev, sh);  // s->scene is `scene` from here on
    if (old != scene) {
        drSceneRetain(scene);
        drSceneDrop(old);
    }
    return rc;
}

dr_status dr_hip_session_estimate_bytes(const dr_scene *scene, const dr_settings *st, uint64_t *bytes) {
    if (!scene || !st || !bytes) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    SessionShape sh;
    dr_status rc = sessionShape(scene, st, sh);
    if (rc != DR_OK) return rc;
    *bytes = sh.lay.bytes + sh.outBytes;
    if (st->denoiser == DR_DENOISER_LEARNED) *bytes += drUnetWorkBytes(st->width, st->height);
    return DR_OK;
}

}  // extern "C"

dr_status drSessionReserveUnet(dr_session *s) {
    const size_t need = drUnetWorkBytes(s->settings.width, s->settings.height);
    if (s->unetWorkBytes >= need) return DR_OK;
    DR_DEVICE_LOCK;
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    devFree(s->unetWork);
    s->unetWork = nullptr;
    s->unetWorkBytes = 0;
    DR_HIP(hipMalloc(&s->unetWork, need), "denoiser workspace");
    s->unetWorkBytes = need;
    return DR_OK;
}

extern "C" {

dr_status dr_hip_session_destroy(dr_session *s) {
    if (!s) return DR_OK;
    freeSession(s);
    dr_scene *sc = s->scene;
    delete s;
    if (sc) drSceneDrop(sc);
    return DR_OK;
}

static DrLaunch makeLaunch(const dr_session *s, uint32_t sampleBegin, uint32_t samples, uint32_t tx, uint32_t ty,
                           uint32_t tw, uint32_t th) {
    const dr_settings &st = s->settings;
    DrLaunch L{};
    L.width = st.width;
    L.height = st.height;
    L.maxBounces = std::max(1u, st.max_bounces);
    L.frameSeed = st.frame_seed;
    L.transparentBackground = st.transparent_background ? 1u : 0u;
    L.showsEnvironmentBackground = st.shows_environment_background ? 1u : 0u;
    L.denoiserEnabled = st.denoiser_enabled || st.denoiser != DR_DENOISER_NONE ? 1u : 0u;  // RenderSession.swift:941
    L.sampleRadianceClamp = st.sample_radiance_clamp < 0.f ? 64.f : st.sample_radiance_clamp;
    L.backgroundColor[0] = st.background_color[0];
    L.backgroundColor[1] = st.background_color[1];
    L.backgroundColor[2] = st.background_color[2];
    L.backgroundColor[3] = 1.f;
    L.tileX = tx;
    L.tileY = ty;
    L.tileWidth = tw;
    L.tileHeight = th;
    L.sampleBegin = sampleBegin;
    L.samplesInBatch = samples;
    L.statsEnabled = st.collects_stats ? 1u : 0u;
    return L;
}

static dr_status enqueueBatch(dr_session *s, const DrLaunch &L) {
    const uint32_t nSlots = L.tileWidth * L.tileHeight * L.samplesInBatch;
    // The persistent tracers get as many workgroups as this batch's paths
    // fill (a batch may hold fewer samples or pixels than the session).
    const uint32_t traceGrid = std::max(1u, std::min(s->traceGrid, (nSlots + DR_TRACE_BLOCK - 1) / DR_TRACE_BLOCK));
    auto run = [&](DrStage stage, uint32_t b) -> hipError_t {
        size_t ev = s->stageEvents.size();
        if (s->stageTiming) {
            DR_DEVICE_LOCK;
            hipEvent_t a, z;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&z) != hipSuccess) return hipErrorOutOfMemory;
            s->stageEvents.push_back(a);
            s->stageEvents.push_back(z);
            (void)hipEventRecord(a, s->stream);
        }
        hipError_t e = dr::launchStage(stage, b, s->sceneDev, L, s->camera, s->previousCamera, s->paths, s->out,
                                       nSlots, traceGrid, s->stream);
        if (s->stageTiming) (void)hipEventRecord(s->stageEvents[ev + 1], s->stream);
        s->stageLaunches[stage]++;
        return e;
    };
    if (s->settings.quality < 2u) {
        // Preview and interactive qualities: one kernel writes the sample
        // (DenrimRenderer.primaryPipeline, DenrimRenderer.swift:197-215).
        DR_HIP(run(DR_STAGE_PREVIEW, 0), "preview launch");
        return DR_OK;
    }
    // Triangle scenes: k_sss traces the shadow rays k_shade queued and the
    // next bounce's k_extend the rest, so k_resolve of bounce b follows
    // k_extend of b + 1 (k_shadow runs only after the last bounce). With
    // volumes every bounce ends with k_shadow and k_resolve.
    const bool vol = s->sceneDev.sdf.volumeCount != 0u;
    DR_HIP(run(DR_STAGE_RAYGEN, 0), "raygen launch");
    for (uint32_t b = 0; b < L.maxBounces; ++b) {
        DR_HIP(run(DR_STAGE_EXTEND, b), "extend launch");
        if (!vol && b > 0) DR_HIP(run(DR_STAGE_RESOLVE, b - 1), "resolve launch");
        DR_HIP(run(DR_STAGE_SHADE, b), "shade launch");
        DR_HIP(run(DR_STAGE_SSS, b), "subsurface launch");
        DR_HIP(run(DR_STAGE_RESUME, b), "shade resume launch");
        if (vol) {
            DR_HIP(run(DR_STAGE_SHADOW, b), "shadow launch");
            DR_HIP(run(DR_STAGE_RESOLVE, b), "resolve launch");
        }
    }
    if (!vol && L.maxBounces > 0) {
        DR_HIP(run(DR_STAGE_SHADOW, L.maxBounces - 1), "shadow launch");
        DR_HIP(run(DR_STAGE_RESOLVE, L.maxBounces - 1), "resolve launch");
    }
    DR_HIP(run(DR_STAGE_FINALIZE, 0), "finalize launch");
    return DR_OK;
}

// Stage order of events pushed by enqueueBatch, to attribute timings.
static void collectStageTimes(dr_session *s) {
    if (!s->stageTiming || s->stageEvents.empty()) return;
    (void)hipStreamSynchronize(s->stream);
    // Events come in (begin, end) pairs in launch order; replay the order.
    size_t i = 0;
    const uint32_t mb = std::max(1u, s->settings.max_bounces);
    const bool preview = s->settings.quality < 2u;
    const bool vol = s->sceneDev.sdf.volumeCount != 0u;
    while (i + 1 < s->stageEvents.size()) {
        auto add = [&](DrStage st) {
            float ms = 0.f;
            if (i + 1 < s->stageEvents.size() && hipEventElapsedTime(&ms, s->stageEvents[i], s->stageEvents[i + 1]) == hipSuccess)
                s->stageMs[st] += ms;
            i += 2;
        };
        if (preview) {
            add(DR_STAGE_PREVIEW);
            continue;
        }
        add(DR_STAGE_RAYGEN);
        for (uint32_t b = 0; b < mb; ++b) {  // the launch order of enqueueBatch
            add(DR_STAGE_EXTEND);
            if (!vol && b > 0) add(DR_STAGE_RESOLVE);
            add(DR_STAGE_SHADE);
            add(DR_STAGE_SSS);
            add(DR_STAGE_RESUME);
            if (vol) {
                add(DR_STAGE_SHADOW);
                add(DR_STAGE_RESOLVE);
            }
        }
        if (!vol) {
            add(DR_STAGE_SHADOW);
            add(DR_STAGE_RESOLVE);
        }
        add(DR_STAGE_FINALIZE);
    }
    DR_DEVICE_LOCK;
    for (auto ev : s->stageEvents) (void)hipEventDestroy(ev);
    s->stageEvents.clear();
}

// The graph of one full-frame batch of k samples (captured on first use):
// the stage launches of the batch, then the advance of sampleBase by k.
static dr_status batchGraph(dr_session *s, uint32_t k, hipGraphExec_t *out) {
    for (auto &g : s->graphs)
        if (g.first == k) {
            *out = g.second;
            return DR_OK;
        }
    DrLaunch L = makeLaunch(s, 0, k, 0, 0, s->settings.width, s->settings.height);
    L.sampleBase = s->sampleBase;
    DR_DEVICE_LOCK;  // no other thread allocates or frees while this stream captures
    DR_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    dr_status r = enqueueBatch(s, L);
    hipError_t e = r == DR_OK ? dr::launchAdvanceSampleBase(s->sampleBase, k, s->stream) : hipSuccess;
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(s->stream, &graph);
    if ((r != DR_OK || e != hipSuccess || ec != hipSuccess) && graph) (void)hipGraphDestroy(graph);
    if (r != DR_OK) return r;
    DR_HIP(e, "advance launch");
    DR_HIP(ec, "end capture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DR_HIP(e, "graph instantiate");
    s->graphs.emplace_back(k, exec);
    *out = exec;
    return DR_OK;
}

static dr_status encodeSamples(dr_session *s, uint32_t samples) {
    DR_HIP(hipSetDevice(s->scene->device->ordinal), "hipSetDevice");
    if (samples > 0) s->denoisedCurrent = false;
    DR_HIP(hipEventRecord(s->evBegin, s->stream), "event");
    const bool graphs = s->useGraphs && !s->stageTiming;
    if (graphs && samples > 0) DR_HIP(dr::launchSetSampleBase(s->sampleBase, s->sampleCount, s->stream), "set base");
    uint32_t left = samples;
    while (left > 0) {
        uint32_t k = std::min(left, s->samplesPerLaunch);
        if (graphs) {
            hipGraphExec_t exec = nullptr;
            dr_status r = batchGraph(s, k, &exec);
            if (r != DR_OK) return r;
            DR_HIP(hipGraphLaunch(exec, s->stream), "graph launch");
        } else {
            DrLaunch L = makeLaunch(s, s->sampleCount, k, 0, 0, s->settings.width, s->settings.height);
            dr_status r = enqueueBatch(s, L);
            if (r != DR_OK) return r;
        }
        s->sampleCount += k;
        left -= k;
    }
    DR_HIP(hipEventRecord(s->evEnd, s->stream), "event");
    return DR_OK;
}

dr_status dr_hip_session_encode(dr_session *s, uint32_t samples) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    return encodeSamples(s, samples);
}

dr_status dr_hip_session_synchronize(dr_session *s) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s->evBegin, s->evEnd) == hipSuccess) s->lastMs = ms;
    collectStageTimes(s);
    return DR_OK;
}

dr_status dr_hip_session_render(dr_session *s, uint32_t samples) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    uint32_t before = s->sampleCount;
    dr_status r = encodeSamples(s, samples);
    if (r == DR_OK) r = dr_hip_session_synchronize(s);
    if (r != DR_OK) s->sampleCount = before;  // RenderSession.swift:844-847
    return r;
}

dr_status dr_hip_session_render_tile(dr_session *s, uint32_t x, uint32_t y, uint32_t w, uint32_t h, int completes) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    // Clamp like RenderSession.encodeNextSample (RenderSession.swift:900-907).
    const uint32_t W = s->settings.width, H = s->settings.height;
    uint32_t tx = std::min(x, W), ty = std::min(y, H);
    uint32_t tw = std::min(w, W - tx), th = std::min(h, H - ty);
    if (tw == 0 || th == 0) return DR_OK;
    DR_HIP(hipSetDevice(s->scene->device->ordinal), "hipSetDevice");
    s->denoisedCurrent = false;
    DR_HIP(hipEventRecord(s->evBegin, s->stream), "event");
    DrLaunch L = makeLaunch(s, s->sampleCount, 1, tx, ty, tw, th);
    dr_status r = enqueueBatch(s, L);
    if (r != DR_OK) return r;
    DR_HIP(hipEventRecord(s->evEnd, s->stream), "event");
    r = dr_hip_session_synchronize(s);
    if (r == DR_OK && completes) s->sampleCount += 1;
    return r;
}

dr_status dr_hip_session_reset(dr_session *s) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    s->sampleCount = 0;
    s->denoisedCurrent = false;
    return DR_OK;
}

dr_status dr_hip_session_update_camera(dr_session *s, const dr_gpu_camera *camera, const dr_gpu_camera *prev) {
    if (!s || !camera) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    DrCamera old = s->camera;
    s->camera = toCam(*camera);
    s->previousCamera = prev ? toCam(*prev) : old;
    s->sampleCount = 0;
    s->denoisedCurrent = false;
    (void)hipStreamSynchronize(s->stream);
    dropGraphs(s);  // the graphs hold the old camera
    return DR_OK;
}

dr_status dr_hip_session_sample_count(const dr_session *s, uint32_t *out) {
    if (!s || !out) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = s->sampleCount;
    return DR_OK;
}

dr_status dr_hip_session_denoise(dr_session *s) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    if (s->settings.denoiser == DR_DENOISER_NONE || s->denoisedCurrent) return DR_OK;
    DR_HIP(hipSetDevice(s->scene->device->ordinal), "hipSetDevice");
    const dr_settings &st = s->settings;
    if (st.denoiser == DR_DENOISER_SIMPLE_SPATIAL) {
        // SimpleSpatialDenoiser.encode (Denoising/Denoiser.swift:112-164): 1-5
        // passes with step 1 << pass, the last one into the denoised plane.
        const uint32_t n = std::max(1u, std::min(st.denoise_iterations, 5u));
        const float4 *src = s->out.accumulation;
        for (uint32_t i = 0; i < n; ++i) {
            float4 *dst = (n - 1 - i) % 2 == 0 ? s->denoised : s->denoiseTemp;
            dr::AtrousParams p;
            p.width = st.width, p.height = st.height;
            p.radius = std::max(1u, std::min(st.denoise_radius, 2u));
            p.step = 1u << i;
            p.normalSigma = std::max(st.denoise_normal_sigma, 0.0001f);
            p.depthSigma = std::max(st.denoise_depth_sigma, 0.0001f);
            p.albedoSigma = std::max(st.denoise_albedo_sigma, 0.0001f);
            p.colorSigma = std::max(st.denoise_color_sigma, 0.0001f);
            DR_HIP(dr::launchAtrous(src, s->out.depth, s->out.normal, s->out.albedo, dst, p, s->stream), "a-trous pass");
            src = dst;
        }
    } else {
        if (!s->network) return fail(DR_ERR_INVALID_ARGUMENT, "no denoiser network set on this session");
        dr_status rc = drUnetDenoiseSession(s->network, s, &s->unetWork, &s->unetWorkBytes);
        if (rc != DR_OK) return rc;
    }
    s->denoisedCurrent = true;
    return DR_OK;
}

dr_status dr_hip_session_set_denoiser_network(dr_session *s, dr_unet *net) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    s->network = net;
    s->denoisedCurrent = false;
    return DR_OK;
}

dr_status dr_hip_session_read_output(dr_session *s, dr_output kind, float *rgba, size_t floats) {
    if (!s || !rgba) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t pixels = size_t(s->settings.width) * s->settings.height;
    if (floats < pixels * 4) return fail(DR_ERR_INVALID_ARGUMENT, "output buffer too small");
    if (int(kind) < 0 || int(kind) > 7) return fail(DR_ERR_INVALID_ARGUMENT, "unknown output kind");
    // The beauty output is the denoised plane when a denoiser is set
    // (RenderSession.texture(for:), RenderSession.swift:1072-1081).
    const bool denoised = kind == DR_OUTPUT_BEAUTY && s->settings.denoiser != DR_DENOISER_NONE;
    if (denoised) {
        dr_status rc = dr_hip_session_denoise(s);
        if (rc != DR_OK) return rc;
    }
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    const float4 *planes[8] = {denoised ? s->denoised : s->out.accumulation, s->out.depth, s->out.normal,
                               s->out.albedo, s->out.materialID, s->out.objectID, s->out.motion, s->out.triangleID};
    DR_HIP(drCopy(rgba, planes[int(kind)], pixels * 16, hipMemcpyDeviceToHost, s->stream), "readback");
    return DR_OK;
}

dr_status dr_hip_session_stats(dr_session *s, uint64_t *counters, size_t count) {
    if (!s || !counters) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    unsigned long long tmp[DR_CNT_SLOTS];
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    DR_HIP(drCopy(tmp, s->paths.stats, sizeof(tmp), hipMemcpyDeviceToHost, s->stream), "stats readback");
    for (size_t i = 0; i < count && i < DR_CNT_SLOTS; ++i) counters[i] = tmp[i];
    return DR_OK;
}

dr_status dr_hip_session_reset_stats(dr_session *s) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    DR_HIP(hipMemsetAsync(s->paths.stats, 0, DR_CNT_SLOTS * 8, s->stream), "stats reset");
    DR_HIP(hipStreamSynchronize(s->stream), "stream synchronize");
    for (int i = 0; i < DR_STAGE_COUNT; ++i) {
        s->stageMs[i] = 0.0;
        s->stageLaunches[i] = 0;
    }
    return DR_OK;
}

dr_status dr_hip_session_set_stage_timing(dr_session *s, int enabled) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    s->stageTiming = enabled != 0;
    return DR_OK;
}

dr_status dr_hip_session_set_graphs(dr_session *s, int enabled) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    if (!enabled && s->useGraphs) {
        (void)hipStreamSynchronize(s->stream);
        dropGraphs(s);
    }
    s->useGraphs = enabled != 0;
    return DR_OK;
}

dr_status dr_hip_session_query(dr_session *s, int *done) {
    if (!s || !done) return fail(DR_ERR_INVALID_ARGUMENT, "NULL argument");
    hipError_t e = hipStreamQuery(s->stream);
    if (e == hipErrorNotReady) {
        *done = 0;
        return DR_OK;
    }
    DR_HIP(e, "stream query");
    *done = 1;
    return DR_OK;
}

dr_status dr_hip_session_timing(dr_session *s, double *total_ms, double *stage_ms, uint64_t *launches, size_t stages) {
    if (!s) return fail(DR_ERR_INVALID_ARGUMENT, "session is NULL");
    if (total_ms) *total_ms = s->lastMs;
    for (size_t i = 0; i < stages && i < DR_STAGE_COUNT; ++i) {
        if (stage_ms) stage_ms[i] = s->stageMs[i];
        if (launches) launches[i] = s->stageLaunches[i];
    }
    return DR_OK;
}

}  // extern "C"
