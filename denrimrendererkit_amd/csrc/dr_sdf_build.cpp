// dr_sdf_build.cpp -- host builder of sparse signed-distance brick volumes.
//
// Restates DistanceVolumeBuilder.buildSparse (Scene/SDFModel.swift:256-355):
// the virtual sample grid spans [boundsMin, boundsMax] with `dims` samples
// per axis (sample i at boundsMin + extent * i / (dims - 1),
// SDFModel.swift:461-472); bricks of `brickSize` core samples are visited in
// z, y, x order, each stores its core plus an apron of `apron` samples per
// side (clamped to the grid) and is kept when its stored samples reach the
// narrow band (min <= band && max >= -band, SDFModel.swift:511-513).
// The field is the reference's smooth union fold (combine,
// SDFModel.swift:515-560) over the primitives in order, with one defined
// difference: a primitive whose smooth-min weight is exactly zero (h >= 1)
// leaves the running sample unchanged. That makes it exact to skip, per
// brick, every primitive whose distance bound at the brick centre is at
// least the smooth radius beyond the nearest earlier primitive, and to skip
// whole bricks whose centre distance exceeds band + the stored region's half
// diagonal (the field is 1-Lipschitz).
// Distances are rounded to FP16 and stored as the equal float; blend weights
// are rounded to multiples of 1/65535; both survive the HIP backend's
// compact encoding unchanged.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/denrim_hip.h"

namespace {

struct Prim {
    uint32_t kind, material;
    float c[3];
    float radius;
    float axis[3];
    float minor;
};

float primDistance(const Prim &p, const float x[3]) {
    float q[3] = {x[0] - p.c[0], x[1] - p.c[1], x[2] - p.c[2]};
    if (p.kind == DR_SDF_SPHERE) return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) - p.radius;
    // Torus around `axis` through the centre: major radius `radius`, tube `minor`.
    float h = q[0] * p.axis[0] + q[1] * p.axis[1] + q[2] * p.axis[2];
    float r[3] = {q[0] - p.axis[0] * h, q[1] - p.axis[1] * h, q[2] - p.axis[2] * h};
    float radial = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) - p.radius;
    return std::sqrt(radial * radial + h * h) - p.minor;
}

struct Field {
    float distance;
    uint32_t a, b;
    float blend;
};

// combine (SDFModel.swift:515-560) for material ids only (no field channels).
void combine(Field &f, float d, uint32_t material, float radius) {
    radius = std::max(radius, 0.0f);
    if (radius <= 1e-6f) {
        if (d < f.distance) f = Field{d, material, material, 0.0f};
        return;
    }
    float h = std::min(std::max(0.5f + 0.5f * (d - f.distance) / radius, 0.0f), 1.0f);
    if (h >= 1.0f) return;  // zero weight: unchanged
    float distance = (d + (f.distance - d) * h) - radius * h * (1.0f - h);
    float w = 1.0f - h;
    if (w <= 0.001f) {
        f.distance = distance;
    } else if (w >= 0.999f) {
        f = Field{distance, material, material, 0.0f};
    } else {
        f = Field{distance, f.a, material, w};
    }
}

// IEEE binary16 round-to-nearest-even, back to float.
float roundHalf(float v) {
    if (!std::isfinite(v)) return v;
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    uint32_t sign = bits & 0x80000000u;
    float a = std::fabs(v);
    if (a >= 65520.0f) {  // overflows to inf in half
        float inf = std::numeric_limits<float>::infinity();
        return sign ? -inf : inf;
    }
    if (a < 6.103515625e-05f) {  // subnormal half: multiple of 2^-24
        float q = std::nearbyint(a * 16777216.0f) / 16777216.0f;
        return sign ? -q : q;
    }
    int e;
    std::frexp(a, &e);  // a = m * 2^e, m in [0.5, 1)
    float scale = std::ldexp(1.0f, 11 - e);
    float q = std::nearbyint(a * scale) / scale;  // 11 significant bits
    return sign ? -q : q;
}

}  // namespace

struct dr_sdf_volume {
    uint32_t dims[3];
    uint32_t brickSize, apron;
    float boundsMin[3], boundsMax[3];
    std::vector<int32_t> storedOrigin, storedDims, coreOrigin, coreDims;  // 3 per brick
    std::vector<uint64_t> sampleOffset;                                    // per brick
    std::vector<float> distance, blend;
    std::vector<uint32_t> materialA, materialB;
    uint64_t candidates = 0;
};

extern "C" {

dr_status dr_sdf_build_sparse(const dr_sdf_primitive *prims, uint32_t count, const dr_sdf_build_settings *st,
                              dr_sdf_volume **out) {
    if (!prims || !st || !out || count == 0) return DR_ERR_INVALID_SCENE;  // SDFModel.swift:401-403
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (!(st->bounds_max[a] > st->bounds_min[a])) return DR_ERR_INVALID_SCENE;
    std::vector<Prim> P(count);
    for (uint32_t i = 0; i < count; ++i) {
        const dr_sdf_primitive &s = prims[i];
        P[i] = Prim{s.kind, s.material, {s.center[0], s.center[1], s.center[2]}, s.radius,
                    {s.axis[0], s.axis[1], s.axis[2]}, s.minor_radius};
        float n = std::sqrt(P[i].axis[0] * P[i].axis[0] + P[i].axis[1] * P[i].axis[1] + P[i].axis[2] * P[i].axis[2]);
        if (P[i].kind == DR_SDF_TORUS) {
            if (!(n > 0.0f)) return DR_ERR_INVALID_SCENE;
            for (float &c : P[i].axis) c /= n;
        }
    }
    auto *v = new dr_sdf_volume;
    uint32_t dims[3];
    for (int a = 0; a < 3; ++a) {
        dims[a] = std::max(st->dims[a], 2u);
        v->dims[a] = dims[a];
        v->boundsMin[a] = st->bounds_min[a];
        v->boundsMax[a] = st->bounds_max[a];
    }
    const uint32_t B = std::max(st->brick_size, 1u);
    const uint32_t apron = st->apron;
    v->brickSize = B;
    v->apron = apron;
    const float band = st->narrow_band;
    const float k = st->smooth_radius;
    float extent[3], step[3];
    for (int a = 0; a < 3; ++a) {
        extent[a] = st->bounds_max[a] - st->bounds_min[a];
        step[a] = extent[a] / float(dims[a] - 1);
    }
    auto samplePos = [&](int32_t i, int a) { return st->bounds_min[a] + extent[a] * (float(i) / float(dims[a] - 1)); };

    uint32_t grid[3];
    for (int a = 0; a < 3; ++a) grid[a] = (dims[a] + B - 1) / B;
    const uint64_t nBricks = uint64_t(grid[0]) * grid[1] * grid[2];

    // Per brick result, filled by worker threads, then compacted in z, y, x order.
    struct Out {
        bool keep = false;
        int32_t so[3], sd[3], co[3], cd[3];
        std::vector<float> dist, blend;
        std::vector<uint32_t> a, b;
    };
    std::vector<Out> res(nBricks);
    std::atomic<uint64_t> next{0}, candidates{0};
    unsigned T = std::max(1u, std::thread::hardware_concurrency());
    auto work = [&]() {
        std::vector<float> centreD(count);
        std::vector<uint32_t> relevant;
        relevant.reserve(count);
        while (true) {
            uint64_t bi = next.fetch_add(1);
            if (bi >= nBricks) break;
            uint32_t bz = uint32_t(bi / (uint64_t(grid[0]) * grid[1]));
            uint32_t by = uint32_t((bi / grid[0]) % grid[1]);
            uint32_t bx = uint32_t(bi % grid[0]);
            int32_t origin[3] = {int32_t(bx * B), int32_t(by * B), int32_t(bz * B)};
            Out &o = res[bi];
            int32_t cd[3], so[3], se[3];
            for (int a = 0; a < 3; ++a) {
                cd[a] = std::min<int32_t>(int32_t(B), int32_t(dims[a]) - origin[a]);
                so[a] = std::max<int32_t>(origin[a] - int32_t(apron), 0);
                se[a] = std::min<int32_t>(origin[a] + cd[a] + int32_t(apron), int32_t(dims[a]));
            }
            // Bounds of the stored region: centre and half diagonal.
            float c[3], half2 = 0.0f;
            for (int a = 0; a < 3; ++a) {
                float lo = samplePos(so[a], a), hi = samplePos(se[a] - 1, a);
                c[a] = 0.5f * (lo + hi);
                half2 += 0.25f * (hi - lo) * (hi - lo);
            }
            float R = std::sqrt(half2) * 1.0001f + 1e-6f;
            float runMin = std::numeric_limits<float>::max();
            float fieldLow = std::numeric_limits<float>::max();  // lower bound of the field over the brick
            float nearest = std::numeric_limits<float>::max();
            relevant.clear();
            for (uint32_t i = 0; i < count; ++i) {
                float d = primDistance(P[i], c);
                centreD[i] = d;
                // Skip exactly when h >= 1 at every sample (see the file comment).
                if (i > 0 && d - R >= runMin + R + k + 1e-6f) continue;
                relevant.push_back(i);
                runMin = std::min(runMin, d);
                nearest = std::min(nearest, d);
            }
            // Smooth union lies in [min - k/4, min]; centre field bound +- R.
            fieldLow = nearest - 0.25f * std::max(k, 0.0f) - R;
            float fieldHigh = nearest + R;
            if (fieldLow > band || fieldHigh < -band) continue;
            candidates.fetch_add(1);
            int32_t n[3] = {se[0] - so[0], se[1] - so[1], se[2] - so[2]};
            size_t ns = size_t(n[0]) * n[1] * n[2];
            o.dist.resize(ns);
            o.blend.resize(ns);
            o.a.resize(ns);
            o.b.resize(ns);
            float mn = std::numeric_limits<float>::max(), mx = -std::numeric_limits<float>::max();
            size_t idx = 0;
            for (int32_t z = 0; z < n[2]; ++z)
                for (int32_t y = 0; y < n[1]; ++y)
                    for (int32_t x = 0; x < n[0]; ++x, ++idx) {
                        float p[3] = {samplePos(so[0] + x, 0), samplePos(so[1] + y, 1), samplePos(so[2] + z, 2)};
                        Field f{std::numeric_limits<float>::max(), P[0].material, P[0].material, 0.0f};
                        for (uint32_t i : relevant) combine(f, primDistance(P[i], p), P[i].material, k);
                        float d = roundHalf(f.distance);
                        uint32_t q = uint32_t(std::lround(std::min(std::max(f.blend, 0.0f), 1.0f) * 65535.0f));
                        o.dist[idx] = d;
                        o.blend[idx] = float(q) / 65535.0f;
                        o.a[idx] = f.a;
                        o.b[idx] = f.b;
                        mn = std::min(mn, d);
                        mx = std::max(mx, d);
                    }
            if (!(mn <= band && mx >= -band)) {
                o.dist.clear();
                o.blend.clear();
                o.a.clear();
                o.b.clear();
                continue;
            }
            o.keep = true;
            for (int a = 0; a < 3; ++a) {
                o.so[a] = so[a];
                o.sd[a] = n[a];
                o.co[a] = origin[a];
                o.cd[a] = cd[a];
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t) pool.emplace_back(work);
    for (auto &t : pool) t.join();
    v->candidates = candidates.load();
    uint64_t offset = 0;
    for (uint64_t bi = 0; bi < nBricks; ++bi) {
        Out &o = res[bi];
        if (!o.keep) continue;
        for (int a = 0; a < 3; ++a) {
            v->storedOrigin.push_back(o.so[a]);
            v->storedDims.push_back(o.sd[a]);
            v->coreOrigin.push_back(o.co[a]);
            v->coreDims.push_back(o.cd[a]);
        }
        v->sampleOffset.push_back(offset);
        offset += o.dist.size();
        v->distance.insert(v->distance.end(), o.dist.begin(), o.dist.end());
        v->blend.insert(v->blend.end(), o.blend.begin(), o.blend.end());
        v->materialA.insert(v->materialA.end(), o.a.begin(), o.a.end());
        v->materialB.insert(v->materialB.end(), o.b.begin(), o.b.end());
        std::vector<float>().swap(o.dist);
        std::vector<float>().swap(o.blend);
        std::vector<uint32_t>().swap(o.a);
        std::vector<uint32_t>().swap(o.b);
    }
    *out = v;
    return DR_OK;
}

dr_status dr_sdf_volume_info(const dr_sdf_volume *v, uint32_t *bricks, uint64_t *samples, uint64_t *candidates) {
    if (!v) return DR_ERR_INVALID_ARGUMENT;
    if (bricks) *bricks = uint32_t(v->sampleOffset.size());
    if (samples) *samples = v->distance.size();
    if (candidates) *candidates = v->candidates;
    return DR_OK;
}

dr_status dr_sdf_volume_export(const dr_sdf_volume *v, int32_t *storedOrigin, int32_t *storedDims, int32_t *coreOrigin,
                               int32_t *coreDims, float *distance, uint32_t *materialA, uint32_t *materialB,
                               float *blend) {
    if (!v) return DR_ERR_INVALID_ARGUMENT;
    auto put = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(storedOrigin, v->storedOrigin);
    put(storedDims, v->storedDims);
    put(coreOrigin, v->coreOrigin);
    put(coreDims, v->coreDims);
    put(distance, v->distance);
    put(materialA, v->materialA);
    put(materialB, v->materialB);
    put(blend, v->blend);
    return DR_OK;
}

dr_status dr_sdf_volume_free(dr_sdf_volume *v) {
    delete v;
    return DR_OK;
}

}  // extern "C"
