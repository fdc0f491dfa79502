// dr_sdf_build.cpp -- host builder of signed-distance volumes (dense grids
// and sparse brick sets).
//
// Restates DistanceVolumeBuilder (Scene/SDFModel.swift):
//   * primitives (CompiledPrimitive.distance, :685-713): sphere, rounded box
//     and capped cylinder in primitive-local space reached through the
//     inverse primitive transform, scaled by the transform's smallest axis
//     scale; plus a torus (an addition of this backend);
//   * the fold over primitives in order (sampleField, :480-505): union by
//     the smooth union `combine` (:515-580) with each primitive's own radius,
//     subtract as max(field, -d);
//   * sample positions boundsMin + extent * i / (dims - 1) (:461-472);
//   * build (dense, every sample) and buildSparse (:256-355): bricks of
//     `brickSize` core samples in z, y, x order, stored with an apron of
//     `apron` samples per side (the reference uses 2) clamped to the grid,
//     kept when min <= band && max >= -band (:511-513).
//
// With `quantize` set (the FP16 brick format of the HIP backend) three
// things change, by design: distances are rounded to FP16 (RNE) and blends
// to multiples of 1/65535; a union step whose smooth-min weight is exactly
// zero (h >= 1) leaves the sample unchanged (the reference recomputes
// mix(d, current, 1), which can differ from current by rounding); and, when
// every primitive is a union, bricks and primitives that provably cannot
// change a brick are skipped (exact, by the Lipschitz bound of each
// primitive's distance and the h >= 1 rule).
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
    dr_sdf_primitive s;
    float k;  // effective smooth radius
};

// Column-major 4x4 times (p, 1): ((c0 x + c1 y) + c2 z) + c3.
void toLocal(const float *m, const float p[3], float out[3]) {
    for (int r = 0; r < 3; ++r) out[r] = ((m[0 + r] * p[0] + m[4 + r] * p[1]) + m[8 + r] * p[2]) + m[12 + r];
}

float len3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }
float len2(float x, float y) { return std::sqrt(x * x + y * y); }

float primDistance(const Prim &P, const float x[3]) {
    const dr_sdf_primitive &p = P.s;
    float l[3] = {x[0], x[1], x[2]};
    if (p.has_transform) toLocal(p.world_to_primitive, x, l);
    float q[3] = {l[0] - p.center[0], l[1] - p.center[1], l[2] - p.center[2]};
    float d;
    switch (p.kind) {
    case DR_SDF_SPHERE:
        d = len3(q[0], q[1], q[2]) - p.radius;
        break;
    case DR_SDF_TORUS: {
        float h = q[0] * p.axis[0] + q[1] * p.axis[1] + q[2] * p.axis[2];
        float r[3] = {q[0] - p.axis[0] * h, q[1] - p.axis[1] * h, q[2] - p.axis[2] * h};
        float radial = len3(r[0], r[1], r[2]) - p.radius;
        d = len2(radial, h) - p.minor_radius;
        break;
    }
    case DR_SDF_BOX: {
        float b[3] = {std::fabs(q[0]) - p.half_extents[0], std::fabs(q[1]) - p.half_extents[1],
                      std::fabs(q[2]) - p.half_extents[2]};
        d = len3(std::max(b[0], 0.0f), std::max(b[1], 0.0f), std::max(b[2], 0.0f)) +
            std::min(std::max(b[0], std::max(b[1], b[2])), 0.0f) - std::max(p.corner_radius, 0.0f);
        break;
    }
    default: {  // DR_SDF_CYLINDER, axis = local y
        float dx = len2(q[0], q[2]) - p.radius, dy = std::fabs(q[1]) - p.half_height;
        d = std::min(std::max(dx, dy), 0.0f) + len2(std::max(dx, 0.0f), std::max(dy, 0.0f));
        break;
    }
    }
    return p.has_transform ? d * p.distance_scale : d;
}

struct Field {
    float distance;
    uint32_t a, b;
    float blend;
};

// combine (SDFModel.swift:515-560) for material ids (no field channels).
void combine(Field &f, float d, uint32_t material, float radius, bool quantize) {
    radius = std::max(radius, 0.0f);
    if (radius <= 1e-6f) {
        if (d < f.distance) f = Field{d, material, material, 0.0f};
        return;
    }
    float h = std::min(std::max(0.5f + 0.5f * (d - f.distance) / radius, 0.0f), 1.0f);
    if (quantize && h >= 1.0f) return;  // zero weight: unchanged
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

struct Model {
    std::vector<Prim> P;
    bool unionOnly = true;
    float kMax = 0.0f;
    bool quantize = false;

    // sampleField (SDFModel.swift:480-505) over the primitives in `idx`
    // (all of them when idx is null).
    Field sample(const float p[3], const std::vector<uint32_t> *idx) const {
        Field f{std::numeric_limits<float>::max(), P[0].s.material, P[0].s.material, 0.0f};
        auto step = [&](const Prim &q) {
            float d = primDistance(q, p);
            if (q.s.operation == DR_SDF_SUBTRACT) {
                if (std::isfinite(f.distance)) f.distance = std::max(f.distance, -d);
            } else {
                combine(f, d, q.s.material, q.k, quantize);
            }
        };
        if (idx)
            for (uint32_t i : *idx) step(P[i]);
        else
            for (const Prim &q : P) step(q);
        return f;
    }
    void store(const Field &f, float &d, uint32_t &a, uint32_t &b, float &w) const {
        a = f.a;
        b = f.b;
        if (quantize) {
            d = roundHalf(f.distance);
            uint32_t q = uint32_t(std::lround(std::min(std::max(f.blend, 0.0f), 1.0f) * 65535.0f));
            w = float(q) / 65535.0f;
        } else {
            d = f.distance;
            w = f.blend;
        }
    }
};

dr_status makeModel(const dr_sdf_primitive *prims, uint32_t count, const dr_sdf_build_settings *st, Model &m) {
    if (!prims || !st || count == 0) return DR_ERR_INVALID_SCENE;  // SDFModel.swift:401-403
    for (int a = 0; a < 3; ++a)
        if (!(st->bounds_max[a] > st->bounds_min[a])) return DR_ERR_INVALID_SCENE;
    m.quantize = st->quantize != 0;
    m.P.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        Prim &q = m.P[i];
        q.s = prims[i];
        if (q.s.kind > DR_SDF_CYLINDER || q.s.operation > DR_SDF_SUBTRACT) return DR_ERR_INVALID_ARGUMENT;
        q.k = q.s.smooth_radius < 0.0f ? st->smooth_radius : q.s.smooth_radius;
        if (q.s.kind == DR_SDF_TORUS) {
            float n = len3(q.s.axis[0], q.s.axis[1], q.s.axis[2]);
            if (!(n > 0.0f)) return DR_ERR_INVALID_SCENE;
            for (float &c : q.s.axis) c /= n;
        }
        if (!q.s.has_transform) q.s.lipschitz = 1.0f;
        if (!(q.s.lipschitz >= 1.0f)) q.s.lipschitz = 1.0f;
        if (q.s.operation == DR_SDF_SUBTRACT) m.unionOnly = false;
        m.kMax = std::max(m.kMax, std::max(q.k, 0.0f));
    }
    return DR_OK;
}

}  // namespace

struct dr_sdf_volume {
    std::vector<int32_t> storedOrigin, storedDims, coreOrigin, coreDims;  // 3 per brick
    std::vector<uint64_t> sampleOffset;                                    // per brick
    std::vector<float> distance, blend;
    std::vector<uint32_t> materialA, materialB;
    uint64_t candidates = 0;
};

extern "C" {

dr_status dr_sdf_build_sparse(const dr_sdf_primitive *prims, uint32_t count, const dr_sdf_build_settings *st,
                              dr_sdf_volume **out) {
    if (!out) return DR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    Model M;
    dr_status rc = makeModel(prims, count, st, M);
    if (rc != DR_OK) return rc;
    uint32_t dims[3];
    for (int a = 0; a < 3; ++a) dims[a] = std::max(st->dims[a], 2u);
    const uint32_t B = std::max(st->brick_size, 1u);
    const uint32_t apron = st->apron;
    const float band = st->narrow_band;
    float extent[3];
    for (int a = 0; a < 3; ++a) extent[a] = st->bounds_max[a] - st->bounds_min[a];
    auto samplePos = [&](int32_t i, int a) { return st->bounds_min[a] + extent[a] * (float(i) / float(dims[a] - 1)); };
    const bool cull = M.quantize && M.unionOnly;

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
            const std::vector<uint32_t> *idx = nullptr;
            if (cull) {
                // Stored region: centre c and half diagonal R. Primitive i
                // satisfies |d_i(p) - d_i(c)| <= L_i R inside it.
                float c[3], half2 = 0.0f;
                for (int a = 0; a < 3; ++a) {
                    float lo = samplePos(so[a], a), hi = samplePos(se[a] - 1, a);
                    c[a] = 0.5f * (lo + hi);
                    half2 += 0.25f * (hi - lo) * (hi - lo);
                }
                float R = std::sqrt(half2) * 1.0001f + 1e-6f;
                float upper = std::numeric_limits<float>::max();  // >= the field at every sample so far
                float lower = std::numeric_limits<float>::max();  // min_i (d_i(c) - L_i R)
                relevant.clear();
                for (uint32_t i = 0; i < count; ++i) {
                    const Prim &q = M.P[i];
                    float d = primDistance(q, c);
                    float LR = q.s.lipschitz * R;
                    // h >= 1 at every sample: d_i(p) - field(p) >= k_i.
                    if (i > 0 && (d - LR) - upper >= q.k + 1e-6f) continue;
                    relevant.push_back(i);
                    upper = std::min(upper, d + LR);
                    lower = std::min(lower, d - LR);
                }
                // The smooth union lies in [min - kMax / 4, min].
                if (lower - 0.25f * M.kMax > band || upper < -band) continue;
                idx = &relevant;
            }
            candidates.fetch_add(1);
            int32_t n[3] = {se[0] - so[0], se[1] - so[1], se[2] - so[2]};
            size_t ns = size_t(n[0]) * n[1] * n[2];
            o.dist.resize(ns);
            o.blend.resize(ns);
            o.a.resize(ns);
            o.b.resize(ns);
            float mn = std::numeric_limits<float>::max(), mx = -std::numeric_limits<float>::max();
            size_t k = 0;
            for (int32_t z = 0; z < n[2]; ++z)
                for (int32_t y = 0; y < n[1]; ++y)
                    for (int32_t x = 0; x < n[0]; ++x, ++k) {
                        float p[3] = {samplePos(so[0] + x, 0), samplePos(so[1] + y, 1), samplePos(so[2] + z, 2)};
                        M.store(M.sample(p, idx), o.dist[k], o.a[k], o.b[k], o.blend[k]);
                        mn = std::min(mn, o.dist[k]);
                        mx = std::max(mx, o.dist[k]);
                    }
            if (!(mn <= band && mx >= -band)) {
                std::vector<float>().swap(o.dist);
                std::vector<float>().swap(o.blend);
                std::vector<uint32_t>().swap(o.a);
                std::vector<uint32_t>().swap(o.b);
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
    auto *v = new dr_sdf_volume;
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

dr_status dr_sdf_build_dense(const dr_sdf_primitive *prims, uint32_t count, const dr_sdf_build_settings *st,
                             float *distance, uint32_t *materialA, uint32_t *materialB, float *blend) {
    Model M;
    dr_status rc = makeModel(prims, count, st, M);
    if (rc != DR_OK) return rc;
    if (!distance || !materialA || !materialB || !blend) return DR_ERR_INVALID_ARGUMENT;
    uint32_t dims[3];
    float extent[3];
    for (int a = 0; a < 3; ++a) {
        dims[a] = std::max(st->dims[a], 2u);
        extent[a] = st->bounds_max[a] - st->bounds_min[a];
    }
    auto samplePos = [&](uint32_t i, int a) { return st->bounds_min[a] + extent[a] * (float(i) / float(dims[a] - 1)); };
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        while (true) {
            uint32_t z = next.fetch_add(1);
            if (z >= dims[2]) break;
            for (uint32_t y = 0; y < dims[1]; ++y)
                for (uint32_t x = 0; x < dims[0]; ++x) {
                    size_t k = x + size_t(y) * dims[0] + size_t(z) * dims[0] * dims[1];
                    float p[3] = {samplePos(x, 0), samplePos(y, 1), samplePos(z, 2)};
                    M.store(M.sample(p, nullptr), distance[k], materialA[k], materialB[k], blend[k]);
                }
        }
    };
    unsigned T = std::max(1u, std::min(std::thread::hardware_concurrency(), dims[2]));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t) pool.emplace_back(work);
    for (auto &t : pool) t.join();
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
