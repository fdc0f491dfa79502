/*
 * oracle.cpp -- CPU oracle for the progressive path tracer.
 *
 * TEST INFRASTRUCTURE ONLY: a plain C++ restatement of the reference path
 * tracer, used by tests/, smoke() and bench.py's cpu_baseline leg as the
 * checker. The product (libdenrim_hip.so) never links or calls it.
 *
 * It follows, function by function, the Metal source
 *   Sources/DenrimRendererKit/Resources/Shaders/PathTrace.metal
 * and the count-median BVH of
 *   Sources/DenrimRendererKit/Acceleration/BVHBuilder.swift:4-135 and
 *   Sources/DenrimRendererKit/Acceleration/FlatBVH.swift:13-33.
 * Each function cites the lines it restates. It is templated on the real type
 * so the same text gives the FP32 oracle (per-pixel parity) and the FP64
 * oracle (converged ground truth).
 *
 * Deliberate, documented deviations from the reference text:
 *   * closest-hit ties (equal t) go to the lowest triangle index, so that
 *     the result does not depend on BVH layout (the reference keeps the
 *     first triangle visited, PathTrace.metal:3371);
 *   * every random draw is its own statement, in the left-to-right order the
 *     reference's clang-based compiler uses for argument lists;
 *   * math is IEEE (no fast-math): normalize(v) = v * (1/sqrt(dot(v,v))),
 *     mix(a,b,t) = a + (b-a)*t, min/max follow fmin/fmax (Metal semantics).
 * Build with -ffp-contract=off so no multiply-add is fused.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_bsdf.h"

namespace {


template <class R> bool projectToScreen(const or_camera &cam, V3<R> p, uint32_t w, uint32_t h, V2<R> &out) {  // :1666-1701
    V3<R> hor = v3<R>(cam.horizontal), ver = v3<R>(cam.vertical), ll = v3<R>(cam.lowerLeft), org = v3<R>(cam.origin);
    V3<R> forward = normalize(cross(ver, hor));
    V3<R> planeCenter = ll + hor * R(0.5) + ver * R(0.5);
    V3<R> planePoint;
    if (R(cam.origin.w) > R(0.5)) {
        planePoint = p - forward * dot(p - planeCenter, forward);
    } else {
        V3<R> toPoint = p - org;
        R denominator = dot(toPoint, forward);
        if (std::fabs(denominator) <= R(1e-5)) return false;
        R planeDistance = dot(planeCenter - org, forward);
        R planeT = planeDistance / denominator;
        if (planeT <= R(0)) return false;
        planePoint = org + toPoint * planeT;
    }
    V3<R> rel = planePoint - ll;
    R u = dot(rel, hor) / fmx(dot(hor, hor), R(1e-5));
    R v = dot(rel, ver) / fmx(dot(ver, ver), R(1e-5));
    out = V2<R>{u * R(w), (R(1) - v) * R(h)};
    return true;
}

template <class R> V3<R> sampleTriangle(const or_triangle &t, R ux, R uy) {  // :1715-1721
    R su0 = std::sqrt(ux);
    R b0 = R(1) - su0;
    R b1 = su0 * (R(1) - uy);
    R b2 = su0 * uy;
    return b0 * v3<R>(t.v0) + b1 * v3<R>(t.v1) + b2 * v3<R>(t.v2);
}

// ---------------------------------------------------- triangle and boxes --
template <class R>
bool intersectTriangle(const Ray<R> &ray, const or_triangle &tri, R &tOut, R &uOut, R &vOut) {  // :1801-1852 (test part)
    V3<R> v0 = v3<R>(tri.v0), v1 = v3<R>(tri.v1), v2 = v3<R>(tri.v2);
    V3<R> edge1 = v1 - v0;
    V3<R> edge2 = v2 - v0;
    V3<R> p = cross(ray.direction, edge2);
    R determinant = dot(edge1, p);
    if (std::fabs(determinant) < R(1e-7)) return false;
    R invDeterminant = R(1) / determinant;
    V3<R> s = ray.origin - v0;
    R u = invDeterminant * dot(s, p);
    if (u < R(0) || u > R(1)) return false;
    V3<R> q = cross(s, edge1);
    R v = invDeterminant * dot(ray.direction, q);
    if (v < R(0) || u + v > R(1)) return false;
    R candidateT = invDeterminant * dot(edge2, q);
    if (candidateT <= R(0.0005)) return false;
    tOut = candidateT;
    uOut = u;
    vOut = v;
    return true;
}
// The watertight test of this build's DR_INTERSECTION_WATERTIGHT mode (no
// reference counterpart): Woop, Benthin and Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013, backface culling off, with the edge
// functions re-evaluated in double when one is exactly zero; the reference's
// t > 0.0005 acceptance and u, v = weights of v1, v2.
template <class R> R comp3(const V3<R> &a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
template <class R>
bool intersectTriangleWatertight(const Ray<R> &ray, const or_triangle &tri, R &tOut, R &uOut, R &vOut) {
    const V3<R> &d = ray.direction;
    const R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    const int kz = ay > ax ? (az > ay ? 2 : 1) : (az > ax ? 2 : 0);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    const R dz = comp3(d, kz);
    if (dz < R(0)) std::swap(kx, ky);
    const R Sx = comp3(d, kx) / dz, Sy = comp3(d, ky) / dz, Sz = R(1) / dz;
    const V3<R> A = v3<R>(tri.v0) - ray.origin, B = v3<R>(tri.v1) - ray.origin, C = v3<R>(tri.v2) - ray.origin;
    const R Az = comp3(A, kz), Bz = comp3(B, kz), Cz = comp3(C, kz);
    const R Ax = comp3(A, kx) - Sx * Az, Ay = comp3(A, ky) - Sy * Az;
    const R Bx = comp3(B, kx) - Sx * Bz, By = comp3(B, ky) - Sy * Bz;
    const R Cx = comp3(C, kx) - Sx * Cz, Cy = comp3(C, ky) - Sy * Cz;
    R U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == R(0) || V == R(0) || W == R(0)) {
        U = R(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = R(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = R(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    const R det = U + V + W;
    if (det == R(0)) return false;
    const R T = U * (Sz * Az) + V * (Sz * Bz) + W * (Sz * Cz);
    const R rcp = R(1) / det;
    const R t = T * rcp;
    if (!(t > R(0.0005))) return false;
    tOut = t;
    uOut = V * rcp;
    vOut = W * rcp;
    return true;
}

// Attribute part of intersectTriangle (:1842-1850), evaluated for the winner.
template <class R> void fillTriangleHit(const Ray<R> &ray, const or_triangle &tri, uint32_t index, R t, R u, R v, Hit<R> &h) {
    R w = R(1) - u - v;
    V3<R> normal = normalize(w * v3<R>(tri.n0) + u * v3<R>(tri.n1) + v * v3<R>(tri.n2));
    h.frontFacing = dot(normal, ray.direction) <= R(0);
    if (dot(normal, ray.direction) > R(0)) normal = -normal;
    h.hit = true;
    h.t = t;
    h.position = ray.origin + t * ray.direction;
    h.normal = normal;
    h.uv = V2<R>{w * R(tri.uv0.x) + u * R(tri.uv1.x) + v * R(tri.uv2.x), w * R(tri.uv0.y) + u * R(tri.uv1.y) + v * R(tri.uv2.y)};
    h.tangent = normalize(v3<R>(tri.tangent));
    h.bitangent = normalize(v3<R>(tri.bitangent));
    h.materialID = tri.materialID;
    h.materialID2 = tri.materialID;
    h.materialBlend = R(0);
    h.objectID = tri.objectID;
    h.primitiveID = index;
}

struct Node {  // GPUAccelerationNode (PathTrace.metal:52-56), 48 B
    float bmin[4];
    float bmax[4];
    uint32_t left, right, firstPrim, primCount;
};
static_assert(sizeof(Node) == 48, "node layout");

template <class R> bool intersectBounds(const Ray<R> &ray, V3<R> inv, const Node &n, R closestT) {  // :1854-1863
    V3<R> bmin(R(n.bmin[0]), R(n.bmin[1]), R(n.bmin[2]));
    V3<R> bmax(R(n.bmax[0]), R(n.bmax[1]), R(n.bmax[2]));
    V3<R> t0 = (bmin - ray.origin) * inv;
    V3<R> t1 = (bmax - ray.origin) * inv;
    V3<R> tn = vmin(t0, t1);
    V3<R> tf = vmax(t0, t1);
    R tNear = fmx(fmx(tn.x, tn.y), fmx(tn.z, R(0)));
    R tFar = fmn(fmn(tf.x, tf.y), fmn(tf.z, closestT));
    return tNear <= tFar;
}

#include "oracle_sdf.h"

} // namespace

// ------------------------------------------------------------------ scene --
struct or_scene {
    std::vector<or_triangle> triangles;
    std::vector<or_material> materials;
    std::vector<or_light> lights;
    std::vector<or_env_sample> env;
    std::vector<float> envAliasProb;      // alias table of the environment CDF (this build's sampling option)
    std::vector<uint32_t> envAliasIndex;
    std::vector<or_texture_desc> textures;
    std::vector<or_float4> texels;
    std::vector<Node> nodes;
    std::vector<uint32_t> primIndices;
    std::vector<or_volume> volumes;
    std::vector<or_brick> bricks;
    std::vector<or_brick_sample> brickSamples;
    std::vector<or_volume_sample> volumeSamples;
    std::vector<or_brick_grid> brickGrids;
    std::vector<uint32_t> brickGridIndices;
};

namespace {

struct AABBf {
    float mn[3], mx[3];
};

// BVHBuilder.buildNode (BVHBuilder.swift:33-85): pre-order node emission,
// count-median split on the largest centroid axis (ties x >= y >= z,
// :126-134), leaf when count <= maxLeaf. The reference sorts with Swift's
// sort; here ties in centroid order fall back to the triangle index so the
// layout is deterministic.
struct Builder {
    const std::vector<AABBf> &bounds;
    std::vector<uint32_t> &idx;
    std::vector<Node> &nodes;
    uint32_t maxLeaf;

    uint32_t build(uint32_t lo, uint32_t hi) {
        uint32_t nodeIndex = uint32_t(nodes.size());
        AABBf b;
        for (int a = 0; a < 3; ++a) { b.mn[a] = std::numeric_limits<float>::max(); b.mx[a] = -std::numeric_limits<float>::max(); }
        AABBf cb = b;
        for (uint32_t i = lo; i < hi; ++i) {
            const AABBf &p = bounds[idx[i]];
            for (int a = 0; a < 3; ++a) {
                b.mn[a] = std::fmin(b.mn[a], p.mn[a]);
                b.mx[a] = std::fmax(b.mx[a], p.mx[a]);
                b.mn[a] = std::fmin(b.mn[a], p.mx[a]);  // include(bounds) includes both corners
                b.mx[a] = std::fmax(b.mx[a], p.mn[a]);
                float c = (p.mn[a] + p.mx[a]) * 0.5f;
                cb.mn[a] = std::fmin(cb.mn[a], c);
                cb.mx[a] = std::fmax(cb.mx[a], c);
            }
        }
        Node n;
        for (int a = 0; a < 3; ++a) { n.bmin[a] = b.mn[a]; n.bmax[a] = b.mx[a]; }
        n.bmin[3] = n.bmax[3] = 0.f;
        n.left = n.right = 0;
        n.firstPrim = lo;
        n.primCount = hi - lo;
        nodes.push_back(n);
        uint32_t count = hi - lo;
        if (count <= std::max<uint32_t>(1u, maxLeaf)) return nodeIndex;
        float ex = cb.mx[0] - cb.mn[0], ey = cb.mx[1] - cb.mn[1], ez = cb.mx[2] - cb.mn[2];
        int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
        std::sort(idx.begin() + lo, idx.begin() + hi, [&](uint32_t l, uint32_t r) {
            float cl = (bounds[l].mn[axis] + bounds[l].mx[axis]) * 0.5f;
            float cr = (bounds[r].mn[axis] + bounds[r].mx[axis]) * 0.5f;
            if (cl != cr) return cl < cr;
            return l < r;
        });
        uint32_t mid = lo + count / 2;
        uint32_t left = build(lo, mid);
        uint32_t right = build(mid, hi);
        Node &me = nodes[nodeIndex];
        me.left = left;
        me.right = right;
        me.firstPrim = 0;  // FlatBVH stores max(0, -1) = 0 for interior nodes
        me.primCount = 0;
        return nodeIndex;
    }
};

// --------------------------------------------------------------- context --
template <class R> struct Ctx {
    const or_scene *s;
    const or_settings *st;
    const or_camera *cam;
    const or_camera *prevCam;
    int traversal;
    uint64_t cnt[OR_CNT_COUNT];
    SdfScene sdf;
};

template <class R> void initCtx(Ctx<R> &c) {
    const or_scene &s = *c.s;
    c.sdf.volumes = s.volumes.data();
    c.sdf.volumeCount = uint32_t(s.volumes.size());
    c.sdf.bricks = s.bricks.data();
    c.sdf.brickCount = uint32_t(s.bricks.size());
    c.sdf.dense = s.volumeSamples.data();
    c.sdf.denseCount = uint32_t(s.volumeSamples.size());
    c.sdf.samples = s.brickSamples.data();
    c.sdf.sampleCount = uint32_t(s.brickSamples.size());
    c.sdf.grids = s.brickGrids.data();
    c.sdf.gridCount = uint32_t(s.brickGrids.size());
    c.sdf.gridIndices = s.brickGridIndices.empty() ? nullptr : s.brickGridIndices.data();
    c.sdf.gridIndexCount = uint32_t(s.brickGridIndices.size());
    c.sdf.renderQuality = c.st->renderQuality;  // constants.renderQuality (RenderSettings.swift:106-115)
}

// intersectScene (PathTrace.metal:3294-3421, triangle part) and
// intersectSceneLinear (:3265-3292), with the lowest-index tie-break.
template <class R> Hit<R> intersectScene(Ctx<R> &c, const Ray<R> &ray) {
    const or_scene &s = *c.s;
    const uint32_t triCount = uint32_t(s.triangles.size());
    R bestT = kInf<R>(), bestU = 0, bestV = 0;
    uint32_t best = 0xffffffffu;
    auto consider = [&](uint32_t ti) {
        R t, u, v;
        c.cnt[OR_CNT_TRIANGLE_TESTS]++;
        const bool hit = c.st->intersection == 1u ? intersectTriangleWatertight(ray, s.triangles[ti], t, u, v)
                                                  : intersectTriangle(ray, s.triangles[ti], t, u, v);
        if (hit && (t < bestT || (t == bestT && ti < best))) {
            bestT = t; bestU = u; bestV = v; best = ti;
        }
    };
    if (c.traversal == 1 || s.nodes.empty()) {
        for (uint32_t i = 0; i < triCount; ++i) consider(i);
    } else {
        V3<R> inv = V3<R>(1) / ray.direction;
        uint32_t stack[64];
        uint32_t sp = 0;
        stack[sp++] = 0;
        const uint32_t nodeCount = uint32_t(s.nodes.size());
        while (sp > 0) {
            uint32_t ni = stack[--sp];
            if (ni >= nodeCount) continue;
            const Node &n = s.nodes[ni];
            c.cnt[OR_CNT_NODE_VISITS]++;
            if (!intersectBounds(ray, inv, n, bestT)) continue;
            if (n.primCount > 0) {
                for (uint32_t o = 0; o < n.primCount; ++o) {
                    uint32_t ti = s.primIndices[n.firstPrim + o];
                    if (ti >= triCount) continue;
                    consider(ti);
                }
            } else if (sp + 2u <= 64u) {
                stack[sp++] = n.right;
                stack[sp++] = n.left;
            }
        }
    }
    Hit<R> h = emptyHit<R>();
    if (best != 0xffffffffu) fillTriangleHit(ray, s.triangles[best], best, bestT, bestU, bestV, h);
    if (!s.volumes.empty()) {
        // closestVolumeHit after the triangles, limited by their closest t (:3397-3418).
        VolHit<R> vh = closestVolumeHit(c.sdf, ray, h.t, c.cnt + OR_CNT_SDF_BASE);
        if (vh.hit) {
            h.hit = true;
            h.t = vh.t;
            h.position = vh.position;
            h.normal = vh.normal;
            h.uv = V2<R>{R(0), R(0)};
            V3<R> helper = std::fabs(vh.normal.y) < R(0.999) ? V3<R>(0, 1, 0) : V3<R>(1, 0, 0);
            h.tangent = normalize(cross(helper, vh.normal));
            h.bitangent = normalize(cross(vh.normal, h.tangent));
            h.materialID = vh.materialA;
            h.materialID2 = vh.materialB;
            h.materialBlend = vh.blend;
            h.objectID = vh.objectID;
            h.primitiveID = vh.primitiveID;
            h.frontFacing = vh.frontFacing;
            if (vh.fields) {
                h.fieldBaseColorOpacity = v4<R>(vh.fields->baseColorOpacity);
                h.fieldEmissionTransmission = v4<R>(vh.fields->emissionTransmission);
                h.fieldSurface = v4<R>(vh.fields->surface);
                h.fieldFlags = vh.fields->materialFieldFlags[0];
            }
        }
    }
    return h;
}

template <class R> Mat<R> blendMaterials(Mat<R> a, const Mat<R> &b, R blend) {  // :549-569
    R t = clampv(blend, R(0), R(1));
    V4<R> *pa = &a.baseColor;
    const V4<R> *pb = &b.baseColor;
    for (int k = 0; k < 17; ++k) pa[k] = mix4(pa[k], pb[k], t);
    return a;
}

// applyVolumeMaterialFields (:571-601); flag bits volumeMaterialField*
// (:255-262).
template <class R> void applyVolumeMaterialFields(Mat<R> &m, const Hit<R> &h) {
    const uint32_t f = h.fieldFlags;
    if (f == 0u) return;
    const V4<R> &bc = h.fieldBaseColorOpacity, &et = h.fieldEmissionTransmission, &su = h.fieldSurface;
    if (f & 1u) m.baseColor.x = fmx(bc.x, R(0)), m.baseColor.y = fmx(bc.y, R(0)), m.baseColor.z = fmx(bc.z, R(0));
    if (f & 2u) m.baseColor.w = clampv(bc.w, R(0), R(1));
    if (f & 4u) m.emission.x = fmx(et.x, R(0)), m.emission.y = fmx(et.y, R(0)), m.emission.z = fmx(et.z, R(0));
    if ((f & 4u) && (f & 128u)) {
        R s = fmx(su.w, R(0));
        m.emission.x = fmx(et.x, R(0)) * s, m.emission.y = fmx(et.y, R(0)) * s, m.emission.z = fmx(et.z, R(0)) * s;
    } else if (f & 128u) {
        R s = fmx(su.w, R(0));
        m.emission.x = fmx(m.emission.x, R(0)) * s, m.emission.y = fmx(m.emission.y, R(0)) * s,
        m.emission.z = fmx(m.emission.z, R(0)) * s;
    }
    if (f & 32u) m.emission.w = clampv(et.w, R(0), R(1));
    if (f & 8u) m.parameters.x = clampv(su.x, R(0.02), R(1));
    if (f & 16u) m.parameters.y = clampv(su.y, R(0), R(1));
    if (f & 64u) m.parameters2.x = clampv(su.z, R(0), R(1));
}

template <class R> Mat<R> materialForHit(const Ctx<R> &c, const Hit<R> &h) {  // :1096-1137
    uint32_t count = uint32_t(c.s->materials.size());
    uint32_t a = std::min<uint32_t>(h.materialID, count - 1u);
    Mat<R> m = toMat<R>(c.s->materials[a]);
    if (h.materialBlend > R(0)) {
        uint32_t b = std::min<uint32_t>(h.materialID2, count - 1u);
        m = blendMaterials(m, toMat<R>(c.s->materials[b]), h.materialBlend);
    }
    applyVolumeMaterialFields(m, h);
    return m;  // no material programs on this path
}

template <class R> V4<R> texel(const or_scene &s, uint32_t i) { return v4<R>(s.texels[i]); }

template <class R> V4<R> sampleTexture(const or_scene &s, uint32_t indexPlusOne, V2<R> uv, V4<R> fallback) {  // :3723-3763
    if (indexPlusOne == 0u || indexPlusOne > s.textures.size()) return fallback;
    const or_texture_desc &d = s.textures[indexPlusOne - 1u];
    uint32_t offset = d.offset, width = d.width, height = d.height, mode = d.mode;
    if (width == 0u || height == 0u) return fallback;
    V2<R> wrapped{fractv(uv.x), fractv(uv.y)};
    if (mode == 1u && width > 1u && height > 1u) {
        R px = wrapped.x * R(width) - R(0.5), py = wrapped.y * R(height) - R(0.5);
        R bx = std::floor(px), by = std::floor(py);
        R fx = fractv(px), fy = fractv(py);
        uint32_t x0 = uint32_t(std::fmod(bx + R(width), R(width)));
        uint32_t y0 = uint32_t(std::fmod(by + R(height), R(height)));
        uint32_t x1 = (x0 + 1u) % width, y1 = (y0 + 1u) % height;
        V4<R> c00 = texel<R>(s, offset + y0 * width + x0), c10 = texel<R>(s, offset + y0 * width + x1);
        V4<R> c01 = texel<R>(s, offset + y1 * width + x0), c11 = texel<R>(s, offset + y1 * width + x1);
        return mix4(mix4(c00, c10, fx), mix4(c01, c11, fx), fy);
    }
    uint32_t x = std::min<uint32_t>(uint32_t(wrapped.x * R(width)), width - 1u);
    uint32_t y = std::min<uint32_t>(uint32_t(wrapped.y * R(height)), height - 1u);
    return texel<R>(s, offset + y * width + x);
}

template <class R> V3<R> fallbackSky(V3<R> d) {  // :3765-3768
    R sky = R(0.5) * (d.y + R(1));
    return mixv(V3<R>(R(0.02), R(0.025), R(0.035)), V3<R>(R(0.25), R(0.32), R(0.45)), sky);
}

template <class R> V3<R> sampleEnvironment(const Ctx<R> &c, V3<R> d) {  // :3770-3803
    const or_settings &st = *c.st;
    R intensity = fmx(R(st.environmentIntensity), R(0));
    if (st.environmentTextureIndexPlusOne == 0u) return fallbackSky(d) * intensity;
    R s = std::sin(R(st.environmentRotationY));
    R co = std::cos(R(st.environmentRotationY));
    V3<R> r = normalize(V3<R>(co * d.x - s * d.z, d.y, s * d.x + co * d.z));
    R u = std::atan2(r.z, r.x) * R(0.15915494309) + R(0.5);
    R v = std::acos(clampv(r.y, R(-1), R(1))) * R(0.31830988618);
    v = clampv(v, R(0.00001), R(0.99999));
    V3<R> sky = fallbackSky(d);
    V3<R> color = sampleTexture(*c.s, st.environmentTextureIndexPlusOne, V2<R>{u, v}, V4<R>{sky.x, sky.y, sky.z, R(1)}).xyz() * intensity;
    R maxRadiance = fmx(R(st.environmentMaxRadiance), R(0));
    if (maxRadiance > R(0)) color = vmin(color, V3<R>(maxRadiance));
    return color;
}

template <class R> uint32_t environmentPixelIndexForDirection(const Ctx<R> &c, V3<R> d) {  // :3882-3910
    const or_settings &st = *c.st;
    if (st.environmentTextureIndexPlusOne == 0u || st.environmentTextureIndexPlusOne > c.s->textures.size()) return 0u;
    const or_texture_desc &td = c.s->textures[st.environmentTextureIndexPlusOne - 1u];
    uint32_t width = td.width, height = td.height;
    if (width == 0u || height == 0u) return 0u;
    R s = std::sin(R(st.environmentRotationY));
    R co = std::cos(R(st.environmentRotationY));
    V3<R> r = normalize(V3<R>(co * d.x - s * d.z, d.y, s * d.x + co * d.z));
    R u = fractv(std::atan2(r.z, r.x) * R(0.15915494309) + R(0.5));
    R v = clampv(std::acos(clampv(r.y, R(-1), R(1))) * R(0.31830988618), R(0), R(0.99999994f));
    uint32_t x = std::min<uint32_t>(uint32_t(u * R(width)), width - 1u);
    uint32_t y = std::min<uint32_t>(uint32_t(v * R(height)), height - 1u);
    return y * width + x;
}

template <class R> R environmentPDF(const Ctx<R> &c, V3<R> d) {  // :3912-3927
    uint32_t count = uint32_t(c.s->env.size());
    if (count == 0u) return R(0);
    uint32_t index = environmentPixelIndexForDirection(c, d);
    if (index >= count) return R(0);
    return fmx(R(c.s->env[index].pdf), R(0));
}

template <class R> uint32_t selectEnvironmentSampleIndex(const or_env_sample *e, uint32_t count, R sample) {  // :3929-3950
    if (count <= 1u) return 0u;
    R target = clampv(sample, R(0), R(0.99999994f));
    uint32_t low = 0u, high = count - 1u;
    while (low < high) {
        uint32_t mid = (low + high) >> 1u;
        if (target <= R(e[mid].cdf)) high = mid; else low = mid + 1u;
    }
    return low;
}

// Walker's alias method with Vose's construction (M. D. Vose, "A linear
// algorithm for generating random numbers with a given distribution", IEEE
// TSE 17(9), 1991) over the texel probabilities cdf[i] - cdf[i-1]: worklists
// of under- and over-full texels in index order, processed from the back,
// in double; leftovers get probability 1. The same construction as the
// library's, so the tables agree bit for bit.
void buildEnvironmentAlias(const std::vector<or_env_sample> &env, std::vector<float> &prob, std::vector<uint32_t> &alias) {
    const size_t n = env.size();
    prob.assign(n, 1.0f);
    alias.resize(n);
    std::vector<double> scaled(n);
    std::vector<uint32_t> small, large;
    double prev = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double c = double(env[i].cdf);
        scaled[i] = std::max(c - prev, 0.0) * double(n);
        prev = std::max(prev, c);
        alias[i] = uint32_t(i);
        (scaled[i] < 1.0 ? small : large).push_back(uint32_t(i));
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t l = small.back(), g = large.back();
        small.pop_back();
        large.pop_back();
        prob[l] = float(scaled[l]);
        alias[l] = g;
        scaled[g] = (scaled[g] + scaled[l]) - 1.0;
        (scaled[g] < 1.0 ? small : large).push_back(g);
    }
}

template <class R> V3<R> sampleEnvironmentDirection(const Ctx<R> &c, R px, R py, R cdfSample, R coin, R &pdf) {  // :3952-4003
    pdf = R(0);
    const or_settings &st = *c.st;
    uint32_t count = uint32_t(c.s->env.size());
    if (st.environmentTextureIndexPlusOne == 0u || count == 0u || st.environmentTextureIndexPlusOne > c.s->textures.size())
        return V3<R>(0, 1, 0);
    const or_texture_desc &td = c.s->textures[st.environmentTextureIndexPlusOne - 1u];
    uint32_t width = td.width, height = td.height;
    if (width == 0u || height == 0u) return V3<R>(0, 1, 0);
    uint32_t index;
    if (st.environmentSampling == 1u && !c.s->envAliasProb.empty()) {
        // Alias table: a column from the CDF sample, the coin picks it or its alias.
        index = std::min(uint32_t(float(cdfSample) * float(count)), count - 1u);
        if (!(float(coin) < c.s->envAliasProb[index])) index = c.s->envAliasIndex[index];
    } else {
        index = selectEnvironmentSampleIndex(c.s->env.data(), count, cdfSample);
    }
    index = std::min(index, count - 1u);
    uint32_t x = index % width, y = index / width;
    R u = (R(x) + px) / R(width);
    R v = (R(y) + py) / R(height);
    R phi = (u - R(0.5)) * R(6.28318530718);
    R theta = clampv(v, R(0.00001), R(0.99999)) * R(3.14159265359);
    R sinTheta = std::sin(theta);
    V3<R> local = normalize(V3<R>(std::cos(phi) * sinTheta, std::cos(theta), std::sin(phi) * sinTheta));
    R s = std::sin(R(st.environmentRotationY));
    R co = std::cos(R(st.environmentRotationY));
    pdf = fmx(R(c.s->env[index].pdf), R(0));
    return normalize(V3<R>(co * local.x + s * local.z, local.y, -s * local.x + co * local.z));
}

template <class R> void applyBaseColorTexture(const Ctx<R> &c, Mat<R> &m, const Hit<R> &h) {  // :4005-4019
    uint32_t ti = uint32_t(m.parameters.z);
    if (ti == 0u) return;
    V4<R> t = sampleTexture(*c.s, ti, h.uv, V4<R>{1, 1, 1, 1});
    m.baseColor = V4<R>{m.baseColor.x * t.x, m.baseColor.y * t.y, m.baseColor.z * t.z, m.baseColor.w * t.w};
}

template <class R> void applyNormalMap(const Ctx<R> &c, Hit<R> &h, const Mat<R> &m, V3<R> rayDirection) {  // :4021-4054
    uint32_t ti = uint32_t(m.parameters.w);
    if (ti == 0u) return;
    V3<R> mapped = sampleTexture(*c.s, ti, h.uv, V4<R>{R(0.5), R(0.5), R(1), R(1)}).xyz();
    mapped = normalize(mapped * R(2) - V3<R>(1));
    V3<R> tangent = h.tangent - h.normal * dot(h.tangent, h.normal);
    if (length(tangent) <= R(1e-5)) {
        V3<R> helper = std::fabs(h.normal.y) < R(0.999) ? V3<R>(0, 1, 0) : V3<R>(1, 0, 0);
        tangent = cross(helper, h.normal);
    }
    tangent = normalize(tangent);
    V3<R> bitangent = h.bitangent - h.normal * dot(h.bitangent, h.normal);
    if (length(bitangent) <= R(1e-5)) bitangent = cross(h.normal, tangent);
    bitangent = normalize(bitangent);
    h.normal = normalize(mapped.x * tangent + mapped.y * bitangent + mapped.z * h.normal);
    if (dot(h.normal, rayDirection) > R(0)) h.normal = -h.normal;
}

template <class R> Ray<R> makeCameraRay(const or_camera &cam, uint32_t width, uint32_t height, uint32_t gx, uint32_t gy,
                                        R jx, R jy, R lx, R ly) {  // :4056-4091
    R u = (R(float(gx)) + jx) / R(float(width));
    R v = R(1) - ((R(float(gy)) + jy) / R(float(height)));
    V3<R> ll = v3<R>(cam.lowerLeft), hor = v3<R>(cam.horizontal), ver = v3<R>(cam.vertical);
    Ray<R> ray;
    V3<R> planePoint = ll + u * hor + v * ver;
    if (R(cam.origin.w) > R(0.5)) {
        ray.origin = planePoint;
        ray.direction = normalize(cross(ver, hor));
    } else {
        ray.origin = v3<R>(cam.origin);
        ray.direction = normalize(planePoint - ray.origin);
    }
    R aperture = fmx(R(cam.lens.x), R(0));
    if (R(cam.origin.w) <= R(0.5) && aperture > R(0)) {
        V3<R> forward = normalize(cross(ver, hor));
        R focusDistance = fmx(R(cam.lens.y), R(0.0001));
        R focalT = focusDistance / fmx(dot(ray.direction, forward), R(0.0001));
        V3<R> focalPoint = ray.origin + ray.direction * focalT;
        V2<R> disk = uniformDisk(lx, ly);
        disk.x *= aperture;
        disk.y *= aperture;
        V3<R> right = normalize(hor);
        V3<R> trueUp = normalize(ver);
        ray.origin += right * disk.x + trueUp * disk.y;
        ray.direction = normalize(focalPoint - ray.origin);
    }
    return ray;
}

template <class R> bool tracePrimaryAOV(Ctx<R> &c, Ray<R> ray, Hit<R> &ph, Mat<R> &pm) {  // :4093-4160
    for (uint32_t pass = 0; pass < 32u; ++pass) {
        c.cnt[OR_CNT_PRIMARY_QUERIES]++;
        Hit<R> h = intersectScene(c, ray);
        if (!h.hit) return false;
        Mat<R> m = materialForHit(c, h);
        applyNormalMap(c, h, m, ray.direction);
        applyBaseColorTexture(c, m, h);
        if (m.baseColor.w <= R(0.001)) {
            ray.origin = h.position + ray.direction * R(0.001);
            continue;
        }
        ph = h;
        pm = m;
        return true;
    }
    return false;
}

template <class R> bool shadowOccluded(Ctx<R> &c, Ray<R> ray, R maxDistance, V3<R> &transmittance) {  // :4196-4274
    R traveled = R(0);
    for (uint32_t step = 0; step < 8u; ++step) {
        c.cnt[OR_CNT_SHADOW_QUERIES]++;
        Hit<R> h = intersectScene(c, ray);
        if (!h.hit || traveled + h.t >= maxDistance) return false;
        Mat<R> m = materialForHit(c, h);
        if (m.baseColor.w <= R(0.001)) {
            traveled += h.t;
            ray.origin = h.position + ray.direction * R(0.001);
            continue;
        }
        R transmission = materialTransmission(m);
        if (transmission <= R(0.001)) return true;
        bool insideSolid = !h.frontFacing && !materialThinWalled(m);
        V3<R> medium = materialHasVolumeScattering(m) && insideSolid ? materialVolumeTransmittance(m, h.t)
                                                                      : materialTransmissionAbsorption(m, h.t, insideSolid);
        transmittance = transmittance * materialTransmissionTint(m, R(0.35)) * medium * transmission;
        if (maxComponent(transmittance) < R(0.02)) return true;
        traveled += h.t;
        ray.origin = h.position + ray.direction * R(0.001);
    }
    return true;
}

template <class R> R lightPDFForHit(const Ctx<R> &c, const Hit<R> &h, V3<R> incoming) {  // :1580-1613
    const or_scene &s = *c.s;
    if (h.primitiveID >= s.triangles.size()) return R(0);
    uint32_t lp1 = s.triangles[h.primitiveID].padding2;
    if (lp1 == 0u) return R(0);
    uint32_t index = lp1 - 1u;
    if (index >= s.lights.size()) return R(0);
    const or_light &l = s.lights[index];
    if (l.triangleIndex != h.primitiveID || l.area <= 0.f) return R(0);
    R cosLight = fmx(R(0), dot(v3<R>(l.normal), -incoming));
    if (cosLight <= R(0)) return R(0);
    return lightSelectionPDF<R>(s.lights.data(), index) * lightSolidAnglePDF(h.t * h.t, cosLight, R(l.area));
}

template <class R> V3<R> sampleDirectLighting(Ctx<R> &c, const Hit<R> &h, const Mat<R> &m, V3<R> view, uint32_t &state) {  // :4407-4585
    const or_scene &s = *c.s;
    V3<R> direct(0);
    const uint32_t lightCount = uint32_t(s.lights.size());
    if (lightCount > 0u) {
        R r0 = randomFloat<R>(state);
        uint32_t index = selectLightIndex<R>(s.lights.data(), lightCount, r0);
        const or_light &light = s.lights[index];
        R selectionPDF = lightSelectionPDF<R>(s.lights.data(), index);
        uint32_t lti = light.triangleIndex;
        if (lti < s.triangles.size() && selectionPDF > R(0)) {
            const or_triangle &lt = s.triangles[lti];
            Mat<R> lm = toMat<R>(s.materials[std::min<uint32_t>(light.materialIndex, uint32_t(s.materials.size()) - 1u)]);
            V3<R> emission = lm.emission.xyz();
            R area = R(light.area);
            if (area > R(0)) {
                R ux = randomFloat<R>(state);
                R uy = randomFloat<R>(state);
                V3<R> lightPoint = sampleTriangle<R>(lt, ux, uy);
                V3<R> toLight = lightPoint - h.position;
                R distanceSquared = dot(toLight, toLight);
                R distanceToLight = std::sqrt(distanceSquared);
                V3<R> lightDirection = toLight / distanceToLight;
                R cosSurface = fmx(R(0), dot(h.normal, lightDirection));
                V3<R> lightNormal = v3<R>(light.normal);
                R cosLight = fmx(R(0), dot(lightNormal, -lightDirection));
                if (cosSurface > R(0) && cosLight > R(0)) {
                    R lightPDF = selectionPDF * lightSolidAnglePDF(distanceSquared, cosLight, area);
                    R bsdfPDF = materialBSDFPDF(m, h.normal, h.tangent, h.bitangent, view, lightDirection);
                    R mis = powerHeuristic(lightPDF, bsdfPDF);
                    Ray<R> sr{h.position + h.normal * R(0.001), lightDirection};
                    V3<R> T(1);
                    if (!shadowOccluded(c, sr, distanceToLight - R(0.002), T)) {
                        R geometryTerm = cosSurface * cosLight * area / fmx(distanceSquared, R(1e-6));
                        V3<R> brdf = evaluateMaterialBRDF(m, h.normal, h.tangent, h.bitangent, view, lightDirection);
                        direct += brdf * emission * geometryTerm * mis * T / selectionPDF;
                    }
                }
            }
        }
    }
    for (uint32_t e = 0; e < 2u; ++e) {
        R px = randomFloat<R>(state);
        R py = randomFloat<R>(state);
        R cdf = randomFloat<R>(state);
        R coin = c.st->environmentSampling == 1u ? randomFloat<R>(state) : R(0);  // alias mode draws one more
        R envPDF;
        V3<R> dir = sampleEnvironmentDirection(c, px, py, cdf, coin, envPDF);
        R cosS = fmx(R(0), dot(h.normal, dir));
        if (envPDF > R(0) && cosS > R(0)) {
            R bsdfPDF = materialBSDFPDF(m, h.normal, h.tangent, h.bitangent, view, dir);
            R mis = powerHeuristic(envPDF, bsdfPDF);
            Ray<R> sr{h.position + h.normal * R(0.001), dir};
            V3<R> T(1);
            if (!shadowOccluded(c, sr, R(1.0e20f), T)) {
                V3<R> brdf = evaluateMaterialBRDF(m, h.normal, h.tangent, h.bitangent, view, dir);
                V3<R> env = sampleEnvironment(c, dir);
                direct += brdf * env * cosS * mis * T / (envPDF * R(2));
            }
        }
    }
    return direct;
}


template <class R>
bool continueSubsurfaceRandomWalk(Ctx<R> &c, const Hit<R> &entry, const Mat<R> &m, Ray<R> &ray, V3<R> &T, R &prevPdf,
                                  uint32_t &state) {  // :3525-3621
    R subsurface = materialSubsurface(m);
    if (subsurface <= R(0) || !entry.frontFacing) return false;
    R e0 = randomFloat<R>(state);
    R e1 = randomFloat<R>(state);
    V3<R> entryDirection = orientHemisphere(cosineHemisphere(e0, e1), -entry.normal);
    Ray<R> medium{entry.position + entryDirection * R(0.001), entryDirection};
    V3<R> sigmaT = subsurfaceSigmaT(m);
    V3<R> albedo = materialSubsurfaceColor(m);
    R g = materialSubsurfaceAnisotropy(m);
    for (uint32_t step = 0u; step < 8u; ++step) {
        c.cnt[OR_CNT_BOUNCE_QUERIES]++;
        Hit<R> b = intersectScene(c, medium);
        if (!b.hit) return false;
        R f0 = randomFloat<R>(state);
        R f1 = randomFloat<R>(state);
        R freeFlight = sampleSubsurfaceFreeFlight(sigmaT, f0, f1);
        if (freeFlight < b.t) {
            medium.origin += medium.direction * freeFlight;
            R h0 = randomFloat<R>(state);
            R h1 = randomFloat<R>(state);
            medium.direction = sampleHenyeyGreenstein(h0, h1, medium.direction, g);
            T *= subsurfaceCollisionWeight(sigmaT, albedo, freeFlight);
            if (maxComponent(T) <= R(1e-4)) return false;
            continue;
        }
        V3<R> exitNormal = b.frontFacing ? b.normal : -b.normal;
        exitNormal = normalize(exitNormal);
        R x0 = randomFloat<R>(state);
        R x1 = randomFloat<R>(state);
        V3<R> local = cosineHemisphere(x0, x1);
        ray.direction = orientHemisphere(local, exitNormal);
        ray.origin = b.position + exitNormal * R(0.001);
        T *= subsurfaceBoundaryWeight(sigmaT, b.t);
        prevPdf = fmx(dot(exitNormal, ray.direction), R(0)) * R(0.31830988618);
        return true;
    }
    return false;
}


// One progressive sample of one pixel: pathTraceKernel (PathTrace.metal:5167-5586).
template <class R>
void pathTracePixel(Ctx<R> &c, uint32_t px, uint32_t py, uint32_t sampleIndex, R *acc, R *aov, size_t planeStride) {
    const or_settings &st = *c.st;
    const uint32_t width = st.width, height = st.height;
    uint32_t state = st.frameSeed ^ hashU32(px + py * width) ^ hashU32(sampleIndex + 1u);
    R jx = randomFloat<R>(state);
    R jy = randomFloat<R>(state);
    R lx = randomFloat<R>(state);
    R ly = randomFloat<R>(state);
    Ray<R> ray = makeCameraRay<R>(*c.cam, width, height, px, py, jx, jy, lx, ly);
    V3<R> radiance(0), T(1);
    R outputAlpha = R(1);
    R prevPdf = R(0);
    Hit<R> ph = emptyHit<R>();
    Mat<R> pm = zeroMat<R>();
    const R clampValue = R(st.sampleRadianceClamp);

    for (uint32_t bounce = 0; bounce < st.maxBounces; ++bounce) {
        c.cnt[bounce == 0u ? OR_CNT_PRIMARY_QUERIES : OR_CNT_BOUNCE_QUERIES]++;
        Hit<R> h = intersectScene(c, ray);
        if (!h.hit) {
            if (st.transparentBackground != 0u && !ph.hit) {
                outputAlpha = R(0);
            } else if (st.showsEnvironmentBackground == 0u && !ph.hit) {
                outputAlpha = R(1);
                radiance += vmax(V3<R>(R(st.backgroundColor[0]), R(st.backgroundColor[1]), R(st.backgroundColor[2])), V3<R>(0));
            } else {
                R w = R(1);
                if (bounce > 0u && prevPdf > R(0)) w = powerHeuristic(prevPdf, environmentPDF(c, ray.direction));
                V3<R> contribution = T * sampleEnvironment(c, ray.direction) * w;
                radiance += clampSampleContribution(contribution, clampValue);
            }
            break;
        }
        Mat<R> m = materialForHit(c, h);
        applyNormalMap(c, h, m, ray.direction);
        applyBaseColorTexture(c, m, h);
        if (m.baseColor.w <= R(0.001)) {
            ray.origin = h.position + ray.direction * R(0.001);
            continue;
        }
        if (!ph.hit) {
            ph = h;
            pm = m;
        }
        if (continueTransmissionVolumeRandomWalk(h, m, ray, T, prevPdf, state)) {
            if (terminatePathWithRussianRoulette(bounce, T, state)) break;
            continue;
        }
        R subsurface = materialSubsurface(m);
        if (subsurface > R(0)) {
            R gate = randomFloat<R>(state);
            if (gate < subsurface && continueSubsurfaceRandomWalk(c, h, m, ray, T, prevPdf, state)) {
                if (terminatePathWithRussianRoulette(bounce, T, state)) break;
                continue;
            }
        }
        if (continueTransmissiveSurface(h, m, ray, T, prevPdf, state)) {
            if (terminatePathWithRussianRoulette(bounce, T, state)) break;
            continue;
        }
        if (maxComponent(m.emission.xyz()) > R(0)) {
            R w = R(1);
            if (bounce > 0u && prevPdf > R(0)) w = powerHeuristic(prevPdf, lightPDFForHit(c, h, ray.direction));
            V3<R> contribution = T * m.emission.xyz() * w;
            radiance += clampSampleContribution(contribution, clampValue);
            break;
        }
        V3<R> view = -ray.direction;
        V3<R> direct = sampleDirectLighting(c, h, m, view, state);
        radiance += clampSampleContribution(T * direct, clampValue);

        V3<R> next = sampleMaterialBounce(m, h.normal, h.tangent, h.bitangent, ray.direction, state, T);
        prevPdf = materialBSDFPDF(m, h.normal, h.tangent, h.bitangent, view, next);
        ray.origin = h.position + h.normal * R(0.001);
        ray.direction = next;
        if (terminatePathWithRussianRoulette(bounce, T, state)) break;
    }

    // Running-mean accumulation (:5520-5524), mix(prev, L, 1/(s+1)).
    size_t pix = (size_t(py) * width + px) * 4;
    R sampleWeight = R(1) / (R(float(sampleIndex)) + R(1));
    acc[pix + 0] = mixs(acc[pix + 0], radiance.x, sampleWeight);
    acc[pix + 1] = mixs(acc[pix + 1], radiance.y, sampleWeight);
    acc[pix + 2] = mixs(acc[pix + 2], radiance.z, sampleWeight);
    acc[pix + 3] = mixs(acc[pix + 3], outputAlpha, sampleWeight);

    if (!aov) return;
    Hit<R> ah = ph;
    Mat<R> am = pm;
    bool has = ph.hit;
    if (st.denoiserEnabled != 0u) {
        has = tracePrimaryAOV(c, makeCameraRay<R>(*c.cam, width, height, px, py, R(0.5), R(0.5), R(0), R(0)), ah, am);
    }
    R *P[7];
    for (int k = 0; k < 7; ++k) P[k] = aov + planeStride * k + pix;
    auto put = [](R *p, R a, R b, R cc, R d) { p[0] = a; p[1] = b; p[2] = cc; p[3] = d; };
    if (has) {
        V3<R> en = ah.normal * R(0.5) + V3<R>(R(0.5));
        put(P[0], ah.t, ah.t, ah.t, R(1));
        put(P[1], en.x, en.y, en.z, R(1));
        put(P[2], am.baseColor.x, am.baseColor.y, am.baseColor.z, am.baseColor.w);
        R mid = R(float(ah.materialID + 1u));
        R oid = R(float(ah.objectID + 1u));
        put(P[3], mid, mid, mid, R(1));
        put(P[4], oid, oid, oid, R(1));
        V2<R> cur, prv;
        if (projectToScreen(*c.cam, ah.position, width, height, cur) && projectToScreen(*c.prevCam, ah.position, width, height, prv)) {
            put(P[5], prv.x - cur.x, prv.y - cur.y, R(0), R(1));
        } else {
            put(P[5], R(0), R(0), R(0), R(0));
        }
        // Internal triangle-ID plane: the global materialized index + 1
        // (Hit.primitiveID, PathTrace.metal:3383), stored exactly as bits.
        uint32_t tid = ah.primitiveID + 1u;
        R tv;
        if (sizeof(R) == 4) { float f; std::memcpy(&f, &tid, 4); tv = R(f); }
        else { tv = R(double(tid)); }
        put(P[6], tv, tv, tv, tv);
    } else {
        put(P[0], 0, 0, 0, 0);
        put(P[1], R(0.5), R(0.5), R(0.5), R(0));
        put(P[2], 0, 0, 0, 0);
        put(P[3], 0, 0, 0, 0);
        put(P[4], 0, 0, 0, 0);
        put(P[5], 0, 0, 0, 0);
        put(P[6], 0, 0, 0, 0);
    }
}


// The preview and interactive qualities' kernels: flatPreviewKernel
// (PathTrace.metal:4728-4868) and interactiveMaterialKernel (:4870-5165),
// selected by quality in DenrimRenderer.primaryPipeline
// (DenrimRenderer.swift:197-215). Pixel-centre ray; accumulation and AOVs
// are written, not averaged.
template <class R> R smoothstepv(R e0, R e1, R x) {
    R t = clampv((x - e0) / (e1 - e0), R(0), R(1));
    return t * t * (R(3) - R(2) * t);
}
template <class R> V3<R> interactiveDiffuseEnvironment(const Ctx<R> &c, V3<R> n) {  // :3805-3823
    V3<R> up = sampleEnvironment(c, V3<R>(0, 1, 0)), down = sampleEnvironment(c, V3<R>(0, -1, 0));
    V3<R> side = (sampleEnvironment(c, V3<R>(1, 0, 0)) + sampleEnvironment(c, V3<R>(-1, 0, 0)) +
                  sampleEnvironment(c, V3<R>(0, 0, 1)) + sampleEnvironment(c, V3<R>(0, 0, -1))) * R(0.25);
    R upW = smoothstepv(R(-0.15), R(0.85), n.y), downW = smoothstepv(R(0.15), R(-0.85), n.y);
    R sideW = fmx(R(0), R(1) - upW - downW);
    return up * upW + down * downW + side * sideW;
}
template <class R> V3<R> interactiveHemisphereDirection(V3<R> d, V3<R> n) {  // :3825-3831
    R nd = dot(n, d);
    if (nd < R(0)) d = d - n * (R(2) * nd);
    return normalize(d);
}
template <class R> V3<R> interactiveSpecularEnvironment(const Ctx<R> &c, V3<R> r, V3<R> n, R roughness) {  // :3833-3880
    r = interactiveHemisphereDirection(normalize(r), n);
    roughness = clampv(roughness, R(0), R(1));
    R spread = roughness * roughness;
    V3<R> helper = std::fabs(r.y) < R(0.999) ? V3<R>(0, 1, 0) : V3<R>(1, 0, 0);
    V3<R> t = normalize(cross(helper, r)), b = cross(r, t);
    V3<R> cone = sampleEnvironment(c, r) * R(0.42);
    cone = cone + sampleEnvironment(c, interactiveHemisphereDirection(r + t * spread * R(0.95) + n * spread * R(0.18), n)) * R(0.145);
    cone = cone + sampleEnvironment(c, interactiveHemisphereDirection(r - t * spread * R(0.95) + n * spread * R(0.18), n)) * R(0.145);
    cone = cone + sampleEnvironment(c, interactiveHemisphereDirection(r + b * spread * R(0.95) + n * spread * R(0.18), n)) * R(0.145);
    cone = cone + sampleEnvironment(c, interactiveHemisphereDirection(r - b * spread * R(0.95) + n * spread * R(0.18), n)) * R(0.145);
    R w = smoothstepv(R(0.62), R(0.95), roughness);
    return cone + (interactiveDiffuseEnvironment(c, n) - cone) * w;
}
template <class R> R interactiveContactOcclusion(Ctx<R> &c, const Hit<R> &h, const Mat<R> &m, V3<R> n) {  // :3882-3944
    Frame<R> f = makeTangentFrame(n, h.tangent, h.bitangent);
    const V3<R> local[4] = {V3<R>(R(0.000f), R(0.000f), R(1.000f)), V3<R>(R(0.780f), R(0.180f), R(0.600f)),
                            V3<R>(R(-0.420f), R(0.720f), R(0.552f)), V3<R>(R(-0.620f), R(-0.520f), R(0.587f))};
    const R radius = R(0.34f);
    R occ = R(0);
    for (int k = 0; k < 4; ++k) {
        V3<R> l = normalize(local[k]);
        V3<R> dir = normalize(l.x * f.tangent + l.y * f.bitangent + l.z * n);
        Ray<R> ray{h.position + n * R(0.003), dir};
        c.cnt[OR_CNT_SHADOW_QUERIES]++;
        Hit<R> o = intersectScene(c, ray);
        if (!o.hit || o.t >= radius) continue;
        Mat<R> om = materialForHit(c, o);
        R blocker = clampv(om.baseColor.w, R(0), R(1)) * (R(1) - materialTransmission(om) * R(0.72));
        occ += blocker * (R(1) - smoothstepv(R(0), radius, o.t));
    }
    R strength = R(0.72) + (R(0.36) - R(0.72)) * materialTransmission(m);
    return R(1) - clampv((occ / R(4)) * strength, R(0), R(0.72));
}

template <class R>
void previewPixel(Ctx<R> &c, uint32_t px, uint32_t py, bool interactive, R *acc, R *aov, size_t planeStride) {
    const or_settings &st = *c.st;
    const uint32_t width = st.width, height = st.height;
    Ray<R> ray = makeCameraRay<R>(*c.cam, width, height, px, py, R(0.5), R(0.5), R(0), R(0));
    c.cnt[OR_CNT_PRIMARY_QUERIES]++;
    Hit<R> h = intersectScene(c, ray);
    size_t pix = (size_t(py) * width + px) * 4;
    R *P[7] = {nullptr};
    if (aov) for (int k = 0; k < 7; ++k) P[k] = aov + planeStride * k + pix;
    auto put = [](R *p, R a, R b, R cc, R d) { if (p) { p[0] = a; p[1] = b; p[2] = cc; p[3] = d; } };
    if (!h.hit) {
        V3<R> color = vmax(V3<R>(R(st.backgroundColor[0]), R(st.backgroundColor[1]), R(st.backgroundColor[2])), V3<R>(0));
        if (st.transparentBackground == 0u && st.showsEnvironmentBackground != 0u) color = sampleEnvironment(c, ray.direction);
        put(acc + pix, color.x, color.y, color.z, st.transparentBackground != 0u ? R(0) : R(1));
        put(P[0], 0, 0, 0, 0);
        put(P[1], R(0.5), R(0.5), R(0.5), R(0));
        for (int k = 2; k < 7; ++k) put(P[k], 0, 0, 0, 0);
        return;
    }
    Mat<R> m = materialForHit(c, h);
    applyNormalMap(c, h, m, ray.direction);
    applyBaseColorTexture(c, m, h);
    V3<R> n = normalize(h.normal), view = normalize(-ray.direction);
    V3<R> baseColor = vmax(m.baseColor.xyz(), V3<R>(0));
    V3<R> color;
    R alpha;
    V4<R> albedo;
    if (!interactive) {  // :4838-4845
        V3<R> key = normalize(V3<R>(R(0.35f), R(0.72f), R(0.58f)));
        R lambert = fmx(dot(n, key), R(0));
        R facing = fmx(dot(n, view), R(0));
        R rim = std::pow(R(1) - facing, R(2)) * R(0.08);
        R shade = R(0.32) + R(0.68) * lambert + rim;
        color = baseColor * shade + vmax(m.emission.xyz(), V3<R>(0));
        alpha = st.transparentBackground != 0u ? clampv(m.baseColor.w, R(0), R(1)) : R(1);
        albedo = m.baseColor;
    } else {  // :4967-5140
        R roughness = clampv(m.parameters.x, R(0.02), R(1));
        R metallic = clampv(m.parameters.y, R(0), R(1));
        R opacity = clampv(m.baseColor.w, R(0), R(1));
        R specular = clampv(m.parameters2.x, R(0), R(1));
        color = vmax(m.emission.xyz(), V3<R>(0));
        const or_scene &s = *c.s;
        uint32_t nl = std::min<uint32_t>(uint32_t(s.lights.size()), 12u);
        for (uint32_t li = 0; li < nl; ++li) {
            const or_light &light = s.lights[li];
            if (light.triangleIndex >= s.triangles.size() || light.materialIndex >= s.materials.size()) continue;
            const or_triangle &lt = s.triangles[light.triangleIndex];
            V3<R> emission = vmax(v4<R>(s.materials[light.materialIndex].emission).xyz(), V3<R>(0));
            if (maxComponent(emission) <= R(0) || R(light.area) <= R(0)) continue;
            V3<R> lp = (v3<R>(lt.v0) + v3<R>(lt.v1) + v3<R>(lt.v2)) / R(3);
            V3<R> toLight = lp - h.position;
            R dist2 = fmx(dot(toLight, toLight), R(1e-5f));
            R dist = std::sqrt(dist2);
            V3<R> ldir = toLight / dist;
            R cosS = fmx(dot(n, ldir), R(0));
            R cosL = fmx(dot(v4<R>(light.normal).xyz(), -ldir), R(0));
            if (cosS <= R(0) || cosL <= R(0)) continue;
            V3<R> tr(1);
            if (shadowOccluded(c, Ray<R>{h.position + n * R(0.001), ldir}, dist - R(0.002), tr)) continue;
            V3<R> brdf = evaluateMaterialBRDF(m, n, h.tangent, h.bitangent, view, ldir);
            R ccGloss = materialClearcoat(m) * std::pow(R(1) - materialClearcoatRoughness(m), R(2));
            R gloss = clampv(metallic + ccGloss + (R(1) - metallic) * std::pow(R(1) - roughness, R(4)) * specular, R(0), R(1));
            V3<R> diffuseOnly = (R(1) - metallic) * baseColor * vmax(V3<R>(0), V3<R>(1) - materialF0(m)) * R(0.31830988618f);
            brdf = diffuseOnly + (brdf - diffuseOnly) * gloss;
            color = color + brdf * emission * (cosS * cosL * R(light.area) / dist2) * tr;
        }
        V3<R> envDiffuse = interactiveDiffuseEnvironment(c, n);
        V3<R> refl = normalize(reflect(-view, n));
        V3<R> baseRefl = interactiveSpecularEnvironment(c, refl, n, roughness);
        R ccRough = materialClearcoatRoughness(m);
        V3<R> ccRefl = interactiveSpecularEnvironment(c, refl, n, ccRough);
        R nDotV = fmx(dot(n, view), R(0));
        V3<R> fresnel = fresnelSchlickThinFilm(nDotV, materialF0(m), m);
        R glossyW = std::pow(R(1) - roughness, R(4));
        R metalW = metallic * std::pow(R(1) - roughness, R(2));
        R ccW = materialClearcoat(m) * std::pow(R(1) - ccRough, R(3));
        R dielW = (R(1) - metallic) * glossyW * specular;
        R rw = smoothstepv(R(0.45), R(0.95), roughness);
        V3<R> roughFresnel = fresnel + (materialF0(m) - fresnel) * rw;
        V3<R> ccFresnel = fresnelSchlickThinFilm(nDotV, materialClearcoatF0Color(m), m);
        R horizon = R(0.58) + R(0.42) * clampv(n.y * R(0.5) + R(0.5), R(0), R(1));
        R contact = interactiveContactOcclusion(c, h, m, n);
        R ao = horizon * contact;
        color = color + baseColor * envDiffuse * (R(0.58) * (R(1) - metallic) * ao);
        color = color + baseRefl * roughFresnel * (metalW * R(1.05) + dielW * R(0.32));
        color = color + ccRefl * ccFresnel * ccW * R(0.42);
        V3<R> key = normalize(V3<R>(R(-0.32f), R(0.78f), R(0.52f)));
        color = color + baseColor * ((R(1) - metallic) * fmx(dot(n, key), R(0)) * R(0.16) * contact);
        R sss = materialSubsurface(m);
        if (sss > R(0)) {
            R wrap = std::pow(R(1) - nDotV, R(2));
            color = color + materialSubsurfaceColor(m) * baseColor * envDiffuse * (sss * (R(0.12) + R(0.22) * wrap) * contact);
        }
        R transmission = materialTransmission(m);
        if (transmission > R(0)) {
            V3<R> transmitted = sampleEnvironment(c, ray.direction);
            V3<R> tt = transmitted * materialTransmissionTint(m, R(0.65));
            color = color + (tt - color) * (transmission * R(0.55));
            color = color + baseRefl * fresnel * (transmission * R(0.35));
        }
        color = vclamp(color, V3<R>(0), V3<R>(fmx(R(st.sampleRadianceClamp), R(1))));
        alpha = st.transparentBackground != 0u ? opacity : R(1);
        albedo = V4<R>{baseColor.x, baseColor.y, baseColor.z, opacity};
    }
    put(acc + pix, color.x, color.y, color.z, alpha);
    if (!aov) return;
    V3<R> en = n * R(0.5) + V3<R>(R(0.5));
    put(P[0], h.t, h.t, h.t, R(1));
    put(P[1], en.x, en.y, en.z, R(1));
    put(P[2], albedo.x, albedo.y, albedo.z, albedo.w);
    R mid = R(float(h.materialID + 1u)), oid = R(float(h.objectID + 1u));
    put(P[3], mid, mid, mid, R(1));
    put(P[4], oid, oid, oid, R(1));
    V2<R> cur, prv;
    if (projectToScreen(*c.cam, h.position, width, height, cur) && projectToScreen(*c.prevCam, h.position, width, height, prv))
        put(P[5], prv.x - cur.x, prv.y - cur.y, R(0), R(1));
    else
        put(P[5], 0, 0, 0, 0);
    uint32_t tid = h.primitiveID + 1u;
    R tv;
    if (sizeof(R) == 4) { float fv; std::memcpy(&fv, &tid, 4); tv = R(fv); }
    else { tv = R(double(tid)); }
    put(P[6], tv, tv, tv, tv);
}

template <class R>
int renderT(const or_scene *s, const or_settings *st, const or_camera *cam, const or_camera *prev, uint32_t sampleBegin,
            uint32_t sampleCount, int traversal, R *acc, R *aov, uint64_t *counters, int threads) {
    const uint32_t W = st->width, H = st->height;
    const uint32_t tx0 = std::min(st->tileX, W), ty0 = std::min(st->tileY, H);
    const uint32_t tw = std::min(st->tileWidth, W - tx0), th = std::min(st->tileHeight, H - ty0);
    const size_t planeStride = size_t(W) * H * 4;
    std::atomic<uint64_t> total[OR_CNT_COUNT];
    for (auto &t : total) t = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int64_t rows = th;
#pragma omp parallel
    {
        Ctx<R> c{s, st, cam, prev, traversal, {0}, {}};
        initCtx(c);
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < rows; ++r) {
            uint32_t py = ty0 + uint32_t(r);
            for (uint32_t px = tx0; px < tx0 + tw; ++px) {
                if (st->renderQuality < 2u) {  // preview / interactive: every sample writes the same image
                    if (sampleCount) previewPixel<R>(c, px, py, st->renderQuality == 1u, acc, aov, planeStride);
                    continue;
                }
                for (uint32_t k = 0; k < sampleCount; ++k) {
                    bool last = (k + 1 == sampleCount);
                    pathTracePixel<R>(c, px, py, sampleBegin + k, acc, last ? aov : nullptr, planeStride);
                }
            }
        }
        for (int k = 0; k < OR_CNT_COUNT; ++k) total[k] += c.cnt[k];
    }
    if (counters)
        for (int k = 0; k < OR_CNT_COUNT; ++k) counters[k] += total[k];
    return 0;
}

} // namespace

// ------------------------------------------------------------------ C ABI --
extern "C" {

uint32_t or_hash(uint32_t value) { return hashU32(value); }
float or_random_float(uint32_t *state) { return randomFloat<float>(*state); }

int or_scene_create(const or_scene_desc *d, uint32_t maxLeaf, or_scene **out) {
    if (!d || !out) return 1;
    auto s = std::make_unique<or_scene>();
    s->triangles.assign(d->triangles, d->triangles + d->triangleCount);
    s->materials.assign(d->materials, d->materials + d->materialCount);
    if (d->lightCount) s->lights.assign(d->lights, d->lights + d->lightCount);
    if (d->envSampleCount) s->env.assign(d->envSamples, d->envSamples + d->envSampleCount);
    if (s->env.size() > 1) buildEnvironmentAlias(s->env, s->envAliasProb, s->envAliasIndex);
    if (d->textureCount) s->textures.assign(d->textureDescs, d->textureDescs + d->textureCount);
    if (d->texturePixelCount) s->texels.assign(d->texturePixels, d->texturePixels + d->texturePixelCount);
    if (d->volumeCount) s->volumes.assign(d->volumes, d->volumes + d->volumeCount);
    if (d->brickCount) s->bricks.assign(d->bricks, d->bricks + d->brickCount);
    if (d->brickSampleCount) s->brickSamples.assign(d->brickSamples, d->brickSamples + d->brickSampleCount);
    if (d->volumeSampleCount) s->volumeSamples.assign(d->volumeSamples, d->volumeSamples + d->volumeSampleCount);
    if (d->brickGridCount) s->brickGrids.assign(d->brickGrids, d->brickGrids + d->brickGridCount);
    if (d->brickGridIndexCount)
        s->brickGridIndices.assign(d->brickGridIndices, d->brickGridIndices + d->brickGridIndexCount);
    if (s->materials.empty()) return 2;
    const uint32_t n = d->triangleCount;
    if (n > 0) {
        std::vector<AABBf> bounds(n);
        for (uint32_t i = 0; i < n; ++i) {  // BVHBuilder.bounds(for:) (:87-93)
            const or_triangle &t = s->triangles[i];
            const float *v[3] = {&t.v0.x, &t.v1.x, &t.v2.x};
            for (int a = 0; a < 3; ++a) {
                bounds[i].mn[a] = std::fmin(std::fmin(v[0][a], v[1][a]), v[2][a]);
                bounds[i].mx[a] = std::fmax(std::fmax(v[0][a], v[1][a]), v[2][a]);
            }
        }
        s->primIndices.resize(n);
        std::iota(s->primIndices.begin(), s->primIndices.end(), 0u);
        s->nodes.reserve(size_t(n) / 2 + 16);
        Builder b{bounds, s->primIndices, s->nodes, maxLeaf == 0 ? 4u : maxLeaf};
        b.build(0, n);
    }
    *out = s.release();
    return 0;
}

void or_scene_destroy(or_scene *s) { delete s; }

uint32_t or_scene_node_count(const or_scene *s) { return s ? uint32_t(s->nodes.size()) : 0u; }

int or_scene_export_bvh(const or_scene *s, void *nodes48, uint32_t *prim) {
    if (!s) return 1;
    if (nodes48) std::memcpy(nodes48, s->nodes.data(), s->nodes.size() * sizeof(Node));
    if (prim) std::memcpy(prim, s->primIndices.data(), s->primIndices.size() * 4);
    return 0;
}

int or_render(const or_scene *s, const or_settings *st, const or_camera *cam, const or_camera *prev, uint32_t sampleBegin,
              uint32_t sampleCount, int precision, int traversal, void *acc, void *aovs, uint64_t *counters, int threads) {
    if (!s || !st || !cam || !acc) return 1;
    if (st->width == 0 || st->height == 0) return 2;
    const or_camera *pc = prev ? prev : cam;
    if (precision == 64)
        return renderT<double>(s, st, cam, pc, sampleBegin, sampleCount, traversal, static_cast<double *>(acc),
                               static_cast<double *>(aovs), counters, threads);
    return renderT<float>(s, st, cam, pc, sampleBegin, sampleCount, traversal, static_cast<float *>(acc),
                          static_cast<float *>(aovs), counters, threads);
}

int or_environment_alias(const or_scene *s, float *prob, uint32_t *alias, uint32_t count) {
    if (!s || count != s->envAliasProb.size()) return 1;
    std::copy(s->envAliasProb.begin(), s->envAliasProb.end(), prob);
    std::copy(s->envAliasIndex.begin(), s->envAliasIndex.end(), alias);
    return 0;
}

int or_intersect(const or_scene *s, const float *rays, uint32_t count, int traversal, int intersection, float *tOut,
                 uint32_t *idOut, float *hitOut) {
    if (!s) return 1;
    or_settings dummy{};
    dummy.renderQuality = 2u;
    dummy.intersection = uint32_t(intersection);
    or_camera cam{};
#pragma omp parallel
    {
        Ctx<float> c{s, &dummy, &cam, &cam, traversal, {0}, {}};
        initCtx(c);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < int64_t(count); ++i) {
            const float *r = rays + 8 * i;
            Ray<float> ray{V3<float>(r[0], r[1], r[2]), V3<float>(r[4], r[5], r[6])};
            Hit<float> h = intersectScene(c, ray);
            tOut[i] = h.hit ? h.t : std::numeric_limits<float>::infinity();
            idOut[i] = h.hit ? h.primitiveID : 0xffffffffu;
            if (hitOut) {
                float *o = hitOut + 8 * i;
                o[0] = h.normal.x, o[1] = h.normal.y, o[2] = h.normal.z;
                o[3] = float(h.materialID), o[4] = float(h.materialID2), o[5] = h.materialBlend;
                o[6] = float(h.objectID), o[7] = h.frontFacing ? 1.0f : 0.0f;
            }
        }
    }
    return 0;
}

}  // extern "C"
