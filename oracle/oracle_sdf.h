// oracle_sdf.h -- TEST INFRASTRUCTURE ONLY (part of the CPU oracle).
//
// Restatement of the reference's sparse signed-distance brick tracing:
//   makeVolumeBrickSampleContext      PathTrace.metal:2074-2094
//   sampleVolumeBrickDistancePrepared :2096-2127
//   gradient (trilinear / central / Catmull-Rom) :2141-2314
//   sampleVolumeBrickMaterial         :2328-2372
//   volumeBrickNormal                 :2420-2469
//   intersectVolumeBrickPrepared      :2608-2807
//   closestSparseVolumeBrickDDAHit    :2881-3088
//   closestVolumeHit (dense loop, brick-grid path) :3090-3163
//   dense volumes: sampleVolumeDistance :1915-1955, sampleVolumeMaterial
//   :1957-1990, volumeNormal :2029-2055, intersectVolume :2471-2606
// Included by oracle.cpp inside its anonymous namespace; uses its vector
// types, the or_* records and Hit<R>.
#pragma once

template <class R> struct Mat4 { V4<R> c[4]; };  // columns
template <class R> Mat4<R> mat4(const or_float4 *cols) {
    Mat4<R> m;
    for (int k = 0; k < 4; ++k) m.c[k] = v4<R>(cols[k]);
    return m;
}
// matrix * (v, w): ((c0*x + c1*y) + c2*z) + c3*w
template <class R> V3<R> mul(const Mat4<R> &m, V3<R> v, R w) {
    return V3<R>(m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z + m.c[3].x * w,
                 m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z + m.c[3].y * w,
                 m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z + m.c[3].z * w);
}

template <class R>
bool intersectAABB(const Ray<R> &ray, V3<R> bmin, V3<R> bmax, R closestT, R &tNear, R &tFar) {  // :1865-1881
    V3<R> inv = V3<R>(1) / ray.direction;
    V3<R> t0 = (bmin - ray.origin) * inv;
    V3<R> t1 = (bmax - ray.origin) * inv;
    V3<R> tn = vmin(t0, t1), tf = vmax(t0, t1);
    tNear = fmx(fmx(tn.x, tn.y), fmx(tn.z, R(0.0005)));
    tFar = fmn(fmn(tf.x, tf.y), fmn(tf.z, closestT));
    return tNear <= tFar;
}

template <class R> struct BrickCtx {  // VolumeBrickSampleContext
    uint32_t dim[3];
    uint32_t sampleOffset, strideY, strideZ, sampleCount;
    V3<R> sampleMin, sampleMax, sampleExtent;
    bool valid;
};

struct SdfScene {
    const or_volume *volumes = nullptr;
    uint32_t volumeCount = 0;
    const or_volume_sample *dense = nullptr;
    uint32_t denseCount = 0;
    const or_brick *bricks = nullptr;
    uint32_t brickCount = 0;
    const or_brick_sample *samples = nullptr;
    uint32_t sampleCount = 0;
    const or_brick_grid *grids = nullptr;
    uint32_t gridCount = 0;
    const uint32_t *gridIndices = nullptr;
    uint32_t gridIndexCount = 0;
    uint32_t renderQuality = 2;
};

template <class R> BrickCtx<R> makeCtx(const or_brick &b, uint32_t brickSampleCount) {  // :2074-2094
    BrickCtx<R> c;
    for (int a = 0; a < 3; ++a) c.dim[a] = (&b.dimensionsAndSampleOffset.x)[a];
    c.sampleOffset = b.dimensionsAndSampleOffset.w;
    uint32_t n = c.dim[0] * c.dim[1] * c.dim[2];
    c.strideY = c.dim[0];
    c.strideZ = c.dim[0] * c.dim[1];
    c.sampleCount = n;
    c.sampleMin = v3<R>(b.sampleBoundsMin);
    c.sampleMax = v3<R>(b.sampleBoundsMax);
    c.sampleExtent = vmax(c.sampleMax - c.sampleMin, V3<R>(R(1e-6)));
    bool rangeValid = c.sampleOffset < brickSampleCount && n <= brickSampleCount - c.sampleOffset;
    c.valid = c.dim[0] >= 2u && c.dim[1] >= 2u && c.dim[2] >= 2u && rangeValid;
    return c;
}

template <class R> R sdAt(const SdfScene &S, const BrickCtx<R> &c, uint32_t x, uint32_t y, uint32_t z) {
    return R(S.samples[c.sampleOffset + x + y * c.strideY + z * c.strideZ].distance);
}

template <class R> void gridCoord(const BrickCtx<R> &c, V3<R> p, uint32_t base[3], uint32_t next[3], V3<R> &frac) {
    V3<R> uvw = vclamp((p - c.sampleMin) / c.sampleExtent, V3<R>(0), V3<R>(1));
    V3<R> g(uvw.x * R(c.dim[0] - 1u), uvw.y * R(c.dim[1] - 1u), uvw.z * R(c.dim[2] - 1u));
    R gv[3] = {g.x, g.y, g.z};
    R fv[3];
    for (int a = 0; a < 3; ++a) {
        base[a] = uint32_t(std::floor(gv[a]));
        next[a] = std::min(base[a] + 1u, c.dim[a] - 1u);
        fv[a] = gv[a] - R(base[a]);
    }
    frac = V3<R>(fv[0], fv[1], fv[2]);
}

template <class R> R sampleDistance(const SdfScene &S, const BrickCtx<R> &c, V3<R> p) {  // :2096-2127
    if (!c.valid) return kInf<R>();
    uint32_t b[3], n[3];
    V3<R> f;
    gridCoord(c, p, b, n, f);
    R c00 = mixs(sdAt(S, c, b[0], b[1], b[2]), sdAt(S, c, n[0], b[1], b[2]), f.x);
    R c10 = mixs(sdAt(S, c, b[0], n[1], b[2]), sdAt(S, c, n[0], n[1], b[2]), f.x);
    R c01 = mixs(sdAt(S, c, b[0], b[1], n[2]), sdAt(S, c, n[0], b[1], n[2]), f.x);
    R c11 = mixs(sdAt(S, c, b[0], n[1], n[2]), sdAt(S, c, n[0], n[1], n[2]), f.x);
    R c0 = mixs(c00, c10, f.y);
    R c1 = mixs(c01, c11, f.y);
    return mixs(c0, c1, f.z);
}

template <class R> struct DG { R distance; V3<R> gradient; };

template <class R> DG<R> sampleDistanceGradient(const SdfScene &S, const BrickCtx<R> &c, V3<R> p) {  // :2161-2219
    DG<R> r{kInf<R>(), V3<R>(0)};
    if (!c.valid) return r;
    uint32_t b[3], n[3];
    V3<R> f;
    gridCoord(c, p, b, n, f);
    R d000 = sdAt(S, c, b[0], b[1], b[2]), d100 = sdAt(S, c, n[0], b[1], b[2]);
    R d010 = sdAt(S, c, b[0], n[1], b[2]), d110 = sdAt(S, c, n[0], n[1], b[2]);
    R d001 = sdAt(S, c, b[0], b[1], n[2]), d101 = sdAt(S, c, n[0], b[1], n[2]);
    R d011 = sdAt(S, c, b[0], n[1], n[2]), d111 = sdAt(S, c, n[0], n[1], n[2]);
    R c00 = mixs(d000, d100, f.x), c10 = mixs(d010, d110, f.x);
    R c01 = mixs(d001, d101, f.x), c11 = mixs(d011, d111, f.x);
    R c0 = mixs(c00, c10, f.y), c1 = mixs(c01, c11, f.y);
    r.distance = mixs(c0, c1, f.z);
    R dx = mixs(mixs(d100 - d000, d110 - d010, f.y), mixs(d101 - d001, d111 - d011, f.y), f.z);
    R dy = mixs(mixs(d010 - d000, d110 - d100, f.x), mixs(d011 - d001, d111 - d101, f.x), f.z);
    R dz = c1 - c0;
    V3<R> g2l = V3<R>(R(c.dim[0] - 1u), R(c.dim[1] - 1u), R(c.dim[2] - 1u)) / c.sampleExtent;
    r.gradient = V3<R>(dx, dy, dz) * g2l;
    return r;
}

template <class R> V3<R> sampleCentralGradient(const SdfScene &S, const BrickCtx<R> &c, V3<R> p, V3<R> sp) {  // :2221-2255
    V3<R> mn = c.sampleMin, mx = c.sampleMax;
    V3<R> x0 = p, x1 = p;
    x0.x = fmx(p.x - sp.x, mn.x);
    x1.x = fmn(p.x + sp.x, mx.x);
    R xw = fmx(x1.x - x0.x, R(1e-6));
    R dx = (sampleDistance(S, c, x1) - sampleDistance(S, c, x0)) / xw;
    V3<R> y0 = p, y1 = p;
    y0.y = fmx(p.y - sp.y, mn.y);
    y1.y = fmn(p.y + sp.y, mx.y);
    R yw = fmx(y1.y - y0.y, R(1e-6));
    R dy = (sampleDistance(S, c, y1) - sampleDistance(S, c, y0)) / yw;
    V3<R> z0 = p, z1 = p;
    z0.z = fmx(p.z - sp.z, mn.z);
    z1.z = fmn(p.z + sp.z, mx.z);
    R zw = fmx(z1.z - z0.z, R(1e-6));
    R dz = (sampleDistance(S, c, z1) - sampleDistance(S, c, z0)) / zw;
    return V3<R>(dx, dy, dz);
}

template <class R> R catmullRom(R p0, R p1, R p2, R p3, R t) {  // :2141-2150
    R t2 = t * t;
    R t3 = t2 * t;
    return R(0.5) * (R(2) * p1 + (-p0 + p2) * t + (R(2) * p0 - R(5) * p1 + R(4) * p2 - p3) * t2 +
                     (-p0 + R(3) * p1 - R(3) * p2 + p3) * t3);
}
template <class R> R catmullRomD(R p0, R p1, R p2, R p3, R t) {  // :2152-2159
    R t2 = t * t;
    return R(0.5) * ((-p0 + p2) + R(2) * (R(2) * p0 - R(5) * p1 + R(4) * p2 - p3) * t +
                     R(3) * (-p0 + R(3) * p1 - R(3) * p2 + p3) * t2);
}

template <class R> DG<R> sampleDistanceGradientCubic(const SdfScene &S, const BrickCtx<R> &c, V3<R> p) {  // :2257-2314
    DG<R> inv{kInf<R>(), V3<R>(0)};
    if (!c.valid) return inv;
    V3<R> uvw = vclamp((p - c.sampleMin) / c.sampleExtent, V3<R>(0), V3<R>(1));
    V3<R> g(uvw.x * R(c.dim[0] - 1u), uvw.y * R(c.dim[1] - 1u), uvw.z * R(c.dim[2] - 1u));
    uint32_t b[3] = {uint32_t(std::floor(g.x)), uint32_t(std::floor(g.y)), uint32_t(std::floor(g.z))};
    if (b[0] < 1u || b[1] < 1u || b[2] < 1u || b[0] + 2u >= c.dim[0] || b[1] + 2u >= c.dim[1] || b[2] + 2u >= c.dim[2])
        return sampleDistanceGradient(S, c, p);
    V3<R> f(g.x - R(b[0]), g.y - R(b[1]), g.z - R(b[2]));
    R xr[4][4], dxr[4][4];
    for (uint32_t z = 0; z < 4; ++z)
        for (uint32_t y = 0; y < 4; ++y) {
            uint32_t yy = b[1] + y - 1u, zz = b[2] + z - 1u;
            R p0 = sdAt(S, c, b[0] - 1u, yy, zz), p1 = sdAt(S, c, b[0], yy, zz);
            R p2 = sdAt(S, c, b[0] + 1u, yy, zz), p3 = sdAt(S, c, b[0] + 2u, yy, zz);
            xr[y][z] = catmullRom(p0, p1, p2, p3, f.x);
            dxr[y][z] = catmullRomD(p0, p1, p2, p3, f.x);
        }
    R yr[4], dyr[4], dxyr[4];
    for (uint32_t z = 0; z < 4; ++z) {
        yr[z] = catmullRom(xr[0][z], xr[1][z], xr[2][z], xr[3][z], f.y);
        dyr[z] = catmullRomD(xr[0][z], xr[1][z], xr[2][z], xr[3][z], f.y);
        dxyr[z] = catmullRom(dxr[0][z], dxr[1][z], dxr[2][z], dxr[3][z], f.y);
    }
    DG<R> r;
    r.distance = catmullRom(yr[0], yr[1], yr[2], yr[3], f.z);
    R dx = catmullRom(dxyr[0], dxyr[1], dxyr[2], dxyr[3], f.z);
    R dy = catmullRom(dyr[0], dyr[1], dyr[2], dyr[3], f.z);
    R dz = catmullRomD(yr[0], yr[1], yr[2], yr[3], f.z);
    V3<R> g2l = V3<R>(R(c.dim[0] - 1u), R(c.dim[1] - 1u), R(c.dim[2] - 1u)) / c.sampleExtent;
    r.gradient = V3<R>(dx, dy, dz) * g2l;
    return r;
}

template <class R>
V3<R> volumeBrickNormal(const SdfScene &S, const or_volume &vol, const BrickCtx<R> &c, V3<R> p, bool cubic, bool coarse) {  // :2420-2469
    if (!c.valid) return V3<R>(0, 1, 0);
    V3<R> ln = cubic ? sampleDistanceGradientCubic(S, c, p).gradient : sampleDistanceGradient(S, c, p).gradient;
    if (coarse) {
        V3<R> ext = vmax(v3<R>(vol.localBoundsMax) - v3<R>(vol.localBoundsMin), V3<R>(R(1e-6)));
        V3<R> dims(R(vol.dimensions.x), R(vol.dimensions.y), R(vol.dimensions.z));
        V3<R> spacing = ext / vmax(dims - V3<R>(1), V3<R>(1));
        V3<R> cn = sampleCentralGradient(S, c, p, spacing * R(1.5));
        if (dot(cn, cn) > R(1e-10)) ln = normalize(mixv(normalize(ln), normalize(cn), R(0.35)));
    }
    if (dot(ln, ln) <= R(1e-10)) ln = V3<R>(0, 1, 0);
    Mat4<R> nt = mat4<R>(&vol.normalTransform0);
    return normalize(mul(nt, normalize(ln), R(0)));
}

template <class R> struct VolHit {
    bool hit;
    R t;
    V3<R> position, normal;
    uint32_t materialA, materialB;
    R blend;
    uint32_t objectID, primitiveID;
    bool frontFacing;
    const or_volume_sample *fields;  // dense hits: the nearest sample (material fields)
};

// Counter slots (the reference's sdfCounter* constants, PathTrace.metal:191-200).
enum { SC_DENSE_TESTS = 0, SC_DENSE_STEPS = 1, SC_GRID_CELLS = 2, SC_MACRO_SKIPS = 3, SC_BRICK_TESTS = 4, SC_BRICK_INVALID = 5, SC_RANGE_CULLS = 6,
       SC_MARCHES = 7, SC_MARCH_STEPS = 8, SC_BRICK_HITS = 9 };

template <class R>
bool intersectVolumeBrickPrepared(const SdfScene &S, const Ray<R> &ray, const or_volume &vol, const or_brick &brick,
                                  R closestT, V3<R> lo, V3<R> ld, R lupw, R cellSize, uint64_t *sc,
                                  VolHit<R> &hit) {  // :2608-2807
    R tNear, tFar;
    Ray<R> lr{lo, ld};
    if (!intersectAABB(lr, v3<R>(brick.localBoundsMin), v3<R>(brick.localBoundsMax), closestT, tNear, tFar)) return false;
    BrickCtx<R> ctx = makeCtx<R>(brick, S.sampleCount);
    if (!ctx.valid) {
        sc[SC_BRICK_INVALID]++;
        return false;
    }
    if (brick.sampleBoundsMin.w > 0.0f || brick.sampleBoundsMax.w < 0.0f) {
        sc[SC_RANGE_CULLS]++;
        return false;
    }
    sc[SC_MARCHES]++;
    R minimumStep = fmx(cellSize * R(0.10) / lupw, R(0.0005));
    R prevT = tNear;
    V3<R> prevP = lo + ld * prevT;
    R prevD = sampleDistance(S, ctx, prevP);
    if (!std::isfinite(prevD)) return false;
    R originalEntryT = prevT;
    if (prevD <= R(0)) {
        R backStep = fmx(cellSize / lupw * R(0.5), minimumStep);
        R searchT = prevT;
        for (uint32_t bt = 0; bt < 12u; ++bt) {
            searchT = fmx(searchT - backStep, R(0.0005));
            V3<R> sp = lo + ld * searchT;
            R sd = sampleDistance(S, ctx, sp);
            if (!std::isfinite(sd)) break;
            if (sd > R(0)) {
                prevT = searchT;
                prevP = sp;
                prevD = sd;
                break;
            }
            if (searchT <= R(0.0005)) break;
        }
    }
    R t = fmx(originalEntryT, prevT + fmx(std::fabs(prevD) / lupw * R(0.9), minimumStep));
    for (uint32_t step = 0; step < 192u && t <= tFar && t < closestT; ++step) {
        sc[SC_MARCH_STEPS]++;
        V3<R> p = lo + ld * t;
        R d = sampleDistance(S, ctx, p);
        if (!std::isfinite(d)) return false;
        bool crossed = (prevD > R(0) && d <= R(0)) || (prevD < R(0) && d >= R(0));
        if (crossed) {
            R lowT = prevT, highT = t, lowD = prevD;
            R refinedT = t;
            for (uint32_t k = 0; k < 8u; ++k) {
                R midT = R(0.5) * (lowT + highT);
                V3<R> mp = lo + ld * midT;
                R md = sampleDistance(S, ctx, mp);
                if ((lowD > R(0) && md > R(0)) || (lowD < R(0) && md < R(0))) {
                    lowT = midT;
                    lowD = md;
                } else {
                    highT = midT;
                }
                refinedT = R(0.5) * (lowT + highT);
            }
            bool cubicPolish = cellSize <= R(0.03);
            bool coarse = S.renderQuality > 0u && cellSize > R(0.03);
            V3<R> rp = lo + ld * refinedT;
            if (cubicPolish) {
                uint32_t polishCount = S.renderQuality > 1u ? 2u : 1u;
                for (uint32_t k = 0; k < polishCount; ++k) {
                    DG<R> pg = sampleDistanceGradientCubic(S, ctx, rp);
                    R slope = dot(pg.gradient, ld);
                    if (!std::isfinite(pg.distance) || std::fabs(slope) <= R(1e-6)) break;
                    R projected = refinedT - pg.distance / slope;
                    if (projected <= lowT || projected >= highT) break;
                    refinedT = projected;
                    rp = lo + ld * refinedT;
                }
            }
            V3<R> wp = ray.origin + ray.direction * refinedT;
            V3<R> n = volumeBrickNormal(S, vol, ctx, rp, cubicPolish, coarse);
            bool ff = dot(n, ray.direction) <= R(0);
            if (!ff) n = -n;
            // sampleVolumeBrickMaterial (:2328-2372): nearest sample.
            V3<R> uvw = vclamp((rp - ctx.sampleMin) / ctx.sampleExtent, V3<R>(0), V3<R>(1));
            uint32_t nx = std::min(uint32_t(std::round(uvw.x * R(ctx.dim[0] - 1u))), ctx.dim[0] - 1u);
            uint32_t ny = std::min(uint32_t(std::round(uvw.y * R(ctx.dim[1] - 1u))), ctx.dim[1] - 1u);
            uint32_t nz = std::min(uint32_t(std::round(uvw.z * R(ctx.dim[2] - 1u))), ctx.dim[2] - 1u);
            const or_brick_sample &ms = S.samples[ctx.sampleOffset + nx + ny * ctx.strideY + nz * ctx.strideZ];
            hit.hit = true;
            hit.t = refinedT;
            hit.position = wp;
            hit.normal = n;
            hit.materialA = ms.materialA;
            hit.materialB = ms.materialB;
            hit.blend = clampv(R(ms.materialBlend), R(0), R(1));
            hit.objectID = vol.metadata.y;
            hit.primitiveID = vol.metadata.z;
            hit.frontFacing = ff;
            sc[SC_BRICK_HITS]++;
            return true;
        }
        prevT = t;
        prevD = d;
        prevP = p;
        t += fmx(std::fabs(d) / lupw * R(0.8), minimumStep);
    }
    return false;
}

template <class R>
VolHit<R> closestSparseVolumeBrickDDAHit(const SdfScene &S, uint32_t volumeIndex, const Ray<R> &ray, R closestT,
                                         uint64_t *sc) {  // :2881-3088
    VolHit<R> closest{};
    closest.hit = false;
    closest.t = closestT;
    const or_volume &vol = S.volumes[volumeIndex];
    const or_brick_grid *grid = nullptr;
    for (uint32_t g = 0; g < S.gridCount; ++g)  // findVolumeBrickGrid (:2862-2879)
        if (S.grids[g].brickSizeAndVolume.w == volumeIndex) {
            grid = &S.grids[g];
            break;
        }
    if (!grid || !S.gridIndices) return closest;
    uint32_t gd[3] = {grid->dimensionsAndIndexOffset.x, grid->dimensionsAndIndexOffset.y, grid->dimensionsAndIndexOffset.z};
    uint32_t gridIndexOffset = grid->dimensionsAndIndexOffset.w;
    uint32_t bs[3] = {std::max(grid->brickSizeAndVolume.x, 1u), std::max(grid->brickSizeAndVolume.y, 1u),
                      std::max(grid->brickSizeAndVolume.z, 1u)};
    if (gd[0] == 0u || gd[1] == 0u || gd[2] == 0u) return closest;
    R tNear, tFar;
    if (!intersectAABB(ray, v3<R>(vol.worldBoundsMin), v3<R>(vol.worldBoundsMax), closestT, tNear, tFar)) return closest;
    Mat4<R> w2l = mat4<R>(&vol.worldToLocal0);
    V3<R> lo = mul(w2l, ray.origin, R(1));
    V3<R> ld = mul(w2l, ray.direction, R(0));
    V3<R> lbmin = v3<R>(vol.localBoundsMin);
    V3<R> extent = vmax(v3<R>(vol.localBoundsMax) - lbmin, V3<R>(R(1e-6)));
    V3<R> den = vmax(V3<R>(R(vol.dimensions.x), R(vol.dimensions.y), R(vol.dimensions.z)) - V3<R>(1), V3<R>(1));
    V3<R> bsf = vmax(V3<R>(R(bs[0]), R(bs[1]), R(bs[2])), V3<R>(1));
    V3<R> spacing = extent / den;
    R cellSize = fmn(spacing.x, fmn(spacing.y, spacing.z));
    R lupw = fmx(length(ld), R(1e-6));

    R t = tNear;
    V3<R> lp = lo + ld * t;
    V3<R> samplePos = ((lp - lbmin) / extent) * den;
    V3<R> brickPos = (samplePos + V3<R>(R(0.5))) / bsf;
    V3<R> gdm1(R(gd[0]) - R(1), R(gd[1]) - R(1), R(gd[2]) - R(1));
    V3<R> fc = vclamp(V3<R>(std::floor(brickPos.x), std::floor(brickPos.y), std::floor(brickPos.z)), V3<R>(0), gdm1);
    int32_t cell[3] = {int32_t(fc.x), int32_t(fc.y), int32_t(fc.z)};
    V3<R> vel = (ld / extent) * den / bsf;
    R velv[3] = {vel.x, vel.y, vel.z};
    R bpv[3] = {brickPos.x, brickPos.y, brickPos.z};
    int32_t stepDir[3];
    R tMax[3], tDelta[3];
    for (int a = 0; a < 3; ++a) {
        stepDir[a] = velv[a] > R(0) ? 1 : (velv[a] < R(0) ? -1 : 0);
        R nb = stepDir[a] > 0 ? R(cell[a]) + R(1) : R(cell[a]);
        tMax[a] = stepDir[a] == 0 ? kInf<R>() : t + (nb - bpv[a]) / velv[a];
        tDelta[a] = stepDir[a] == 0 ? kInf<R>() : std::fabs(R(1) / velv[a]);
    }
    uint32_t md[3] = {grid->macroDimensionsAndIndexOffset.x, grid->macroDimensionsAndIndexOffset.y,
                      grid->macroDimensionsAndIndexOffset.z};
    uint32_t macroOffset = grid->macroDimensionsAndIndexOffset.w;
    uint32_t ms[3] = {std::max(grid->macroSizeAndReserved.x, 1u), std::max(grid->macroSizeAndReserved.y, 1u),
                      std::max(grid->macroSizeAndReserved.z, 1u)};
    bool hasMacro = md[0] > 0u && md[1] > 0u && md[2] > 0u && macroOffset < S.gridIndexCount;
    uint32_t maxSteps = std::min(gd[0] + gd[1] + gd[2] + 8u, 1024u);
    for (uint32_t step = 0; step < maxSteps && t <= tFar && t < closest.t; ++step) {
        sc[SC_GRID_CELLS]++;
        if (cell[0] < 0 || cell[1] < 0 || cell[2] < 0 || cell[0] >= int32_t(gd[0]) || cell[1] >= int32_t(gd[1]) ||
            cell[2] >= int32_t(gd[2]))
            break;
        if (hasMacro) {
            uint32_t mc[3];
            for (int a = 0; a < 3; ++a) mc[a] = std::min(uint32_t(cell[a]) / ms[a], md[a] - 1u);
            uint32_t slot = macroOffset + mc[0] + mc[1] * md[0] + mc[2] * md[0] * md[1];
            if (slot < S.gridIndexCount && S.gridIndices[slot] == 0u) {
                V3<R> cur = brickPos + vel * (t - tNear);
                R cv[3] = {cur.x, cur.y, cur.z};
                R mt[3];
                for (int a = 0; a < 3; ++a) {
                    R boundary = stepDir[a] > 0 ? R(std::min((mc[a] + 1u) * ms[a], gd[a])) : R(mc[a] * ms[a]);
                    mt[a] = stepDir[a] == 0 ? kInf<R>() : t + (boundary - cv[a]) / velv[a];
                }
                R macroNextT = fmn(mt[0], fmn(mt[1], mt[2]));
                if (macroNextT > t + R(1e-6)) {
                    sc[SC_MACRO_SKIPS]++;
                    if (macroNextT > tFar || macroNextT >= closest.t) break;
                    t = macroNextT + R(1e-5);
                    cur = brickPos + vel * (t - tNear);
                    V3<R> fc2 = vclamp(V3<R>(std::floor(cur.x), std::floor(cur.y), std::floor(cur.z)), V3<R>(0), gdm1);
                    cell[0] = int32_t(fc2.x);
                    cell[1] = int32_t(fc2.y);
                    cell[2] = int32_t(fc2.z);
                    R cv2[3] = {cur.x, cur.y, cur.z};
                    for (int a = 0; a < 3; ++a) {
                        R nb = stepDir[a] > 0 ? R(cell[a]) + R(1) : R(cell[a]);
                        tMax[a] = stepDir[a] == 0 ? kInf<R>() : t + (nb - cv2[a]) / velv[a];
                    }
                    continue;
                }
            }
        }
        uint32_t slot = gridIndexOffset + uint32_t(cell[0]) + uint32_t(cell[1]) * gd[0] + uint32_t(cell[2]) * gd[0] * gd[1];
        if (slot < S.gridIndexCount) {
            uint32_t bi = S.gridIndices[slot];
            if (bi != 0xffffffffu && bi < S.brickCount) {
                sc[SC_BRICK_TESTS]++;
                VolHit<R> cand{};
                if (intersectVolumeBrickPrepared(S, ray, vol, S.bricks[bi], closest.t, lo, ld, lupw, cellSize, sc, cand) &&
                    cand.t < closest.t)
                    closest = cand;
            }
        }
        R nextT = fmn(tMax[0], fmn(tMax[1], tMax[2]));
        if (closest.hit && closest.t <= nextT + R(0.0005)) return closest;
        if (nextT > tFar || nextT >= closest.t) break;
        if (tMax[0] <= tMax[1] && tMax[0] <= tMax[2]) {
            cell[0] += stepDir[0];
            t = tMax[0];
            tMax[0] += tDelta[0];
        } else if (tMax[1] <= tMax[2]) {
            cell[1] += stepDir[1];
            t = tMax[1];
            tMax[1] += tDelta[1];
        } else {
            cell[2] += stepDir[2];
            t = tMax[2];
            tMax[2] += tDelta[2];
        }
    }
    return closest;
}

// ------------------------------------------------------- dense volumes --
// sampleVolumeDistance (:1915-1955): trilinear over the dense grid.
template <class R> R denseDistance(const SdfScene &S, const or_volume &vol, V3<R> lp) {
    uint32_t dim[3] = {vol.dimensions.x, vol.dimensions.y, vol.dimensions.z};
    uint32_t off = vol.metadata.x;
    uint32_t n = dim[0] * dim[1] * dim[2];
    if (off >= S.denseCount || off + n > S.denseCount) return kInf<R>();
    V3<R> lmin = v3<R>(vol.localBoundsMin);
    V3<R> ext = vmax(v3<R>(vol.localBoundsMax) - lmin, V3<R>(R(1e-6)));
    V3<R> uvw = vclamp((lp - lmin) / ext, V3<R>(0), V3<R>(1));
    R g[3] = {uvw.x * R(dim[0] - 1u), uvw.y * R(dim[1] - 1u), uvw.z * R(dim[2] - 1u)};
    uint32_t b[3], nx[3];
    R f[3];
    for (int a = 0; a < 3; ++a) {
        b[a] = uint32_t(std::floor(g[a]));
        nx[a] = std::min(b[a] + 1u, dim[a] - 1u);
        f[a] = g[a] - R(b[a]);
    }
    uint32_t sy = dim[0], sz = dim[0] * dim[1];
    auto d = [&](uint32_t x, uint32_t y, uint32_t z) { return R(S.dense[off + x + y * sy + z * sz].distance); };
    R c00 = mixs(d(b[0], b[1], b[2]), d(nx[0], b[1], b[2]), f[0]);
    R c10 = mixs(d(b[0], nx[1], b[2]), d(nx[0], nx[1], b[2]), f[0]);
    R c01 = mixs(d(b[0], b[1], nx[2]), d(nx[0], b[1], nx[2]), f[0]);
    R c11 = mixs(d(b[0], nx[1], nx[2]), d(nx[0], nx[1], nx[2]), f[0]);
    R c0 = mixs(c00, c10, f[1]);
    R c1 = mixs(c01, c11, f[1]);
    return mixs(c0, c1, f[2]);
}

// sampleVolumeMaterial (:1957-1990): the nearest sample, or the volume's
// material when the sample range is invalid (nullptr here).
template <class R> const or_volume_sample *denseNearest(const SdfScene &S, const or_volume &vol, V3<R> lp) {
    uint32_t dim[3] = {vol.dimensions.x, vol.dimensions.y, vol.dimensions.z};
    uint32_t off = vol.metadata.x;
    uint32_t n = dim[0] * dim[1] * dim[2];
    if (off >= S.denseCount || off + n > S.denseCount) return nullptr;
    V3<R> ext = vmax(v3<R>(vol.localBoundsMax) - v3<R>(vol.localBoundsMin), V3<R>(R(1e-6)));
    V3<R> uvw = vclamp((lp - v3<R>(vol.localBoundsMin)) / ext, V3<R>(0), V3<R>(1));
    R u[3] = {uvw.x, uvw.y, uvw.z};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = std::min(uint32_t(std::round(u[a] * R(dim[a] - 1u))), dim[a] - 1u);
    return &S.dense[off + q[0] + q[1] * dim[0] + q[2] * dim[0] * dim[1]];
}

// volumeNormal (:2029-2055): central differences one sample apart.
template <class R> V3<R> denseNormal(const SdfScene &S, const or_volume &vol, V3<R> lp) {
    V3<R> ext = vmax(v3<R>(vol.localBoundsMax) - v3<R>(vol.localBoundsMin), V3<R>(R(1e-6)));
    V3<R> sp = ext / vmax(V3<R>(R(vol.dimensions.x), R(vol.dimensions.y), R(vol.dimensions.z)) - V3<R>(1), V3<R>(1));
    R dx = denseDistance(S, vol, lp + V3<R>(sp.x, 0, 0)) - denseDistance(S, vol, lp - V3<R>(sp.x, 0, 0));
    R dy = denseDistance(S, vol, lp + V3<R>(0, sp.y, 0)) - denseDistance(S, vol, lp - V3<R>(0, sp.y, 0));
    R dz = denseDistance(S, vol, lp + V3<R>(0, 0, sp.z)) - denseDistance(S, vol, lp - V3<R>(0, 0, sp.z));
    V3<R> ln(dx, dy, dz);
    if (dot(ln, ln) <= R(1e-10)) ln = V3<R>(0, 1, 0);
    Mat4<R> nt = mat4<R>(&vol.normalTransform0);
    return normalize(mul(nt, normalize(ln), R(0)));
}

// intersectVolume (:2471-2606): march in world t, <= 512 steps, 10 bisections.
template <class R>
bool intersectDenseVolume(const SdfScene &S, const or_volume &vol, const Ray<R> &ray, R closestT, uint64_t *sc,
                          VolHit<R> &hit) {
    R tNear, tFar;
    if (!intersectAABB(ray, v3<R>(vol.worldBoundsMin), v3<R>(vol.worldBoundsMax), closestT, tNear, tFar)) return false;
    Mat4<R> w2l = mat4<R>(&vol.worldToLocal0);
    V3<R> localStep = mul(w2l, ray.direction, R(0));
    R lupw = fmx(length(localStep), R(1e-6));
    V3<R> ext = vmax(v3<R>(vol.localBoundsMax) - v3<R>(vol.localBoundsMin), V3<R>(R(1e-6)));
    R cellSize = fmn(ext.x / fmx(R(vol.dimensions.x) - R(1), R(1)),
                     fmn(ext.y / fmx(R(vol.dimensions.y) - R(1), R(1)), ext.z / fmx(R(vol.dimensions.z) - R(1), R(1))));
    R minimumStep = fmx(cellSize * R(0.10) / lupw, R(0.0005));
    R prevT = tNear;
    V3<R> prevW = ray.origin + ray.direction * prevT;
    R prevD = denseDistance(S, vol, mul(w2l, prevW, R(1)));
    if (!std::isfinite(prevD)) return false;
    R t = prevT + fmx(std::fabs(prevD) / lupw * R(0.9), minimumStep);
    for (uint32_t step = 0; step < 512u && t <= tFar && t < closestT; ++step) {
        sc[SC_DENSE_STEPS]++;
        V3<R> w = ray.origin + ray.direction * t;
        R d = denseDistance(S, vol, mul(w2l, w, R(1)));
        if (!std::isfinite(d)) return false;
        bool crossed = (prevD > R(0) && d <= R(0)) || (prevD < R(0) && d >= R(0));
        if (crossed) {
            R lowT = prevT, highT = t, lowD = prevD, refinedT = t;
            for (uint32_t k = 0; k < 10u; ++k) {
                R midT = R(0.5) * (lowT + highT);
                R md = denseDistance(S, vol, mul(w2l, ray.origin + ray.direction * midT, R(1)));
                if ((lowD > R(0) && md > R(0)) || (lowD < R(0) && md < R(0))) {
                    lowT = midT;
                    lowD = md;
                } else {
                    highT = midT;
                }
                refinedT = R(0.5) * (lowT + highT);
            }
            V3<R> rw = ray.origin + ray.direction * refinedT;
            V3<R> rl = mul(w2l, rw, R(1));
            V3<R> n = denseNormal(S, vol, rl);
            bool ff = dot(n, ray.direction) <= R(0);
            if (!ff) n = -n;
            const or_volume_sample *ms = denseNearest(S, vol, rl);
            hit.hit = true;
            hit.t = refinedT;
            hit.position = rw;
            hit.normal = n;
            hit.materialA = ms ? ms->materialA : vol.dimensions.w;
            hit.materialB = ms ? ms->materialB : vol.dimensions.w;
            hit.blend = ms ? clampv(R(ms->materialBlend), R(0), R(1)) : R(0);
            hit.fields = ms;
            hit.objectID = vol.metadata.y;
            hit.primitiveID = vol.metadata.z;
            hit.frontFacing = ff;
            return true;
        }
        prevT = t;
        prevD = d;
        t += fmx(std::fabs(d) / lupw * R(0.8), minimumStep);
    }
    return false;
}

// closestVolumeHit (:3090-3163): dense volumes, then the brick-grid branch
// for sparse volumes.
template <class R> VolHit<R> closestVolumeHit(const SdfScene &S, const Ray<R> &ray, R closestT, uint64_t *sc) {
    VolHit<R> closest{};
    closest.hit = false;
    closest.t = closestT;
    for (uint32_t v = 0; v < S.volumeCount; ++v) {
        if (S.volumes[v].metadata.w != 0u) continue;
        sc[SC_DENSE_TESTS]++;
        VolHit<R> cand{};
        if (intersectDenseVolume(S, S.volumes[v], ray, closest.t, sc, cand) && cand.t < closest.t) closest = cand;
    }
    if (S.gridCount == 0u || S.gridIndexCount == 0u) return closest;
    for (uint32_t v = 0; v < S.volumeCount; ++v) {
        if (S.volumes[v].metadata.w != 1u) continue;
        VolHit<R> cand = closestSparseVolumeBrickDDAHit(S, v, ray, closest.t, sc);
        if (cand.hit && cand.t < closest.t) closest = cand;
    }
    return closest;
}
