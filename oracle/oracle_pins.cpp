/*
 * oracle_pins.cpp -- C entry points into the oracle's BSDF, phase-function
 * and free-flight functions (oracle_bsdf.h), so that tests can check them
 * against what the mathematics fixes: normalisation by quadrature, sampling
 * histograms against the pdfs, reciprocity, white-furnace energy, closed-form
 * Fresnel, Henyey-Greenstein moments, unbiased free-flight weights.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Every function takes `prec` (32 or
 * 64) and runs the oracle template at that precision; inputs and outputs are
 * doubles. `frame` is the shading frame as 9 doubles: normal, tangent,
 * bitangent. Per-sample random streams are the reference's lowbias32 chain
 * seeded with seed ^ hash(i + 1). Returns 0, or -1 for an unknown precision.
 */
#include "oracle_bsdf.h"

#include <cstddef>

namespace {

template <class R> V3<R> d3(const double *p) { return V3<R>(R(p[0]), R(p[1]), R(p[2])); }
template <class R> void put3(double *p, V3<R> v) { p[0] = double(v.x); p[1] = double(v.y); p[2] = double(v.z); }
inline uint32_t streamSeed(uint32_t seed, size_t i) { return seed ^ hashU32(uint32_t(i) + 1u); }

template <class R> int brdf(const or_material *om, const double *fr, const double *v, const double *l, size_t n, double *out) {
    Mat<R> m = toMat<R>(*om);
    for (size_t i = 0; i < n; ++i)
        put3(out + 3 * i, evaluateMaterialBRDF(m, d3<R>(fr), d3<R>(fr + 3), d3<R>(fr + 6), d3<R>(v + 3 * i), d3<R>(l + 3 * i)));
    return 0;
}
template <class R> int bsdfPdf(const or_material *om, const double *fr, const double *v, const double *l, size_t n, double *out) {
    Mat<R> m = toMat<R>(*om);
    for (size_t i = 0; i < n; ++i)
        out[i] = double(materialBSDFPDF(m, d3<R>(fr), d3<R>(fr + 3), d3<R>(fr + 6), d3<R>(v + 3 * i), d3<R>(l + 3 * i)));
    return 0;
}
template <class R> int diffuseWeight(const or_material *om, const double *fr, const double *v, const double *l, size_t n, double *out) {
    Mat<R> m = toMat<R>(*om);
    for (size_t i = 0; i < n; ++i)
        put3(out + 3 * i, materialDiffuseBounceWeight(m, d3<R>(fr), d3<R>(v + 3 * i), d3<R>(l + 3 * i)));
    return 0;
}
template <class R> int lobes(const or_material *om, double *out) {
    Mat<R> m = toMat<R>(*om);
    R pd, ps, pc;
    materialLobeProbabilities(m, pd, ps, pc);
    out[0] = double(pd); out[1] = double(ps); out[2] = double(pc);
    return 0;
}
template <class R> int ggxPdf(const double *fr, double rough, double aniso, const double *v, const double *l, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i)
        out[i] = double(ggxReflectionPDF(d3<R>(fr), d3<R>(fr + 3), d3<R>(fr + 6), d3<R>(v + 3 * i), d3<R>(l + 3 * i), R(rough), R(aniso)));
    return 0;
}
template <class R> int ggxD(int aniso, const double *fr, double rough, double an, const double *h, size_t n, double *out) {
    V3<R> nn = d3<R>(fr), t = d3<R>(fr + 3), b = d3<R>(fr + 6);
    for (size_t i = 0; i < n; ++i) {
        V3<R> hh = d3<R>(h + 3 * i);
        out[i] = aniso ? double(distributionGGXAnisotropic(hh, nn, t, b, R(rough), R(an)))
                       : double(distributionGGX(fmx(dot(nn, hh), R(0)), R(rough)));
    }
    return 0;
}
template <class R> int ggxG1(int aniso, const double *fr, double rough, double an, const double *d, size_t n, double *out) {
    V3<R> nn = d3<R>(fr), t = d3<R>(fr + 3), b = d3<R>(fr + 6);
    for (size_t i = 0; i < n; ++i) {
        V3<R> dd = d3<R>(d + 3 * i);
        out[i] = aniso ? double(geometrySmithG1GGXAnisotropic(dd, nn, t, b, R(rough), R(an)))
                       : double(geometrySmithG1GGX(fmx(dot(nn, dd), R(0)), R(rough)));
    }
    return 0;
}
template <class R> int vndf(const double *fr, double rough, double aniso, const double *v, const double *u, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i)
        put3(out + 3 * i, sampleGGXVisibleNormal(R(u[2 * i]), R(u[2 * i + 1]), d3<R>(fr), d3<R>(fr + 3), d3<R>(fr + 6),
                                                 d3<R>(v), R(rough), R(aniso)));
    return 0;
}
template <class R> int ggxHalf(const double *normal, double rough, const double *u, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i)
        put3(out + 3 * i, sampleGGXHalfVector(R(u[2 * i]), R(u[2 * i + 1]), d3<R>(normal), R(rough)));
    return 0;
}
template <class R> int bounce(const or_material *om, const double *fr, const double *incoming, uint32_t seed, size_t n,
                              double *dir, double *weight) {
    Mat<R> m = toMat<R>(*om);
    for (size_t i = 0; i < n; ++i) {
        uint32_t state = streamSeed(seed, i);
        V3<R> T(1);
        V3<R> next = sampleMaterialBounce(m, d3<R>(fr), d3<R>(fr + 3), d3<R>(fr + 6), d3<R>(incoming), state, T);
        put3(dir + 3 * i, next);
        put3(weight + 3 * i, T);
    }
    return 0;
}
template <class R> int fresnel(const double *c, double etaI, double etaT, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) out[i] = double(dielectricFresnel(R(c[i]), R(etaI), R(etaT)));
    return 0;
}
template <class R> int hg(double g, const double *dir, const double *u, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) put3(out + 3 * i, sampleHenyeyGreenstein(R(u[2 * i]), R(u[2 * i + 1]), d3<R>(dir), R(g)));
    return 0;
}
template <class R> int freeFlight(const double *sigmaT, const double *u, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) out[i] = double(sampleSubsurfaceFreeFlight(d3<R>(sigmaT), R(u[2 * i]), R(u[2 * i + 1])));
    return 0;
}
template <class R> int sssPdfs(const double *sigmaT, const double *d, size_t n, double *coll, double *bound) {
    for (size_t i = 0; i < n; ++i) {
        coll[i] = double(subsurfaceCollisionPDF(d3<R>(sigmaT), R(d[i])));
        bound[i] = double(subsurfaceBoundaryPDF(d3<R>(sigmaT), R(d[i])));
    }
    return 0;
}
template <class R> int sssWeights(const double *sigmaT, const double *albedo, const double *d, size_t n, double *coll, double *bound) {
    for (size_t i = 0; i < n; ++i) {
        put3(coll + 3 * i, subsurfaceCollisionWeight(d3<R>(sigmaT), d3<R>(albedo), R(d[i])));
        put3(bound + 3 * i, subsurfaceBoundaryWeight(d3<R>(sigmaT), R(d[i])));
    }
    return 0;
}
template <class R> Hit<R> surfaceHit(const double *normal, int frontFacing, double t) {
    Hit<R> h = emptyHit<R>();
    h.hit = true;
    h.t = R(t);
    h.position = V3<R>(0);
    h.normal = d3<R>(normal);
    h.frontFacing = frontFacing != 0;
    return h;
}
template <class R> int transmissive(const or_material *om, const double *normal, int frontFacing, double t, const double *incoming,
                                    uint32_t seed, size_t n, double *dir, double *weight, int32_t *taken) {
    Mat<R> m = toMat<R>(*om);
    Hit<R> h = surfaceHit<R>(normal, frontFacing, t);
    for (size_t i = 0; i < n; ++i) {
        uint32_t state = streamSeed(seed, i);
        Ray<R> ray{V3<R>(0), d3<R>(incoming)};
        V3<R> T(1);
        R prevPdf = R(-1);
        bool ok = continueTransmissiveSurface(h, m, ray, T, prevPdf, state);
        taken[i] = ok ? 1 : 0;
        put3(dir + 3 * i, ray.direction);
        put3(weight + 3 * i, T);
    }
    return 0;
}
template <class R> int volumeWalk(const or_material *om, double t, const double *incoming, uint32_t seed, size_t n, double *weight,
                                  int32_t *scattered) {
    Mat<R> m = toMat<R>(*om);
    const double up[3] = {0, 0, 1};
    Hit<R> h = surfaceHit<R>(up, 0, t);
    for (size_t i = 0; i < n; ++i) {
        uint32_t state = streamSeed(seed, i);
        Ray<R> ray{V3<R>(0), d3<R>(incoming)};
        V3<R> T(1);
        R prevPdf = R(-1);
        V3<R> before = ray.origin;
        bool cont = continueTransmissionVolumeRandomWalk(h, m, ray, T, prevPdf, state);
        // A scattering event moved the origin; an escape left it in place.
        bool moved = ray.origin.x != before.x || ray.origin.y != before.y || ray.origin.z != before.z;
        scattered[i] = moved ? (cont ? 1 : 2) : 0;
        put3(weight + 3 * i, T);
    }
    return 0;
}

}  // namespace

// The library is built with -fvisibility=hidden; only these entry points are exported.
#pragma GCC visibility push(default)
extern "C" {

int orp_brdf(int prec, const or_material *m, const double *frame, const double *v, const double *l, size_t n, double *out) {
    return prec == 32 ? brdf<float>(m, frame, v, l, n, out) : prec == 64 ? brdf<double>(m, frame, v, l, n, out) : -1;
}
int orp_bsdf_pdf(int prec, const or_material *m, const double *frame, const double *v, const double *l, size_t n, double *out) {
    return prec == 32 ? bsdfPdf<float>(m, frame, v, l, n, out) : prec == 64 ? bsdfPdf<double>(m, frame, v, l, n, out) : -1;
}
int orp_diffuse_bounce_weight(int prec, const or_material *m, const double *frame, const double *v, const double *l, size_t n,
                              double *out) {
    return prec == 32 ? diffuseWeight<float>(m, frame, v, l, n, out) : prec == 64 ? diffuseWeight<double>(m, frame, v, l, n, out) : -1;
}
int orp_lobe_probabilities(int prec, const or_material *m, double *out3) {
    return prec == 32 ? lobes<float>(m, out3) : prec == 64 ? lobes<double>(m, out3) : -1;
}
int orp_ggx_reflection_pdf(int prec, const double *frame, double roughness, double anisotropy, const double *v, const double *l,
                           size_t n, double *out) {
    return prec == 32 ? ggxPdf<float>(frame, roughness, anisotropy, v, l, n, out)
         : prec == 64 ? ggxPdf<double>(frame, roughness, anisotropy, v, l, n, out) : -1;
}
int orp_ggx_d(int prec, int anisotropic, const double *frame, double roughness, double anisotropy, const double *h, size_t n, double *out) {
    return prec == 32 ? ggxD<float>(anisotropic, frame, roughness, anisotropy, h, n, out)
         : prec == 64 ? ggxD<double>(anisotropic, frame, roughness, anisotropy, h, n, out) : -1;
}
int orp_ggx_g1(int prec, int anisotropic, const double *frame, double roughness, double anisotropy, const double *d, size_t n, double *out) {
    return prec == 32 ? ggxG1<float>(anisotropic, frame, roughness, anisotropy, d, n, out)
         : prec == 64 ? ggxG1<double>(anisotropic, frame, roughness, anisotropy, d, n, out) : -1;
}
int orp_sample_vndf(int prec, const double *frame, double roughness, double anisotropy, const double *v, const double *u, size_t n,
                    double *outH) {
    return prec == 32 ? vndf<float>(frame, roughness, anisotropy, v, u, n, outH)
         : prec == 64 ? vndf<double>(frame, roughness, anisotropy, v, u, n, outH) : -1;
}
int orp_sample_ggx_half_vector(int prec, const double *normal, double roughness, const double *u, size_t n, double *outH) {
    return prec == 32 ? ggxHalf<float>(normal, roughness, u, n, outH) : prec == 64 ? ggxHalf<double>(normal, roughness, u, n, outH) : -1;
}
int orp_sample_bounce(int prec, const or_material *m, const double *frame, const double *incoming, uint32_t seed, size_t n,
                      double *outDir, double *outWeight) {
    return prec == 32 ? bounce<float>(m, frame, incoming, seed, n, outDir, outWeight)
         : prec == 64 ? bounce<double>(m, frame, incoming, seed, n, outDir, outWeight) : -1;
}
int orp_dielectric_fresnel(int prec, const double *cosI, double etaI, double etaT, size_t n, double *out) {
    return prec == 32 ? fresnel<float>(cosI, etaI, etaT, n, out) : prec == 64 ? fresnel<double>(cosI, etaI, etaT, n, out) : -1;
}
int orp_sample_hg(int prec, double g, const double *dir, const double *u, size_t n, double *out3) {
    return prec == 32 ? hg<float>(g, dir, u, n, out3) : prec == 64 ? hg<double>(g, dir, u, n, out3) : -1;
}
int orp_sss_free_flight(int prec, const double *sigmaT, const double *u, size_t n, double *out) {
    return prec == 32 ? freeFlight<float>(sigmaT, u, n, out) : prec == 64 ? freeFlight<double>(sigmaT, u, n, out) : -1;
}
int orp_sss_pdfs(int prec, const double *sigmaT, const double *d, size_t n, double *coll, double *bound) {
    return prec == 32 ? sssPdfs<float>(sigmaT, d, n, coll, bound) : prec == 64 ? sssPdfs<double>(sigmaT, d, n, coll, bound) : -1;
}
int orp_sss_weights(int prec, const double *sigmaT, const double *albedo, const double *d, size_t n, double *coll3, double *bound3) {
    return prec == 32 ? sssWeights<float>(sigmaT, albedo, d, n, coll3, bound3)
         : prec == 64 ? sssWeights<double>(sigmaT, albedo, d, n, coll3, bound3) : -1;
}
int orp_transmissive_surface(int prec, const or_material *m, const double *normal, int frontFacing, double t, const double *incoming,
                             uint32_t seed, size_t n, double *outDir, double *outWeight, int32_t *taken) {
    return prec == 32 ? transmissive<float>(m, normal, frontFacing, t, incoming, seed, n, outDir, outWeight, taken)
         : prec == 64 ? transmissive<double>(m, normal, frontFacing, t, incoming, seed, n, outDir, outWeight, taken) : -1;
}
int orp_transmission_volume_walk(int prec, const or_material *m, double t, const double *incoming, uint32_t seed, size_t n,
                                 double *outWeight, int32_t *scattered) {
    return prec == 32 ? volumeWalk<float>(m, t, incoming, seed, n, outWeight, scattered)
         : prec == 64 ? volumeWalk<double>(m, t, incoming, seed, n, outWeight, scattered) : -1;
}

}  // extern "C"
#pragma GCC visibility pop
