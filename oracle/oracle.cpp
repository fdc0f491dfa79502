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

namespace {

// ---------------------------------------------------------------- vectors --
template <class R> struct V2 { R x, y; };
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R a) : x(a), y(a), z(a) {}
};
template <class R> struct V4 {
    R x, y, z, w;
    V3<R> xyz() const { return V3<R>(x, y, z); }
};

template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return V3<R>(s * a.x, s * a.y, s * a.z); }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template <class R> inline V3<R> operator/(R s, V3<R> a) { return V3<R>(s / a.x, s / a.y, s / a.z); }
template <class R> inline V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }
template <class R> inline V3<R> &operator+=(V3<R> &a, V3<R> b) { a = a + b; return a; }
template <class R> inline V3<R> &operator*=(V3<R> &a, V3<R> b) { a = a * b; return a; }
template <class R> inline V3<R> &operator*=(V3<R> &a, R s) { a = a * s; return a; }
template <class R> inline V3<R> &operator/=(V3<R> &a, R s) { a = a / s; return a; }

template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / std::sqrt(dot(a, a))); }
template <class R> inline R fmn(R a, R b) { return std::fmin(a, b); }
template <class R> inline R fmx(R a, R b) { return std::fmax(a, b); }
template <class R> inline R clampv(R x, R lo, R hi) { return fmn(fmx(x, lo), hi); }
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return V3<R>(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return V3<R>(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
template <class R> inline V3<R> vclamp(V3<R> x, V3<R> lo, V3<R> hi) { return vmin(vmax(x, lo), hi); }
template <class R> inline R mixs(R a, R b, R t) { return a + (b - a) * t; }
template <class R> inline V3<R> mixv(V3<R> a, V3<R> b, R t) { return a + (b - a) * t; }
template <class R> inline V4<R> mix4(V4<R> a, V4<R> b, R t) {
    return V4<R>{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t};
}
template <class R> inline V3<R> vexp(V3<R> a) { return V3<R>(std::exp(a.x), std::exp(a.y), std::exp(a.z)); }
template <class R> inline V3<R> vlog(V3<R> a) { return V3<R>(std::log(a.x), std::log(a.y), std::log(a.z)); }
template <class R> inline V3<R> vcos(V3<R> a) { return V3<R>(std::cos(a.x), std::cos(a.y), std::cos(a.z)); }
template <class R> inline R maxComponent(V3<R> v) { return fmx(v.x, fmx(v.y, v.z)); }  // :326
template <class R> inline bool allFinite(V3<R> v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - R(2) * dot(n, i) * n; }
template <class R> inline V3<R> refract(V3<R> i, V3<R> n, R eta) {
    R d = dot(n, i);
    R k = R(1) - eta * eta * (R(1) - d * d);
    if (k < R(0)) return V3<R>(R(0));
    return eta * i - (eta * d + std::sqrt(k)) * n;
}
template <class R> inline R fractv(R x) {
    const R below1 = R(1) - std::numeric_limits<R>::epsilon() / R(2);
    return fmn(x - std::floor(x), below1);
}
template <class R> inline V3<R> v3(const or_float4 &f) { return V3<R>(R(f.x), R(f.y), R(f.z)); }
template <class R> inline V4<R> v4(const or_float4 &f) { return V4<R>{R(f.x), R(f.y), R(f.z), R(f.w)}; }

template <class R> constexpr R kInf() { return std::numeric_limits<R>::infinity(); }

// -------------------------------------------------------------------- RNG --
// PathTrace.metal:264-276 (Wellons lowbias32, 24-bit float).
inline uint32_t hashU32(uint32_t value) {
    value ^= value >> 16;
    value *= 0x7feb352du;
    value ^= value >> 15;
    value *= 0x846ca68bu;
    value ^= value >> 16;
    return value;
}
template <class R> inline R randomFloat(uint32_t &state) {
    state = hashU32(state);
    return R(float(state & 0x00ffffffu) / float(0x01000000u));
}

// --------------------------------------------------------------- material --
template <class R> struct Mat {
    V4<R> baseColor, emission, parameters, parameters2, specularColor, sheenColor,
        transmissionColor, parameters3, clearcoatColor, clearcoatAttenuation,
        transmissionAbsorption, thinFilm, subsurfaceColor, subsurfaceRadius,
        subsurfaceParameters, volumeScattering, volumeParameters;
};
template <class R> Mat<R> toMat(const or_material &m) {
    return Mat<R>{v4<R>(m.baseColor), v4<R>(m.emission), v4<R>(m.parameters), v4<R>(m.parameters2),
                  v4<R>(m.specularColor), v4<R>(m.sheenColor), v4<R>(m.transmissionColor),
                  v4<R>(m.parameters3), v4<R>(m.clearcoatColor), v4<R>(m.clearcoatAttenuation),
                  v4<R>(m.transmissionAbsorption), v4<R>(m.thinFilm), v4<R>(m.subsurfaceColor),
                  v4<R>(m.subsurfaceRadius), v4<R>(m.subsurfaceParameters), v4<R>(m.volumeScattering),
                  v4<R>(m.volumeParameters)};
}
template <class R> Mat<R> zeroMat() { Mat<R> m; std::memset(&m, 0, sizeof(m)); return m; }

// ----------------------------------------------------------- ray and hit --
template <class R> struct Ray { V3<R> origin, direction; };
template <class R> struct Hit {  // PathTrace.metal:229-253 (triangle and sparse-volume fields)
    bool hit;
    R t;
    V3<R> position, normal;
    V2<R> uv;
    V3<R> tangent, bitangent;
    uint32_t materialID, materialID2;
    R materialBlend;
    uint32_t objectID, primitiveID;
    bool frontFacing;
    // Dense-volume material fields (volumeBaseColorOpacity ... materialFieldFlags).
    V4<R> fieldBaseColorOpacity, fieldEmissionTransmission, fieldSurface;
    uint32_t fieldFlags;
};
template <class R> Hit<R> emptyHit() {  // :1883-1909
    Hit<R> h;
    h.hit = false;
    h.t = kInf<R>();
    h.position = V3<R>(0);
    h.normal = V3<R>(0, 1, 0);
    h.uv = V2<R>{0, 0};
    h.tangent = V3<R>(1, 0, 0);
    h.bitangent = V3<R>(0, 0, 1);
    h.materialID = 0;
    h.materialID2 = 0;
    h.materialBlend = R(0);
    h.objectID = 0;
    h.primitiveID = 0;
    h.frontFacing = true;
    h.fieldBaseColorOpacity = h.fieldEmissionTransmission = h.fieldSurface = V4<R>{0, 0, 0, 0};
    h.fieldFlags = 0u;
    return h;
}

// -------------------------------------------------------------- sampling --
template <class R> V3<R> cosineHemisphere(R ux, R uy) {  // :278-282
    R r = std::sqrt(ux);
    R phi = R(6.28318530718) * uy;
    return V3<R>(r * std::cos(phi), r * std::sin(phi), std::sqrt(fmx(R(0), R(1) - ux)));
}
template <class R> V2<R> uniformDisk(R ux, R uy) {  // :284-288
    R r = std::sqrt(ux);
    R phi = R(6.28318530718) * uy;
    return V2<R>{r * std::cos(phi), r * std::sin(phi)};
}
template <class R> V3<R> uniformSphere(R ux, R uy) {  // :290-295
    R z = R(1) - R(2) * ux;
    R r = std::sqrt(fmx(R(0), R(1) - z * z));
    R phi = R(6.28318530718) * uy;
    return V3<R>(r * std::cos(phi), r * std::sin(phi), z);
}
template <class R> V3<R> orientHemisphere(V3<R> local, V3<R> normal) {  // :297-302
    V3<R> helper = std::fabs(normal.y) < R(0.999) ? V3<R>(0, 1, 0) : V3<R>(1, 0, 0);
    V3<R> tangent = normalize(cross(helper, normal));
    V3<R> bitangent = cross(normal, tangent);
    return normalize(local.x * tangent + local.y * bitangent + local.z * normal);
}
template <class R> V3<R> orientAroundDirection(V3<R> local, V3<R> direction) {  // :304-309
    return orientHemisphere(local, direction);
}
template <class R> V3<R> sampleHenyeyGreenstein(R ux, R uy, V3<R> direction, R anisotropy) {  // :311-324
    R g = clampv(anisotropy, R(-0.95), R(0.95));
    if (std::fabs(g) < R(1e-4)) return uniformSphere(ux, uy);
    R oneMinusGSquared = R(1) - g * g;
    R denominator = R(1) - g + R(2) * g * ux;
    R q = oneMinusGSquared / fmx(denominator, R(1e-5));
    R cosTheta = (R(1) + g * g - q * q) / (R(2) * g);
    cosTheta = clampv(cosTheta, R(-1), R(1));
    R sinTheta = std::sqrt(fmx(R(0), R(1) - cosTheta * cosTheta));
    R phi = R(6.28318530718) * uy;
    return orientAroundDirection(V3<R>(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta), direction);
}

template <class R> R luminance(V3<R> v) { return dot(v, V3<R>(R(0.2126), R(0.7152), R(0.0722))); }  // :348

template <class R> V3<R> clampSampleContribution(V3<R> c, R sampleRadianceClamp) {  // :330-346
    if (!allFinite(c)) return V3<R>(0);
    R clampValue = fmx(sampleRadianceClamp, R(0));
    if (clampValue <= R(0)) return c;
    R peak = maxComponent(c);
    if (peak <= clampValue) return c;
    return c * (clampValue / fmx(peak, R(1e-5)));
}

// --------------------------------------------------------------- fresnel --
template <class R> V3<R> fresnelSchlick(R cosTheta, V3<R> f0) {  // :352-355
    R factor = std::pow(clampv(R(1) - cosTheta, R(0), R(1)), R(5));
    return f0 + (V3<R>(1) - f0) * factor;
}
template <class R> V3<R> thinFilmTint(const Mat<R> &m, R cosTheta) {  // :357-373
    R strength = clampv(m.thinFilm.x, R(0), R(1));
    R thickness = fmx(m.thinFilm.y, R(0));
    if (strength <= R(0) || thickness <= R(0)) return V3<R>(1);
    R filmIOR = fmx(m.thinFilm.z, R(1.0001));
    R angle = std::pow(clampv(R(1) - cosTheta, R(0), R(1)), R(0.65));
    R opticalPath = R(2) * filmIOR * thickness * (R(0.62) + R(0.42) * angle);
    V3<R> wavelengths(R(650), R(510), R(475));
    V3<R> phase = (R(6.28318530718) * opticalPath) / wavelengths;
    V3<R> interference = V3<R>(R(0.62)) + R(0.38) * vcos(phase);
    R normalizedLuminance = fmx(luminance(interference), R(0.35));
    V3<R> tint = vclamp(interference / normalizedLuminance, V3<R>(R(0.35)), V3<R>(R(1.8)));
    return mixv(V3<R>(1), tint, strength);
}
template <class R> V3<R> fresnelSchlickThinFilm(R cosTheta, V3<R> f0, const Mat<R> &m) {  // :375-377
    return fresnelSchlick(cosTheta, f0) * thinFilmTint(m, cosTheta);
}
template <class R> R dielectricFresnel(R cosThetaI, R etaI, R etaT) {  // :379-393
    cosThetaI = clampv(cosThetaI, R(0), R(1));
    R eta = etaI / etaT;
    R sinThetaTSquared = eta * eta * fmx(R(0), R(1) - cosThetaI * cosThetaI);
    if (sinThetaTSquared >= R(1)) return R(1);
    R cosThetaT = std::sqrt(fmx(R(0), R(1) - sinThetaTSquared));
    R parallel = ((etaT * cosThetaI) - (etaI * cosThetaT)) / fmx((etaT * cosThetaI) + (etaI * cosThetaT), R(1e-5));
    R perpendicular = ((etaI * cosThetaI) - (etaT * cosThetaT)) / fmx((etaI * cosThetaI) + (etaT * cosThetaT), R(1e-5));
    return clampv(R(0.5) * (parallel * parallel + perpendicular * perpendicular), R(0), R(1));
}
template <class R> R dielectricF0(R ior, R specularWeight) {  // :395-399
    R i = fmx(ior, R(1));
    R f0 = (i - R(1)) / (i + R(1));
    return f0 * f0 * clampv(specularWeight, R(0), R(1));
}

// ----------------------------------------------------- material accessors --
template <class R> V3<R> materialF0(const Mat<R> &m) {  // :401-406
    R metallic = clampv(m.parameters.y, R(0), R(1));
    R dielectric = dielectricF0(m.parameters2.y, m.parameters2.x);
    V3<R> dielectricColor = dielectric * vclamp(m.specularColor.xyz(), V3<R>(0), V3<R>(1));
    return mixv(dielectricColor, m.baseColor.xyz(), metallic);
}
template <class R> R materialClearcoat(const Mat<R> &m) { return clampv(m.parameters2.z, R(0), R(1)); }  // :408
template <class R> R materialClearcoatRoughness(const Mat<R> &m) { return clampv(m.parameters2.w, R(0.02), R(1)); }  // :412
template <class R> R materialClearcoatThickness(const Mat<R> &m) { return fmx(m.clearcoatAttenuation.w, R(0)); }  // :416
template <class R> R materialClearcoatF0(const Mat<R> &m) { return dielectricF0(m.specularColor.w, R(1)); }  // :420
template <class R> V3<R> materialClearcoatF0Color(const Mat<R> &m) {  // :424-427
    return materialClearcoatF0(m) * vclamp(m.clearcoatColor.xyz(), V3<R>(0), V3<R>(1));
}
template <class R> V3<R> materialClearcoatAttenuationColor(const Mat<R> &m) {  // :429-431
    return vclamp(m.clearcoatAttenuation.xyz(), V3<R>(R(0.0001)), V3<R>(1));
}
template <class R> V3<R> materialClearcoatDepthAttenuation(const Mat<R> &m, V3<R> n, V3<R> v, V3<R> l) {  // :433-448
    R thickness = materialClearcoatThickness(m) * materialClearcoat(m);
    if (thickness <= R(1e-5)) return V3<R>(1);
    R nDotV = fmx(dot(n, v), R(0.05));
    R nDotL = fmx(dot(n, l), R(0.05));
    R pathLength = thickness * R(0.5) * (R(1) / nDotV + R(1) / nDotL);
    return vexp(vlog(materialClearcoatAttenuationColor(m)) * pathLength);
}
template <class R> R materialSpecularAnisotropy(const Mat<R> &m) { return clampv(m.parameters3.w, R(-0.95), R(0.95)); }  // :450
template <class R> R materialTransmission(const Mat<R> &m) { return clampv(m.emission.w, R(0), R(1)); }  // :454
template <class R> V3<R> materialTransmissionColor(const Mat<R> &m) { return vclamp(m.transmissionColor.xyz(), V3<R>(0), V3<R>(1)); }
template <class R> bool materialThinWalled(const Mat<R> &m) { return m.transmissionColor.w > R(0.5); }  // :462
template <class R> R materialTransmissionRoughness(const Mat<R> &m) { return clampv(m.parameters3.y, R(0.001), R(1)); }
template <class R> R materialTransmissionIOR(const Mat<R> &m) { return fmx(m.parameters3.z, R(1.0001)); }
template <class R> R materialSheen(const Mat<R> &m) { return clampv(m.sheenColor.w, R(0), R(1)); }  // :474
template <class R> V3<R> materialSheenColor(const Mat<R> &m) { return vclamp(m.sheenColor.xyz(), V3<R>(0), V3<R>(1)); }
template <class R> R materialSheenRoughness(const Mat<R> &m) { return clampv(m.parameters3.x, R(0.02), R(1)); }
template <class R> R materialSubsurface(const Mat<R> &m) {  // :486-490
    R metallic = clampv(m.parameters.y, R(0), R(1));
    R transmission = materialTransmission(m);
    return clampv(m.subsurfaceColor.w, R(0), R(1)) * (R(1) - metallic) * (R(1) - transmission);
}
template <class R> V3<R> materialSubsurfaceColor(const Mat<R> &m) { return vclamp(m.subsurfaceColor.xyz(), V3<R>(0), V3<R>(R(0.999))); }
template <class R> V3<R> materialSubsurfaceRadius(const Mat<R> &m) {  // :496-499
    R scale = fmx(m.subsurfaceRadius.w, R(0));
    return vmax(m.subsurfaceRadius.xyz() * scale, V3<R>(R(1e-4)));
}
template <class R> R materialSubsurfaceAnisotropy(const Mat<R> &m) { return clampv(m.subsurfaceParameters.x, R(-0.95), R(0.95)); }
template <class R> V3<R> materialTransmissionTint(const Mat<R> &m, R strength) {  // :505-507
    return mixv(V3<R>(1), materialTransmissionColor(m), strength);
}
template <class R> V3<R> materialTransmissionAbsorptionCoefficient(const Mat<R> &m) {  // :509-517
    R referenceDistance = m.transmissionAbsorption.w;
    if (referenceDistance <= R(1e-5)) return V3<R>(0);
    V3<R> referenceColor = vclamp(m.transmissionAbsorption.xyz(), V3<R>(R(0.0001)), V3<R>(1));
    return (-vlog(referenceColor)) / referenceDistance;
}
template <class R> V3<R> materialTransmissionAbsorption(const Mat<R> &m, R distance, bool applies) {  // :519-525
    if (!applies || distance <= R(0)) return V3<R>(1);
    return vexp((-materialTransmissionAbsorptionCoefficient(m)) * distance);
}
template <class R> R materialVolumeScattering(const Mat<R> &m) { return clampv(m.volumeScattering.w, R(0), R(1)); }
template <class R> V3<R> materialVolumeScatteringColor(const Mat<R> &m) { return vclamp(m.volumeScattering.xyz(), V3<R>(0), V3<R>(1)); }
template <class R> R materialVolumeScatteringDistance(const Mat<R> &m) { return fmx(m.volumeParameters.x, R(1e-4)); }
template <class R> R materialVolumeAnisotropy(const Mat<R> &m) { return clampv(m.volumeParameters.y, R(-0.95), R(0.95)); }
template <class R> bool materialHasVolumeScattering(const Mat<R> &m) {  // :543-547
    return materialTransmission(m) > R(0) && !materialThinWalled(m) && materialVolumeScattering(m) > R(0);
}
template <class R> V3<R> materialVolumeSigmaS(const Mat<R> &m) {  // :1139-1143
    return materialVolumeScatteringColor(m) * materialVolumeScattering(m) / materialVolumeScatteringDistance(m);
}
template <class R> V3<R> materialVolumeSigmaT(const Mat<R> &m) {  // :1145-1147
    return materialTransmissionAbsorptionCoefficient(m) + materialVolumeSigmaS(m);
}
template <class R> V3<R> materialVolumeTransmittance(const Mat<R> &m, R distance) {  // :1149-1151
    return vexp((-materialVolumeSigmaT(m)) * distance);
}

// ------------------------------------------------------------------ GGX --
template <class R> struct Frame { V3<R> tangent, bitangent; };
template <class R> Frame<R> makeTangentFrame(V3<R> normal, V3<R> rawTangent, V3<R> rawBitangent) {  // :1158-1175
    Frame<R> f;
    V3<R> tangent = rawTangent - normal * dot(rawTangent, normal);
    if (dot(tangent, tangent) <= R(1e-8)) {
        V3<R> helper = std::fabs(normal.y) < R(0.999) ? V3<R>(0, 1, 0) : V3<R>(1, 0, 0);
        tangent = cross(helper, normal);
    }
    f.tangent = normalize(tangent);
    V3<R> bitangent = rawBitangent - normal * dot(rawBitangent, normal) - f.tangent * dot(rawBitangent, f.tangent);
    if (dot(bitangent, bitangent) <= R(1e-8)) bitangent = cross(normal, f.tangent);
    f.bitangent = normalize(bitangent);
    return f;
}
template <class R> V2<R> anisotropicGGXAlpha(R roughness, R anisotropy) {  // :1177-1186
    R alpha = fmx(roughness * roughness, R(0.001));
    R amount = clampv(anisotropy, R(-0.95), R(0.95));
    R aspect = std::sqrt(fmx(R(0.1), R(1) - R(0.9) * std::fabs(amount)));
    R alphaLong = alpha / aspect;
    R alphaShort = alpha * aspect;
    return amount >= R(0) ? V2<R>{alphaLong, alphaShort} : V2<R>{alphaShort, alphaLong};
}
template <class R> R distributionGGX(R nDotH, R roughness) {  // :1188-1193
    R alpha = roughness * roughness;
    R alpha2 = alpha * alpha;
    R denominator = nDotH * nDotH * (alpha2 - R(1)) + R(1);
    return alpha2 / fmx(R(3.14159265359) * denominator * denominator, R(1e-5));
}
template <class R> R geometrySmithG1GGX(R nDotV, R roughness) {  // :1195-1200
    R alpha = roughness * roughness;
    R alpha2 = alpha * alpha;
    R denominator = nDotV + std::sqrt(fmx(alpha2 + (R(1) - alpha2) * nDotV * nDotV, R(0)));
    return (R(2) * nDotV) / fmx(denominator, R(1e-5));
}
template <class R> R geometrySmith(R nDotV, R nDotL, R roughness) {  // :1202-1204
    return geometrySmithG1GGX(nDotV, roughness) * geometrySmithG1GGX(nDotL, roughness);
}
template <class R> R distributionGGXAnisotropic(V3<R> h, V3<R> n, V3<R> t, V3<R> b, R roughness, R anisotropy) {  // :1206-1226
    R nDotH = fmx(dot(n, h), R(0));
    if (nDotH <= R(0)) return R(0);
    V2<R> alpha = anisotropicGGXAlpha(roughness, anisotropy);
    R tDotH = dot(t, h);
    R bDotH = dot(b, h);
    R denominator = (tDotH * tDotH) / fmx(alpha.x * alpha.x, R(1e-6)) + (bDotH * bDotH) / fmx(alpha.y * alpha.y, R(1e-6)) +
                    nDotH * nDotH;
    return R(1) / fmx(R(3.14159265359) * alpha.x * alpha.y * denominator * denominator, R(1e-6));
}
template <class R> R geometrySmithG1GGXAnisotropic(V3<R> d, V3<R> n, V3<R> t, V3<R> b, R roughness, R anisotropy) {  // :1228-1251
    R nDotD = fmx(dot(n, d), R(0));
    if (nDotD <= R(0)) return R(0);
    V2<R> alpha = anisotropicGGXAlpha(roughness, anisotropy);
    R tDotD = dot(t, d);
    R bDotD = dot(b, d);
    R root = std::sqrt(fmx(alpha.x * alpha.x * tDotD * tDotD + alpha.y * alpha.y * bDotD * bDotD + nDotD * nDotD, R(0)));
    return (R(2) * nDotD) / fmx(nDotD + root, R(1e-5));
}
template <class R> R geometrySmithAnisotropic(V3<R> v, V3<R> l, V3<R> n, V3<R> t, V3<R> b, R roughness, R anisotropy) {  // :1253-1264
    return geometrySmithG1GGXAnisotropic(v, n, t, b, roughness, anisotropy) *
           geometrySmithG1GGXAnisotropic(l, n, t, b, roughness, anisotropy);
}
template <class R> V3<R> evaluateMaterialSheenBRDF(const Mat<R> &m, V3<R> n, V3<R> v, V3<R> l) {  // :1266-1300
    R sheen = materialSheen(m);
    if (sheen <= R(0)) return V3<R>(0);
    R metallic = clampv(m.parameters.y, R(0), R(1));
    V3<R> halfDirection = v + l;
    if (dot(halfDirection, halfDirection) <= R(1e-8)) return V3<R>(0);
    V3<R> h = normalize(halfDirection);
    R nDotL = fmx(dot(n, l), R(0));
    R nDotV = fmx(dot(n, v), R(0));
    R nDotH = fmx(dot(n, h), R(0));
    if (nDotL <= R(0) || nDotV <= R(0)) return V3<R>(0);
    R sinThetaH = std::sqrt(fmx(R(0), R(1) - nDotH * nDotH));
    R inverseAlpha = mixs(R(20), R(0.5), materialSheenRoughness(m));
    R distribution = (R(2) + inverseAlpha) * std::pow(sinThetaH, inverseAlpha) * R(0.15915494309);
    R visibility = R(1) / fmx(R(4) * (nDotL + nDotV - nDotL * nDotV), R(1e-5));
    return sheen * (R(1) - metallic) * materialSheenColor(m) * distribution * visibility;
}
template <class R> V3<R> materialClearcoatAttenuationForDirection(const Mat<R> &m, V3<R> n, V3<R> v, V3<R> l) {  // :1302-1328
    R clearcoat = materialClearcoat(m);
    if (clearcoat <= R(0)) return V3<R>(1);
    V3<R> halfDirection = v + l;
    if (dot(halfDirection, halfDirection) <= R(1e-8)) return V3<R>(1);
    V3<R> h = normalize(halfDirection);
    R vDotH = fmx(dot(v, h), R(0));
    V3<R> clearcoatFresnel = fresnelSchlickThinFilm(vDotH, materialClearcoatF0Color(m), m);
    V3<R> fresnelAttenuation = vmax(V3<R>(0), V3<R>(1) - clearcoatFresnel * clearcoat);
    return fresnelAttenuation * materialClearcoatDepthAttenuation(m, n, v, l);
}
template <class R> V3<R> evaluateMaterialBRDF(const Mat<R> &m, V3<R> n, V3<R> t, V3<R> b, V3<R> v, V3<R> l) {  // :1330-1390
    const R inversePi = R(0.31830988618);
    R roughness = clampv(m.parameters.x, R(0.02), R(1));
    R metallic = clampv(m.parameters.y, R(0), R(1));
    V3<R> baseColor = m.baseColor.xyz();
    V3<R> h = normalize(v + l);
    R nDotL = fmx(dot(n, l), R(0));
    R nDotV = fmx(dot(n, v), R(0));
    R nDotH = fmx(dot(n, h), R(0));
    R vDotH = fmx(dot(v, h), R(0));
    if (nDotL <= R(0) || nDotV <= R(0)) return V3<R>(0);
    V3<R> f0 = materialF0(m);
    V3<R> fresnel = fresnelSchlickThinFilm(vDotH, f0, m);
    Frame<R> frame = makeTangentFrame(n, t, b);
    R anisotropy = materialSpecularAnisotropy(m);
    R distribution = distributionGGXAnisotropic(h, n, frame.tangent, frame.bitangent, roughness, anisotropy);
    R geometry = geometrySmithAnisotropic(v, l, n, frame.tangent, frame.bitangent, roughness, anisotropy);
    V3<R> specular = distribution * geometry * fresnel / fmx(R(4) * nDotV * nDotL, R(1e-5));
    R clearcoat = materialClearcoat(m);
    R clearcoatRoughness = materialClearcoatRoughness(m);
    V3<R> clearcoatFresnel = fresnelSchlickThinFilm(vDotH, materialClearcoatF0Color(m), m);
    R clearcoatDistribution = distributionGGX(nDotH, clearcoatRoughness);
    R clearcoatGeometry = geometrySmith(nDotV, nDotL, clearcoatRoughness);
    V3<R> clearcoatSpecular =
        clearcoatDistribution * clearcoatGeometry * clearcoatFresnel / fmx(R(4) * nDotV * nDotL, R(1e-5));
    V3<R> clearcoatAttenuation = materialClearcoatAttenuationForDirection(m, n, v, l);
    V3<R> diffuse = clearcoatAttenuation * (V3<R>(1) - fresnel) * (R(1) - metallic) * baseColor * inversePi;
    V3<R> sheen = evaluateMaterialSheenBRDF(m, n, v, l);
    return diffuse + clearcoatAttenuation * (specular + sheen) + clearcoatSpecular * clearcoat;
}
template <class R> R powerHeuristic(R a, R b) {  // :1392-1396
    R a2 = a * a;
    R b2 = b * b;
    return a2 / fmx(a2 + b2, R(1e-8));
}
template <class R> void materialLobeProbabilities(const Mat<R> &m, R &pd, R &ps, R &pc) {  // :1398-1426
    R metallic = clampv(m.parameters.y, R(0), R(1));
    V3<R> baseColor = m.baseColor.xyz();
    V3<R> f0 = materialF0(m);
    R clearcoat = materialClearcoat(m);
    R sheenWeight = materialSheen(m) * luminance(materialSheenColor(m)) * (R(1) - metallic);
    R diffuseWeight = luminance(baseColor * (R(1) - metallic)) + sheenWeight;
    R specularWeight = luminance(f0);
    R clearcoatWeight = clearcoat * luminance(materialClearcoatF0Color(m));
    R totalLobeWeight = fmx(diffuseWeight + specularWeight + clearcoatWeight, R(1e-5));
    pc = clearcoatWeight > R(1e-5) ? clampv(clearcoatWeight / totalLobeWeight, R(0.02), R(0.35)) : R(0);
    R remaining = fmx(R(0), R(1) - pc);
    if (specularWeight <= R(1e-5)) {
        ps = R(0);
    } else if (diffuseWeight <= R(1e-5)) {
        ps = remaining;
    } else {
        ps = clampv(specularWeight / totalLobeWeight, R(0.08), remaining);
    }
    pd = fmx(R(0), R(1) - ps - pc);
}
template <class R> R ggxReflectionPDF(V3<R> n, V3<R> t, V3<R> b, V3<R> v, V3<R> l, R roughness, R anisotropy) {  // :1428-1448
    V3<R> h = normalize(v + l);
    R nDotH = fmx(dot(n, h), R(0));
    R nDotV = fmx(dot(n, v), R(0));
    R vDotH = fmx(dot(v, h), R(0));
    if (nDotH <= R(0) || nDotV <= R(0) || vDotH <= R(0)) return R(0);
    Frame<R> f = makeTangentFrame(n, t, b);
    return distributionGGXAnisotropic(h, n, f.tangent, f.bitangent, roughness, anisotropy) *
           geometrySmithG1GGXAnisotropic(v, n, f.tangent, f.bitangent, roughness, anisotropy) / fmx(R(4) * nDotV, R(1e-5));
}
template <class R> R materialBSDFPDF(const Mat<R> &m, V3<R> n, V3<R> t, V3<R> b, V3<R> v, V3<R> l) {  // :1450-1495
    const R inversePi = R(0.31830988618);
    R nDotL = fmx(dot(n, l), R(0));
    R nDotV = fmx(dot(n, v), R(0));
    if (nDotL <= R(0) || nDotV <= R(0)) return R(0);
    R pd, ps, pc;
    materialLobeProbabilities(m, pd, ps, pc);
    R roughness = clampv(m.parameters.x, R(0.02), R(1));
    R clearcoatRoughness = materialClearcoatRoughness(m);
    R anisotropy = materialSpecularAnisotropy(m);
    R diffusePDF = nDotL * inversePi;
    R specularPDF = ggxReflectionPDF(n, t, b, v, l, roughness, anisotropy);
    R clearcoatPDF = ggxReflectionPDF(n, t, b, v, l, clearcoatRoughness, R(0));
    return pd * diffusePDF + ps * specularPDF + pc * clearcoatPDF;
}
template <class R> V3<R> materialDiffuseBounceWeight(const Mat<R> &m, V3<R> n, V3<R> v, V3<R> l) {  // :1497-1523
    R metallic = clampv(m.parameters.y, R(0), R(1));
    V3<R> baseColor = m.baseColor.xyz();
    V3<R> halfDirection = v + l;
    if (dot(halfDirection, halfDirection) <= R(1e-8)) return V3<R>(0);
    V3<R> h = normalize(halfDirection);
    R vDotH = fmx(dot(v, h), R(0));
    V3<R> fresnel = fresnelSchlickThinFilm(vDotH, materialF0(m), m);
    V3<R> clearcoatAttenuation = materialClearcoatAttenuationForDirection(m, n, v, l);
    V3<R> diffuse = (V3<R>(1) - fresnel) * (R(1) - metallic) * baseColor;
    V3<R> sheen = evaluateMaterialSheenBRDF(m, n, v, l) * R(3.14159265359);
    return clearcoatAttenuation * (diffuse + sheen);
}
template <class R> bool terminatePathWithRussianRoulette(uint32_t bounce, V3<R> &throughput, uint32_t &state) {  // :1525-1546
    R continuationWeight = maxComponent(throughput);
    if (continuationWeight <= R(0)) return true;
    if (bounce < 2u) return false;
    R survival = clampv(continuationWeight, R(0.05), R(0.95));
    if (randomFloat<R>(state) > survival) return true;
    throughput /= survival;
    return false;
}

// ------------------------------------------------------------------ lights --
template <class R> R lightSolidAnglePDF(R distanceSquared, R cosLight, R area) {  // :1548-1550
    return distanceSquared / fmx(cosLight * area, R(1e-6));
}
template <class R> R lightSelectionCDF(const or_light *lights, uint32_t i) { return clampv(R(lights[i].selectionCDF), R(0), R(1)); }
template <class R> R lightSelectionPDF(const or_light *lights, uint32_t i) {  // :1556-1559
    R prev = i == 0u ? R(0) : lightSelectionCDF<R>(lights, i - 1u);
    return fmx(lightSelectionCDF<R>(lights, i) - prev, R(0));
}
template <class R> uint32_t selectLightIndex(const or_light *lights, uint32_t count, R sample) {  // :1561-1578
    if (count <= 1u) return 0u;
    R target = clampv(sample, R(0), R(0.99999994f));
    uint32_t low = 0u, high = count - 1u;
    while (low < high) {
        uint32_t mid = (low + high) >> 1u;
        if (target <= lightSelectionCDF<R>(lights, mid)) high = mid; else low = mid + 1u;
    }
    return low;
}

template <class R> V3<R> sampleGGXHalfVector(R ux, R uy, V3<R> normal, R roughness) {  // :1615-1623
    R alpha = roughness * roughness;
    R alpha2 = alpha * alpha;
    R phi = R(6.28318530718) * ux;
    R cosTheta = std::sqrt((R(1) - uy) / fmx(R(1) + (alpha2 - R(1)) * uy, R(1e-5)));
    R sinTheta = std::sqrt(fmx(R(0), R(1) - cosTheta * cosTheta));
    V3<R> local(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);
    return orientHemisphere(local, normal);
}
template <class R> V3<R> sampleGGXVisibleNormal(R ux, R uy, V3<R> n, V3<R> t, V3<R> b, V3<R> v, R roughness, R anisotropy) {  // :1625-1664
    Frame<R> f = makeTangentFrame(n, t, b);
    V2<R> alpha = anisotropicGGXAlpha(roughness, anisotropy);
    V3<R> localView = normalize(V3<R>(dot(v, f.tangent), dot(v, f.bitangent), fmx(dot(v, n), R(1e-5))));
    V3<R> sv = normalize(V3<R>(alpha.x * localView.x, alpha.y * localView.y, localView.z));
    R lensq = sv.x * sv.x + sv.y * sv.y;
    V3<R> t1 = lensq > R(1e-7) ? V3<R>(-sv.y, sv.x, R(0)) * (R(1) / std::sqrt(lensq)) : V3<R>(1, 0, 0);
    V3<R> t2 = cross(sv, t1);
    R radius = std::sqrt(ux);
    R phi = R(6.28318530718) * uy;
    R p1 = radius * std::cos(phi);
    R p2 = radius * std::sin(phi);
    R blend = R(0.5) * (R(1) + sv.z);
    p2 = mixs(std::sqrt(fmx(R(0), R(1) - p1 * p1)), p2, blend);
    V3<R> vn = p1 * t1 + p2 * t2 + std::sqrt(fmx(R(0), R(1) - p1 * p1 - p2 * p2)) * sv;
    V3<R> lh = normalize(V3<R>(alpha.x * vn.x, alpha.y * vn.y, fmx(vn.z, R(0))));
    return normalize(lh.x * f.tangent + lh.y * f.bitangent + lh.z * n);
}

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

// Subsurface helpers (:3490-3523).
template <class R> V3<R> subsurfaceSigmaT(const Mat<R> &m) { return V3<R>(1) / materialSubsurfaceRadius(m); }
template <class R> R sampleSubsurfaceFreeFlight(V3<R> sigmaT, R ux, R uy) {
    R sigma = sigmaT.z;
    if (ux < R(0.33333333333)) sigma = sigmaT.x;
    else if (ux < R(0.66666666667)) sigma = sigmaT.y;
    return -std::log(fmx(R(1) - uy, R(1e-6))) / fmx(sigma, R(1e-4));
}
template <class R> R subsurfaceCollisionPDF(V3<R> sigmaT, R d) {
    V3<R> tr = vexp((-sigmaT) * d);
    return fmx((sigmaT.x * tr.x + sigmaT.y * tr.y + sigmaT.z * tr.z) * R(0.33333333333), R(1e-6));
}
template <class R> R subsurfaceBoundaryPDF(V3<R> sigmaT, R d) {
    V3<R> tr = vexp((-sigmaT) * d);
    return fmx((tr.x + tr.y + tr.z) * R(0.33333333333), R(1e-6));
}
template <class R> V3<R> subsurfaceCollisionWeight(V3<R> sigmaT, V3<R> albedo, R d) {
    V3<R> tr = vexp((-sigmaT) * d);
    return albedo * sigmaT * tr / subsurfaceCollisionPDF(sigmaT, d);
}
template <class R> V3<R> subsurfaceBoundaryWeight(V3<R> sigmaT, R d) {
    V3<R> tr = vexp((-sigmaT) * d);
    return tr / subsurfaceBoundaryPDF(sigmaT, d);
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

template <class R>
bool continueTransmissionVolumeRandomWalk(const Hit<R> &b, const Mat<R> &m, Ray<R> &ray, V3<R> &T, R &prevPdf,
                                          uint32_t &state) {  // :3685-3721
    if (b.frontFacing || !materialHasVolumeScattering(m)) return false;
    V3<R> sigmaT = materialVolumeSigmaT(m);
    V3<R> sigmaS = materialVolumeSigmaS(m);
    if (maxComponent(sigmaT) <= R(1e-5) || maxComponent(sigmaS) <= R(1e-5)) return false;
    R f0 = randomFloat<R>(state);
    R f1 = randomFloat<R>(state);
    R freeFlight = sampleSubsurfaceFreeFlight(sigmaT, f0, f1);
    if (freeFlight >= b.t) {
        T *= materialVolumeTransmittance(m, b.t) / subsurfaceBoundaryPDF(sigmaT, b.t);
        return false;
    }
    V3<R> tr = vexp((-sigmaT) * freeFlight);
    T *= sigmaS * tr / subsurfaceCollisionPDF(sigmaT, freeFlight);
    ray.origin += ray.direction * freeFlight;
    R h0 = randomFloat<R>(state);
    R h1 = randomFloat<R>(state);
    ray.direction = sampleHenyeyGreenstein(h0, h1, ray.direction, materialVolumeAnisotropy(m));
    ray.origin += ray.direction * R(0.001);
    prevPdf = R(0);
    return maxComponent(T) > R(1e-4);
}

template <class R>
bool continueTransmissiveSurface(const Hit<R> &h, const Mat<R> &m, Ray<R> &ray, V3<R> &T, R &prevPdf, uint32_t &state) {  // :1723-1799
    R transmission = materialTransmission(m);
    if (transmission <= R(0)) return false;
    R gate = randomFloat<R>(state);
    if (gate > transmission) return false;
    V3<R> incoming = ray.direction;
    V3<R> view = -incoming;
    R roughness = materialTransmissionRoughness(m);
    R ior = materialTransmissionIOR(m);
    V3<R> micro = h.normal;
    if (roughness > R(0.01)) {
        R a = randomFloat<R>(state);
        R b = randomFloat<R>(state);
        micro = sampleGGXHalfVector(a, b, h.normal, roughness);
        if (dot(micro, incoming) > R(0)) micro = -micro;
    }
    R etaI = h.frontFacing ? R(1) : ior;
    R etaT = h.frontFacing ? ior : R(1);
    R eta = etaI / etaT;
    R cosThetaI = fmx(dot(view, micro), R(0));
    R fresnel = dielectricFresnel(cosThetaI, etaI, etaT);
    if (materialThinWalled(m)) {
        fresnel = dielectricFresnel(cosThetaI, R(1), ior);
        R reflectProbability = clampv(fresnel, R(0.02), R(0.98));
        R choice = randomFloat<R>(state);
        if (choice < reflectProbability) {
            ray.direction = normalize(reflect(incoming, micro));
            ray.origin = h.position + micro * R(0.001);
            T *= fresnel / fmx(reflectProbability, R(1e-5));
        } else {
            ray.direction = incoming;
            ray.origin = h.position + ray.direction * R(0.001);
            R grazing = R(1) - std::fabs(dot(view, h.normal));
            R tintStrength = mixs(R(0.35), R(0.95), grazing * grazing);
            R w = (R(1) - fresnel) / fmx(R(1) - reflectProbability, R(1e-5));
            T *= materialTransmissionTint(m, tintStrength) * w;
        }
        prevPdf = R(0);
        return true;
    }
    V3<R> refracted = refract(incoming, micro, eta);
    bool tir = dot(refracted, refracted) <= R(1e-7);
    R reflectProbability = tir ? R(1) : clampv(fresnel, R(0.02), R(0.98));
    R choice = randomFloat<R>(state);
    if (choice < reflectProbability) {
        ray.direction = normalize(reflect(incoming, micro));
        ray.origin = h.position + micro * R(0.001);
        T *= tir ? R(1) : fresnel / fmx(reflectProbability, R(1e-5));
    } else {
        ray.direction = normalize(refracted);
        ray.origin = h.position + ray.direction * R(0.001);
        R grazing = R(1) - std::fabs(dot(view, h.normal));
        R tintStrength = mixs(R(0.35), R(0.95), grazing * grazing);
        R w = (R(1) - fresnel) / fmx(R(1) - reflectProbability, R(1e-5));
        V3<R> absorption = materialTransmissionAbsorption(m, h.t, !h.frontFacing && !materialHasVolumeScattering(m));
        T *= materialTransmissionTint(m, tintStrength) * absorption * w;
    }
    prevPdf = R(0);
    return true;
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

        R roughness = clampv(m.parameters.x, R(0.02), R(1));
        R anisotropy = materialSpecularAnisotropy(m);
        V3<R> f0 = materialF0(m);
        R clearcoat = materialClearcoat(m);
        R clearcoatRoughness = materialClearcoatRoughness(m);
        R pd, ps, pc;
        materialLobeProbabilities(m, pd, ps, pc);
        R d0 = randomFloat<R>(state);
        R d1 = randomFloat<R>(state);
        V3<R> local = cosineHemisphere(d0, d1);
        V3<R> diffuseDirection = orientHemisphere(local, h.normal);
        V3<R> next;
        R lobe = randomFloat<R>(state);
        if (lobe < pc) {
            R a = randomFloat<R>(state);
            R b = randomFloat<R>(state);
            V3<R> hv = sampleGGXVisibleNormal(a, b, h.normal, h.tangent, h.bitangent, view, clearcoatRoughness, R(0));
            next = normalize(reflect(ray.direction, hv));
            if (dot(next, h.normal) <= R(0)) {
                next = diffuseDirection;
                T *= materialDiffuseBounceWeight(m, h.normal, view, next) / fmx(pd, R(1e-5));
            } else {
                R nDotL = fmx(dot(h.normal, next), R(1e-5));
                R vDotH = fmx(dot(view, hv), R(1e-5));
                T *= clearcoat * fresnelSchlickThinFilm(vDotH, materialClearcoatF0Color(m), m) *
                     geometrySmithG1GGX(nDotL, clearcoatRoughness) / fmx(pc, R(1e-5));
            }
        } else if (lobe < pc + ps) {
            R a = randomFloat<R>(state);
            R b = randomFloat<R>(state);
            V3<R> hv = sampleGGXVisibleNormal(a, b, h.normal, h.tangent, h.bitangent, view, roughness, anisotropy);
            next = normalize(reflect(ray.direction, hv));
            if (dot(next, h.normal) <= R(0)) {
                next = diffuseDirection;
                T *= materialDiffuseBounceWeight(m, h.normal, view, next) / fmx(pd, R(1e-5));
            } else {
                R vDotH = fmx(dot(view, hv), R(1e-5));
                V3<R> fresnel = fresnelSchlickThinFilm(vDotH, f0, m);
                V3<R> cca = materialClearcoatAttenuationForDirection(m, h.normal, view, next);
                T *= cca * fresnel * geometrySmithG1GGXAnisotropic(next, h.normal, h.tangent, h.bitangent, roughness, anisotropy) /
                     fmx(ps, R(1e-5));
            }
        } else {
            next = diffuseDirection;
            T *= materialDiffuseBounceWeight(m, h.normal, view, next) / fmx(pd, R(1e-5));
        }
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
