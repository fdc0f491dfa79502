/*
 * oracle_bsdf.h -- the scene-independent part of the CPU oracle: vector
 * math, the RNG, material accessors, Fresnel, GGX, the layered BSDF
 * (evaluation, lobe probabilities, pdfs, lobe sampling), Henyey-Greenstein,
 * subsurface free flight and the transmissive-surface / transmission-volume
 * continuations.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): included by oracle.cpp (the
 * renderer oracle) and by oracle_pins.cpp (the entry points through which
 * tests/test_oracle_bsdf_pins.py checks these functions against what the
 * mathematics fixes). Every function restates the PathTrace.metal lines it
 * cites. Everything is in an anonymous namespace: each library that includes
 * this file gets its own copy, so a library built from a modified copy (the
 * mutation checks of the pin tests) can never bind to another's symbols.
 *
 * Pinned by tests/test_oracle_bsdf_pins.py (closed forms, quadrature,
 * chi-square histograms, Monte Carlo unbiasedness, reciprocity; each check
 * also shown to catch a planted mistake): dielectricFresnel, distributionGGX
 * and its anisotropic form, both Smith G1 forms, ggxReflectionPDF,
 * sampleGGXVisibleNormal, sampleGGXHalfVector, materialBSDFPDF with
 * materialLobeProbabilities' sum, sampleMaterialBounce's directions and
 * weights, evaluateMaterialBRDF (reciprocity, Lambertian normalisation, white
 * metal energy), sampleHenyeyGreenstein, the subsurface free flight, pdfs and
 * weights, continueTransmissiveSurface (Snell, Fresnel share, unbiased
 * weights, total internal reflection, Beer-Lambert) and
 * continueTransmissionVolumeRandomWalk's weights.
 * Parity unpinned (no law fixes them; they are held to the reference text
 * line by line and to the HIP kernel by GPU parity): thinFilmTint's
 * empirical tint (only its reciprocity is checked), the sheen lobe's shape,
 * the clearcoat thickness attenuation's path length, the lobe-probability
 * clamps (0.02-0.35, 0.08), the transmission tint mix(0.35, 0.95, grazing^2).
 */
#ifndef DENRIM_ORACLE_BSDF_H
#define DENRIM_ORACLE_BSDF_H

#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

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

// The lobe choice and throughput update that end one bounce
// (PathTrace.metal:5423-5513): a cosine direction is always drawn, then the
// lobe sample; the clearcoat and specular lobes draw a visible normal and fall
// back to the cosine direction (with the diffuse weight) when the reflection
// points below the surface. `incoming` is the ray direction (-view); T is
// multiplied by the lobe's weight; returns the next direction.
template <class R>
V3<R> sampleMaterialBounce(const Mat<R> &m, V3<R> n, V3<R> t, V3<R> b, V3<R> incoming, uint32_t &state, V3<R> &T) {
    V3<R> view = -incoming;
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
    V3<R> diffuseDirection = orientHemisphere(local, n);
    V3<R> next;
    R lobe = randomFloat<R>(state);
    if (lobe < pc) {
        R a = randomFloat<R>(state);
        R c = randomFloat<R>(state);
        V3<R> hv = sampleGGXVisibleNormal(a, c, n, t, b, view, clearcoatRoughness, R(0));
        next = normalize(reflect(incoming, hv));
        if (dot(next, n) <= R(0)) {
            next = diffuseDirection;
            T *= materialDiffuseBounceWeight(m, n, view, next) / fmx(pd, R(1e-5));
        } else {
            R nDotL = fmx(dot(n, next), R(1e-5));
            R vDotH = fmx(dot(view, hv), R(1e-5));
            T *= clearcoat * fresnelSchlickThinFilm(vDotH, materialClearcoatF0Color(m), m) *
                 geometrySmithG1GGX(nDotL, clearcoatRoughness) / fmx(pc, R(1e-5));
        }
    } else if (lobe < pc + ps) {
        R a = randomFloat<R>(state);
        R c = randomFloat<R>(state);
        V3<R> hv = sampleGGXVisibleNormal(a, c, n, t, b, view, roughness, anisotropy);
        next = normalize(reflect(incoming, hv));
        if (dot(next, n) <= R(0)) {
            next = diffuseDirection;
            T *= materialDiffuseBounceWeight(m, n, view, next) / fmx(pd, R(1e-5));
        } else {
            R vDotH = fmx(dot(view, hv), R(1e-5));
            V3<R> fresnel = fresnelSchlickThinFilm(vDotH, f0, m);
            V3<R> cca = materialClearcoatAttenuationForDirection(m, n, view, next);
            T *= cca * fresnel * geometrySmithG1GGXAnisotropic(next, n, t, b, roughness, anisotropy) /
                 fmx(ps, R(1e-5));
        }
    } else {
        next = diffuseDirection;
        T *= materialDiffuseBounceWeight(m, n, view, next) / fmx(pd, R(1e-5));
    }
    return next;
}

}  // namespace

#endif
