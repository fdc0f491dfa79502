// dr_shade.h -- layered BSDF, light and environment sampling, textures and
// camera rays for the HIP path tracer.
//
// Each function states the reference lines (PathTrace.metal) whose behaviour
// it reproduces; formulas keep the reference's operation order so the FP32
// result tracks an IEEE evaluation of the reference to the last few ulps.
#pragma once

#include "dr_math.h"
#include "dr_scene.h"

// A material is read in place from the scene's GPUMaterial array; only the
// base colour can be changed per hit (base-colour texture), so it is held
// by value.
// A hit's material: record p, or p blended toward q by t when t > 0
// (blendMaterials, PathTrace.metal:549-569, as materialForHit applies it for
// sparse-volume samples). base holds the (textured) base colour.
// ovr marks slots replaced by dense-volume material fields (emis: 1,
// par: 2, par2: 4; applyVolumeMaterialFields, PathTrace.metal:571-601).
struct Mat {
    const DrMaterial *p;
    const DrMaterial *q;
    float t;
    float4 base;
    uint32_t ovr;
    float4 emis, par, par2;
};
template <int S> __device__ __forceinline__ float4 ms(const Mat &m) {
    if constexpr (S == M_BASE) return m.base;
    else {
        if constexpr (S == M_EMISSION) {
            if (m.ovr & 1u) return m.emis;
        } else if constexpr (S == M_PARAMS) {
            if (m.ovr & 2u) return m.par;
        } else if constexpr (S == M_PARAMS2) {
            if (m.ovr & 4u) return m.par2;
        }
        float4 v = m.p->s[S];
        if (m.t > 0.0f) v = mix4(v, m.q->s[S], m.t);
        return v;
    }
}

// applyVolumeMaterialFields (PathTrace.metal:571-601) with the flag bits
// volumeMaterialField* (:255-262).
__device__ __forceinline__ void applyVolumeMaterialFields(Mat &m, const DrVolumeSample &s) {
    const uint32_t f = s.materialFieldFlags.x;
    if (f == 0u) return;
    float4 e = ms<M_EMISSION>(m), pa = ms<M_PARAMS>(m), p2 = ms<M_PARAMS2>(m);
    const float4 bc = s.baseColorOpacity, et = s.emissionTransmission, su = s.surface;
    if (f & 1u) m.base.x = fmaxf(bc.x, 0.0f), m.base.y = fmaxf(bc.y, 0.0f), m.base.z = fmaxf(bc.z, 0.0f);
    if (f & 2u) m.base.w = fminf(fmaxf(bc.w, 0.0f), 1.0f);
    if (f & 4u) e.x = fmaxf(et.x, 0.0f), e.y = fmaxf(et.y, 0.0f), e.z = fmaxf(et.z, 0.0f);
    if ((f & 4u) && (f & 128u)) {
        float k = fmaxf(su.w, 0.0f);
        e.x = fmaxf(et.x, 0.0f) * k, e.y = fmaxf(et.y, 0.0f) * k, e.z = fmaxf(et.z, 0.0f) * k;
    } else if (f & 128u) {
        float k = fmaxf(su.w, 0.0f);
        e.x = fmaxf(e.x, 0.0f) * k, e.y = fmaxf(e.y, 0.0f) * k, e.z = fmaxf(e.z, 0.0f) * k;
    }
    if (f & 32u) e.w = fminf(fmaxf(et.w, 0.0f), 1.0f);
    if (f & 8u) pa.x = fminf(fmaxf(su.x, 0.02f), 1.0f);
    if (f & 16u) pa.y = fminf(fmaxf(su.y, 0.0f), 1.0f);
    if (f & 64u) p2.x = fminf(fmaxf(su.z, 0.0f), 1.0f);
    m.emis = e;
    m.par = pa;
    m.par2 = p2;
    m.ovr = 7u;
}
__device__ __forceinline__ vec3 xyz(float4 f) { return vec3{f.x, f.y, f.z}; }

// ------------------------------------------------------------ sampling ----
__device__ __forceinline__ vec3 cosineHemisphere(float ux, float uy) {  // :278-282
    float r = sqrtf(ux);
    float phi = 6.28318530718f * uy;
    return vec3{r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - ux))};
}
__device__ __forceinline__ vec2 uniformDisk(float ux, float uy) {  // :284-288
    float r = sqrtf(ux);
    float phi = 6.28318530718f * uy;
    return vec2{r * cosf(phi), r * sinf(phi)};
}
__device__ __forceinline__ vec3 uniformSphere(float ux, float uy) {  // :290-295
    float z = 1.0f - 2.0f * ux;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = 6.28318530718f * uy;
    return vec3{r * cosf(phi), r * sinf(phi), z};
}
__device__ __forceinline__ vec3 orientHemisphere(vec3 l, vec3 n) {  // :297-309
    vec3 helper = fabsf(n.y) < 0.999f ? v3(0, 1, 0) : v3(1, 0, 0);
    vec3 t = normalize(cross(helper, n));
    vec3 b = cross(n, t);
    return normalize(l.x * t + l.y * b + l.z * n);
}
__device__ __forceinline__ vec3 sampleHenyeyGreenstein(float ux, float uy, vec3 dir, float anisotropy) {  // :311-324
    float g = clampf(anisotropy, -0.95f, 0.95f);
    if (fabsf(g) < 1e-4f) return uniformSphere(ux, uy);
    float omg2 = 1.0f - g * g;
    float den = 1.0f - g + 2.0f * g * ux;
    float q = omg2 / fmaxf(den, 1e-5f);
    float cosT = (1.0f + g * g - q * q) / (2.0f * g);
    cosT = clampf(cosT, -1.0f, 1.0f);
    float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
    float phi = 6.28318530718f * uy;
    return orientHemisphere(vec3{cosf(phi) * sinT, sinf(phi) * sinT, cosT}, dir);
}
__device__ __forceinline__ vec3 clampSampleContribution(vec3 c, float clampValueIn) {  // :330-346
    if (!allFinite(c)) return v3s(0.0f);
    float clampValue = fmaxf(clampValueIn, 0.0f);
    if (clampValue <= 0.0f) return c;
    float peak = maxComponent(c);
    if (peak <= clampValue) return c;
    return c * (clampValue / fmaxf(peak, 1e-5f));
}

// ------------------------------------------------------------- fresnel ----
__device__ __forceinline__ vec3 fresnelSchlick(float cosTheta, vec3 f0) {  // :352-355
    float factor = powf(clampf(1.0f - cosTheta, 0.0f, 1.0f), 5.0f);
    return f0 + (v3s(1.0f) - f0) * factor;
}
__device__ __forceinline__ vec3 thinFilmTint(const Mat &m, float cosTheta) {  // :357-373
    float4 tf = ms<M_THINFILM>(m);
    float strength = clampf(tf.x, 0.0f, 1.0f);
    float thickness = fmaxf(tf.y, 0.0f);
    if (strength <= 0.0f || thickness <= 0.0f) return v3s(1.0f);
    float filmIOR = fmaxf(tf.z, 1.0001f);
    float angle = powf(clampf(1.0f - cosTheta, 0.0f, 1.0f), 0.65f);
    float opticalPath = 2.0f * filmIOR * thickness * (0.62f + 0.42f * angle);
    vec3 phase = (6.28318530718f * opticalPath) / vec3{650.0f, 510.0f, 475.0f};
    vec3 interference = v3s(0.62f) + 0.38f * vcos(phase);
    float nl = fmaxf(luminance(interference), 0.35f);
    vec3 tint = vclamp(interference / nl, v3s(0.35f), v3s(1.8f));
    return mixv(v3s(1.0f), tint, strength);
}
__device__ __forceinline__ vec3 fresnelSchlickThinFilm(float cosTheta, vec3 f0, const Mat &m) {
    return fresnelSchlick(cosTheta, f0) * thinFilmTint(m, cosTheta);
}
__device__ __forceinline__ float dielectricFresnel(float cosThetaI, float etaI, float etaT) {  // :379-393
    cosThetaI = clampf(cosThetaI, 0.0f, 1.0f);
    float eta = etaI / etaT;
    float s2 = eta * eta * fmaxf(0.0f, 1.0f - cosThetaI * cosThetaI);
    if (s2 >= 1.0f) return 1.0f;
    float cosThetaT = sqrtf(fmaxf(0.0f, 1.0f - s2));
    float par = ((etaT * cosThetaI) - (etaI * cosThetaT)) / fmaxf((etaT * cosThetaI) + (etaI * cosThetaT), 1e-5f);
    float perp = ((etaI * cosThetaI) - (etaT * cosThetaT)) / fmaxf((etaI * cosThetaI) + (etaT * cosThetaT), 1e-5f);
    return clampf(0.5f * (par * par + perp * perp), 0.0f, 1.0f);
}
__device__ __forceinline__ float dielectricF0(float ior, float specularWeight) {  // :395-399
    float i = fmaxf(ior, 1.0f);
    float f0 = (i - 1.0f) / (i + 1.0f);
    return f0 * f0 * clampf(specularWeight, 0.0f, 1.0f);
}

// --------------------------------------------------- material accessors ----
__device__ __forceinline__ float matMetallic(const Mat &m) { return clampf(ms<M_PARAMS>(m).y, 0.0f, 1.0f); }
__device__ __forceinline__ vec3 materialF0(const Mat &m) {  // :401-406
    float4 p2 = ms<M_PARAMS2>(m);
    float dielectric = dielectricF0(p2.y, p2.x);
    vec3 dc = dielectric * vclamp(xyz(ms<M_SPECULAR>(m)), v3s(0.0f), v3s(1.0f));
    return mixv(dc, xyz(m.base), matMetallic(m));
}
__device__ __forceinline__ float materialClearcoat(const Mat &m) { return clampf(ms<M_PARAMS2>(m).z, 0.0f, 1.0f); }
__device__ __forceinline__ float materialClearcoatRoughness(const Mat &m) { return clampf(ms<M_PARAMS2>(m).w, 0.02f, 1.0f); }
__device__ __forceinline__ vec3 materialClearcoatF0Color(const Mat &m) {  // :420-427
    return dielectricF0(ms<M_SPECULAR>(m).w, 1.0f) * vclamp(xyz(ms<M_CLEARCOAT>(m)), v3s(0.0f), v3s(1.0f));
}
__device__ __forceinline__ vec3 materialClearcoatDepthAttenuation(const Mat &m, vec3 n, vec3 v, vec3 l) {  // :433-448
    float4 att = ms<M_CC_ATTEN>(m);
    float thickness = fmaxf(att.w, 0.0f) * materialClearcoat(m);
    if (thickness <= 1e-5f) return v3s(1.0f);
    float nDotV = fmaxf(dot(n, v), 0.05f);
    float nDotL = fmaxf(dot(n, l), 0.05f);
    float pathLength = thickness * 0.5f * (1.0f / nDotV + 1.0f / nDotL);
    return vexp(vlog(vclamp(xyz(att), v3s(0.0001f), v3s(1.0f))) * pathLength);
}
__device__ __forceinline__ float materialSpecularAnisotropy(const Mat &m) { return clampf(ms<M_PARAMS3>(m).w, -0.95f, 0.95f); }
__device__ __forceinline__ float materialTransmission(const Mat &m) { return clampf(ms<M_EMISSION>(m).w, 0.0f, 1.0f); }
__device__ __forceinline__ vec3 materialTransmissionColor(const Mat &m) { return vclamp(xyz(ms<M_TRANS>(m)), v3s(0.0f), v3s(1.0f)); }
__device__ __forceinline__ bool materialThinWalled(const Mat &m) { return ms<M_TRANS>(m).w > 0.5f; }
__device__ __forceinline__ float materialTransmissionRoughness(const Mat &m) { return clampf(ms<M_PARAMS3>(m).y, 0.001f, 1.0f); }
__device__ __forceinline__ float materialTransmissionIOR(const Mat &m) { return fmaxf(ms<M_PARAMS3>(m).z, 1.0001f); }
__device__ __forceinline__ float materialSheen(const Mat &m) { return clampf(ms<M_SHEEN>(m).w, 0.0f, 1.0f); }
__device__ __forceinline__ vec3 materialSheenColor(const Mat &m) { return vclamp(xyz(ms<M_SHEEN>(m)), v3s(0.0f), v3s(1.0f)); }
__device__ __forceinline__ float materialSheenRoughness(const Mat &m) { return clampf(ms<M_PARAMS3>(m).x, 0.02f, 1.0f); }
__device__ __forceinline__ float materialSubsurface(const Mat &m) {  // :486-490
    return clampf(ms<M_SSS_COLOR>(m).w, 0.0f, 1.0f) * (1.0f - matMetallic(m)) * (1.0f - materialTransmission(m));
}
__device__ __forceinline__ vec3 materialSubsurfaceColor(const Mat &m) { return vclamp(xyz(ms<M_SSS_COLOR>(m)), v3s(0.0f), v3s(0.999f)); }
__device__ __forceinline__ vec3 materialSubsurfaceRadius(const Mat &m) {  // :496-499
    float4 r = ms<M_SSS_RADIUS>(m);
    return vmax(xyz(r) * fmaxf(r.w, 0.0f), v3s(1e-4f));
}
__device__ __forceinline__ float materialSubsurfaceAnisotropy(const Mat &m) { return clampf(ms<M_SSS_PARAMS>(m).x, -0.95f, 0.95f); }
__device__ __forceinline__ vec3 materialTransmissionTint(const Mat &m, float strength) {
    return mixv(v3s(1.0f), materialTransmissionColor(m), strength);
}
__device__ __forceinline__ vec3 materialTransmissionAbsorptionCoefficient(const Mat &m) {  // :509-517
    float4 a = ms<M_TRANS_ABS>(m);
    if (a.w <= 1e-5f) return v3s(0.0f);
    vec3 rc = vclamp(xyz(a), v3s(0.0001f), v3s(1.0f));
    return (-vlog(rc)) / a.w;
}
__device__ __forceinline__ vec3 materialTransmissionAbsorption(const Mat &m, float distance, bool applies) {
    if (!applies || distance <= 0.0f) return v3s(1.0f);
    return vexp((-materialTransmissionAbsorptionCoefficient(m)) * distance);
}
__device__ __forceinline__ float materialVolumeScattering(const Mat &m) { return clampf(ms<M_VOL_SCATTER>(m).w, 0.0f, 1.0f); }
__device__ __forceinline__ bool materialHasVolumeScattering(const Mat &m) {
    return materialTransmission(m) > 0.0f && !materialThinWalled(m) && materialVolumeScattering(m) > 0.0f;
}
__device__ __forceinline__ vec3 materialVolumeSigmaS(const Mat &m) {  // :1139-1143
    return vclamp(xyz(ms<M_VOL_SCATTER>(m)), v3s(0.0f), v3s(1.0f)) * materialVolumeScattering(m) /
           fmaxf(ms<M_VOL_PARAMS>(m).x, 1e-4f);
}
__device__ __forceinline__ vec3 materialVolumeSigmaT(const Mat &m) {
    return materialTransmissionAbsorptionCoefficient(m) + materialVolumeSigmaS(m);
}
__device__ __forceinline__ vec3 materialVolumeTransmittance(const Mat &m, float d) {
    return vexp((-materialVolumeSigmaT(m)) * d);
}

// ----------------------------------------------------------------- GGX ----
struct Frame {
    vec3 t, b;
};
__device__ __forceinline__ Frame makeTangentFrame(vec3 n, vec3 rt, vec3 rb) {  // :1158-1175
    Frame f;
    vec3 t = rt - n * dot(rt, n);
    if (dot(t, t) <= 1e-8f) {
        vec3 helper = fabsf(n.y) < 0.999f ? v3(0, 1, 0) : v3(1, 0, 0);
        t = cross(helper, n);
    }
    f.t = normalize(t);
    vec3 b = rb - n * dot(rb, n) - f.t * dot(rb, f.t);
    if (dot(b, b) <= 1e-8f) b = cross(n, f.t);
    f.b = normalize(b);
    return f;
}
__device__ __forceinline__ vec2 anisotropicGGXAlpha(float roughness, float anisotropy) {  // :1177-1186
    float alpha = fmaxf(roughness * roughness, 0.001f);
    float amount = clampf(anisotropy, -0.95f, 0.95f);
    float aspect = sqrtf(fmaxf(0.1f, 1.0f - 0.9f * fabsf(amount)));
    float al = alpha / aspect, as = alpha * aspect;
    return amount >= 0.0f ? vec2{al, as} : vec2{as, al};
}
__device__ __forceinline__ float distributionGGX(float nDotH, float roughness) {  // :1188-1193
    float a = roughness * roughness;
    float a2 = a * a;
    float den = nDotH * nDotH * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(3.14159265359f * den * den, 1e-5f);
}
__device__ __forceinline__ float geometrySmithG1GGX(float nDotV, float roughness) {  // :1195-1200
    float a = roughness * roughness;
    float a2 = a * a;
    float den = nDotV + sqrtf(fmaxf(a2 + (1.0f - a2) * nDotV * nDotV, 0.0f));
    return (2.0f * nDotV) / fmaxf(den, 1e-5f);
}
__device__ __forceinline__ float distributionGGXAnisotropic(vec3 h, vec3 n, vec3 t, vec3 b, float roughness, float anisotropy) {  // :1206-1226
    float nDotH = fmaxf(dot(n, h), 0.0f);
    if (nDotH <= 0.0f) return 0.0f;
    vec2 alpha = anisotropicGGXAlpha(roughness, anisotropy);
    float tDotH = dot(t, h), bDotH = dot(b, h);
    float den = (tDotH * tDotH) / fmaxf(alpha.x * alpha.x, 1e-6f) + (bDotH * bDotH) / fmaxf(alpha.y * alpha.y, 1e-6f) +
                nDotH * nDotH;
    return 1.0f / fmaxf(3.14159265359f * alpha.x * alpha.y * den * den, 1e-6f);
}
__device__ __forceinline__ float geometrySmithG1GGXAnisotropic(vec3 d, vec3 n, vec3 t, vec3 b, float roughness, float anisotropy) {  // :1228-1251
    float nDotD = fmaxf(dot(n, d), 0.0f);
    if (nDotD <= 0.0f) return 0.0f;
    vec2 alpha = anisotropicGGXAlpha(roughness, anisotropy);
    float tDotD = dot(t, d), bDotD = dot(b, d);
    float root = sqrtf(fmaxf(alpha.x * alpha.x * tDotD * tDotD + alpha.y * alpha.y * bDotD * bDotD + nDotD * nDotD, 0.0f));
    return (2.0f * nDotD) / fmaxf(nDotD + root, 1e-5f);
}
__device__ __forceinline__ vec3 evaluateMaterialSheenBRDF(const Mat &m, vec3 n, vec3 v, vec3 l) {  // :1266-1300
    float sheen = materialSheen(m);
    if (sheen <= 0.0f) return v3s(0.0f);
    float metallic = matMetallic(m);
    vec3 hd = v + l;
    if (dot(hd, hd) <= 1e-8f) return v3s(0.0f);
    vec3 h = normalize(hd);
    float nDotL = fmaxf(dot(n, l), 0.0f), nDotV = fmaxf(dot(n, v), 0.0f), nDotH = fmaxf(dot(n, h), 0.0f);
    if (nDotL <= 0.0f || nDotV <= 0.0f) return v3s(0.0f);
    float sinThetaH = sqrtf(fmaxf(0.0f, 1.0f - nDotH * nDotH));
    float inverseAlpha = mixf(20.0f, 0.5f, materialSheenRoughness(m));
    float distribution = (2.0f + inverseAlpha) * powf(sinThetaH, inverseAlpha) * 0.15915494309f;
    float visibility = 1.0f / fmaxf(4.0f * (nDotL + nDotV - nDotL * nDotV), 1e-5f);
    return sheen * (1.0f - metallic) * materialSheenColor(m) * distribution * visibility;
}
__device__ __forceinline__ vec3 materialClearcoatAttenuationForDirection(const Mat &m, vec3 n, vec3 v, vec3 l) {  // :1302-1328
    float clearcoat = materialClearcoat(m);
    if (clearcoat <= 0.0f) return v3s(1.0f);
    vec3 hd = v + l;
    if (dot(hd, hd) <= 1e-8f) return v3s(1.0f);
    vec3 h = normalize(hd);
    float vDotH = fmaxf(dot(v, h), 0.0f);
    vec3 ccF = fresnelSchlickThinFilm(vDotH, materialClearcoatF0Color(m), m);
    vec3 fa = vmax(v3s(0.0f), v3s(1.0f) - ccF * clearcoat);
    return fa * materialClearcoatDepthAttenuation(m, n, v, l);
}
__device__ __forceinline__ vec3 evaluateMaterialBRDF(const Mat &m, vec3 n, vec3 t, vec3 b, vec3 v, vec3 l) {  // :1330-1390
    const float inversePi = 0.31830988618f;
    float roughness = clampf(ms<M_PARAMS>(m).x, 0.02f, 1.0f);
    float metallic = matMetallic(m);
    vec3 baseColor = xyz(m.base);
    vec3 h = normalize(v + l);
    float nDotL = fmaxf(dot(n, l), 0.0f);
    float nDotV = fmaxf(dot(n, v), 0.0f);
    float nDotH = fmaxf(dot(n, h), 0.0f);
    float vDotH = fmaxf(dot(v, h), 0.0f);
    if (nDotL <= 0.0f || nDotV <= 0.0f) return v3s(0.0f);
    vec3 f0 = materialF0(m);
    vec3 fresnel = fresnelSchlickThinFilm(vDotH, f0, m);
    Frame f = makeTangentFrame(n, t, b);
    float anisotropy = materialSpecularAnisotropy(m);
    float D = distributionGGXAnisotropic(h, n, f.t, f.b, roughness, anisotropy);
    float G = geometrySmithG1GGXAnisotropic(v, n, f.t, f.b, roughness, anisotropy) *
              geometrySmithG1GGXAnisotropic(l, n, f.t, f.b, roughness, anisotropy);
    vec3 specular = D * G * fresnel / fmaxf(4.0f * nDotV * nDotL, 1e-5f);
    float clearcoat = materialClearcoat(m);
    float ccr = materialClearcoatRoughness(m);
    vec3 ccF = fresnelSchlickThinFilm(vDotH, materialClearcoatF0Color(m), m);
    float ccD = distributionGGX(nDotH, ccr);
    float ccG = geometrySmithG1GGX(nDotV, ccr) * geometrySmithG1GGX(nDotL, ccr);
    vec3 ccSpec = ccD * ccG * ccF / fmaxf(4.0f * nDotV * nDotL, 1e-5f);
    vec3 cca = materialClearcoatAttenuationForDirection(m, n, v, l);
    vec3 diffuse = cca * (v3s(1.0f) - fresnel) * (1.0f - metallic) * baseColor * inversePi;
    vec3 sheen = evaluateMaterialSheenBRDF(m, n, v, l);
    return diffuse + cca * (specular + sheen) + ccSpec * clearcoat;
}
__device__ __forceinline__ float powerHeuristic(float a, float b) {  // :1392-1396
    float a2 = a * a, b2 = b * b;
    return a2 / fmaxf(a2 + b2, 1e-8f);
}
__device__ __forceinline__ void materialLobeProbabilities(const Mat &m, float &pd, float &ps, float &pc) {  // :1398-1426
    float metallic = matMetallic(m);
    vec3 baseColor = xyz(m.base);
    vec3 f0 = materialF0(m);
    float clearcoat = materialClearcoat(m);
    float sheenWeight = materialSheen(m) * luminance(materialSheenColor(m)) * (1.0f - metallic);
    float dw = luminance(baseColor * (1.0f - metallic)) + sheenWeight;
    float sw = luminance(f0);
    float cw = clearcoat * luminance(materialClearcoatF0Color(m));
    float total = fmaxf(dw + sw + cw, 1e-5f);
    pc = cw > 1e-5f ? clampf(cw / total, 0.02f, 0.35f) : 0.0f;
    float remaining = fmaxf(0.0f, 1.0f - pc);
    if (sw <= 1e-5f) ps = 0.0f;
    else if (dw <= 1e-5f) ps = remaining;
    else ps = clampf(sw / total, 0.08f, remaining);
    pd = fmaxf(0.0f, 1.0f - ps - pc);
}
__device__ __forceinline__ float ggxReflectionPDF(vec3 n, vec3 t, vec3 b, vec3 v, vec3 l, float roughness, float anisotropy) {  // :1428-1448
    vec3 h = normalize(v + l);
    float nDotH = fmaxf(dot(n, h), 0.0f), nDotV = fmaxf(dot(n, v), 0.0f), vDotH = fmaxf(dot(v, h), 0.0f);
    if (nDotH <= 0.0f || nDotV <= 0.0f || vDotH <= 0.0f) return 0.0f;
    Frame f = makeTangentFrame(n, t, b);
    return distributionGGXAnisotropic(h, n, f.t, f.b, roughness, anisotropy) *
           geometrySmithG1GGXAnisotropic(v, n, f.t, f.b, roughness, anisotropy) / fmaxf(4.0f * nDotV, 1e-5f);
}
__device__ __forceinline__ float materialBSDFPDF(const Mat &m, vec3 n, vec3 t, vec3 b, vec3 v, vec3 l) {  // :1450-1495
    const float inversePi = 0.31830988618f;
    float nDotL = fmaxf(dot(n, l), 0.0f), nDotV = fmaxf(dot(n, v), 0.0f);
    if (nDotL <= 0.0f || nDotV <= 0.0f) return 0.0f;
    float pd, ps, pc;
    materialLobeProbabilities(m, pd, ps, pc);
    float roughness = clampf(ms<M_PARAMS>(m).x, 0.02f, 1.0f);
    float ccr = materialClearcoatRoughness(m);
    float anisotropy = materialSpecularAnisotropy(m);
    float diffusePDF = nDotL * inversePi;
    float specularPDF = ggxReflectionPDF(n, t, b, v, l, roughness, anisotropy);
    float clearcoatPDF = ggxReflectionPDF(n, t, b, v, l, ccr, 0.0f);
    return pd * diffusePDF + ps * specularPDF + pc * clearcoatPDF;
}
__device__ __forceinline__ vec3 materialDiffuseBounceWeight(const Mat &m, vec3 n, vec3 v, vec3 l) {  // :1497-1523
    float metallic = matMetallic(m);
    vec3 hd = v + l;
    if (dot(hd, hd) <= 1e-8f) return v3s(0.0f);
    vec3 h = normalize(hd);
    float vDotH = fmaxf(dot(v, h), 0.0f);
    vec3 fresnel = fresnelSchlickThinFilm(vDotH, materialF0(m), m);
    vec3 cca = materialClearcoatAttenuationForDirection(m, n, v, l);
    vec3 diffuse = (v3s(1.0f) - fresnel) * (1.0f - metallic) * xyz(m.base);
    vec3 sheen = evaluateMaterialSheenBRDF(m, n, v, l) * 3.14159265359f;
    return cca * (diffuse + sheen);
}
__device__ __forceinline__ bool terminatePathWithRussianRoulette(uint32_t bounce, vec3 &T, uint32_t &state) {  // :1525-1546
    float w = maxComponent(T);
    if (w <= 0.0f) return true;
    if (bounce < 2u) return false;
    float survival = clampf(w, 0.05f, 0.95f);
    if (randomFloat(state) > survival) return true;
    T /= survival;
    return false;
}
__device__ __forceinline__ vec3 sampleGGXHalfVector(float ux, float uy, vec3 n, float roughness) {  // :1615-1623
    float a = roughness * roughness;
    float a2 = a * a;
    float phi = 6.28318530718f * ux;
    float cosT = sqrtf((1.0f - uy) / fmaxf(1.0f + (a2 - 1.0f) * uy, 1e-5f));
    float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
    return orientHemisphere(vec3{cosf(phi) * sinT, sinf(phi) * sinT, cosT}, n);
}
__device__ __forceinline__ vec3 sampleGGXVisibleNormal(float ux, float uy, vec3 n, vec3 t, vec3 b, vec3 v, float roughness,
                                                    float anisotropy) {  // :1625-1664
    Frame f = makeTangentFrame(n, t, b);
    vec2 alpha = anisotropicGGXAlpha(roughness, anisotropy);
    vec3 lv = normalize(vec3{dot(v, f.t), dot(v, f.b), fmaxf(dot(v, n), 1e-5f)});
    vec3 sv = normalize(vec3{alpha.x * lv.x, alpha.y * lv.y, lv.z});
    float lensq = sv.x * sv.x + sv.y * sv.y;
    vec3 t1 = lensq > 1e-7f ? vec3{-sv.y, sv.x, 0.0f} * (1.0f / sqrtf(lensq)) : v3(1, 0, 0);
    vec3 t2 = cross(sv, t1);
    float radius = sqrtf(ux);
    float phi = 6.28318530718f * uy;
    float p1 = radius * cosf(phi);
    float p2 = radius * sinf(phi);
    float blend = 0.5f * (1.0f + sv.z);
    p2 = mixf(sqrtf(fmaxf(0.0f, 1.0f - p1 * p1)), p2, blend);
    vec3 vn = p1 * t1 + p2 * t2 + sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2)) * sv;
    vec3 lh = normalize(vec3{alpha.x * vn.x, alpha.y * vn.y, fmaxf(vn.z, 0.0f)});
    return normalize(lh.x * f.t + lh.y * f.b + lh.z * n);
}

// --------------------------------------------------------------- lights ----
__device__ __forceinline__ float lightSolidAnglePDF(float d2, float cosLight, float area) {  // :1548-1550
    return d2 / fmaxf(cosLight * area, 1e-6f);
}
__device__ __forceinline__ float lightSelectionCDF(const DrLight *L, uint32_t i) { return clampf(L[i].selectionCDF, 0.0f, 1.0f); }
__device__ __forceinline__ float lightSelectionPDF(const DrLight *L, uint32_t i) {  // :1556-1559
    float prev = i == 0u ? 0.0f : lightSelectionCDF(L, i - 1u);
    return fmaxf(lightSelectionCDF(L, i) - prev, 0.0f);
}
__device__ __forceinline__ uint32_t selectLightIndex(const DrLight *L, uint32_t count, float sample) {  // :1561-1578
    if (count <= 1u) return 0u;
    float target = clampf(sample, 0.0f, 0.99999994f);
    uint32_t lo = 0u, hi = count - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1u;
        if (target <= lightSelectionCDF(L, mid)) hi = mid; else lo = mid + 1u;
    }
    return lo;
}

// ----------------------------------------------------- textures and env ----
__device__ __forceinline__ float4 sampleTexture(const DrSceneDev &S, uint32_t ip1, vec2 uv, float4 fallback) {  // :3723-3763
    if (ip1 == 0u || ip1 > S.texCount) return fallback;
    uint4 d = S.texDesc[ip1 - 1u];
    uint32_t offset = d.x, width = d.y, height = d.z, mode = d.w;
    if (width == 0u || height == 0u) return fallback;
    float wx = fractf(uv.x), wy = fractf(uv.y);
    if (mode == 1u && width > 1u && height > 1u) {
        float px = wx * (float)width - 0.5f, py = wy * (float)height - 0.5f;
        float bx = floorf(px), by = floorf(py);
        float fx = fractf(px), fy = fractf(py);
        uint32_t x0 = (uint32_t)fmodf(bx + (float)width, (float)width);
        uint32_t y0 = (uint32_t)fmodf(by + (float)height, (float)height);
        uint32_t x1 = (x0 + 1u) % width, y1 = (y0 + 1u) % height;
        float4 c00 = S.texels[offset + y0 * width + x0], c10 = S.texels[offset + y0 * width + x1];
        float4 c01 = S.texels[offset + y1 * width + x0], c11 = S.texels[offset + y1 * width + x1];
        return mix4(mix4(c00, c10, fx), mix4(c01, c11, fx), fy);
    }
    uint32_t x = min((uint32_t)(wx * (float)width), width - 1u);
    uint32_t y = min((uint32_t)(wy * (float)height), height - 1u);
    return S.texels[offset + y * width + x];
}
__device__ __forceinline__ vec3 fallbackSky(vec3 d) {  // :3765-3768
    float sky = 0.5f * (d.y + 1.0f);
    return mixv(vec3{0.02f, 0.025f, 0.035f}, vec3{0.25f, 0.32f, 0.45f}, sky);
}
__device__ __forceinline__ vec3 sampleEnvironment(const DrSceneDev &S, vec3 d) {  // :3770-3803
    float intensity = fmaxf(S.envIntensity, 0.0f);
    if (S.envTexIndexPlusOne == 0u) return fallbackSky(d) * intensity;
    float s = sinf(S.envRotationY), c = cosf(S.envRotationY);
    vec3 r = normalize(vec3{c * d.x - s * d.z, d.y, s * d.x + c * d.z});
    float u = atan2f(r.z, r.x) * 0.15915494309f + 0.5f;
    float v = acosf(clampf(r.y, -1.0f, 1.0f)) * 0.31830988618f;
    v = clampf(v, 0.00001f, 0.99999f);
    vec3 sky = fallbackSky(d);
    vec3 color = xyz(sampleTexture(S, S.envTexIndexPlusOne, vec2{u, v}, make_float4(sky.x, sky.y, sky.z, 1.0f))) * intensity;
    float maxRadiance = fmaxf(S.envMaxRadiance, 0.0f);
    if (maxRadiance > 0.0f) color = vmin(color, v3s(maxRadiance));
    return color;
}
__device__ __forceinline__ float environmentPDF(const DrSceneDev &S, vec3 d) {  // :3882-3927
    if (S.envCount == 0u) return 0.0f;
    uint32_t index = 0u;
    if (S.envTexIndexPlusOne != 0u && S.envTexIndexPlusOne <= S.texCount) {
        uint4 td = S.texDesc[S.envTexIndexPlusOne - 1u];
        uint32_t width = td.y, height = td.z;
        if (width != 0u && height != 0u) {
            float s = sinf(S.envRotationY), c = cosf(S.envRotationY);
            vec3 r = normalize(vec3{c * d.x - s * d.z, d.y, s * d.x + c * d.z});
            float u = fractf(atan2f(r.z, r.x) * 0.15915494309f + 0.5f);
            float v = clampf(acosf(clampf(r.y, -1.0f, 1.0f)) * 0.31830988618f, 0.0f, 0.99999994f);
            uint32_t x = min((uint32_t)(u * (float)width), width - 1u);
            uint32_t y = min((uint32_t)(v * (float)height), height - 1u);
            index = y * width + x;
        }
    }
    if (index >= S.envCount) return 0.0f;
    return fmaxf(S.env[index].y, 0.0f);
}
__device__ __forceinline__ vec3 sampleEnvironmentDirection(const DrSceneDev &S, float px, float py, float cdfSample,
                                                           float coin, float &pdf) {  // :3952-4003
    pdf = 0.0f;
    if (S.envTexIndexPlusOne == 0u || S.envCount == 0u || S.envTexIndexPlusOne > S.texCount) return v3(0, 1, 0);
    uint4 td = S.texDesc[S.envTexIndexPlusOne - 1u];
    uint32_t width = td.y, height = td.z;
    if (width == 0u || height == 0u) return v3(0, 1, 0);
    // CDF inversion (:3929-3950): the lower bound of the target in the CDF.
    // A guide table over target space brackets it -- guide[k] is the first
    // texel whose cdf >= k/G, so the answer lies in [guide[k], guide[k+1]]
    // for k = floor(target * G) -- and the reference's binary search runs
    // only inside that bracket. G is a power of two, so target * G and k / G
    // are exact and the index is the one the full search returns.
    uint32_t index = 0u;
    if (S.envCount > 1u && S.envAliasSampling) {
        // Alias table (this backend's option): the CDF sample picks a
        // column, the coin keeps it or takes its alias -- the same texel
        // probabilities in O(1).
        index = min((uint32_t)(cdfSample * (float)S.envCount), S.envCount - 1u);
        const uint2 a = S.envAlias[index];
        if (!(coin < __uint_as_float(a.x))) index = a.y;
    } else if (S.envCount > 1u) {
        float target = clampf(cdfSample, 0.0f, 0.99999994f);
        uint32_t lo = 0u, hi = S.envCount - 1u;
        if (S.envGuideSize > 0u) {
            uint32_t k = min((uint32_t)(target * (float)S.envGuideSize), S.envGuideSize - 1u);
            lo = S.envGuide[k];
            hi = S.envGuide[k + 1u];
        }
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1u;
            if (target <= S.env[mid].x) hi = mid; else lo = mid + 1u;
        }
        index = lo;
    }
    index = min(index, S.envCount - 1u);
    uint32_t x = index % width, y = index / width;
    float u = ((float)x + px) / (float)width;
    float v = ((float)y + py) / (float)height;
    float phi = (u - 0.5f) * 6.28318530718f;
    float theta = clampf(v, 0.00001f, 0.99999f) * 3.14159265359f;
    float sinT = sinf(theta);
    vec3 local = normalize(vec3{cosf(phi) * sinT, cosf(theta), sinf(phi) * sinT});
    float s = sinf(S.envRotationY), c = cosf(S.envRotationY);
    pdf = fmaxf(S.env[index].y, 0.0f);
    return normalize(vec3{c * local.x + s * local.z, local.y, -s * local.x + c * local.z});
}

// --------------------------------------------------------------- camera ----
__device__ __forceinline__ void makeCameraRay(const DrCamera &cam, uint32_t width, uint32_t height, uint32_t gx,
                                              uint32_t gy, float jx, float jy, float lx, float ly, vec3 &o,
                                              vec3 &d) {  // :4056-4091
    float u = ((float)gx + jx) / (float)width;
    float v = 1.0f - (((float)gy + jy) / (float)height);
    vec3 ll = xyz(cam.lowerLeft), hor = xyz(cam.horizontal), ver = xyz(cam.vertical);
    vec3 planePoint = ll + u * hor + v * ver;
    if (cam.origin.w > 0.5f) {
        o = planePoint;
        d = normalize(cross(ver, hor));
    } else {
        o = xyz(cam.origin);
        d = normalize(planePoint - o);
    }
    float aperture = fmaxf(cam.lens.x, 0.0f);
    if (cam.origin.w <= 0.5f && aperture > 0.0f) {
        vec3 forward = normalize(cross(ver, hor));
        float focus = fmaxf(cam.lens.y, 0.0001f);
        float focalT = focus / fmaxf(dot(d, forward), 0.0001f);
        vec3 fp = o + d * focalT;
        vec2 disk = uniformDisk(lx, ly);
        disk.x *= aperture;
        disk.y *= aperture;
        o += normalize(hor) * disk.x + normalize(ver) * disk.y;
        d = normalize(fp - o);
    }
}
__device__ __forceinline__ bool projectToScreen(const DrCamera &cam, vec3 p, uint32_t w, uint32_t h, vec2 &out) {  // :1666-1701
    vec3 hor = xyz(cam.horizontal), ver = xyz(cam.vertical), ll = xyz(cam.lowerLeft), org = xyz(cam.origin);
    vec3 forward = normalize(cross(ver, hor));
    vec3 pc = ll + hor * 0.5f + ver * 0.5f;
    vec3 pp;
    if (cam.origin.w > 0.5f) {
        pp = p - forward * dot(p - pc, forward);
    } else {
        vec3 tp = p - org;
        float den = dot(tp, forward);
        if (fabsf(den) <= 1e-5f) return false;
        float planeT = dot(pc - org, forward) / den;
        if (planeT <= 0.0f) return false;
        pp = org + tp * planeT;
    }
    vec3 rel = pp - ll;
    float u = dot(rel, hor) / fmaxf(dot(hor, hor), 1e-5f);
    float v = dot(rel, ver) / fmaxf(dot(ver, ver), 1e-5f);
    out = vec2{u * (float)w, (1.0f - v) * (float)h};
    return true;
}
