// dr_math.h -- small-vector math for the HIP path tracer (gfx950).
//
// Operation order is fixed and written out (no fused multiply-add; the
// library is compiled with -ffp-contract=off) so that camera rays and the
// triangle test produce the same bits as an IEEE evaluation of the reference
// formulas in PathTrace.metal. Conventions:
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v)  = v * (1 / sqrt(dot(v,v)))
//   mix(a,b,t)    = a + (b - a) * t
//   min/max       = fminf/fmaxf (NaN-ignoring, the Metal semantics)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DR_HD __host__ __device__ __forceinline__

struct vec2 {
    float x, y;
};
struct vec3 {
    float x, y, z;
};

DR_HD vec3 v3(float x, float y, float z) { return vec3{x, y, z}; }
DR_HD vec3 v3s(float s) { return vec3{s, s, s}; }
DR_HD vec3 v3f(const float4 &f) { return vec3{f.x, f.y, f.z}; }
DR_HD vec3 operator+(vec3 a, vec3 b) { return vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DR_HD vec3 operator-(vec3 a, vec3 b) { return vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DR_HD vec3 operator*(vec3 a, vec3 b) { return vec3{a.x * b.x, a.y * b.y, a.z * b.z}; }
DR_HD vec3 operator/(vec3 a, vec3 b) { return vec3{a.x / b.x, a.y / b.y, a.z / b.z}; }
DR_HD vec3 operator*(vec3 a, float s) { return vec3{a.x * s, a.y * s, a.z * s}; }
DR_HD vec3 operator*(float s, vec3 a) { return vec3{s * a.x, s * a.y, s * a.z}; }
DR_HD vec3 operator/(vec3 a, float s) { return vec3{a.x / s, a.y / s, a.z / s}; }
DR_HD vec3 operator/(float s, vec3 a) { return vec3{s / a.x, s / a.y, s / a.z}; }
DR_HD vec3 operator-(vec3 a) { return vec3{-a.x, -a.y, -a.z}; }
DR_HD vec3 &operator+=(vec3 &a, vec3 b) { a = a + b; return a; }
DR_HD vec3 &operator*=(vec3 &a, vec3 b) { a = a * b; return a; }
DR_HD vec3 &operator*=(vec3 &a, float s) { a = a * s; return a; }
DR_HD vec3 &operator/=(vec3 &a, float s) { a = a / s; return a; }

DR_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DR_HD vec3 cross(vec3 a, vec3 b) {
    return vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DR_HD float length(vec3 a) { return sqrtf(dot(a, a)); }
DR_HD vec3 normalize(vec3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
DR_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
DR_HD vec3 vmin(vec3 a, vec3 b) { return vec3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
DR_HD vec3 vmax(vec3 a, vec3 b) { return vec3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
DR_HD vec3 vclamp(vec3 x, vec3 lo, vec3 hi) { return vmin(vmax(x, lo), hi); }
DR_HD float mixf(float a, float b, float t) { return a + (b - a) * t; }
DR_HD vec3 mixv(vec3 a, vec3 b, float t) { return a + (b - a) * t; }
DR_HD float4 mix4(float4 a, float4 b, float t) {
    return make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t);
}
DR_HD vec3 vexp(vec3 a) { return vec3{expf(a.x), expf(a.y), expf(a.z)}; }
DR_HD vec3 vlog(vec3 a) { return vec3{logf(a.x), logf(a.y), logf(a.z)}; }
DR_HD vec3 vcos(vec3 a) { return vec3{cosf(a.x), cosf(a.y), cosf(a.z)}; }
DR_HD float maxComponent(vec3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
DR_HD bool allFinite(vec3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }
DR_HD vec3 reflect(vec3 i, vec3 n) { return i - 2.0f * dot(n, i) * n; }
DR_HD vec3 refract(vec3 i, vec3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3s(0.0f);
    return eta * i - (eta * d + sqrtf(k)) * n;
}
DR_HD float fractf(float x) { return fminf(x - floorf(x), 0x1.fffffep-1f); }
DR_HD float luminance(vec3 v) { return dot(v, vec3{0.2126f, 0.7152f, 0.0722f}); }

// Wellons' lowbias32 hash and the 24-bit float draw (PathTrace.metal:264-276).
DR_HD uint32_t dr_hash(uint32_t value) {
    value ^= value >> 16;
    value *= 0x7feb352du;
    value ^= value >> 15;
    value *= 0x846ca68bu;
    value ^= value >> 16;
    return value;
}
DR_HD float randomFloat(uint32_t &state) {
    state = dr_hash(state);
    return (float)(state & 0x00ffffffu) / (float)0x01000000u;
}
