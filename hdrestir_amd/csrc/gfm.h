// gfm.h -- device vector math with OpenUSD Gf rounding semantics.
//
// The reference's arithmetic runs through Gf types (source/**/*.cpp use GfVec3f, GfMatrix4f,
// GfMatrix4d). To keep per-pixel RNG streams aligned with the CPU path, every operation here
// rounds exactly like Gf: float add/sub/dot/cross evaluated left to right, scalar products via
// double (float(double(x) * s)), v / s == v * (1.0 / s) in double, GetNormalized() with
// GF_MIN_VECTOR_LENGTH = 1e-10. The library is compiled with -ffp-contract=off so no
// multiply-add is fused. std::max / std::min / std::clamp are spelled out so NaN handling
// matches libstdc++.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __host__ __device__ __forceinline__

namespace hdr {

constexpr float kPiF = 3.14159265358979323846f;
constexpr double kPiD = 3.14159265358979323846;

HD float smax(float a, float b) { return (a < b) ? b : a; }   // std::max
HD float smin(float a, float b) { return (b < a) ? b : a; }   // std::min
HD float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
HD int iclamp(int v, int lo, int hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

struct V3 {
    float x, y, z;
    HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
HD V3 v3(float a, float b, float c) { return V3{a, b, c}; }
HD V3 v3s(float a) { return V3{a, a, a}; }
HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
// GfVec3f * float: the double product of two floats is exact, so one float rounding.
HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
HD V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
// GfVec3f * double
HD V3 muld(V3 a, double s) {
    return {float(double(a.x) * s), float(double(a.y) * s), float(double(a.z) * s)};
}
HD V3 divd(V3 a, double s) { return muld(a, 1.0 / s); }  // GfVec3f / s
HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
HD float length(V3 a) { return sqrtf(dot(a, a)); }
HD V3 normalized(V3 a) {
    const float len = length(a);
    const float eps = 1e-10f;
    return divd(a, double(len > eps ? len : eps));
}
HD V3 cmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }

// Row-vector float 4x4 (GfMatrix4f) stored row-major.
struct M4 {
    float m[16];
};
HD V3 xform_point(const float *m, V3 p) {  // GfMatrix4f::Transform + GfProject(GfVec4f)
    const float x = p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12];
    const float y = p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13];
    const float z = p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14];
    const float w = p.x * m[3] + p.y * m[7] + p.z * m[11] + m[15];
    const float inv = (w != 0.0f) ? 1.0f / w : 1.0f;
    return {inv * x, inv * y, inv * z};
}
HD V3 xform_dir(const float *m, V3 p) {  // GfMatrix4f::TransformDir
    return {p.x * m[0] + p.y * m[4] + p.z * m[8], p.x * m[1] + p.y * m[5] + p.z * m[9],
            p.x * m[2] + p.y * m[6] + p.z * m[10]};
}

struct V3D {
    double x, y, z;
};
HD V3D xform_point_d(const double *m, V3D p) {  // GfMatrix4d::Transform + GfProject(GfVec4d)
    const double x = p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12];
    const double y = p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13];
    const double z = p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14];
    const double w = p.x * m[3] + p.y * m[7] + p.z * m[11] + m[15];
    const double inv = (w != 0.0) ? 1.0 / w : 1.0;
    return {inv * x, inv * y, inv * z};
}

// Transcendentals evaluated in double and rounded once to float: these agree with glibc's
// expf / powf(x, 5) / sinf / cosf on all but ~1e-4 of float inputs, so branch decisions
// that depend on them (and with them the RNG streams) line up with the CPU path.
__device__ __forceinline__ float exp_f(float x) { return float(exp(double(x))); }
__device__ __forceinline__ float pow5_f(float x) {
    const double d = double(x);
    const double d2 = d * d;
    return float(d2 * d2 * d);
}
__device__ __forceinline__ float sin_f(float x) { return float(sin(double(x))); }
__device__ __forceinline__ float cos_f(float x) { return float(cos(double(x))); }
__device__ __forceinline__ float exp2_f(float x) { return float(exp2(double(x))); }
__device__ __forceinline__ float log_f(float x) { return float(log(double(x))); }
__device__ __forceinline__ float acos_f(float x) { return float(acos(double(x))); }
__device__ __forceinline__ float atan2_f(float y, float x) { return float(atan2(double(y), double(x))); }

}  // namespace hdr
