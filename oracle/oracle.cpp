// oracle.cpp -- CPU oracle for the hdRestir per-pixel hot path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). A line-by-line C++20 restatement of the
// reference's CPU path; every function cites the reference file:line it follows
// (paths relative to the hdRestir checkout). Arithmetic follows OpenUSD Gf semantics:
//   * GfVec3f (+,-,Dot,Cross) in float, left to right;
//   * GfVec3f * s and s * GfVec3f go through double (float(double(x)*s)); for a float
//     scalar that equals the float product, for a double scalar (e.g. 1.0/len) it does not;
//   * GfVec3f / s == v * (1.0 / s) in double;
//   * GetNormalized(): v / (len > 1e-10f ? len : 1e-10f) with len = sqrtf(Dot(v, v));
//   * GfMatrix4d::Transform in double with a w-divide (GfProject); GfMatrix4f in float.
// Function arguments that each draw from the RNG are sequenced left to right (the
// evaluation order of the reference's primary platform, clang); the reference leaves it
// unspecified (e.g. SampleGGX(rng.NextFloat(), rng.NextFloat(), r)).
// Compiled with -ffp-contract=off so no multiply-add is fused.
//
// What pins each part (tests/test_oracle.py, tests/test_oracle_pins.py; a planted mistake in any
// of these functions fails its pin):
//   Rng / NextFloat ............ the mt19937_64 KAT and the exact canonical mapping
//   CameraRay .................. pinhole rays in closed form (identity and look-at); CameraRayParams:
//                                the DoF centre ray meets the focus plane (test_camera.cpp:40-60)
//   Lambda, RGB<->spectrum ..... the round-trip expectation invariant, flat spectra E-balanced
//   GGXEval / GGXPdfFull ....... reciprocity, pdf integral <= 1, energy bound
//   PreviewSampleBounce ........ direction histogram vs the GGX / cosine mixture at normal
//                                incidence, reported pdf per lobe, unit directions
//   GetClosure ................. parameter mapping (PreviewSurface and DefaultMaterial)
//   Light::SampleLight ......... rect: irradiance vs the rectangle view factor, uniform points,
//                                back faces; point: no 1/d^2 and cone smoothstep; distant: the -Z
//                                axis; dome: texel histogram vs lum sin(theta), pdf, EvalPdf,
//                                E[L/pdf] vs quadrature
//   LightIndex ................. pinned values and equal-width intervals
//   Reservoir, RandomIndex ..... the KATs of test_reservoir.cpp:18-106, proportional selection,
//                                the RIS-1D closed form, PowerHeuristic normalisation
//   Scene::Intersect ........... float64 brute force over every instanced triangle
//   whole estimators ........... exact linearity in the emitted radiance (exposure + 1 doubles
//                                every pixel, streams unchanged); MIS and RIS agree in mean
// Parity unpinned: the last-ulp rounding of real Gf (GetInverse, GfRange3f) and of the reference
// platforms' libm; the reference's own images (not available).

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <optional>
#include <random>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

constexpr float kPiF = 3.14159265358979323846f;
constexpr double kPiD = 3.14159265358979323846;

// ---------------------------------------------------------------- Gf layer
struct V2 {
    float x = 0, y = 0;
};
struct V3 {
    float v[3] = {0, 0, 0};
    V3() = default;
    V3(float a, float b, float c) : v{a, b, c} {}
    explicit V3(float a) : v{a, a, a} {}
    float &operator[](int i) { return v[i]; }
    const float &operator[](int i) const { return v[i]; }
};
inline V3 operator+(const V3 &a, const V3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 operator-(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 operator-(const V3 &a) { return {-a[0], -a[1], -a[2]}; }
// GfVec3f * double (float(double(x) * s)); exact float product when s is a float.
inline V3 muld(const V3 &a, double s) {
    return {float(double(a[0]) * s), float(double(a[1]) * s), float(double(a[2]) * s)};
}
inline V3 operator*(const V3 &a, float s) { return muld(a, double(s)); }
inline V3 operator*(float s, const V3 &a) { return muld(a, double(s)); }
inline V3 divd(const V3 &a, double s) { return muld(a, 1.0 / s); }  // GfVec3f / s
inline float Dot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 Cross(const V3 &a, const V3 &b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline float Length(const V3 &a) { return std::sqrt(Dot(a, a)); }
inline V3 Normalized(const V3 &a) {
    const float len = Length(a);
    const float eps = float(1e-10);  // GF_MIN_VECTOR_LENGTH
    return divd(a, double(len > eps ? len : eps));
}
inline V3 CompMult(const V3 &a, const V3 &b) { return {a[0] * b[0], a[1] * b[1], a[2] * b[2]}; }

struct V3d {
    double v[3];
};

struct M4d {
    double m[4][4];
    static M4d From(const double *p) {
        M4d r;
        for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = p[i];
        return r;
    }
    // GfMatrix4d::Transform(GfVec3d) with GfProject.
    V3d Transform(const V3d &p) const {
        const double x = p.v[0] * m[0][0] + p.v[1] * m[1][0] + p.v[2] * m[2][0] + m[3][0];
        const double y = p.v[0] * m[0][1] + p.v[1] * m[1][1] + p.v[2] * m[2][1] + m[3][1];
        const double z = p.v[0] * m[0][2] + p.v[1] * m[1][2] + p.v[2] * m[2][2] + m[3][2];
        const double w = p.v[0] * m[0][3] + p.v[1] * m[1][3] + p.v[2] * m[2][3] + m[3][3];
        const double inv = (w != 0.0) ? 1.0 / w : 1.0;
        return {{inv * x, inv * y, inv * z}};
    }
};

// 4x4 inverse by 2x2 sub-determinants (cofactor expansion), in double; singular -> FLT_MAX scale.
// Stands for GfMatrix4d::GetInverse / GfMatrix4f::GetInverse (last-ulp parity with real Gf unpinned).
static void Inverse4(const double a[4][4], double out[4][4]) {
    const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const double s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const double s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const double s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const double c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const double c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const double c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (det == 0.0) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) out[i][j] = (i == j) ? double(std::numeric_limits<float>::max()) : 0.0;
        return;
    }
    const double r = 1.0 / det;
    out[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * r;
    out[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * r;
    out[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * r;
    out[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * r;
    out[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * r;
    out[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * r;
    out[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * r;
    out[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * r;
    out[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * r;
    out[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * r;
    out[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * r;
    out[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * r;
    out[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * r;
    out[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * r;
    out[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * r;
    out[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * r;
}

inline M4d InverseD(const M4d &a) {
    M4d r;
    Inverse4(a.m, r.m);
    return r;
}

struct M4f {
    float m[4][4];
    static M4f From(const float *p) {
        M4f r;
        for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = p[i];
        return r;
    }
    static M4f FromD(const M4d &d) {  // GfMatrix4f(GfMatrix4d)
        M4f r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = float(d.m[i][j]);
        return r;
    }
    // GfMatrix4f::Transform(GfVec3f): float, with GfProject(GfVec4f).
    V3 Transform(const V3 &p) const {
        const float x = p[0] * m[0][0] + p[1] * m[1][0] + p[2] * m[2][0] + m[3][0];
        const float y = p[0] * m[0][1] + p[1] * m[1][1] + p[2] * m[2][1] + m[3][1];
        const float z = p[0] * m[0][2] + p[1] * m[1][2] + p[2] * m[2][2] + m[3][2];
        const float w = p[0] * m[0][3] + p[1] * m[1][3] + p[2] * m[2][3] + m[3][3];
        const float inv = (w != 0.0f) ? 1.0f / w : 1.0f;
        return {inv * x, inv * y, inv * z};
    }
    V3 TransformDir(const V3 &p) const {
        return {p[0] * m[0][0] + p[1] * m[1][0] + p[2] * m[2][0], p[0] * m[0][1] + p[1] * m[1][1] + p[2] * m[2][1],
                p[0] * m[0][2] + p[1] * m[1][2] + p[2] * m[2][2]};
    }
    V3 ExtractTranslation() const { return {m[3][0], m[3][1], m[3][2]}; }
    M4f Inverse() const {
        double a[4][4], o[4][4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) a[i][j] = m[i][j];
        Inverse4(a, o);
        M4f r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = float(o[i][j]);
        return r;
    }
};

struct Range3 {  // GfRange3f
    V3 mn{std::numeric_limits<float>::max()};
    V3 mx{-std::numeric_limits<float>::max()};
    void SetEmpty() {
        mn = V3(std::numeric_limits<float>::max());
        mx = V3(-std::numeric_limits<float>::max());
    }
    bool IsEmpty() const { return mn[0] > mx[0] || mn[1] > mx[1] || mn[2] > mx[2]; }
    void ExtendBy(const V3 &p) {
        for (int i = 0; i < 3; ++i) {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
    V3 GetSize() const { return mx - mn; }
};

// ---------------------------------------------------------------- RNG (source/core/rng.h:10-56)
struct Rng {
    std::mt19937_64 engine;
    std::uniform_real_distribution<float> dist;
    uint32_t draws = 0;
    explicit Rng(uint32_t seed) : engine(seed) {}
    float NextFloat() {
        ++draws;
        return dist(engine);
    }
};

inline uint32_t Fnv(uint32_t h, uint32_t v) {
    for (int b = 0; b < 4; ++b) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 16777619u;
    }
    return h;
}
constexpr uint32_t kFnvInit = 2166136261u;

struct PixelTrace {  // parity instrumentation (does not alter any arithmetic)
    uint32_t lightHash = kFnvInit;
    uint32_t selectHash = kFnvInit;
    int32_t lightPicks = 0;
};

struct Counters {
    uint64_t primary = 0, bounce = 0, shadow = 0, candidate = 0, nodes = 0, tris = 0, instances = 0;
};

// ---------------------------------------------------------------- spectrum (source/core/spectrum.{h,cpp})
constexpr int kSpec = 4;
constexpr float kLambdaMin = 360.0f, kLambdaMax = 830.0f;

struct Lambda {
    float l[kSpec] = {0, 0, 0, 0};
    // SampledWavelengths::SampleUniform, spectrum.h:24-37
    static Lambda SampleUniform(float u) {
        Lambda s;
        s.l[0] = kLambdaMin + u * (kLambdaMax - kLambdaMin);
        for (int i = 1; i < kSpec; ++i) {
            s.l[i] = s.l[0] + (float(i) * (kLambdaMax - kLambdaMin) / float(kSpec));
            if (s.l[i] > kLambdaMax) s.l[i] -= (kLambdaMax - kLambdaMin);
        }
        return s;
    }
};

struct Spec {
    float v[kSpec];
    Spec() {
        for (float &e : v) e = 0.0f;
    }
    explicit Spec(float x) {
        for (float &e : v) e = x;
    }
    Spec operator+(const Spec &o) const {
        Spec r;
        for (int i = 0; i < kSpec; ++i) r.v[i] = v[i] + o.v[i];
        return r;
    }
    Spec &operator+=(const Spec &o) {
        for (int i = 0; i < kSpec; ++i) v[i] += o.v[i];
        return *this;
    }
    Spec operator*(const Spec &o) const {
        Spec r;
        for (int i = 0; i < kSpec; ++i) r.v[i] = v[i] * o.v[i];
        return r;
    }
    Spec &operator*=(const Spec &o) {
        for (int i = 0; i < kSpec; ++i) v[i] *= o.v[i];
        return *this;
    }
    Spec operator*(float s) const {
        Spec r;
        for (int i = 0; i < kSpec; ++i) r.v[i] = v[i] * s;
        return r;
    }
    Spec &operator*=(float s) {
        for (float &e : v) e *= s;
        return *this;
    }
    float Max() const {
        float m = v[0];
        for (int i = 1; i < kSpec; ++i) m = std::max(m, v[i]);
        return m;
    }
};

// Wyman 2013 fits, spectrum.cpp:4-29
float xFit(float w) {
    float t1 = (w - 442.0f) * ((w < 442.0f) ? 0.0624f : 0.0374f);
    float t2 = (w - 599.8f) * ((w < 599.8f) ? 0.0264f : 0.0323f);
    float t3 = (w - 501.1f) * ((w < 501.1f) ? 0.0490f : 0.0382f);
    return 0.362f * std::exp(-0.5f * t1 * t1) + 1.056f * std::exp(-0.5f * t2 * t2) - 0.065f * std::exp(-0.5f * t3 * t3);
}
float yFit(float w) {
    float t1 = (w - 568.8f) * ((w < 568.8f) ? 0.0213f : 0.0247f);
    float t2 = (w - 530.9f) * ((w < 530.9f) ? 0.0613f : 0.0322f);
    return 0.821f * std::exp(-0.5f * t1 * t1) + 0.286f * std::exp(-0.5f * t2 * t2);
}
float zFit(float w) {
    float t1 = (w - 437.0f) * ((w < 437.0f) ? 0.0845f : 0.0278f);
    float t2 = (w - 459.0f) * ((w < 459.0f) ? 0.0385f : 0.0725f);
    return 1.217f * std::exp(-0.5f * t1 * t1) + 0.681f * std::exp(-0.5f * t2 * t2);
}

// SpectrumToRGB, spectrum.cpp:31-68
V3 SpectrumToRGB(const Spec &s, const Lambda &lam) {
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    float weight = (kLambdaMax - kLambdaMin) / kSpec;
    weight /= 106.856f;
    for (int i = 0; i < kSpec; ++i) {
        const float l = lam.l[i];
        const float val = s.v[i];
        X += val * xFit(l);
        Y += val * yFit(l);
        Z += val * zFit(l);
    }
    X *= weight;
    Y *= weight;
    Z *= weight;
    float r = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
    float g = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    float b = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
    r /= 1.198185f;
    g /= 0.950366f;
    b /= 0.907827f;
    return {std::max(0.0f, r), std::max(0.0f, g), std::max(0.0f, b)};
}

float Gaussian(float x, float c, float sigma) {  // spectrum.cpp:71-75
    const float t = (x - c) / sigma;
    return std::exp(-0.5f * t * t);
}

// RGBToSpectrum, spectrum.cpp:77-103
Spec RGBToSpectrum(const V3 &rgb, const Lambda &lam) {
    Spec s;
    const float white = std::min(rgb[0], std::min(rgb[1], rgb[2]));
    const V3 rem(rgb[0] - white, rgb[1] - white, rgb[2] - white);
    for (int i = 0; i < kSpec; ++i) {
        const float l = lam.l[i];
        float val = white;
        val += rem[0] * Gaussian(l, 615.0f, 20.0f) * 1.295f;
        val += rem[1] * Gaussian(l, 540.0f, 20.0f) * 1.556f;
        val += rem[2] * Gaussian(l, 460.0f, 20.0f) * 1.382f;
        s.v[i] = std::max(0.0f, val);
    }
    return s;
}

float SpectrumLuminance(const Spec &s, const Lambda &lam) {  // spectrum.cpp:105-109
    const V3 rgb = SpectrumToRGB(s, lam);
    return std::max(0.0f, 0.2126f * rgb[0] + 0.7152f * rgb[1] + 0.0722f * rgb[2]);
}

// ---------------------------------------------------------------- core types
struct Ray {
    V3 o, d;
};
struct HitRecord {  // source/core/hit_record.h:15-27 (+ original triangle index for parity)
    V3 Position, Normal, SmoothNormal, Dpdu{1, 0, 0}, Dpdv{0, 1, 0};
    V2 Uv;
    V3 Albedo;
    float Depth = std::numeric_limits<float>::max();
    int PrimId = -1;
    int MatId = -1;
    int TriId = -1;
};
enum class PdfSpace { Area, SolidAngle };
struct Pdf {  // source/math/pdf.h:13-28
    float value = 0.f;
    PdfSpace space = PdfSpace::Area;
    Pdf ConvertTo(PdfSpace target, float dist2, float cosY) const {
        if (space == target) return *this;
        if (space == PdfSpace::Area) return {value * dist2 / cosY, PdfSpace::SolidAngle};
        return {value * cosY / dist2, PdfSpace::Area};
    }
};

struct Closure {  // BSDFClosure, source/core/shading_types.h:18-42
    V3 BaseColor{0.18f, 0.18f, 0.18f};
    V3 Emission{0, 0, 0};
    V3 Normal{0, 0, 1};
    V3 SpecularColor{1, 1, 1};
    V3 CoatColor{1, 1, 1};
    V3 SheenColor{1, 1, 1};
    V3 SubsurfaceColor{0.7f, 0.2f, 0.2f};
    V3 TransmissionColor{1, 1, 1};
    float Metallic = 0.0f, Roughness = 0.5f, Ior = 1.5f, Specular = 0.5f, Transmission = 0.0f,
          TransmissionDepth = 0.0f, Coat = 0.0f, CoatRoughness = 0.0f, CoatIor = 1.5f, Sheen = 0.0f,
          SheenRoughness = 0.3f, Subsurface = 0.0f, Opacity = 1.0f;
    bool ThinWalled = false;
};

struct BounceSample {  // BsdfBounceSample, shading_types.h:59-67
    Ray NextRay;
    Spec ThroughputMul{1.0f};
    Spec ThroughputIntegrandMul{0.0f};
    Pdf BsdfPdf{0.f, PdfSpace::SolidAngle};
    bool ImpossibleNEEConnection = false;
    bool SkipRoulette = false;
};
struct BounceConfig {
    int maxRefl, maxRefr;
};
struct BounceState {
    int refl = 0, refr = 0;
};
struct ShadingPoint {
    const Closure *c;
    V3 n;  // shading normal
    const Lambda *lambda;
    bool isInside;
};
struct Connection {  // BsdfBounceConnection
    BounceSample Bounce;
    std::optional<HitRecord> Hit;
};

// ---------------------------------------------------------------- shading helpers (source/materials/shading_helpers.h)
V3 SampleCosineHemisphere(float u1, float u2) {  // :26-31
    const float r = std::sqrt(u1);
    const float theta = 2.0f * kPiF * u2;
    return {r * std::cos(theta), std::sqrt(1.0f - u1), r * std::sin(theta)};
}
V3 SampleGGX(float u1, float u2, float roughness) {  // :33-41
    const float alpha = std::max(0.001f, roughness * roughness);
    const float alpha2 = alpha * alpha;
    const float phi = 2.0f * kPiF * u1;
    const float cosTheta = std::sqrt((1.0f - u2) / (1.0f + (alpha2 - 1.0f) * u2));
    const float sinTheta = std::sqrt(std::max(0.0f, 1.0f - cosTheta * cosTheta));
    return {sinTheta * std::cos(phi), cosTheta, sinTheta * std::sin(phi)};
}
V3 AlignToNormal(const V3 &s, const V3 &n) {  // :43-49
    const V3 up = std::abs(n[1]) < 0.999f ? V3(0, 1, 0) : V3(1, 0, 0);
    const V3 t = Normalized(Cross(up, n));
    const V3 b = Cross(n, t);
    return s[0] * t + s[1] * n + s[2] * b;
}
float FresnelDielectric(float cosThetaI, float ior) {  // :51-65
    cosThetaI = std::clamp(cosThetaI, -1.0f, 1.0f);
    float etaI = 1.0f, etaT = ior;
    if (cosThetaI > 0) std::swap(etaI, etaT);
    const float sinThetaT = etaI / etaT * std::sqrt(std::max(0.0f, 1.0f - cosThetaI * cosThetaI));
    if (sinThetaT >= 1.0f) return 1.0f;
    const float cosThetaT = std::sqrt(std::max(0.0f, 1.0f - sinThetaT * sinThetaT));
    cosThetaI = std::abs(cosThetaI);
    const float rParl = ((etaT * cosThetaI) - (etaI * cosThetaT)) / ((etaT * cosThetaI) + (etaI * cosThetaT));
    const float rPerp = ((etaI * cosThetaI) - (etaT * cosThetaT)) / ((etaI * cosThetaI) + (etaT * cosThetaT));
    return (rParl * rParl + rPerp * rPerp) / 2.0f;
}
float GGXPdf(const V3 &h, const V3 &n, float alpha) {  // :68-74
    const float cosH = std::max(0.f, Dot(h, n));
    const float a2 = alpha * alpha;
    const float denom = cosH * cosH * (a2 - 1.f) + 1.f;
    return a2 / (kPiF * denom * denom) * cosH;
}
float GGXReflectionWeight(float nDotV, float nDotL, float vDotH, float nDotH, float alpha) {  // :78-93
    nDotV = std::max(0.001f, nDotV);
    nDotL = std::max(0.001f, nDotL);
    vDotH = std::max(0.001f, vDotH);
    nDotH = std::max(0.001f, nDotH);
    alpha = std::max(0.001f, alpha);
    const float alpha2 = alpha * alpha;
    const float g1V = 2.0f * nDotV / (nDotV + std::sqrt(alpha2 + (1.0f - alpha2) * nDotV * nDotV));
    const float g1L = 2.0f * nDotL / (nDotL + std::sqrt(alpha2 + (1.0f - alpha2) * nDotL * nDotL));
    const float weight = (g1V * g1L * vDotH) / (nDotV * nDotH);
    return std::min(weight, 10.0f);
}
float PowerHeuristic2(float f, float g) {  // :95-100
    const float f2 = f * f, g2 = g * g;
    return f2 / (f2 + g2);
}
// PowerHeuristic(f, nf, g, ng, h, nh), :121-130 (denominator order: nf f^2 + ng g^2 + (nh h^2 + 0))
float PowerHeuristic3(float f, int nf, float g, int ng, float h, int nh) {
    const float f2 = f * f, g2 = g * g;
    const float rest = float(nh) * (h * h) + 0.0f;
    const float denom = float(nf) * f2 + float(ng) * g2 + rest;
    return denom > 0.0f ? f2 / denom : 0.0f;
}
void BeerAbsorption(Spec &tp, const Closure &c, float t, bool isInside, const Lambda &lam) {  // :133-145
    if (isInside && c.TransmissionDepth > 0.0f && !c.ThinWalled) {
        const Spec ts = RGBToSpectrum(c.TransmissionColor, lam);
        for (int i = 0; i < kSpec; ++i) {
            const float sigma_a = -std::log(std::max(ts.v[i], 1e-4f)) / c.TransmissionDepth;
            tp.v[i] *= std::exp(-sigma_a * t);
        }
    }
}

// ---------------------------------------------------------------- GGXBsdf (source/materials/ggx.cpp:20-67)
V3 GGXEval(const Closure &c, const V3 &n, const V3 &wo, const V3 &wi) {
    const V3 h = Normalized(wo + wi);
    const float nDotL = std::max(0.001f, Dot(n, wi));
    const float nDotV = std::max(0.001f, Dot(n, wo));
    const float nDotH = std::max(0.001f, Dot(n, h));
    const float lDotH = std::max(0.001f, Dot(wi, h));
    const float alpha = std::max(0.001f, c.Roughness * c.Roughness);
    const float alpha2 = alpha * alpha;
    const float denom = nDotH * nDotH * (alpha2 - 1.0f) + 1.0f;
    const float Dval = alpha2 / (kPiF * denom * denom);
    const float k_g = alpha / 2.0f;
    const float G_l = nDotL / (nDotL * (1.0f - k_g) + k_g);
    const float G_v = nDotV / (nDotV * (1.0f - k_g) + k_g);
    const float Gval = G_l * G_v;
    const V3 F0 = c.SpecularColor * (1.0f - c.Metallic) * 0.04f + c.BaseColor * c.Metallic;
    const V3 Fval = F0 + (V3(1.0f, 1.0f, 1.0f) - F0) * std::pow(1.0f - lDotH, 5.0f);
    const V3 specBsdf = divd(Fval * Dval * Gval, double(4.0f * nDotL * nDotV));
    const V3 finalDiffuse = c.BaseColor * (1.0f - c.Subsurface) + c.SubsurfaceColor * c.Subsurface;
    const V3 diffBsdf = divd(finalDiffuse * (1.0f - c.Metallic) * (1.0f - c.Transmission), double(kPiF));
    return diffBsdf + specBsdf;
}
float GGXPdfFull(const Closure &c, const V3 &n, const V3 &wo, const V3 &wi) {
    const V3 h = Normalized(wo + wi);
    const float nDotL = std::max(0.001f, Dot(n, wi));
    const float nDotV = std::max(0.001f, Dot(n, wo));
    const float nDotH = std::max(0.001f, Dot(n, h));
    const float lDotH = std::max(0.001f, Dot(wi, h));
    const float alpha = std::max(0.001f, c.Roughness * c.Roughness);
    const float alpha2 = alpha * alpha;
    const float denom = nDotH * nDotH * (alpha2 - 1.0f) + 1.0f;
    const float Dval = alpha2 / (kPiF * denom * denom);
    float reflectProb = FresnelDielectric(nDotV, c.Ior) * c.Specular;
    if (c.Metallic > 0.0f) reflectProb = std::max(reflectProb, c.Metallic);
    const float diffPdf = nDotL / kPiF;
    const float specPdf = (Dval * nDotH) / (4.0f * lDotH);
    return reflectProb * specPdf + (1.0f - reflectProb) * (1.0f - c.Transmission) * diffPdf;
}

// ---------------------------------------------------------------- scene data
struct MaterialParams {  // PreviewSurfaceParams subset
    V3 DiffuseColor{1.0f};
    float Metallic = 0, Roughness = 0.5f;
    V3 SpecularColor{1.0f};
    float Specular = 1.0f, Opacity = 1.0f, Ior = 1.5f, Transmission = 0;
    V3 TransmissionColor{1.0f};
    float TransmissionDepth = 0;
    V3 EmissionColor{1.0f};
    float Emission = 0, Coat = 0;
    V3 CoatColor{1.0f};
    float CoatRoughness = 0.1f, CoatIor = 1.5f, Sheen = 0;
    V3 SheenColor{1.0f};
    float SheenRoughness = 0.3f, Subsurface = 0;
    V3 SubsurfaceColor{1.0f};
    bool ThinWalled = false;
};

struct BVHNode {  // bvh.h:14-19
    Range3 bounds;
    int leftChild = 0;
    int triangleCount = 0;
};
struct BVHTriangle {  // bvh.h:21-30 (+ original index)
    V3 v0, v1, v2;
    V2 uv0, uv1, uv2;
    V3 n0, n1, n2, c0, c1, c2, dpdu, dpdv, centroid;
    int materialIndex = -1;
    int origIndex = -1;
};

// IntersectTriangle, bvh.cpp:8-30
bool IntersectTriangle(const V3 &o, const V3 &d, const V3 &v0, const V3 &v1, const V3 &v2, float &t, float &outU,
                       float &outV) {
    const V3 e1 = v1 - v0;
    const V3 e2 = v2 - v0;
    const V3 pvec = Cross(d, e2);
    const float det = Dot(e1, pvec);
    if (double(std::abs(det)) < 1e-8) return false;
    const float invDet = 1.0f / det;
    const V3 tvec = o - v0;
    const float u = Dot(tvec, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 qvec = Cross(tvec, e1);
    const float v = Dot(d, qvec) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = Dot(e2, qvec) * invDet;
    outU = u;
    outV = v;
    return double(t) > 1e-4;
}

// IntersectAABB, bvh.cpp:32-64 (double thresholds) and scene.cpp:13-46 (float thresholds)
template <bool kFloatThresholds>
bool IntersectAABB(const V3 &o, const V3 &d, const Range3 &r, float &tMinHit) {
    if (r.IsEmpty()) return false;
    float tmin = -1e30f, tmax = 1e30f;
    for (int i = 0; i < 3; ++i) {
        const bool parallel = kFloatThresholds ? (std::abs(d[i]) < 1e-8f) : (double(std::abs(d[i])) < 1e-8);
        if (parallel) {
            if (o[i] < r.mn[i] || o[i] > r.mx[i]) return false;
        } else {
            const float invDir = 1.0f / d[i];
            float t1 = (r.mn[i] - o[i]) * invDir;
            float t2 = (r.mx[i] - o[i]) * invDir;
            if (t1 > t2) std::swap(t1, t2);
            tmin = std::max(tmin, t1);
            tmax = std::min(tmax, t2);
            if (tmin > tmax) return false;
        }
    }
    tMinHit = tmin;
    if (kFloatThresholds) return tmax > 0.0f && tmax > 1e-4f;
    return tmax > 0 && double(tmax) > 1e-4;
}

struct BLAS {
    std::vector<BVHNode> nodes;
    std::vector<BVHTriangle> tris;
    Range3 range;  // MeshSubset::range (hd_restir_mesh.cpp:521-527)

    // BVH::Build, bvh.cpp:66-182
    void Build(const hdr_mesh_desc &m) {
        tris.clear();
        nodes.clear();
        if (m.num_triangles <= 0 || m.num_points <= 0) return;
        const int nIdx = m.num_triangles;
        auto P = [&](int k) { return V3(m.points[3 * k], m.points[3 * k + 1], m.points[3 * k + 2]); };
        for (int i = 0; i < nIdx; ++i) {
            const int i0 = m.indices[3 * i], i1 = m.indices[3 * i + 1], i2 = m.indices[3 * i + 2];
            if (i0 < 0 || i0 >= m.num_points || i1 < 0 || i1 >= m.num_points || i2 < 0 || i2 >= m.num_points) continue;
            BVHTriangle tri;
            tri.v0 = P(i0);
            tri.v1 = P(i1);
            tri.v2 = P(i2);
            const int ids[3] = {i0, i1, i2};
            auto uvAt = [&](int k) { return V2{m.uvs[2 * k], m.uvs[2 * k + 1]}; };
            auto v3At = [&](const float *a, int k) { return V3(a[3 * k], a[3 * k + 1], a[3 * k + 2]); };
            V2 *uvs[3] = {&tri.uv0, &tri.uv1, &tri.uv2};
            V3 *ns[3] = {&tri.n0, &tri.n1, &tri.n2};
            V3 *cs[3] = {&tri.c0, &tri.c1, &tri.c2};
            for (int k = 0; k < 3; ++k) {
                if (m.num_uvs > 0 && m.uvs) {
                    if (m.num_uvs == nIdx * 3) *uvs[k] = uvAt(i * 3 + k);
                    else *uvs[k] = (ids[k] < m.num_uvs) ? uvAt(ids[k]) : V2{0, 0};
                } else {
                    *uvs[k] = V2{0, 0};
                }
                if (m.num_normals > 0 && m.normals) {
                    if (m.num_normals == nIdx * 3) *ns[k] = v3At(m.normals, i * 3 + k);
                    else *ns[k] = (ids[k] < m.num_normals) ? v3At(m.normals, ids[k]) : V3(0, 1, 0);
                } else {
                    *ns[k] = V3(0, 0, 0);
                }
                if (m.num_colors > 0 && m.colors) {
                    if (m.num_colors == nIdx * 3) *cs[k] = v3At(m.colors, i * 3 + k);
                    else *cs[k] = (ids[k] < m.num_colors) ? v3At(m.colors, ids[k]) : V3(1.0f);
                } else {
                    *cs[k] = V3(1.0f);
                }
            }
            const V3 edge1 = tri.v1 - tri.v0;
            const V3 edge2 = tri.v2 - tri.v0;
            const V2 d1{tri.uv1.x - tri.uv0.x, tri.uv1.y - tri.uv0.y};
            const V2 d2{tri.uv2.x - tri.uv0.x, tri.uv2.y - tri.uv0.y};
            float f = d1.x * d2.y - d2.x * d1.y;
            if (std::abs(f) > 1e-8f) {
                f = 1.0f / f;
                tri.dpdu = Normalized((edge1 * d2.y - edge2 * d1.y) * f);
                tri.dpdv = Normalized((edge2 * d1.x - edge1 * d2.x) * f);
            } else {
                const V3 n = Normalized(Cross(edge1, edge2));
                const V3 up = std::abs(n[1]) < 0.999f ? V3(0, 1, 0) : V3(1, 0, 0);
                tri.dpdu = Normalized(Cross(up, n));
                tri.dpdv = Normalized(Cross(n, tri.dpdu));
            }
            tri.materialIndex = -1;
            tri.centroid = divd(tri.v0 + tri.v1 + tri.v2, double(3.0f));
            tri.origIndex = i;
            tris.push_back(tri);
        }
        range.SetEmpty();
        for (int i = 0; i < nIdx; ++i)
            for (int k = 0; k < 3; ++k) {
                const int id = m.indices[3 * i + k];
                if (id >= 0 && id < m.num_points) range.ExtendBy(P(id));
            }
        if (tris.empty()) return;
        nodes.reserve(size_t(nIdx) * 2);
        nodes.push_back(BVHNode());
        Subdivide(0, 0, int(tris.size()));
    }

    // BVH::_Subdivide, bvh.cpp:184-241
    void Subdivide(int nodeIdx, int start, int end) {
        BVHNode &node = nodes[nodeIdx];
        node.bounds.SetEmpty();
        for (int i = start; i < end; ++i) {
            node.bounds.ExtendBy(tris[i].v0);
            node.bounds.ExtendBy(tris[i].v1);
            node.bounds.ExtendBy(tris[i].v2);
        }
        const int count = end - start;
        if (count <= 4) {
            node.leftChild = -start - 1;
            node.triangleCount = count;
            return;
        }
        const V3 size = node.bounds.GetSize();
        int axis = 0;
        if (size[1] > size[0]) axis = 1;
        if (size[2] > size[axis]) axis = 2;
        const float splitPos = node.bounds.mn[axis] + size[axis] * 0.5f;
        int i = start, j = end - 1;
        while (i <= j) {
            if (tris[i].centroid[axis] < splitPos) {
                i++;
            } else {
                std::swap(tris[i], tris[j]);
                j--;
            }
        }
        const int leftCount = i - start;
        if (leftCount == 0 || leftCount == count) i = start + count / 2;
        const int leftChildIdx = int(nodes.size());
        nodes.push_back(BVHNode());
        nodes.push_back(BVHNode());
        nodes[nodeIdx].leftChild = leftChildIdx;
        nodes[nodeIdx].triangleCount = 0;
        Subdivide(leftChildIdx, start, i);
        Subdivide(leftChildIdx + 1, i, end);
    }

    struct Out {
        float t;
        V3 normal, smoothNormal, dpdu, dpdv, color;
        V2 uv;
        int mat, tri;
    };

    // BVH::_IntersectNode, bvh.cpp:251-305
    bool IntersectNode(int nodeIdx, const V3 &o, const V3 &d, Out &out, Counters &cnt) const {
        const BVHNode &node = nodes[nodeIdx];
        ++cnt.nodes;
        float tAabb;
        if (!IntersectAABB<false>(o, d, node.bounds, tAabb)) return false;
        if (tAabb > out.t) return false;
        if (node.leftChild < 0) {
            bool hit = false;
            const int start = -node.leftChild - 1;
            for (int i = 0; i < node.triangleCount; ++i) {
                const BVHTriangle &tri = tris[start + i];
                ++cnt.tris;
                float triT, triU, triV;
                if (IntersectTriangle(o, d, tri.v0, tri.v1, tri.v2, triT, triU, triV)) {
                    if (triT < out.t) {
                        out.t = triT;
                        out.normal = Normalized(Cross(tri.v1 - tri.v0, tri.v2 - tri.v0));
                        const float w = 1.0f - triU - triV;
                        out.uv = V2{tri.uv0.x * w + tri.uv1.x * triU + tri.uv2.x * triV,
                                    tri.uv0.y * w + tri.uv1.y * triU + tri.uv2.y * triV};
                        if (Dot(tri.n0, tri.n0) > 1e-6f)
                            out.smoothNormal = Normalized(tri.n0 * w + tri.n1 * triU + tri.n2 * triV);
                        else
                            out.smoothNormal = out.normal;
                        out.dpdu = tri.dpdu;
                        out.dpdv = tri.dpdv;
                        out.color = tri.c0 * w + tri.c1 * triU + tri.c2 * triV;
                        out.mat = tri.materialIndex;
                        out.tri = tri.origIndex;
                        hit = true;
                    }
                }
            }
            return hit;
        }
        const bool hitLeft = IntersectNode(node.leftChild, o, d, out, cnt);
        const bool hitRight = IntersectNode(node.leftChild + 1, o, d, out, cnt);
        return hitLeft || hitRight;
    }
};

// ---------------------------------------------------------------- lights
struct Light {
    int type;
    V3 Color;
    float Intensity, Exposure;
    M4d T;
    float Width, Height, ConeAngle, ConeSoftness;
    // dome
    std::vector<float> pixels, rowCdf, colCdf;
    int w = 0, h = 0;
    float totalLum = 0.0f;

    float EffectiveIntensity() const { return Intensity * std::exp2(Exposure); }  // light_params.h:27-30
    bool IsDelta() const { return type == HDR_LIGHT_DISTANT; }  // distant_light.h:19-22; others default false

    // DomeLight::Prepare, dome_light.cpp:21-77
    void PrepareDome(const float *px) {
        pixels.assign(px, px + size_t(w) * h * 3);
        rowCdf.assign(h + 1, 0.0f);
        colCdf.assign(size_t(h) * (w + 1), 0.0f);
        totalLum = 0.0f;
        for (int y = 0; y < h; ++y) {
            float rowLum = 0.0f;
            const float sinTheta = float(std::sin(kPiD * double(float(y + 0.5f)) / double(float(h))));
            for (int x = 0; x < w; ++x) {
                const size_t idx = size_t(y * w + x) * 3;
                float lum = 0.2126f * pixels[idx] + 0.7152f * pixels[idx + 1] + 0.0722f * pixels[idx + 2];
                lum *= sinTheta;
                rowLum += lum;
                colCdf[size_t(y) * (w + 1) + x + 1] = colCdf[size_t(y) * (w + 1) + x] + lum;
            }
            if (rowLum > 0)
                for (int x = 0; x <= w; ++x) colCdf[size_t(y) * (w + 1) + x] /= rowLum;
            totalLum += rowLum;
            rowCdf[y + 1] = rowCdf[y] + rowLum;
        }
        if (totalLum > 0)
            for (int y = 0; y <= h; ++y) rowCdf[y] /= totalLum;
    }
    float DomePdfScale() const {
        return totalLum * (float(kPiD) / float(h)) * (2.0f * float(kPiD) / float(w));
    }
    void DomeTexel(const V3 &dir, int &x, int &y) const {  // dome_light.cpp:123-128 / 139-144
        const float theta = std::acos(std::clamp(dir[1], -1.0f, 1.0f));
        float phi = std::atan2(dir[2], dir[0]);
        if (phi < 0) phi = float(double(phi) + 2.0 * kPiD);
        x = std::clamp(int(double(phi) / (2.0 * kPiD) * w), 0, w - 1);
        y = std::clamp(int(double(theta) / kPiD * h), 0, h - 1);
    }
    float DomeEvalPdf(const V3 &dir) const {  // :119-133
        if (pixels.empty() || totalLum <= 0.f) return 0.f;
        int x, y;
        DomeTexel(dir, x, y);
        const size_t idx = size_t(y * w + x) * 3;
        const float lum = 0.2126f * pixels[idx] + 0.7152f * pixels[idx + 1] + 0.0722f * pixels[idx + 2];
        return std::max(lum / DomePdfScale(), 1e-6f);
    }
    V3 DomeSample(const V3 &dir) const {  // :135-148
        if (pixels.empty()) return V3(0.0f);
        int x, y;
        DomeTexel(dir, x, y);
        const size_t idx = size_t(y * w + x) * 3;
        const V3 c(pixels[idx], pixels[idx + 1], pixels[idx + 2]);
        return CompMult(c, Color) * EffectiveIntensity();
    }

    struct Sample {
        V3 Dir, Color, LightNormal;
        float Dist = 1e30f;
        Pdf pdf;
    };

    std::optional<Sample> SampleLight(const V3 &hitPos, Rng &rng) const {
        switch (type) {
            case HDR_LIGHT_RECT: {  // rect_light.cpp:14-37
                const float u = rng.NextFloat() - 0.5f;
                const float v = rng.NextFloat() - 0.5f;
                const V3 lLocal(u * Width, v * Height, 0.0f);
                const M4f Tf = M4f::FromD(T);
                const V3 lWorld = Tf.Transform(lLocal);
                const V3 toLight = lWorld - hitPos;
                const float dist = Length(toLight);
                const V3 lDir = divd(toLight, double(dist));
                const float area = Width * Height;
                if (area <= 0.0f) return std::nullopt;
                const V3 lNormal = Normalized(Tf.TransformDir(V3(0, 0, -1)));
                const float cosL = std::max(0.0f, Dot(lNormal, -lDir));
                if (cosL <= 0.0f) return std::nullopt;
                const V3 lColor = Color * EffectiveIntensity();
                if (lColor[0] <= 0.0f && lColor[1] <= 0.0f && lColor[2] <= 0.0f) return std::nullopt;
                return Sample{lDir, lColor, lNormal, dist, Pdf{1.f / area, PdfSpace::Area}};
            }
            case HDR_LIGHT_POINT: {  // point_light.cpp:16-55
                const M4f Tf = M4f::FromD(T);
                const V3 lPos = Tf.ExtractTranslation();
                const V3 toLight = lPos - hitPos;
                const float dist = Length(toLight);
                const V3 lDir = divd(toLight, double(dist));
                V3 lColor = Color * EffectiveIntensity();
                const float coneAngle = ConeAngle;
                if (coneAngle < 180.0f) {
                    const V3 lNormal = Normalized(Tf.TransformDir(V3(0, 0, -1)));
                    const float cosTheta = Dot(lNormal, -lDir);
                    const float coneRad = coneAngle * float(kPiD / 180.0);
                    const float cosCone = std::cos(coneRad);
                    if (cosTheta <= cosCone) {
                        lColor = V3(0.0f);
                    } else if (ConeSoftness > 0.0f) {
                        const float innerRad = coneRad * (1.0f - ConeSoftness);
                        const float cosInner = std::cos(innerRad);
                        if (cosTheta < cosInner) {
                            float f = (cosTheta - cosCone) / (cosInner - cosCone);
                            f = f * f * (3.0f - 2.0f * f);
                            lColor = lColor * f;
                        }
                    }
                }
                if (lColor[0] <= 0.0f && lColor[1] <= 0.0f && lColor[2] <= 0.0f) return std::nullopt;
                return Sample{lDir, lColor, -lDir, dist, Pdf{1.f, PdfSpace::SolidAngle}};
            }
            case HDR_LIGHT_DISTANT: {  // distant_light.cpp:10-17
                const V3 lDir = Normalized(M4f::FromD(T).TransformDir(V3(0, 0, -1)));
                const V3 lColor = Color * EffectiveIntensity();
                if (lColor[0] <= 0.0f && lColor[1] <= 0.0f && lColor[2] <= 0.0f) return std::nullopt;
                return Sample{lDir, lColor, -lDir, 1e30f, Pdf{1.f, PdfSpace::SolidAngle}};
            }
            case HDR_LIGHT_DOME: {  // dome_light.cpp:79-111
                if (pixels.empty()) return std::nullopt;
                const float u1 = rng.NextFloat();
                const float u2 = rng.NextFloat();
                const auto itY = std::lower_bound(rowCdf.begin(), rowCdf.end(), u1);
                const int y = std::clamp(int(std::distance(rowCdf.begin(), itY)) - 1, 0, h - 1);
                const float *cc = &colCdf[size_t(y) * (w + 1)];
                const float *itX = std::lower_bound(cc, cc + w + 1, u2);
                const int x = std::clamp(int(itX - cc) - 1, 0, w - 1);
                const float theta = float(kPiD * double(float(y + 0.5f)) / double(float(h)));
                const float phi = float(2.0 * kPiD * double(float(x + 0.5f)) / double(float(w)));
                const V3 lDir(std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi));
                const size_t idx = size_t(y * w + x) * 3;
                const V3 tex(pixels[idx], pixels[idx + 1], pixels[idx + 2]);
                const V3 lColor = CompMult(tex, Color) * EffectiveIntensity();
                float pdf = 1.0f;
                if (totalLum > 0) {
                    const float lum = 0.2126f * tex[0] + 0.7152f * tex[1] + 0.0722f * tex[2];
                    pdf = std::max(lum / DomePdfScale(), 1e-6f);
                }
                if (lColor[0] <= 0.0f && lColor[1] <= 0.0f && lColor[2] <= 0.0f) return std::nullopt;
                return Sample{lDir, lColor, -lDir, 1e30f, Pdf{pdf, PdfSpace::SolidAngle}};
            }
        }
        return std::nullopt;
    }

    Pdf EvalPdf(const V3 &dir) const {  // ILight::EvalPdf for each light type
        switch (type) {
            case HDR_LIGHT_RECT: {
                const float area = Width * Height;
                if (area <= 0.f) return {0.f, PdfSpace::Area};
                return {1.f / area, PdfSpace::Area};
            }
            case HDR_LIGHT_DOME:
                return {DomeEvalPdf(dir), PdfSpace::SolidAngle};
            default:
                return {0.f, PdfSpace::Area};
        }
    }
};

// PreviewSurfaceParams (preview_surface_params.h:12-52, textures excluded) from a material desc.
MaterialParams MaterialFromDesc(const hdr_material_desc &s) {
    MaterialParams p;
    auto v3 = [](const float *a) { return V3(a[0], a[1], a[2]); };
    p.DiffuseColor = v3(s.diffuse_color);
    p.Metallic = s.metallic;
    p.Roughness = s.roughness;
    p.SpecularColor = v3(s.specular_color);
    p.Specular = s.specular;
    p.Opacity = s.opacity;
    p.Ior = s.ior;
    p.Transmission = s.transmission;
    p.TransmissionColor = v3(s.transmission_color);
    p.TransmissionDepth = s.transmission_depth;
    p.EmissionColor = v3(s.emission_color);
    p.Emission = s.emission;
    p.Coat = s.coat;
    p.CoatColor = v3(s.coat_color);
    p.CoatRoughness = s.coat_roughness;
    p.CoatIor = s.coat_ior;
    p.Sheen = s.sheen;
    p.SheenColor = v3(s.sheen_color);
    p.SheenRoughness = s.sheen_roughness;
    p.Subsurface = s.subsurface;
    p.SubsurfaceColor = v3(s.subsurface_color);
    p.ThinWalled = s.thin_walled != 0;
    return p;
}

// LightFactory (light_factory.cpp:11-25) + LightParams (light_params.h:15-31) from a light desc.
void MakeLight(const hdr_light_desc &s, Light &l) {
    l.type = s.type;
    l.Color = V3(s.color[0], s.color[1], s.color[2]);
    l.Intensity = s.intensity;
    l.Exposure = s.exposure;
    l.T = M4d::From(s.transform);
    l.Width = s.width;
    l.Height = s.height;
    l.ConeAngle = s.shaping_cone_angle;
    l.ConeSoftness = s.shaping_cone_softness;
    if (s.type == HDR_LIGHT_DOME && s.dome_pixels && s.dome_width > 0 && s.dome_height > 0) {
        l.w = s.dome_width;
        l.h = s.dome_height;
        l.PrepareDome(s.dome_pixels);
    }
}

// ---------------------------------------------------------------- scene (source/scene/scene.cpp)
struct Instance {
    int mesh;
    const MaterialParams *params;  // nullptr -> DefaultMaterial
    int light;
    M4f transform, invTransform;
    Range3 bounds;
    V3 centroid;
};
struct TLASNode {
    Range3 bounds;
    int leftChild = 0;
    int instanceCount = 0;
};

struct SurfaceHit {  // Scene::SurfaceHit subset used by IntersectScene
    float t = 1e30f;
    V3 normal, smoothNormal, dpdu{1, 0, 0}, dpdv{0, 1, 0};
    V2 uv;
    V3 baseColor{1.0f};
    int instanceIdx = -1;
    int tri = -1;
};

struct Scene {
    std::vector<BLAS> blas;
    std::vector<MaterialParams> materials;
    std::vector<Instance> instances;
    std::vector<Light> lights;
    std::vector<const Light *> lightPtrs;
    int environment = -1;
    std::vector<TLASNode> tlas;
    std::vector<int> tlasIdx;

    static Range3 TransformBounds(const Range3 &b, const M4f &m) {  // scene.cpp:100-116
        if (b.IsEmpty()) return b;
        const V3 mn = b.mn, mx = b.mx;
        const V3 corners[8] = {V3(mn[0], mn[1], mn[2]), V3(mx[0], mn[1], mn[2]), V3(mn[0], mx[1], mn[2]),
                               V3(mx[0], mx[1], mn[2]), V3(mn[0], mn[1], mx[2]), V3(mx[0], mn[1], mx[2]),
                               V3(mn[0], mx[1], mx[2]), V3(mx[0], mx[1], mx[2])};
        Range3 r;
        for (const V3 &c : corners) r.ExtendBy(m.Transform(c));
        return r;
    }

    void Load(const hdr_scene_desc &d) {
        blas.resize(d.num_meshes);
        for (int i = 0; i < d.num_meshes; ++i) blas[i].Build(d.meshes[i]);
        materials.resize(d.num_materials);
        for (int i = 0; i < d.num_materials; ++i) materials[i] = MaterialFromDesc(d.materials[i]);
        lights.resize(d.num_lights);
        for (int i = 0; i < d.num_lights; ++i) MakeLight(d.lights[i], lights[i]);
        for (const Light &l : lights) lightPtrs.push_back(&l);
        environment = d.environment;
        // appendMeshInstance, scene.cpp:178-207
        instances.resize(d.num_instances);
        for (int i = 0; i < d.num_instances; ++i) {
            const hdr_instance_desc &s = d.instances[i];
            Instance &inst = instances[i];
            inst.mesh = s.mesh;
            inst.params = (s.material >= 0 && s.material < d.num_materials) ? &materials[s.material] : nullptr;
            inst.light = s.light;
            inst.transform = M4f::From(s.transform);
            inst.invTransform = inst.transform.Inverse();
            inst.bounds = TransformBounds(blas[s.mesh].range, inst.transform);
            inst.centroid = (inst.bounds.mn + inst.bounds.mx) * 0.5f;
        }
        BuildTLAS();
    }

    void BuildTLAS() {  // scene.cpp:333-350
        tlas.clear();
        tlasIdx.clear();
        if (instances.empty()) return;
        tlasIdx.resize(instances.size());
        for (size_t i = 0; i < instances.size(); ++i) tlasIdx[i] = int(i);
        tlas.reserve(instances.size() * 2);
        tlas.push_back(TLASNode{});
        SubdivideTLAS(0, 0, int(instances.size()));
    }

    void SubdivideTLAS(int nodeIdx, int start, int end) {  // scene.cpp:352-407
        const int leftChildIdx = int(tlas.size());
        tlas.push_back(TLASNode{});
        tlas.push_back(TLASNode{});
        tlas[nodeIdx].leftChild = leftChildIdx;
        tlas[nodeIdx].instanceCount = 0;
        tlas[nodeIdx].bounds.SetEmpty();
        for (int i = start; i < end; ++i) {
            const Range3 &b = instances[tlasIdx[i]].bounds;
            tlas[nodeIdx].bounds.ExtendBy(b.mn);
            tlas[nodeIdx].bounds.ExtendBy(b.mx);
        }
        const int count = end - start;
        if (count <= 2) {
            tlas[nodeIdx].leftChild = -start - 1;
            tlas[nodeIdx].instanceCount = count;
            return;
        }
        const V3 size = tlas[nodeIdx].bounds.GetSize();
        int axis = 0;
        if (size[1] > size[0]) axis = 1;
        if (size[2] > size[axis]) axis = 2;
        const float splitPos = tlas[nodeIdx].bounds.mn[axis] + size[axis] * 0.5f;
        int i = start, j = end - 1;
        while (i <= j) {
            if (instances[tlasIdx[i]].centroid[axis] < splitPos) {
                ++i;
            } else {
                std::swap(tlasIdx[i], tlasIdx[j]);
                --j;
            }
        }
        if (i == start || i == end) i = start + count / 2;
        SubdivideTLAS(leftChildIdx, start, i);
        SubdivideTLAS(leftChildIdx + 1, i, end);
    }

    // Scene::Intersect, scene.cpp:409-513
    bool Intersect(const V3 &o, const V3 &d, SurfaceHit &hit, Counters &cnt) const {
        if (tlas.empty()) return false;
        int stack[64];
        int sp = 0;
        stack[sp++] = 0;
        bool wasHit = false;
        while (sp > 0) {
            const int nodeIdx = stack[--sp];
            const TLASNode &node = tlas[nodeIdx];
            float tAabb = 0.0f;
            if (!IntersectAABB<true>(o, d, node.bounds, tAabb) || tAabb > hit.t) continue;
            if (node.leftChild < 0) {
                const int start = -node.leftChild - 1;
                for (int k = 0; k < node.instanceCount; ++k) {
                    const int instanceIdx = tlasIdx[start + k];
                    const Instance &inst = instances[instanceIdx];
                    ++cnt.instances;
                    const V3 oo = inst.invTransform.Transform(o);
                    const V3 od = inst.invTransform.TransformDir(d);
                    BLAS::Out out{};
                    out.t = hit.t;
                    const BLAS &b = blas[inst.mesh];
                    if (b.nodes.empty() || !b.IntersectNode(0, oo, od, out, cnt)) continue;
                    if (out.t >= hit.t) continue;
                    hit.instanceIdx = instanceIdx;
                    hit.tri = out.tri;
                    hit.t = out.t;
                    hit.normal = Normalized(inst.transform.TransformDir(out.normal));
                    hit.smoothNormal = Normalized(inst.transform.TransformDir(out.smoothNormal));
                    hit.dpdu = Normalized(inst.transform.TransformDir(out.dpdu));
                    hit.dpdv = Normalized(inst.transform.TransformDir(out.dpdv));
                    hit.uv = out.uv;
                    hit.baseColor = out.color;
                    if (inst.params) hit.baseColor = inst.params->DiffuseColor;
                    wasHit = true;
                }
            } else {
                stack[sp++] = node.leftChild + 1;
                stack[sp++] = node.leftChild;
            }
        }
        return wasHit;
    }

    // Scene::IntersectScene, scene.cpp:515-535
    std::optional<HitRecord> IntersectScene(const V3 &o, const V3 &d, Counters &cnt) const {
        SurfaceHit hit{};
        if (!Intersect(o, d, hit, cnt)) return std::nullopt;
        HitRecord r;
        r.Position = o + hit.t * d;
        r.Normal = hit.normal;
        r.SmoothNormal = hit.smoothNormal;
        r.Dpdu = hit.dpdu;
        r.Dpdv = hit.dpdv;
        r.Uv = hit.uv;
        r.Albedo = hit.baseColor;
        r.Depth = hit.t;
        r.PrimId = hit.instanceIdx;
        r.MatId = hit.instanceIdx;
        r.TriId = hit.tri;
        return r;
    }

    const Light *GetLightAtHit(const HitRecord &h) const {  // scene.cpp:300-326
        if (h.PrimId < 0 || h.PrimId >= int(instances.size())) return nullptr;
        const int l = instances[h.PrimId].light;
        return (l >= 0 && l < int(lights.size())) ? &lights[l] : nullptr;
    }
    const Light *GetEnvironment() const {
        return (environment >= 0 && environment < int(lights.size())) ? &lights[environment] : nullptr;
    }
    // Scene::GetMaterial, scene.cpp:273-281: instance shader or DefaultMaterial
    const MaterialParams *GetMaterial(int matId) const {
        if (matId < 0 || matId >= int(instances.size())) return nullptr;
        return instances[matId].params;
    }
};

// ---------------------------------------------------------------- materials
// PreviewSurfaceMaterial::GetClosure (preview_surface.cpp:61-124, no textures);
// DefaultMaterial::GetClosure (default_material.cpp:55-63) when p == nullptr.
Closure GetClosure(const MaterialParams *p, const HitRecord &hit) {
    Closure c;
    if (!p) {
        c.BaseColor = hit.Albedo;
        c.Normal = hit.Normal;
        c.Metallic = 0.0f;
        c.Roughness = 1.0f;
        return c;
    }
    c.BaseColor = p->DiffuseColor;
    c.Emission = p->EmissionColor * p->Emission;
    c.Normal = hit.SmoothNormal;
    c.SpecularColor = p->SpecularColor;
    c.CoatColor = p->CoatColor;
    c.SheenColor = p->SheenColor;
    c.SubsurfaceColor = p->SubsurfaceColor;
    c.TransmissionColor = p->TransmissionColor;
    c.Metallic = p->Metallic;
    c.Roughness = std::max(0.001f, p->Roughness);
    c.Ior = p->Ior;
    c.Specular = p->Specular;
    c.Transmission = p->Transmission;
    c.TransmissionDepth = p->TransmissionDepth;
    c.Coat = p->Coat;
    c.CoatRoughness = p->CoatRoughness;
    c.CoatIor = p->CoatIor;
    c.Sheen = p->Sheen;
    c.SheenRoughness = p->SheenRoughness;
    c.Subsurface = p->Subsurface;
    c.Opacity = p->Opacity;
    c.ThinWalled = p->ThinWalled;
    return c;
}

V3 MirrorDir(const V3 &d, const V3 &n) { return Normalized(d - 2.0f * Dot(d, n) * n); }

// Rough reflection shared by the coat, sheen and base reflect lobes (preview_surface.cpp:146-170 etc.)
struct RoughReflect {
    V3 dir, h;
    float cosH, nDotV, nDotL, nDotH;
};
RoughReflect SampleRoughReflect(const V3 &rayDir, const V3 &n, float roughness, Rng &rng) {
    const float u1 = rng.NextFloat();
    const float u2 = rng.NextFloat();
    RoughReflect r;
    r.h = AlignToNormal(SampleGGX(u1, u2, roughness), n);
    r.dir = Normalized(rayDir - 2.0f * Dot(rayDir, r.h) * r.h);
    if (Dot(r.dir, n) < 0) r.dir = MirrorDir(rayDir, n);
    r.cosH = std::max(0.f, Dot(r.dir, r.h));
    r.nDotV = std::max(0.f, Dot(n, -rayDir));
    r.nDotL = std::max(0.f, Dot(n, r.dir));
    r.nDotH = std::max(0.f, Dot(n, r.h));
    return r;
}

// Coat / sheen layer (preview_surface.cpp:136-236). Returns true when the layer was chosen.
bool LayerBounce(const ShadingPoint &sp, const V3 &hitPos, const V3 &rayDir, const BounceConfig &cfg,
                 BounceState &st, Rng &rng, float prob, float roughness, const V3 &tint, BounceSample &out,
                 bool &error) {
    if (!(rng.NextFloat() < prob)) return false;
    if (st.refl >= cfg.maxRefl) {
        error = true;
        return true;
    }
    ++st.refl;
    const V3 &n = sp.n;
    if (roughness > 0.0f) {
        const RoughReflect r = SampleRoughReflect(rayDir, n, roughness, rng);
        const float alpha = roughness * roughness;
        const float pdf = r.cosH > 0.f ? GGXPdf(r.h, n, alpha) / (4.f * r.cosH) : 0.f;
        const float wgt = GGXReflectionWeight(r.nDotV, r.nDotL, r.cosH, r.nDotH, alpha);
        const Spec tMul = RGBToSpectrum(tint, *sp.lambda) * wgt;
        out = BounceSample{};
        out.NextRay = {hitPos + n * 1e-4f, r.dir};
        out.ThroughputMul = tMul;
        out.ThroughputIntegrandMul = tMul * pdf;
        out.BsdfPdf = {pdf, PdfSpace::SolidAngle};
        out.ImpossibleNEEConnection = false;
        out.SkipRoulette = true;
    } else {
        out = BounceSample{};
        out.NextRay = {hitPos + n * 1e-4f, MirrorDir(rayDir, n)};
        out.ThroughputMul = RGBToSpectrum(tint, *sp.lambda);
        out.BsdfPdf = {0.f, PdfSpace::SolidAngle};
        out.ImpossibleNEEConnection = true;
        out.SkipRoulette = true;
    }
    return true;
}

// PreviewSurfaceMaterial::SampleBounce, preview_surface.cpp:126-381. Returns false on a bounce-limit error.
bool PreviewSampleBounce(const ShadingPoint &sp, const V3 &hitPos, const V3 &rayDir, const BounceConfig &cfg,
                         BounceState &st, Rng &rng, BounceSample &out) {
    const Closure &c = *sp.c;
    const V3 &n = sp.n;
    const Lambda &lam = *sp.lambda;
    bool error = false;
    if (c.Coat > 0.0f && !sp.isInside) {
        const float coatFresnel = c.Coat * FresnelDielectric(Dot(rayDir, n), c.CoatIor);
        if (LayerBounce(sp, hitPos, rayDir, cfg, st, rng, coatFresnel, c.CoatRoughness, c.CoatColor, out, error))
            return !error;
    }
    if (c.Sheen > 0.0f && !sp.isInside) {
        const float cosTheta = std::max(0.0f, Dot(-rayDir, n));
        const float sheenFresnel = c.Sheen * std::pow(1.0f - cosTheta, 5.0f);
        if (LayerBounce(sp, hitPos, rayDir, cfg, st, rng, sheenFresnel, c.SheenRoughness, c.SheenColor, out, error))
            return !error;
    }
    float localIor = c.Ior;
    {
        const float C = 10000.0f;
        const float B = localIor - C / (589.3f * 589.3f);
        localIor = B + C / (lam.l[0] * lam.l[0]);
    }
    const float fresnel = FresnelDielectric(Dot(rayDir, n), localIor);
    float reflectProb = fresnel * c.Specular;
    if (c.Metallic > 0.0f) reflectProb = std::max(reflectProb, c.Metallic);
    const float randVal = rng.NextFloat();
    if (randVal < reflectProb) {
        if (st.refl >= cfg.maxRefl) return false;
        ++st.refl;
        const V3 reflTint = c.SpecularColor * (1.0f - c.Metallic) + c.BaseColor * c.Metallic;
        if (c.Roughness > 0.0f) {
            const RoughReflect r = SampleRoughReflect(rayDir, n, c.Roughness, rng);
            const float baseAlpha = c.Roughness * c.Roughness;
            const float condPdf = r.cosH > 0.f ? GGXPdf(r.h, n, baseAlpha) / (4.f * r.cosH) : 0.f;
            const float wgt = GGXReflectionWeight(r.nDotV, r.nDotL, r.cosH, r.nDotH, baseAlpha);
            const Spec tMul = RGBToSpectrum(reflTint, lam) * wgt;
            const float bsdfPdf = GGXPdfFull(c, n, -rayDir, r.dir);
            out = BounceSample{};
            out.NextRay = {hitPos + n * 1e-4f, r.dir};
            out.ThroughputMul = tMul;
            out.ThroughputIntegrandMul = tMul * condPdf;
            out.BsdfPdf = {bsdfPdf, PdfSpace::SolidAngle};
            out.ImpossibleNEEConnection = false;
        } else {
            out = BounceSample{};
            out.NextRay = {hitPos + n * 1e-4f, MirrorDir(rayDir, n)};
            out.ThroughputMul = RGBToSpectrum(reflTint, lam);
            out.BsdfPdf = {0.f, PdfSpace::SolidAngle};
            out.ImpossibleNEEConnection = true;
        }
        return true;
    }
    const float remainingProb = (randVal - reflectProb) / (1.0f - reflectProb);
    if (c.Transmission > 1e-6f && remainingProb < c.Transmission) {
        float etaI = 1.0f, etaT = localIor;
        if (sp.isInside) std::swap(etaI, etaT);
        const float eta = etaI / etaT;
        const float cosThetaI = Dot(rayDir, n);
        const float k = 1.0f - eta * eta * (1.0f - cosThetaI * cosThetaI);
        if (k >= 0) {
            if (st.refr >= cfg.maxRefr) return false;
            ++st.refr;
            V3 refractDir = Normalized(eta * rayDir - (eta * cosThetaI + std::sqrt(k)) * n);
            if (c.Roughness > 0.0f) {
                const float u1 = rng.NextFloat();
                const float u2 = rng.NextFloat();
                const V3 h = AlignToNormal(SampleGGX(u1, u2, c.Roughness), n);
                const float cosH = Dot(rayDir, h);
                const float kh = 1.0f - eta * eta * (1.0f - cosH * cosH);
                if (kh >= 0.0f)
                    refractDir = Normalized(eta * rayDir - (eta * cosH + std::sqrt(kh)) * h);
                else
                    refractDir = Normalized(rayDir - 2.0f * cosH * h);
            }
            out = BounceSample{};
            out.NextRay = {hitPos - n * 1e-4f, refractDir};
            out.ThroughputMul = RGBToSpectrum(c.TransmissionColor, lam);
            out.BsdfPdf = {0.f, PdfSpace::SolidAngle};
            out.ImpossibleNEEConnection = true;
            return true;
        }
        if (st.refl >= cfg.maxRefl) return false;
        ++st.refl;
        out = BounceSample{};
        out.NextRay = {hitPos + n * 1e-4f, Normalized(rayDir - 2.0f * cosThetaI * n)};
        out.BsdfPdf = {0.f, PdfSpace::SolidAngle};
        out.ImpossibleNEEConnection = true;
        return true;
    }
    const float u1 = rng.NextFloat();
    const float u2 = rng.NextFloat();
    const V3 dir = AlignToNormal(SampleCosineHemisphere(u1, u2), n);
    const float nDotL = std::max(0.0f, Dot(n, dir));
    const float rawPdf = nDotL / kPiF;
    const float safePdf = std::max(rawPdf, 1e-6f);
    const float scale = rawPdf / safePdf;
    const V3 finalDiffuse = c.BaseColor * (1.0f - c.Subsurface) + c.SubsurfaceColor * c.Subsurface;
    out = BounceSample{};
    out.NextRay = {hitPos + n * 1e-4f, dir};
    out.ThroughputMul = RGBToSpectrum(finalDiffuse, lam) * scale;
    out.ThroughputIntegrandMul = RGBToSpectrum(finalDiffuse, lam) * (nDotL / kPiF);
    out.BsdfPdf = {safePdf, PdfSpace::SolidAngle};
    out.ImpossibleNEEConnection = false;
    return true;
}

// DefaultMaterial::SampleBounce, default_material.cpp:65-86 (Duff et al. ONB, z-up hemisphere)
bool DefaultSampleBounce(const ShadingPoint &sp, const V3 &hitPos, Rng &rng, BounceSample &out) {
    const V3 &n = sp.n;
    const float sign = (n[2] >= 0.0f) ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n[2]);
    const float bb = n[0] * n[1] * a;
    const V3 t(1.0f + sign * n[0] * n[0] * a, sign * bb, -sign * n[0]);
    const V3 b(bb, sign + n[1] * n[1] * a, -n[1]);
    const float u1 = rng.NextFloat();
    const float u2 = rng.NextFloat();
    const float sinTheta = std::sqrt(u1);
    const float phi = 2.0f * kPiF * u2;
    const V3 wiL(sinTheta * std::cos(phi), sinTheta * std::sin(phi), std::sqrt(std::max(0.0f, 1.0f - u1)));
    const V3 wi = wiL[0] * t + wiL[1] * b + wiL[2] * n;
    const float nDotL = wiL[2];
    const float rawPdf = nDotL / kPiF;
    const float safePdf = std::max(rawPdf, 1e-6f);
    const float scale = rawPdf / safePdf;
    out = BounceSample{};
    out.NextRay = {hitPos + n * 1e-4f, wi};
    out.ThroughputMul = RGBToSpectrum(sp.c->BaseColor, *sp.lambda) * scale;
    out.ThroughputIntegrandMul = RGBToSpectrum(sp.c->BaseColor, *sp.lambda) * (nDotL / kPiF);
    out.BsdfPdf = {safePdf, PdfSpace::SolidAngle};
    out.ImpossibleNEEConnection = false;
    return true;
}

// ---------------------------------------------------------------- light sampler (uniform_light_sampler.cpp:12-52)
// UniformLightSampler::ProposeCandidate's index (uniform_light_sampler.cpp:30-33): float product u * n.
inline size_t LightIndex(float u, size_t n) { return std::min(size_t(u * float(n)), n - 1u); }

struct LightSampler {
    const std::vector<const Light *> *lights;
    std::unordered_set<const Light *> ownSet;
    const std::unordered_set<const Light *> *set;
    float selectPdf;
    LightSampler(const std::vector<const Light *> &l, const std::unordered_set<const Light *> *hoisted)
        : lights(&l), selectPdf(l.empty() ? 0.0f : 1.0f / float(l.size())) {
        if (hoisted) {
            set = hoisted;
        } else {
            ownSet.reserve(l.size());
            for (const Light *p : l) ownSet.insert(p);
            set = &ownSet;
        }
    }
    struct Candidate {
        const Light *light;
        Light::Sample ls;
    };
    std::optional<Candidate> Propose(const V3 &hitPos, Rng &rng, PixelTrace &tr) const {
        if (lights->empty()) return std::nullopt;
        const size_t n = lights->size();
        const size_t idx = LightIndex(rng.NextFloat(), n);
        tr.lightHash = Fnv(tr.lightHash, uint32_t(idx));
        ++tr.lightPicks;
        const Light &light = *(*lights)[idx];
        auto s = light.SampleLight(hitPos, rng);
        if (!s) return std::nullopt;
        s->pdf.value *= selectPdf;
        return Candidate{&light, *s};
    }
    float EvalPdf(const Light *l) const { return set->count(l) ? selectPdf : 0.0f; }
};

// ---------------------------------------------------------------- RIS types
struct RISCandidate {  // ris_direct_light_integrator.h:26-43
    std::optional<LightSampler::Candidate> Candidate;
    Spec Throughput;
    double risWeight = 0.0;
    double targetFunction = 0.0;
    bool VisibilityTested = true;
};
struct Reservoir {  // WeightedReservoir<RISLightCandidate>, reservoir.h:69-161
    std::optional<RISCandidate> ChosenSample;
    float WSum = 0.0f;
    int CandidateCount = 0;
    float W = 0.0f;
    void Update(const RISCandidate &s, float weight, Rng &rng, int index, int &sel) {
        ++CandidateCount;
        if (weight <= 0.0f) return;
        WSum += weight;
        if (rng.NextFloat() <= weight / WSum) {
            ChosenSample = s;
            sel = index;
        }
    }
    void CountMiss() { ++CandidateCount; }
    void Finalize(int candidatesPerFrame) {
        if (!ChosenSample.has_value() || CandidateCount <= 0) {
            W = 0.0f;
            return;
        }
        const float phat = float(ChosenSample->targetFunction);
        if (phat <= 0.0f) {
            W = 0.0f;
            return;
        }
        const int frameCount = std::max(1, CandidateCount / std::max(1, candidatesPerFrame));
        W = WSum / (float(frameCount) * phat);
        if (!std::isfinite(W)) W = 0.0f;
    }
    void Merge(const Reservoir &other, float mergeWeight, Rng &rng, int &sel) {
        CandidateCount += other.CandidateCount;
        if (!other.ChosenSample.has_value() || mergeWeight <= 0.0f) return;
        WSum += mergeWeight;
        if (rng.NextFloat() <= mergeWeight / WSum) {
            ChosenSample = other.ChosenSample;
            sel = -2;
        }
    }
};

struct MISContrib {
    Spec Emission, ThroughputMul;
    float PNee = 0.0f, PBsdf = 0.0f;
    bool UseMis = false;
    bool VisibilityTested = true;
};

// ---------------------------------------------------------------- the renderer
struct Settings {
    hdr_pipeline_desc p;
};

// Per-pixel record the temporal pass leaves for the spatial-reuse pass (an extension beyond the
// reference, see the spatial pass below): the shading vertex and the path weights around its direct
// light.
struct SpatialAux {
    bool valid = false;
    HitRecord hit;
    Closure cl;
    V3 rayDir, sn;
    bool inside = false;
    Lambda lam;
    Spec throughput, emission;
    // the radiance of the path's later vertices (1 .. maxDepth-1), summed in path order from zero;
    // the spatial pass adds it after the primary vertex's emission and reused direct light
    Spec rest;
};

struct PathCtx {  // per-pixel thread-local state passed through the integrators
    const Scene &scene;
    const hdr_pipeline_desc &pipe;
    Rng &rng;
    const Lambda &lam;
    PixelTrace &tr;
    Counters &cnt;
    const std::unordered_set<const Light *> *hoisted;
    Reservoir *reservoir;  // persistent buffer entry of this pixel
    // spatial reuse: the post-temporal reservoir goes here (and the winner's visibility test moves
    // to the spatial pass); null without spatial reuse
    Reservoir *reservoirOut = nullptr;
    SpatialAux *aux = nullptr;
};

std::optional<HitRecord> Trace(const PathCtx &c, const V3 &o, const V3 &d, uint64_t &classCounter) {
    ++classCounter;
    return c.scene.IntersectScene(o, d, c.cnt);
}

// Detail::SampleBounceWithConnection, direct_light_integrator_interface.h:18-40.
// kind: 0 = path bounce, 1 = BSDF candidate (for ray counters only).
bool SampleBounceWithConnection(const PathCtx &c, const MaterialParams *mat, const ShadingPoint &sp,
                                const HitRecord &hit, const V3 &rayDir, const BounceConfig &cfg, BounceState &st,
                                Connection &out, int kind) {
    bool ok;
    if (mat) ok = PreviewSampleBounce(sp, hit.Position, rayDir, cfg, st, c.rng, out.Bounce);
    else ok = DefaultSampleBounce(sp, hit.Position, c.rng, out.Bounce);
    if (!ok) return false;
    out.Hit = Trace(c, out.Bounce.NextRay.o, out.Bounce.NextRay.d, kind == 0 ? c.cnt.bounce : c.cnt.candidate);
    return true;
}

// ---- MIS (source/integrators/mis_direct_light_integrator.cpp)
struct MisNee {
    Spec Radiance;
    float PNee = 0.0f, PBsdf = 0.0f;
    bool IsDelta = false;
};

MisNee MisEvaluateNEE(const PathCtx &c, const ShadingPoint &sp, const HitRecord &hit, const V3 &rayDir,
                      const Light &light, const Light::Sample &ls) {  // :70-120
    const V3 &n = sp.n;
    const float nDotL = Dot(n, ls.Dir);
    if (nDotL <= 0.0f) return MisNee{};
    const float dist2 = ls.Dist * ls.Dist;
    const float cosY = std::max(0.0f, Dot(-ls.Dir, ls.LightNormal));
    if (cosY <= 0.0f && !light.IsDelta()) return MisNee{};
    const float pNee = ls.pdf.ConvertTo(PdfSpace::SolidAngle, dist2, std::max(cosY, 1e-6f)).value;
    if (pNee <= 0.0f) return MisNee{};
    const V3 shadowOrigin = hit.Position + n * 1e-4f;
    const auto shadowHit = Trace(c, shadowOrigin, ls.Dir, c.cnt.shadow);
    if (shadowHit && shadowHit->Depth < ls.Dist - 1e-3f) return MisNee{};
    const V3 wo = -rayDir;
    const float bsdfPdf = GGXPdfFull(*sp.c, n, wo, ls.Dir);
    const V3 bsdfValue = GGXEval(*sp.c, n, wo, ls.Dir);
    MisNee r;
    r.Radiance = RGBToSpectrum(bsdfValue, *sp.lambda) * RGBToSpectrum(ls.Color, *sp.lambda) * (nDotL / pNee);
    r.PNee = pNee;
    r.PBsdf = bsdfPdf;
    r.IsDelta = light.IsDelta();
    return r;
}

Spec MisLi(const PathCtx &c, const LightSampler &sampler, const ShadingPoint &sp, const HitRecord &hit,
           const V3 &rayDir, const std::optional<Connection> &conn) {  // :257-271
    const bool useBsdf = conn.has_value();
    Spec total(0.0f);
    // _integrateNEE, :122-179
    {
        const Light *env = c.scene.GetEnvironment();
        const float envSelectPdf = env ? sampler.EvalPdf(env) : 0.0f;
        const bool separateEnv = env != nullptr && envSelectPdf <= 0.0f;
        const auto cand = sampler.Propose(hit.Position, c.rng, c.tr);
        if (cand) {
            const MisNee nee = MisEvaluateNEE(c, sp, hit, rayDir, *cand->light, cand->ls);
            if (nee.PNee > 0.0f) {
                float w;
                if (nee.IsDelta || !useBsdf) w = 1.0f;
                else if (separateEnv) w = PowerHeuristic3(nee.PNee, 1, 0.0f, 1, nee.PBsdf, 1);
                else w = PowerHeuristic2(nee.PNee, nee.PBsdf);
                total += nee.Radiance * w;
            }
        }
        if (separateEnv) {
            const auto ls = env->SampleLight(hit.Position, c.rng);
            if (ls) {
                const MisNee nee = MisEvaluateNEE(c, sp, hit, rayDir, *env, *ls);
                if (nee.PNee > 0.0f) {
                    const float w = (nee.IsDelta || !useBsdf) ? 1.0f : PowerHeuristic3(nee.PNee, 1, 0.0f, 1, nee.PBsdf, 1);
                    total += nee.Radiance * w;
                }
            }
        }
    }
    if (conn) {  // _integrateBSDFConnection, :181-255
        const Lambda &lam = *sp.lambda;
        const Connection &cn = *conn;
        Spec part(0.0f);
        bool done = false;
        if (cn.Bounce.ImpossibleNEEConnection || cn.Bounce.BsdfPdf.value <= 0.0f) {
            if (!cn.Hit) {
                const Light *env = c.scene.GetEnvironment();
                part = (env ? RGBToSpectrum(env->DomeSample(cn.Bounce.NextRay.d), lam) : Spec(0.0f)) *
                       cn.Bounce.ThroughputMul;
            } else {
                const Closure cl = GetClosure(c.scene.GetMaterial(cn.Hit->MatId), *cn.Hit);
                part = RGBToSpectrum(cl.Emission, lam) * cn.Bounce.ThroughputMul;
            }
            done = true;
        }
        if (!done) {
            float pNeeSampler = 0.0f, pNeeEnv = 0.0f;
            Spec radiance(0.0f);
            bool early = false;
            if (!cn.Hit) {
                const Light *env = c.scene.GetEnvironment();
                if (!env) {
                    early = true;
                } else {
                    radiance = RGBToSpectrum(env->DomeSample(cn.Bounce.NextRay.d), lam);
                    const float envSelectPdf = sampler.EvalPdf(env);
                    const float envPdf = env->DomeEvalPdf(cn.Bounce.NextRay.d);
                    pNeeSampler = envSelectPdf * envPdf;
                    pNeeEnv = envSelectPdf <= 0.0f ? envPdf : 0.0f;
                }
            } else {
                const Closure cl = GetClosure(c.scene.GetMaterial(cn.Hit->MatId), *cn.Hit);
                radiance = RGBToSpectrum(cl.Emission, lam);
                const Light *hitLight = c.scene.GetLightAtHit(*cn.Hit);
                if (hitLight) {
                    const float sel = sampler.EvalPdf(hitLight);
                    if (sel > 0.0f) {
                        const V3 off = cn.Hit->Position - cn.Bounce.NextRay.o;
                        const float dist2 = Dot(off, off);
                        const float cosY = std::max(0.0f, Dot(-cn.Bounce.NextRay.d, cn.Hit->Normal));
                        if (cosY > 0.0f) {
                            const Pdf areaPdf = hitLight->EvalPdf(cn.Bounce.NextRay.d);
                            pNeeSampler = sel * areaPdf.ConvertTo(PdfSpace::SolidAngle, dist2, cosY).value;
                        }
                    }
                }
            }
            if (!early) {
                if (pNeeSampler + pNeeEnv <= 0.0f) {
                    part = radiance * cn.Bounce.ThroughputMul;
                } else {
                    const float w = PowerHeuristic3(cn.Bounce.BsdfPdf.value, 1, pNeeSampler, 1, pNeeEnv, 1);
                    part = radiance * w * cn.Bounce.ThroughputMul;
                }
            }
        }
        total += part;
    }
    return total;
}

// ---- RIS (source/restir/ris_direct_light_integrator.cpp)
struct Strategy {
    const LightSampler &sampler;
    int nNeeSampler, nNeeSky, nBsdf;
    bool skipVisibility;
};

MISContrib RisEvaluateLightSample(const PathCtx &c, const ShadingPoint &sp, const HitRecord &hit, const V3 &rayDir,
                                  const Light &light, const Light::Sample &ls, bool skipVisibility) {  // :482-527
    const V3 wo = -rayDir;
    const float nDotL = Dot(sp.n, ls.Dir);
    const float dist2 = ls.Dist * ls.Dist;
    const float cosY = std::max(0.0f, Dot(-ls.Dir, ls.LightNormal));
    const float pNee = ls.pdf.ConvertTo(PdfSpace::SolidAngle, dist2, std::max(cosY, 1e-6f)).value;
    const float pBsdf = GGXPdfFull(*sp.c, sp.n, wo, ls.Dir);
    const V3 bsdfValue = GGXEval(*sp.c, sp.n, wo, ls.Dir);
    const bool useMis = !light.IsDelta();
    MISContrib r;
    r.PNee = pNee;
    r.PBsdf = pBsdf;
    r.UseMis = useMis;
    if (nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !light.IsDelta())) return r;
    if (skipVisibility) {
        r.Emission = RGBToSpectrum(ls.Color, *sp.lambda);
        r.ThroughputMul = RGBToSpectrum(bsdfValue * nDotL, *sp.lambda);
        r.VisibilityTested = false;
        return r;
    }
    const V3 shadowOrigin = hit.Position + sp.n * 1e-4f;
    const auto shadowHit = Trace(c, shadowOrigin, ls.Dir, c.cnt.shadow);
    if (shadowHit && shadowHit->Depth < ls.Dist - 1e-3f) {
        r.ThroughputMul = RGBToSpectrum(bsdfValue * nDotL, *sp.lambda);
        return r;
    }
    r.Emission = RGBToSpectrum(ls.Color, *sp.lambda);
    r.ThroughputMul = RGBToSpectrum(bsdfValue * nDotL, *sp.lambda);
    return r;
}

MISContrib RisEvaluateBSDFConnection(const PathCtx &c, const ShadingPoint &sp, const Connection &cn,
                                     const Strategy &strat) {  // :547-634
    const Lambda &lam = *sp.lambda;
    const float pBsdf = cn.Bounce.BsdfPdf.value;
    const V3 &wi = cn.Bounce.NextRay.d;
    const Spec &integrandMul = cn.Bounce.ThroughputIntegrandMul;
    MISContrib r;
    r.ThroughputMul = integrandMul;
    r.PBsdf = pBsdf;
    if (cn.Bounce.ImpossibleNEEConnection || pBsdf <= 0.0f) {
        if (!cn.Hit) {
            const Light *env = c.scene.GetEnvironment();
            r.Emission = env ? RGBToSpectrum(env->DomeSample(wi), lam) : Spec();
        } else {
            r.Emission = RGBToSpectrum(GetClosure(c.scene.GetMaterial(cn.Hit->MatId), *cn.Hit).Emission, lam);
        }
        r.PNee = 0.0f;
        r.UseMis = false;
        return r;
    }
    if (!cn.Hit) {
        const Light *env = c.scene.GetEnvironment();
        if (!env) return r;
        const float sel = strat.sampler.EvalPdf(env);
        const float envPdf = env->DomeEvalPdf(wi);
        const float pNee = sel > 0.0f ? sel * envPdf : (strat.nNeeSky > 0 ? envPdf : 0.0f);
        r.Emission = RGBToSpectrum(env->DomeSample(wi), lam);
        r.PNee = pNee;
        r.UseMis = pNee > 0.0f;
        return r;
    }
    float pNee = 0.f;
    const Closure cl = GetClosure(c.scene.GetMaterial(cn.Hit->MatId), *cn.Hit);
    const Spec emission = RGBToSpectrum(cl.Emission, lam);
    const Light *hitLight = c.scene.GetLightAtHit(*cn.Hit);
    if (hitLight) {
        const float sel = strat.sampler.EvalPdf(hitLight);
        const V3 off = cn.Hit->Position - cn.Bounce.NextRay.o;
        const float dist2 = Dot(off, off);
        const float cosY = std::max(0.0f, Dot(-wi, cn.Hit->Normal));
        if (cosY > 0.0f && sel > 0.0f) {
            const Pdf areaPdf = hitLight->EvalPdf(wi);
            pNee = sel * areaPdf.ConvertTo(PdfSpace::SolidAngle, dist2, cosY).value;
        }
    }
    r.Emission = emission;
    r.PNee = pNee;
    r.UseMis = pNee > 0.0f;
    return r;
}

Spec RisReEvaluate(const ShadingPoint &sp, const V3 &rayDir, const LightSampler::Candidate &cand) {  // :529-545
    const V3 wo = -rayDir;
    const Light::Sample &ls = cand.ls;
    const float nDotL = Dot(sp.n, ls.Dir);
    const float dist2 = ls.Dist * ls.Dist;
    const float cosY = std::max(0.0f, Dot(-ls.Dir, ls.LightNormal));
    const float pNee = ls.pdf.ConvertTo(PdfSpace::SolidAngle, dist2, std::max(cosY, 1e-6f)).value;
    if (nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !cand.light->IsDelta())) return Spec();
    const V3 bsdfValue = GGXEval(*sp.c, sp.n, wo, ls.Dir);
    return RGBToSpectrum(ls.Color, *sp.lambda) * RGBToSpectrum(bsdfValue * nDotL, *sp.lambda);
}

using OptCand = std::optional<RISCandidate>;

Spec RisLi(const PathCtx &c, const LightSampler &sampler, const ShadingPoint &sp, const MaterialParams *mat,
           const HitRecord &hit, const V3 &rayDir, const std::optional<Connection> &conn, bool primary) {  // :84-215
    const Scene &scene = c.scene;
    const int M = c.pipe.candidate_count;
    const Light *envLight = scene.GetEnvironment();
    const float skySel = envLight ? sampler.EvalPdf(envLight) : 0.0f;
    const int nTech = (envLight && skySel <= 0.0f) ? 3 : 2;
    const int nNeeSampler = std::max(1, M / nTech);
    const int nNeeSky = (envLight && skySel <= 0.0f) ? M / nTech : 0;
    const int nBsdf = std::max(0, M - nNeeSampler - nNeeSky);
    const Strategy strat{sampler, nNeeSampler, nNeeSky, nBsdf, c.pipe.skip_visibility != 0};
    const Lambda &lam = *sp.lambda;

    std::vector<OptCand> all;
    // _generateNEECandidates, :407-480
    for (int i = 0; i < strat.nNeeSampler; ++i) {
        const auto cand = sampler.Propose(hit.Position, c.rng, c.tr);
        if (!cand) {
            all.push_back(std::nullopt);
            continue;
        }
        const MISContrib nee = RisEvaluateLightSample(c, sp, hit, rayDir, *cand->light, cand->ls, strat.skipVisibility);
        const Spec contribution = nee.Emission * nee.ThroughputMul;
        const float target = SpectrumLuminance(contribution, lam);
        const float mis = nee.UseMis ? PowerHeuristic3(nee.PNee, strat.nNeeSampler, 0.0f, strat.nNeeSky, nee.PBsdf,
                                                       strat.nBsdf)
                                     : 1.0f;
        const double risWeight = double(mis * target) / (double(nee.PNee) + 1e-9);
        RISCandidate rc;
        rc.Candidate = *cand;
        rc.Throughput = contribution;
        rc.risWeight = risWeight;
        rc.targetFunction = target;
        rc.VisibilityTested = nee.VisibilityTested;
        all.push_back(rc);
    }
    if (strat.nNeeSky > 0 && envLight) {
        for (int i = 0; i < strat.nNeeSky; ++i) {
            const auto ls = envLight->SampleLight(hit.Position, c.rng);
            if (!ls) {
                all.push_back(std::nullopt);
                continue;
            }
            const MISContrib nee = RisEvaluateLightSample(c, sp, hit, rayDir, *envLight, *ls, strat.skipVisibility);
            const Spec contribution = nee.Emission * nee.ThroughputMul;
            const float target = SpectrumLuminance(contribution, lam);
            const float mis = nee.UseMis ? PowerHeuristic3(nee.PNee, strat.nNeeSky, 0.0f, strat.nNeeSampler, nee.PBsdf,
                                                           strat.nBsdf)
                                         : 1.0f;
            RISCandidate rc;
            rc.Candidate = LightSampler::Candidate{envLight, *ls};
            rc.Throughput = contribution;
            rc.risWeight = double(mis * target) / (double(nee.PNee) + 1e-9);
            rc.targetFunction = target;
            rc.VisibilityTested = nee.VisibilityTested;
            all.push_back(rc);
        }
    }
    // _generateBSDFSamplingCandidates, :250-405
    int bsdfCount = 0;
    if (conn && conn->Bounce.BsdfPdf.value > 0.0f) {
        const MISContrib bc = RisEvaluateBSDFConnection(c, sp, *conn, strat);
        const Spec contribution = bc.Emission * bc.ThroughputMul;
        const double target = SpectrumLuminance(contribution, lam);
        float pNeeSampler = bc.PNee, pNeeSky = 0.0f;
        if (!conn->Hit) {
            const Light *env = scene.GetEnvironment();
            if (env == nullptr || strat.sampler.EvalPdf(env) <= 0.0f) {
                pNeeSampler = 0.0f;
                pNeeSky = bc.PNee;
            }
        }
        const double mis =
            PowerHeuristic3(bc.PBsdf, strat.nBsdf, pNeeSampler, strat.nNeeSampler, pNeeSky, strat.nNeeSky);
        RISCandidate rc;
        rc.Throughput = contribution;
        rc.risWeight = mis * target / (double(bc.PBsdf) + 1e-9);
        rc.targetFunction = target;
        all.push_back(rc);
        bsdfCount = 1;
    }
    const int toGenerate = strat.nBsdf - bsdfCount;
    for (int i = 0; i < toGenerate; ++i) {
        BounceState st{};
        const BounceConfig cfg{1, 1};
        Connection sc;
        if (!SampleBounceWithConnection(c, mat, sp, hit, rayDir, cfg, st, sc, 1)) {
            all.push_back(std::nullopt);
            continue;
        }
        if (!sc.Hit) {
            const Light *env = scene.GetEnvironment();
            if (!env) {
                all.push_back(std::nullopt);
                continue;
            }
            const Spec &envTp = sc.Bounce.ThroughputIntegrandMul;
            const V3 wi = sc.Bounce.NextRay.d;
            const V3 envRGB = env->DomeSample(wi);
            const Spec envEm = RGBToSpectrum(envRGB, lam);
            const float pBsdf = sc.Bounce.BsdfPdf.value;
            const float sel = strat.sampler.EvalPdf(env);
            const float envPdf = env->DomeEvalPdf(wi);
            const float pNeeSampler = sel * envPdf;
            const float pNeeSky = (sel <= 0.0f && strat.nNeeSky > 0) ? envPdf : 0.0f;
            const float pNee = pNeeSampler + pNeeSky;
            const double target = SpectrumLuminance(envEm * envTp, lam);
            const double mis = PowerHeuristic3(pBsdf, strat.nBsdf, pNeeSampler, strat.nNeeSampler, pNeeSky, strat.nNeeSky);
            RISCandidate rc;
            if (pNee > 0.0f) {
                Light::Sample ls;
                ls.Dir = wi;
                ls.Color = envRGB;
                ls.LightNormal = -wi;
                ls.Dist = 1e30f;
                ls.pdf = {pNee, PdfSpace::SolidAngle};
                rc.Candidate = LightSampler::Candidate{env, ls};
            }
            rc.Throughput = envEm * envTp;
            rc.risWeight = mis * target / (double(pBsdf) + 1e-9);
            rc.targetFunction = target;
            all.push_back(rc);
            continue;
        }
        const Light *light = scene.GetLightAtHit(*sc.Hit);
        if (!light) {
            all.push_back(std::nullopt);
            continue;
        }
        const MISContrib bc = RisEvaluateBSDFConnection(c, sp, sc, strat);
        const Spec contribution = bc.Emission * bc.ThroughputMul;
        const double target = SpectrumLuminance(contribution, lam);
        const double mis = PowerHeuristic3(bc.PBsdf, strat.nBsdf, bc.PNee, strat.nNeeSampler, 0.0f, strat.nNeeSky);
        const V3 wi = sc.Bounce.NextRay.d;
        const V3 emissionRGB = GetClosure(scene.GetMaterial(sc.Hit->MatId), *sc.Hit).Emission;
        Pdf lightPdf = light->EvalPdf(wi);
        lightPdf.value *= strat.sampler.EvalPdf(light);
        Light::Sample ls;
        ls.Dir = wi;
        ls.Color = emissionRGB;
        ls.LightNormal = sc.Hit->Normal;
        ls.Dist = sc.Hit->Depth;
        ls.pdf = lightPdf;
        RISCandidate rc;
        rc.Candidate = LightSampler::Candidate{light, ls};
        rc.Throughput = contribution;
        rc.risWeight = mis * target / (double(bc.PBsdf) + 1e-9);
        rc.targetFunction = target;
        all.push_back(rc);
    }

    if (c.pipe.use_reservoir && primary) {  // :121-185
        Reservoir res;
        const int perFrame = int(all.size());
        int sel = -1;
        for (int k = 0; k < int(all.size()); ++k) {
            if (all[k]) res.Update(*all[k], float(all[k]->risWeight), c.rng, k, sel);
            else res.CountMiss();
        }
        Reservoir prev = *c.reservoir;
        if (prev.ChosenSample && prev.ChosenSample->Candidate) {
            const LightSampler::Candidate &pc = *prev.ChosenSample->Candidate;
            const Spec reeval = RisReEvaluate(sp, rayDir, pc);
            const float phatCurr = SpectrumLuminance(reeval, lam);
            const float phatGen = float(prev.ChosenSample->targetFunction);
            if (phatGen > 0.0f) {
                prev.ChosenSample->targetFunction = double(phatCurr);
                prev.ChosenSample->Throughput = reeval;
                const float mergeWeight = prev.WSum * (phatCurr / phatGen);
                res.Merge(prev, mergeWeight, c.rng, sel);
            }
        }
        res.Finalize(perFrame);
        c.tr.selectHash = Fnv(c.tr.selectHash, uint32_t(sel));
        if (c.reservoirOut) {  // spatial reuse follows: it tests the final winner
            *c.reservoirOut = res;
            return Spec();
        }
        if (res.ChosenSample && !res.ChosenSample->VisibilityTested && res.ChosenSample->Candidate) {
            RISCandidate &winner = *res.ChosenSample;
            const LightSampler::Candidate &lc = *winner.Candidate;
            const MISContrib nee = RisEvaluateLightSample(c, sp, hit, rayDir, *lc.light, lc.ls, false);
            winner.Throughput = nee.Emission * nee.ThroughputMul;
            winner.VisibilityTested = true;
        }
        *c.reservoir = res;
        if (!res.ChosenSample || res.W <= 0.0f) return Spec();
        return res.ChosenSample->Throughput * res.W;
    }

    // RIS::RandomIndexFromWeights, random_index.h:27-63
    double weightSum = 0.0;
    for (const OptCand &k : all) {
        const double w = k ? k->risWeight : 0.0;
        if (w > 0.0) weightSum += w;
    }
    int chosen = -1;
    if (weightSum > 0.0) {
        double remaining = double(c.rng.NextFloat()) * weightSum;
        for (int i = 0; i < int(all.size()); ++i) {
            const double w = all[i] ? all[i]->risWeight : 0.0;
            if (w <= 0.0) continue;
            remaining -= w;
            if (remaining <= 0.0) {
                chosen = i;
                break;
            }
        }
    }
    c.tr.selectHash = Fnv(c.tr.selectHash, uint32_t(chosen));
    if (chosen < 0) return Spec();
    const RISCandidate &cc = *all[chosen];
    const float W = float(weightSum / cc.targetFunction);
    Spec winnerTp = cc.Throughput;
    if (!cc.VisibilityTested && cc.Candidate) {
        const MISContrib nee = RisEvaluateLightSample(c, sp, hit, rayDir, *cc.Candidate->light, cc.Candidate->ls, false);
        winnerTp = nee.Emission * nee.ThroughputMul;
    }
    return winnerTp * W;
}

// ---- PathIntegrator (source/integrators/path_integrator.cpp:48-209)
Spec DirectLi(const PathCtx &c, const LightSampler &sampler, const ShadingPoint &sp, const MaterialParams *mat,
              const HitRecord &hit, const V3 &rayDir, const std::optional<Connection> &conn, bool primary) {
    if (c.pipe.direct_light == HDR_DIRECT_RIS) return RisLi(c, sampler, sp, mat, hit, rayDir, conn, primary);
    return MisLi(c, sampler, sp, hit, rayDir, conn);
}

Spec PathLiInner(const PathCtx &c, const ShadingPoint &sp0, const HitRecord &hit0, const V3 &ray0Dir) {  // _li :92-209
    const Scene &scene = c.scene;
    const BounceConfig cfg{c.pipe.max_reflection_bounces, c.pipe.max_refraction_bounces};
    const LightSampler sampler(scene.lightPtrs, c.hoisted);  // _factory->Create (:99)
    // With spatial reuse (an extension) the primary vertex's emission and direct light are added by
    // the spatial pass, so the sum here starts from zero and holds the later vertices only.
    Spec total = c.aux ? Spec(0.0f) : RGBToSpectrum(sp0.c->Emission, *sp0.lambda);
    Spec throughput(1.0f);
    BounceState st{};
    auto done = [&]() -> Spec {
        if (c.aux) c.aux->rest = total;
        return total;
    };
    const MaterialParams *firstMat = scene.GetMaterial(hit0.MatId);
    Connection first;
    const bool firstOk = SampleBounceWithConnection(c, firstMat, sp0, hit0, ray0Dir, cfg, st, first, 0);
    std::optional<Connection> firstConn;
    if (firstOk) firstConn = first;
    total += DirectLi(c, sampler, sp0, firstMat, hit0, ray0Dir, firstConn, true);
    if (!firstOk) return done();
    throughput *= first.Bounce.ThroughputMul;
    Ray cur = first.Bounce.NextRay;
    std::optional<HitRecord> next = first.Hit;
    for (int bounce = 1; bounce < c.pipe.max_depth; ++bounce) {
        if (!next) break;
        const HitRecord bh = *next;
        const MaterialParams *mat = scene.GetMaterial(bh.MatId);
        Closure cl = GetClosure(mat, bh);
        if (!c.pipe.enable_subsurface) cl.Subsurface = 0.0f;
        if (cl.Opacity < 0.999f && c.rng.NextFloat() > cl.Opacity) {
            cur.o = bh.Position + cur.d * 1e-4f;
            next = Trace(c, cur.o, cur.d, c.cnt.bounce);
            --bounce;
            continue;
        }
        const bool isInside = Dot(cl.Normal, cur.d) > 0.0f;
        BeerAbsorption(throughput, cl, bh.Depth, isInside, *sp0.lambda);
        V3 sn = cl.Normal;
        if (isInside) sn = -sn;
        const ShadingPoint sp{&cl, sn, sp0.lambda, isInside};
        Connection bc;
        const bool ok = SampleBounceWithConnection(c, mat, sp, bh, cur.d, cfg, st, bc, 0);
        std::optional<Connection> conn;
        if (ok) conn = bc;
        total += throughput * DirectLi(c, sampler, sp, mat, bh, cur.d, conn, false);
        if (!ok) break;
        throughput *= bc.Bounce.ThroughputMul;
        cur = bc.Bounce.NextRay;
        next = bc.Hit;
        if (bc.Bounce.SkipRoulette) continue;
        if (bounce > 3) {
            const float p = std::clamp(throughput.Max() * c.pipe.roulette_aggressiveness, 0.0f, 1.0f);
            if (c.rng.NextFloat() > p) break;
            throughput *= 1.0f / p;
        }
    }
    return done();
}

Spec PathLi(const PathCtx &c, const Ray &ray, std::optional<HitRecord> hit) {  // Li :48-90
    const Scene &scene = c.scene;
    const Light *env = scene.GetEnvironment();
    const V3 rayDir = ray.d;
    Spec throughput(1.0f);
    while (hit) {
        const MaterialParams *mat = scene.GetMaterial(hit->MatId);
        Closure cl = GetClosure(mat, *hit);
        if (!c.pipe.enable_subsurface) cl.Subsurface = 0.0f;
        if (cl.Opacity >= 0.999f || c.rng.NextFloat() <= cl.Opacity) {
            const bool isInside = Dot(cl.Normal, rayDir) > 0.0f;
            BeerAbsorption(throughput, cl, hit->Depth, isInside, c.lam);
            V3 sn = cl.Normal;
            if (isInside) sn = -sn;
            const ShadingPoint sp{&cl, sn, &c.lam, isInside};
            if (c.aux) {
                SpatialAux &x = *c.aux;
                x.valid = true;
                x.hit = *hit;
                x.cl = cl;
                x.rayDir = rayDir;
                x.sn = sn;
                x.inside = isInside;
                x.lam = c.lam;
                x.throughput = throughput;
                x.emission = RGBToSpectrum(cl.Emission, c.lam);
            }
            return throughput * PathLiInner(c, sp, *hit, rayDir);
        }
        hit = Trace(c, hit->Position + rayDir * 1e-4f, rayDir, c.cnt.bounce);
    }
    if (env && c.pipe.render_ibl_background) return RGBToSpectrum(env->DomeSample(rayDir), c.lam);
    return Spec(0.0f);
}

// ---- camera (source/scene/camera_ray.cpp:10-76, camera_params.h:23-30)
V3 ToF(const V3d &d) { return V3(float(d.v[0]), float(d.v[1]), float(d.v[2])); }
V3d ToD(const V3 &f) { return {{double(f[0]), double(f[1]), double(f[2])}}; }

Ray CameraRay(const M4d &invView, const M4d &invProj, float px, float py, int width, int height) {
    const float ndcX = (2.0f * px / float(width)) - 1.0f;
    const float ndcY = (2.0f * py / float(height)) - 1.0f;
    const V3 origin = ToF(invView.Transform({{0.0, 0.0, 0.0}}));
    const V3 nearCam = ToF(invProj.Transform({{double(ndcX), double(ndcY), -1.0}}));
    const V3 nearWorld = ToF(invView.Transform(ToD(nearCam)));
    return {origin, Normalized(nearWorld - origin)};
}

Ray CameraRayParams(const M4d &invView, const M4d &invProj, float px, float py, int width, int height,
                    const hdr_camera_params &p, Rng &rng) {
    float ndcX = (2.0f * px / float(width)) - 1.0f;
    float ndcY = (2.0f * py / float(height)) - 1.0f;
    if (p.lens_distortion != 0.0f) {
        const float r2 = ndcX * ndcX + ndcY * ndcY;
        const float f = 1.0f + p.lens_distortion * r2;
        ndcX *= f;
        ndcY *= f;
    }
    const V3 origin = ToF(invView.Transform({{0.0, 0.0, 0.0}}));
    const V3 nearCam = ToF(invProj.Transform({{double(ndcX), double(ndcY), -1.0}}));
    if (!p.enable_dof) {
        const V3 nearWorld = ToF(invView.Transform(ToD(nearCam)));
        return {origin, Normalized(nearWorld - origin)};
    }
    const float aperture = (p.focal_length / 10.0f) / (2.0f * p.f_stop);
    float lensU, lensV;
    if (p.bokeh_blades < 3) {
        const float r = std::sqrt(rng.NextFloat());
        const float theta = float(2.0 * kPiD * double(rng.NextFloat()));
        lensU = r * std::cos(theta);
        lensV = r * std::sin(theta);
    } else {
        const float theta = float(2.0 * kPiD * double(rng.NextFloat()));
        const float r = std::sqrt(rng.NextFloat());
        const float sectorAngle = float(2.0 * kPiD / double(float(p.bokeh_blades)));
        const float sector = std::floor(theta / sectorAngle);
        const float angleIn = theta - sector * sectorAngle;
        const float d = std::cos(sectorAngle / 2.0f) / std::cos(sectorAngle / 2.0f - angleIn);
        lensU = r * d * std::cos(theta);
        lensV = r * d * std::sin(theta);
    }
    lensU *= aperture;
    lensV *= aperture;
    const V3 lensCam(lensU, lensV, 0.0f);
    const V3 focalCam = nearCam * p.focus_distance;
    const V3 lensWorld = ToF(invView.Transform(ToD(lensCam)));
    const V3 focalWorld = ToF(invView.Transform(ToD(focalCam)));
    return {lensWorld, Normalized(focalWorld - lensWorld)};
}

float ExposureMultiplier(const hdr_camera_params &p) {  // camera_params.h:23-30
    if (!p.enable_exposure) return 1.0f;
    const float safeF = std::max(p.f_stop, 0.1f);
    return (p.iso / 100.0f) * p.shutter_speed / (safeF * safeF) * 100.0f;
}

inline uint32_t PixelSeed(size_t i, int frameIndex) {  // integration_pass.cpp:210-211
    return uint32_t(i) ^ uint32_t((unsigned long long)(long long)frameIndex * 12345ULL);
}

}  // namespace

// ---------------------------------------------------------------- C API
// ---- Spatial reuse: an extension beyond the reference (which reuses temporally only).
// After every pixel's temporal pass, pixel q resamples among its own post-temporal reservoir and
// those of up to spatial_count neighbours picked within spatial_radius pixels that pass a
// similarity test (depth within 10 %, geometric normals within 25 degrees). Each reservoir j
// (confidence c_j = its candidate count, capped at 20 frames) offers its sample y_j, shifted to q
// (RetargetSample, Jacobian ShiftJacobian), with the generalized balance heuristic
//   w_j = c_j W_j p_j(y_j) p_q(y_j) / sum_k c_k p_k(y_j -> k) |J_{q->k}|,
// whose weights sum to one over the domains, so the combination stays unbiased for any BSDF
// (the plain count-weighted merge of the temporal step is heavy-tailed on glossy lobes). The
// winner gets W = sum w / p_q(y) and a shadow ray at q; it is stored for the next frame in the
// reservoir form of reservoir.h (WSum = W * frames * p_q, count = sum of confidences, capped).
// Neighbour picks and selections draw from a splitmix64 stream seeded by (pixel, frame),
// integer-only, so the reference's per-pixel mt19937_64 stream is untouched.
// Any depth: the reuse replaces the primary vertex's direct light only. The path's later vertices
// run as in the reference (same draws) and sum, from zero, into SpatialAux::rest; the pixel's
// radiance is throughput * ((emission + reused direct) + rest).
constexpr int kMaxSpatial = 8;
struct SpatialRng {
    uint64_t s;
    SpatialRng(uint32_t pixel, int32_t frame)
        : s(uint64_t(pixel) * 0x9E3779B97F4A7C15ULL ^ uint64_t(uint32_t(frame)) * 0xC2B2AE3D27D4EB4FULL) {}
    uint64_t Next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    float NextFloat() { return float(Next() >> 40) * 0x1p-24f; }
};

bool PickNeighbour(SpatialRng &r, int x, int y, int W, int H, int R, int &nx, int &ny) {
    const uint64_t D = uint64_t(2 * R + 1);
    for (int t = 0; t < 8; ++t) {
        const uint64_t z = r.Next();
        const int dx = int((z & 0xffffULL) % D) - R;
        const int dy = int(((z >> 16) & 0xffffULL) % D) - R;
        if ((dx == 0 && dy == 0) || dx * dx + dy * dy > R * R) continue;
        nx = x + dx;
        ny = y + dy;
        return nx >= 0 && ny >= 0 && nx < W && ny < H;
    }
    return false;
}

bool SpatiallySimilar(const HitRecord &q, const HitRecord &n) {
    if (std::fabs(n.Depth - q.Depth) > 0.1f * q.Depth) return false;
    return Dot(q.Normal, n.Normal) >= 0.90630779f;  // cos 25 degrees
}

// The neighbour's light sample seen from q: distant and dome samples are directions; a point
// light is sampled again at q (it draws nothing); an area sample keeps its point on the light.
std::optional<Light::Sample> RetargetSample(const Light &L, const Light::Sample &ls, const V3 &posN, const V3 &posQ,
                                            float selectPdf) {
    if (L.type == HDR_LIGHT_DISTANT || L.type == HDR_LIGHT_DOME) return ls;
    if (L.type == HDR_LIGHT_POINT) {
        Rng none(0);
        auto s = L.SampleLight(posQ, none);
        if (!s) return std::nullopt;
        s->pdf.value *= selectPdf;
        return s;
    }
    const V3 P = posN + ls.Dir * ls.Dist;
    const V3 toLight = P - posQ;
    const float dist = Length(toLight);
    Light::Sample o = ls;
    o.Dir = divd(toLight, double(dist));
    o.Dist = dist;
    return o;
}

// |d omega_q / d omega_n| of moving a light sample from n to q: 1 for directions and point lights;
// for a point on an area light, (cos_q / d_q^2) / (cos_n / d_n^2).
float ShiftJacobian(const Light::Sample &atN, const Light::Sample &atQ) {
    if (atN.pdf.space == PdfSpace::SolidAngle) return 1.0f;
    const float cosN = std::max(0.0f, Dot(-atN.Dir, atN.LightNormal));
    const float cosQ = std::max(0.0f, Dot(-atQ.Dir, atQ.LightNormal));
    if (cosN <= 0.0f) return 0.0f;
    return (cosQ / (atQ.Dist * atQ.Dist)) / (cosN / (atN.Dist * atN.Dist));
}

struct oracle_ctx {
    Scene scene;
    int W, H;
    bool hoist;
    int threads;
    bool fireflyInPlace = true;
    std::vector<Reservoir> reservoirs;      // __ris_reservoir
    std::vector<Reservoir> resT;            // spatial reuse: post-temporal reservoirs of this frame
    std::vector<SpatialAux> aux;
    std::vector<float> accum;               // __accum_color (RGBA)
    std::vector<double> lumSum, lumSumSq;   // __accum_lum_sum / __accum_lum_sum_sq (METRICS_ENABLED)
    int sampleCount = 0;                    // frameIndex + 1 of the last frame
    bool accumValid = false;
    std::vector<std::optional<HitRecord>> gbuf;
    std::vector<float> depth, albedo, normal;
    Counters last;
    // Reprojected temporal reuse (an extension, hdr_pipeline_desc::temporal_reprojection): the
    // previous frame's camera and whether it left its records (a spatial frame with the flag);
    // prevAux / prevRes are those records, copied at the start of a frame that reprojects, and
    // reproj[i] the previous frame's pixel that saw pixel i's hit (-1: none).
    bool histCam = false, histRecords = false;
    double histView[16] = {}, histProj[16] = {};
    V3 histOrigin;
    std::vector<SpatialAux> prevAux;
    std::vector<Reservoir> prevRes;
    std::vector<int> reproj;
};

// The pixel of the previous frame that saw hit h, or -1 (the reprojection of the spatial pass's
// extra domain): h's position through the previous view and projection matrices (row vectors, in
// double; clip w > 0), NDC [-1, 1] to [0, W) x [0, H) as CameraRay's inverse, and the pixel
// containing it if it was shaded then and its hit is similar -- depth within 10 % of h's distance
// from the previous camera, geometric normals within 25 degrees (SpatiallySimilar's thresholds).
int ReprojectPixel(const oracle_ctx &c, const HitRecord &h) {
    const double X = double(h.Position[0]), Y = double(h.Position[1]), Z = double(h.Position[2]);
    const double *V = c.histView, *Q = c.histProj;
    const double cx = X * V[0] + Y * V[4] + Z * V[8] + V[12];
    const double cy = X * V[1] + Y * V[5] + Z * V[9] + V[13];
    const double cz = X * V[2] + Y * V[6] + Z * V[10] + V[14];
    const double cw = X * V[3] + Y * V[7] + Z * V[11] + V[15];
    const double hx = cx * Q[0] + cy * Q[4] + cz * Q[8] + cw * Q[12];
    const double hy = cx * Q[1] + cy * Q[5] + cz * Q[9] + cw * Q[13];
    const double hw = cx * Q[3] + cy * Q[7] + cz * Q[11] + cw * Q[15];
    if (!(hw > 0.0)) return -1;
    const double px = (hx / hw + 1.0) * (0.5 * double(c.W));
    const double py = (hy / hw + 1.0) * (0.5 * double(c.H));
    if (!(px >= 0.0 && py >= 0.0 && px < double(c.W) && py < double(c.H))) return -1;
    const int j = int(py) * c.W + int(px);
    const SpatialAux &pj = c.prevAux[size_t(j)];
    if (!pj.valid) return -1;
    const float dist = Length(h.Position - c.histOrigin);
    if (std::fabs(pj.hit.Depth - dist) > 0.1f * dist) return -1;
    return Dot(h.Normal, pj.hit.Normal) >= 0.90630779f ? j : -1;
}

extern "C" {

oracle_ctx *oracle_create(const hdr_scene_desc *scene, int32_t width, int32_t height, int32_t hoist, int32_t threads) {
    auto *c = new oracle_ctx();
    c->scene.Load(*scene);
    c->W = width;
    c->H = height;
    c->hoist = hoist != 0;
    c->threads = threads > 0 ? threads : int(std::max(1u, std::thread::hardware_concurrency()));
    const size_t n = size_t(width) * height;
    c->reservoirs.assign(n, Reservoir{});
    c->resT.assign(n, Reservoir{});
    c->aux.assign(n, SpatialAux{});
    c->accum.assign(n * 4, 0.0f);
    c->lumSum.assign(n, 0.0);
    c->lumSumSq.assign(n, 0.0);
    c->gbuf.assign(n, std::nullopt);
    c->depth.assign(n, 1.0f);
    c->albedo.assign(n * 3, 0.0f);
    c->normal.assign(n * 3, 0.0f);
    return c;
}

void oracle_destroy(oracle_ctx *c) { delete c; }

void oracle_reset(oracle_ctx *c) {
    std::fill(c->reservoirs.begin(), c->reservoirs.end(), Reservoir{});
    std::fill(c->accum.begin(), c->accum.end(), 0.0f);
    std::fill(c->lumSum.begin(), c->lumSum.end(), 0.0);
    std::fill(c->lumSumSq.begin(), c->lumSumSq.end(), 0.0);
    c->sampleCount = 0;
    c->histCam = c->histRecords = false;
}

// Per pixel, the previous frame's pixel whose reservoir the last frame reused as its reprojected
// domain (-1: none, or the last frame did not reproject).
void oracle_read_reprojection(oracle_ctx *c, int32_t *out) {
    const size_t n = size_t(c->W) * c->H;
    for (size_t i = 0; i < n; ++i) out[i] = i < c->reproj.size() ? c->reproj[i] : -1;
}

// hdr_view_reset_accumulation: the accumulation restarts, the reservoirs and the previous frame's
// records stay (a camera move under reprojected temporal reuse).
void oracle_reset_accumulation(oracle_ctx *c) {
    std::fill(c->accum.begin(), c->accum.end(), 0.0f);
    std::fill(c->lumSum.begin(), c->lumSum.end(), 0.0);
    std::fill(c->lumSumSq.begin(), c->lumSumSq.end(), 0.0);
    c->sampleCount = 0;
}

void oracle_set_firefly_in_place(oracle_ctx *c, int32_t v) { c->fireflyInPlace = v != 0; }

static void ParallelTiles(int threads, int numTiles, const std::function<void(int, Counters &)> &fn,
                          Counters &total) {
    std::atomic<int> next{0};
    std::vector<Counters> per(threads);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (;;) {
                const int k = next.fetch_add(1);
                if (k >= numTiles) break;
                fn(k, per[t]);
            }
        });
    for (auto &th : pool) th.join();
    for (const Counters &p : per) {
        total.primary += p.primary;
        total.bounce += p.bounce;
        total.shadow += p.shadow;
        total.candidate += p.candidate;
        total.nodes += p.nodes;
        total.tris += p.tris;
        total.instances += p.instances;
    }
}

double oracle_render_frame(oracle_ctx *c, const hdr_frame_desc *frame, const hdr_pipeline_desc *pipe, int32_t x0,
                           int32_t y0, int32_t x1, int32_t y1, float *color_out, hdr_pixel_debug *debug_out) {
    const int W = c->W, H = c->H;
    x0 = std::max(0, x0);
    y0 = std::max(0, y0);
    x1 = std::min(W, x1);
    y1 = std::min(H, y1);
    const M4d view = M4d::From(frame->view), proj = M4d::From(frame->proj);
    const M4d invView = InverseD(view), invProj = InverseD(proj);
    const bool hasCam = frame->has_camera_params != 0;
    const int frameIndex = frame->frame_index;
    const bool spatial = pipe->spatial_count > 0;
    if (spatial && (pipe->direct_light != HDR_DIRECT_RIS || !pipe->use_reservoir))
        return -1.0;  // spatial reuse is defined for ReSTIR direct lighting (at the primary vertex)
    if (pipe->temporal_reprojection && (!spatial || pipe->spatial_count > kMaxSpatial - 1))
        return -1.0;  // reprojected temporal reuse is a domain of the spatial pass (at most 7 neighbours)
    // reprojected temporal reuse: on a camera move the same-pixel merge is skipped and, with the
    // previous frame's records, the spatial pass adds the reprojected domain
    bool temporalSkip = false, reproject = false;
    if (pipe->temporal_reprojection) {
        const bool moved = c->histCam && (std::memcmp(frame->view, c->histView, sizeof c->histView) != 0 ||
                                          std::memcmp(frame->proj, c->histProj, sizeof c->histProj) != 0);
        temporalSkip = moved;
        reproject = moved && c->histRecords && !(hasCam && frame->camera.lens_distortion != 0.0f);
    }
    c->reproj.assign(reproject ? size_t(W) * H : 0, -1);
    if (reproject) {
        c->prevAux = c->aux;
        c->prevRes = c->reservoirs;
    }
    if (spatial)
        for (SpatialAux &a : c->aux) a.valid = false;
    constexpr int kTile = 16;
    const int ntx = (W + kTile - 1) / kTile, nty = (H + kTile - 1) / kTile;
    Counters cnt;
    std::unordered_set<const Light *> hoistedSet;
    if (c->hoist)
        for (const Light *l : c->scene.lightPtrs) hoistedSet.insert(l);

    // RaycastPass, raycast_pass.cpp:57-162
    ParallelTiles(c->threads, ntx * nty,
                  [&](int t, Counters &lc) {
                      const int sx = (t % ntx) * kTile, sy = (t / ntx) * kTile;
                      for (int y = std::max(sy, y0); y < std::min(sy + kTile, y1); ++y)
                          for (int x = std::max(sx, x0); x < std::min(sx + kTile, x1); ++x) {
                              const size_t i = size_t(y) * W + x;
                              const float px = float(x) + 0.5f, py = float(y) + 0.5f;
                              Ray ray;
                              if (hasCam) {
                                  Rng rng(PixelSeed(i, frameIndex));
                                  ray = CameraRayParams(invView, invProj, px, py, W, H, frame->camera, rng);
                              } else {
                                  ray = CameraRay(invView, invProj, px, py, W, H);
                              }
                              ++lc.primary;
                              c->gbuf[i] = c->scene.IntersectScene(ray.o, ray.d, lc);
                              if (c->gbuf[i]) {
                                  c->depth[i] = c->gbuf[i]->Depth;
                                  for (int k = 0; k < 3; ++k) {
                                      c->albedo[3 * i + k] = c->gbuf[i]->Albedo[k];
                                      c->normal[3 * i + k] = c->gbuf[i]->Normal[k];
                                  }
                              } else {
                                  c->depth[i] = 1.0f;
                                  for (int k = 0; k < 3; ++k) c->albedo[3 * i + k] = c->normal[3 * i + k] = 0.0f;
                              }
                          }
                  },
                  cnt);

    if (reproject)
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) {
                const size_t i = size_t(y) * W + x;
                if (c->gbuf[i]) c->reproj[i] = ReprojectPixel(*c, *c->gbuf[i]);
            }

    // IntegrationPass, integration_pass.cpp:157-252
    std::vector<float> fb(size_t(W) * H * 4, 0.0f);
    const auto t0 = std::chrono::steady_clock::now();
    ParallelTiles(c->threads, ntx * nty,
                  [&](int t, Counters &lc) {
                      const int sx = (t % ntx) * kTile, sy = (t / ntx) * kTile;
                      for (int y = std::max(sy, y0); y < std::min(sy + kTile, y1); ++y)
                          for (int x = std::max(sx, x0); x < std::min(sx + kTile, x1); ++x) {
                              const size_t i = size_t(y) * W + x;
                              Rng rng(PixelSeed(i, frameIndex));
                              const float px = float(x) + rng.NextFloat();
                              const float py = float(y) + rng.NextFloat();
                              const Ray ray = hasCam
                                                  ? CameraRayParams(invView, invProj, px, py, W, H, frame->camera, rng)
                                                  : CameraRay(invView, invProj, px, py, W, H);
                              const Lambda lam = Lambda::SampleUniform(rng.NextFloat());
                              PixelTrace tr;
                              Reservoir none;  // a camera move under reprojection: no same-pixel merge
                              const PathCtx pc{c->scene,
                                               *pipe,
                                               rng,
                                               lam,
                                               tr,
                                               lc,
                                               c->hoist ? &hoistedSet : nullptr,
                                               temporalSkip ? &none : &c->reservoirs[i],
                                               spatial ? &c->resT[i] : nullptr,
                                               spatial ? &c->aux[i] : nullptr};
                              const Spec s = PathLi(pc, ray, c->gbuf[i]);
                              V3 L = SpectrumToRGB(s, lam);
                              if (hasCam && frame->camera.enable_exposure) L = L * ExposureMultiplier(frame->camera);
                              fb[4 * i + 0] = L[0];
                              fb[4 * i + 1] = L[1];
                              fb[4 * i + 2] = L[2];
                              fb[4 * i + 3] = 1.0f;
                              if (debug_out) {
                                  hdr_pixel_debug &d = debug_out[i];
                                  d.draws = rng.draws;
                                  d.light_hash = tr.lightHash;
                                  d.select_hash = tr.selectHash;
                                  d.n_light_picks = tr.lightPicks;
                                  d.prim_instance = c->gbuf[i] ? c->gbuf[i]->PrimId : -1;
                                  d.prim_triangle = c->gbuf[i] ? c->gbuf[i]->TriId : -1;
                                  d.prim_t = c->gbuf[i] ? c->gbuf[i]->Depth : 0.0f;
                                  d.sample_rgb[0] = L[0];
                                  d.sample_rgb[1] = L[1];
                                  d.sample_rgb[2] = L[2];
                              }
                          }
                  },
                  cnt);

    if (spatial) {  // the spatial-reuse pass (see SpatialRng above)
        const int K = std::min(pipe->spatial_count, kMaxSpatial);
        const int R = std::max(1, int(pipe->spatial_radius));
        const int M = std::max(1, pipe->candidate_count);
        const int cap = 20 * M;
        const float selPdf = c->scene.lights.empty() ? 0.0f : 1.0f / float(c->scene.lights.size());
        ParallelTiles(
            c->threads, ntx * nty,
            [&](int t, Counters &lc) {
                const int sx = (t % ntx) * kTile, sy = (t / ntx) * kTile;
                for (int y = std::max(sy, y0); y < std::min(sy + kTile, y1); ++y)
                    for (int x = std::max(sx, x0); x < std::min(sx + kTile, x1); ++x) {
                        const size_t i = size_t(y) * W + x;
                        if (!c->aux[i].valid) continue;
                        // the domains: q itself, the reprojected previous frame's pixel (on a camera
                        // move under reprojected temporal reuse), then the picked neighbours that
                        // pass the similarity test; each with its vertex and reservoir
                        const SpatialAux *dAux[kMaxSpatial + 1];
                        const Reservoir *dRes[kMaxSpatial + 1];
                        int slot[kMaxSpatial + 1];
                        int nd = 0;
                        dAux[nd] = &c->aux[i];
                        dRes[nd] = &c->resT[i];
                        slot[nd++] = -1;
                        if (reproject && c->reproj[i] >= 0) {
                            dAux[nd] = &c->prevAux[size_t(c->reproj[i])];
                            dRes[nd] = &c->prevRes[size_t(c->reproj[i])];
                            slot[nd++] = kMaxSpatial;
                        }
                        SpatialRng srng(uint32_t(i), frameIndex);
                        for (int k = 0; k < K; ++k) {
                            int nx, ny;
                            if (!PickNeighbour(srng, x, y, W, H, R, nx, ny)) continue;
                            const size_t n = size_t(ny) * W + nx;
                            if (!c->aux[n].valid || !SpatiallySimilar(c->aux[i].hit, c->aux[n].hit)) continue;
                            dAux[nd] = &c->aux[n];
                            dRes[nd] = &c->resT[n];
                            slot[nd++] = k;
                        }
                        float conf[kMaxSpatial + 1];
                        int countSum = 0;
                        for (int d = 0; d < nd; ++d) {
                            conf[d] = float(std::min(dRes[d]->CandidateCount, cap));
                            countSum += std::min(dRes[d]->CandidateCount, cap);
                        }
                        const SpatialAux &q = c->aux[i];
                        const ShadingPoint sp{&q.cl, q.sn, &q.lam, q.inside};
                        std::optional<RISCandidate> chosen;
                        float wsum = 0.0f;
                        int ssel = -1;
                        for (int j = 0; j < nd; ++j) {
                            const Reservoir &rj = *dRes[j];
                            if (!rj.ChosenSample || !rj.ChosenSample->Candidate || rj.W <= 0.0f) continue;
                            const float phatJ = float(rj.ChosenSample->targetFunction);
                            if (phatJ <= 0.0f) continue;
                            const Light &light = *rj.ChosenSample->Candidate->light;
                            const Light::Sample &atJ = rj.ChosenSample->Candidate->ls;
                            const auto atQ = RetargetSample(light, atJ, dAux[j]->hit.Position, q.hit.Position,
                                                            selPdf);
                            if (!atQ) continue;
                            const Spec reevalQ = RisReEvaluate(sp, q.rayDir, LightSampler::Candidate{&light, *atQ});
                            const float phatQ = SpectrumLuminance(reevalQ, q.lam);
                            if (phatQ <= 0.0f) continue;
                            const float jacJQ = j == 0 ? 1.0f : ShiftJacobian(atJ, *atQ);
                            if (jacJQ <= 0.0f) continue;
                            // balance-heuristic denominator over every domain, in q's measure
                            float den = 0.0f;
                            for (int k = 0; k < nd; ++k) {
                                float pk;
                                if (k == j) {
                                    pk = phatJ / jacJQ;
                                } else if (k == 0) {
                                    pk = phatQ;
                                } else {
                                    const SpatialAux &ak = *dAux[k];
                                    const auto atK = RetargetSample(light, *atQ, q.hit.Position, ak.hit.Position, selPdf);
                                    pk = 0.0f;
                                    if (atK) {
                                        const ShadingPoint spk{&ak.cl, ak.sn, &ak.lam, ak.inside};
                                        const float ph = SpectrumLuminance(
                                            RisReEvaluate(spk, ak.rayDir, LightSampler::Candidate{&light, *atK}), ak.lam);
                                        pk = ph * ShiftJacobian(*atQ, *atK);
                                    }
                                }
                                den += conf[k] * pk;
                            }
                            if (!(den > 0.0f)) continue;
                            const float w = conf[j] * rj.W * phatJ * phatQ / den;
                            if (!(w > 0.0f)) continue;
                            wsum += w;
                            if (srng.NextFloat() <= w / wsum) {
                                RISCandidate cnd = *rj.ChosenSample;
                                cnd.Candidate->ls = *atQ;
                                cnd.Throughput = reevalQ;
                                cnd.targetFunction = double(phatQ);
                                if (j != 0) cnd.VisibilityTested = false;
                                chosen = cnd;
                                ssel = slot[j];
                            }
                        }
                        Reservoir s;
                        s.CandidateCount = std::min(countSum, cap);
                        if (chosen) {
                            s.ChosenSample = chosen;
                            s.W = wsum / float(chosen->targetFunction);
                            if (!std::isfinite(s.W)) s.W = 0.0f;
                            const int frameCount = std::max(1, s.CandidateCount / M);
                            s.WSum = s.W * float(frameCount) * float(chosen->targetFunction);
                        }
                        if (s.ChosenSample && !s.ChosenSample->VisibilityTested && s.ChosenSample->Candidate) {
                            RISCandidate &wn = *s.ChosenSample;
                            PixelTrace tr;
                            Rng none(0);
                            const PathCtx pc{c->scene, *pipe, none, q.lam, tr, lc, nullptr, nullptr};
                            const MISContrib nee = RisEvaluateLightSample(pc, sp, q.hit, q.rayDir,
                                                                          *wn.Candidate->light, wn.Candidate->ls, false);
                            wn.Throughput = nee.Emission * nee.ThroughputMul;
                            wn.VisibilityTested = true;
                        }
                        const Spec direct =
                            (!s.ChosenSample || s.W <= 0.0f) ? Spec() : s.ChosenSample->Throughput * s.W;
                        c->reservoirs[i] = s;
                        const Spec Ls = q.throughput * ((q.emission + direct) + q.rest);
                        V3 L = SpectrumToRGB(Ls, q.lam);
                        if (hasCam && frame->camera.enable_exposure) L = L * ExposureMultiplier(frame->camera);
                        fb[4 * i + 0] = L[0];
                        fb[4 * i + 1] = L[1];
                        fb[4 * i + 2] = L[2];
                        if (debug_out) {
                            hdr_pixel_debug &d = debug_out[i];
                            d.select_hash = Fnv(d.select_hash, uint32_t(ssel));
                            d.sample_rgb[0] = L[0];
                            d.sample_rgb[1] = L[1];
                            d.sample_rgb[2] = L[2];
                        }
                    }
            },
            cnt);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    // AccumulationPass, accumulation_pass.cpp:64-106 (firefly: :22-60 in place, or out of place)
    auto lum = [](const float *p) { return 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2]; };
    if (pipe->enable_firefly_filter) {
        const std::vector<float> src = fb;
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) {
                const size_t idx = size_t(y) * W + x;
                const float *basep = c->fireflyInPlace ? fb.data() : src.data();
                const float l = lum(&basep[4 * idx]);
                float maxN = 0.0f;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (dx == 0 && dy == 0) continue;
                        const int nx = std::clamp(x + dx, 0, W - 1), ny = std::clamp(y + dy, 0, H - 1);
                        maxN = std::max(maxN, lum(&basep[4 * (size_t(ny) * W + nx)]));
                    }
                const float thr = maxN * 4.0f + 0.5f;
                if (l > thr && l > 0.0f) {
                    const float s = thr / l;
                    for (int k = 0; k < 3; ++k) fb[4 * idx + k] *= s;
                }
            }
    }
    const float invN = 1.0f / float(frameIndex + 1);
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            const size_t i = size_t(y) * W + x;
            // the Variance metrics (METRICS_ENABLED, accumulation_pass.cpp:95-98): the luminance of the
            // (clamped) frame sample, accumulated in double
            const double luminance = double(lum(&fb[4 * i]));
            c->lumSum[i] += luminance;
            c->lumSumSq[i] += luminance * luminance;
            for (int k = 0; k < 4; ++k) {
                c->accum[4 * i + k] += fb[4 * i + k];
                if (color_out) color_out[4 * i + k] = c->accum[4 * i + k] * invN;
            }
        }
    c->sampleCount = frameIndex + 1;
    c->last = cnt;
    c->histCam = true;
    c->histRecords = spatial && pipe->temporal_reprojection;
    std::memcpy(c->histView, frame->view, sizeof c->histView);
    std::memcpy(c->histProj, frame->proj, sizeof c->histProj);
    c->histOrigin = ToF(invView.Transform({{0.0, 0.0, 0.0}}));  // CameraRay's origin
    return secs;
}

// AccumulationPass's Variance metrics (accumulation_pass.cpp:108-136) over the pixel rectangle
// [x0,x1) x [y0,y1): out = {mean estimator variance, max estimator variance, mean luminance}; returns
// 0 (nothing computed) below 2 samples, as the reference logs nothing there. lum_sum / lum_sum_sq
// (optional, width*height doubles) receive the per-pixel sums.
int32_t oracle_read_variance(oracle_ctx *c, int32_t x0, int32_t y0, int32_t x1, int32_t y1, double *out,
                             double *lum_sum, double *lum_sum_sq) {
    if (lum_sum) std::copy(c->lumSum.begin(), c->lumSum.end(), lum_sum);
    if (lum_sum_sq) std::copy(c->lumSumSq.begin(), c->lumSumSq.end(), lum_sum_sq);
    const int sampleCount = c->sampleCount;
    if (sampleCount < 2) return 0;
    double meanEstimatorVariance = 0.0, maxEstimatorVariance = 0.0, meanLuminance = 0.0;
    size_t count = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            const size_t i = size_t(y) * c->W + x;
            const double n = double(sampleCount);
            const double mean = c->lumSum[i] / n;
            const double centeredSumSquares = c->lumSumSq[i] - n * mean * mean;
            const double sampleVariance = std::max(0.0, centeredSumSquares / (n - 1.0));
            const double estimatorVariance = sampleVariance / n;
            meanEstimatorVariance += estimatorVariance;
            maxEstimatorVariance = std::max(maxEstimatorVariance, estimatorVariance);
            meanLuminance += mean;
            ++count;
        }
    out[0] = meanEstimatorVariance / double(count);
    out[1] = maxEstimatorVariance;
    out[2] = meanLuminance / double(count);
    return 1;
}

void oracle_read_reservoirs(oracle_ctx *c, hdr_reservoir *out) {
    for (size_t i = 0; i < c->reservoirs.size(); ++i) {
        const Reservoir &r = c->reservoirs[i];
        hdr_reservoir o;
        std::memset(&o, 0, sizeof o);
        o.wsum = r.WSum;
        o.count = r.CandidateCount;
        o.w = r.W;
        uint32_t flags = 0, light = 0;
        if (r.ChosenSample) {
            flags |= HDR_RES_HAS_SAMPLE;
            const RISCandidate &s = *r.ChosenSample;
            if (s.VisibilityTested) flags |= HDR_RES_VIS_TESTED;
            o.target = float(s.targetFunction);
            for (int k = 0; k < 4; ++k) o.throughput[k] = s.Throughput.v[k];
            if (s.Candidate) {
                flags |= HDR_RES_HAS_CANDIDATE;
                const Light::Sample &ls = s.Candidate->ls;
                for (int k = 0; k < 3; ++k) {
                    o.dir[k] = ls.Dir[k];
                    o.color[k] = ls.Color[k];
                    o.light_normal[k] = ls.LightNormal[k];
                }
                o.dist = ls.Dist;
                o.pdf = ls.pdf.value;
                if (ls.pdf.space == PdfSpace::SolidAngle) flags |= HDR_RES_PDF_SOLID_ANGLE;
                if (s.Candidate->light->IsDelta()) flags |= HDR_RES_LIGHT_DELTA;
                light = uint32_t(s.Candidate->light - c->scene.lights.data()) + 1u;
            }
        }
        o.flags_light = (flags << 24) | (light & 0xffffffu);
        out[i] = o;
    }
}

void oracle_read_aovs(oracle_ctx *c, float *depth, float *albedo, float *normal) {
    if (depth) std::memcpy(depth, c->depth.data(), c->depth.size() * sizeof(float));
    if (albedo) std::memcpy(albedo, c->albedo.data(), c->albedo.size() * sizeof(float));
    if (normal) std::memcpy(normal, c->normal.data(), c->normal.size() * sizeof(float));
}

void oracle_read_counters(oracle_ctx *c, hdr_frame_counters *out) {
    std::memset(out, 0, sizeof *out);
    out->primary_rays = c->last.primary;
    out->bounce_rays = c->last.bounce;
    out->shadow_rays = c->last.shadow;
    out->candidate_rays = c->last.candidate;
    out->node_visits = c->last.nodes;
    out->triangle_tests = c->last.tris;
    out->instance_visits = c->last.instances;
}

// The dome CDFs of light `light` (DomeLight::Prepare): row_cdf h+1 floats, col_cdf h*(w+1), total.
int32_t oracle_dome_cdf(oracle_ctx *c, int32_t light, float *row_cdf, float *col_cdf, float *total_lum) {
    if (light < 0 || light >= int32_t(c->scene.lights.size())) return -1;
    const Light &l = c->scene.lights[size_t(light)];
    if (l.pixels.empty()) return -1;
    std::copy(l.rowCdf.begin(), l.rowCdf.end(), row_cdf);
    std::copy(l.colCdf.begin(), l.colCdf.end(), col_cdf);
    *total_lum = l.totalLum;
    return 0;
}

void oracle_intersect(oracle_ctx *c, const float *origins, const float *dirs, int32_t n, int32_t *inst, int32_t *tri,
                      float *t) {
    Counters cnt;
    for (int k = 0; k < n; ++k) {
        const V3 o(origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]);
        const V3 d(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]);
        const auto h = c->scene.IntersectScene(o, d, cnt);
        inst[k] = h ? h->PrimId : -1;
        tri[k] = h ? h->TriId : -1;
        t[k] = h ? h->Depth : 0.0f;
    }
}

void oracle_rng_raw(const uint32_t *seeds, int32_t n, int32_t n_draws, uint64_t *out) {
    for (int k = 0; k < n; ++k) {
        std::mt19937_64 e(seeds[k]);
        for (int j = 0; j < n_draws; ++j) out[size_t(k) * n_draws + j] = e();
    }
}

void oracle_rng_float(const uint32_t *seeds, int32_t n, int32_t n_draws, float *out) {
    for (int k = 0; k < n; ++k) {
        Rng r(seeds[k]);
        for (int j = 0; j < n_draws; ++j) out[size_t(k) * n_draws + j] = r.NextFloat();
    }
}

void oracle_camera_rays(const hdr_frame_desc *frame, int32_t W, int32_t H, int32_t centre, float *origins,
                        float *dirs) {
    const M4d invView = InverseD(M4d::From(frame->view)), invProj = InverseD(M4d::From(frame->proj));
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = size_t(y) * W + x;
            Rng rng(PixelSeed(i, frame->frame_index));
            float px, py;
            if (centre) {
                px = float(x) + 0.5f;
                py = float(y) + 0.5f;
            } else {
                px = float(x) + rng.NextFloat();
                py = float(y) + rng.NextFloat();
            }
            const Ray r = frame->has_camera_params
                              ? CameraRayParams(invView, invProj, px, py, W, H, frame->camera, rng)
                              : CameraRay(invView, invProj, px, py, W, H);
            for (int k = 0; k < 3; ++k) {
                origins[3 * i + k] = r.o[k];
                dirs[3 * i + k] = r.d[k];
            }
        }
}

}  // extern "C"

// ---------------------------------------------------------------- unit-level exports (tests only)
extern "C" {

// WeightedReservoir over `n` candidate weights drawn from mt19937_64(seed): Update each weight
// (CountMiss for NaN entries), then Finalize(perFrame) with p-hat = phat[selected].
// Outputs: selected index (-1 none), WSum, CandidateCount, W, RNG draws used.
void oracle_reservoir_run(const float *weights, const float *phat, int32_t n, int32_t perFrame, uint32_t seed,
                          int32_t *sel, float *wsum, int32_t *count, float *W, int32_t *draws) {
    Rng rng(seed);
    Reservoir r;
    int s = -1;
    for (int k = 0; k < n; ++k) {
        if (std::isnan(weights[k])) {
            r.CountMiss();
            continue;
        }
        RISCandidate c;
        c.targetFunction = phat[k];
        r.Update(c, weights[k], rng, k, s);
    }
    r.Finalize(perFrame);
    *sel = r.ChosenSample ? s : -1;
    *wsum = r.WSum;
    *count = r.CandidateCount;
    *W = r.W;
    *draws = int32_t(rng.draws);
}

// RIS::RandomIndexFromWeights with a given uniform u (random_index.h:27-63): chosen index or -1.
int32_t oracle_random_index(const double *weights, int32_t n, float u, double *weightSum) {
    double sum = 0.0;
    for (int k = 0; k < n; ++k)
        if (weights[k] > 0.0) sum += weights[k];
    *weightSum = sum;
    if (sum <= 0.0) return -1;
    double rem = double(u) * sum;
    for (int k = 0; k < n; ++k) {
        if (weights[k] <= 0.0) continue;
        rem -= weights[k];
        if (rem <= 0.0) return k;
    }
    return -1;
}

// SpectrumToRGB(RGBToSpectrum(rgb)) at the hero wavelengths of uniform u.
void oracle_spectrum_roundtrip(const float *rgb, float u, float *out) {
    const Lambda lam = Lambda::SampleUniform(u);
    const V3 r = SpectrumToRGB(RGBToSpectrum(V3(rgb[0], rgb[1], rgb[2]), lam), lam);
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

// SpectrumToRGB of an explicit 4-sample spectrum at the wavelengths of u.
void oracle_spectrum_to_rgb(const float *s, float u, float *out) {
    const Lambda lam = Lambda::SampleUniform(u);
    Spec sp;
    for (int k = 0; k < kSpec; ++k) sp.v[k] = s[k];
    const V3 r = SpectrumToRGB(sp, lam);
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

// GGXBsdf::Eval / Pdf for a PreviewSurface closure with the given parameters, shading normal n.
void oracle_ggx(const float *base, float metallic, float roughness, float specular, float ior, const float *n,
                const float *wo, const float *wi, float *f_out, float *pdf_out) {
    Closure c;
    c.BaseColor = V3(base[0], base[1], base[2]);
    c.Metallic = metallic;
    c.Roughness = std::max(0.001f, roughness);
    c.Specular = specular;
    c.Ior = ior;
    c.SpecularColor = V3(1.0f);
    c.Subsurface = 0.0f;
    const V3 nn(n[0], n[1], n[2]), o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    const V3 f = GGXEval(c, nn, o, i);
    f_out[0] = f[0];
    f_out[1] = f[1];
    f_out[2] = f[2];
    *pdf_out = GGXPdfFull(c, nn, o, i);
}

// PowerHeuristic(f, nf, g, ng, h, nh)
float oracle_power_heuristic(float f, int32_t nf, float g, int32_t ng, float h, int32_t nh) {
    return PowerHeuristic3(f, nf, g, ng, h, nh);
}

// PreviewSurfaceMaterial::SampleBounce on a flat surface (normal +Y, ray straight down) with the
// closure above; returns the direction, BsdfPdf and the draws used, for n seeds.
void oracle_sample_bounce(const float *base, float metallic, float roughness, float specular, float ior,
                          const float *rayDir, const uint32_t *seeds, int32_t n, float *dirs, float *pdfs,
                          int32_t *draws) {
    MaterialParams p;
    p.DiffuseColor = V3(base[0], base[1], base[2]);
    p.Metallic = metallic;
    p.Roughness = roughness;
    p.Specular = specular;
    p.Ior = ior;
    HitRecord h;
    h.SmoothNormal = V3(0, 1, 0);
    h.Normal = V3(0, 1, 0);
    const Closure c = GetClosure(&p, h);
    const Lambda lam = Lambda::SampleUniform(0.5f);
    const ShadingPoint sp{&c, c.Normal, &lam, false};
    for (int k = 0; k < n; ++k) {
        Rng rng(seeds[k]);
        BounceState st{};
        BounceSample b;
        const bool ok = PreviewSampleBounce(sp, V3(0, 0, 0), V3(rayDir[0], rayDir[1], rayDir[2]), BounceConfig{8, 8},
                                            st, rng, b);
        dirs[3 * k] = ok ? b.NextRay.d[0] : 0.0f;
        dirs[3 * k + 1] = ok ? b.NextRay.d[1] : 0.0f;
        dirs[3 * k + 2] = ok ? b.NextRay.d[2] : 0.0f;
        pdfs[k] = ok ? b.BsdfPdf.value : -1.0f;
        draws[k] = int32_t(rng.draws);
    }
}

// ILight::SampleLight of one light built from `desc` (as the scene builds it) at hitPos, once per
// seed (a fresh mt19937_64(seed) each). Per sample: ok (0 = nullopt), Dir, Color, LightNormal, Dist,
// the pdf value and its space (0 = area, 1 = solid angle), the draws used, and EvalPdf(Dir).
void oracle_sample_light(const hdr_light_desc *desc, const float *hitPos, const uint32_t *seeds, int32_t n,
                         int32_t *ok, float *dir, float *color, float *normal, float *dist, float *pdf,
                         int32_t *pdf_space, int32_t *draws, float *eval_pdf) {
    Light l;
    MakeLight(*desc, l);
    const V3 p(hitPos[0], hitPos[1], hitPos[2]);
    for (int k = 0; k < n; ++k) {
        Rng rng(seeds[k]);
        const auto s = l.SampleLight(p, rng);
        ok[k] = s ? 1 : 0;
        draws[k] = int32_t(rng.draws);
        const Light::Sample z = s ? *s : Light::Sample{};
        for (int i = 0; i < 3; ++i) {
            dir[3 * k + i] = z.Dir[i];
            color[3 * k + i] = z.Color[i];
            normal[3 * k + i] = z.LightNormal[i];
        }
        dist[k] = z.Dist;
        pdf[k] = z.pdf.value;
        pdf_space[k] = z.pdf.space == PdfSpace::SolidAngle ? 1 : 0;
        eval_pdf[k] = s ? l.EvalPdf(z.Dir).value : 0.0f;
    }
}

// UniformLightSampler's light index for uniform draws u over n lights.
void oracle_light_index(const float *u, int32_t count, int64_t n_lights, int64_t *out) {
    for (int k = 0; k < count; ++k) out[k] = int64_t(LightIndex(u[k], size_t(n_lights)));
}

// GetClosure (PreviewSurface, or DefaultMaterial when mat is null) at a hit with geometric normal
// ng, smooth normal ns and vertex colour albedo. Output floats: BaseColor 3, Emission 3, Normal 3,
// Metallic, Roughness, Specular, Ior, Opacity, Transmission.
void oracle_closure(const hdr_material_desc *mat, const float *ng, const float *ns, const float *albedo,
                    float *out) {
    MaterialParams p;
    if (mat) p = MaterialFromDesc(*mat);
    HitRecord h;
    h.Normal = V3(ng[0], ng[1], ng[2]);
    h.SmoothNormal = V3(ns[0], ns[1], ns[2]);
    h.Albedo = V3(albedo[0], albedo[1], albedo[2]);
    const Closure c = GetClosure(mat ? &p : nullptr, h);
    const float v[15] = {c.BaseColor[0], c.BaseColor[1], c.BaseColor[2], c.Emission[0], c.Emission[1],
                         c.Emission[2], c.Normal[0], c.Normal[1], c.Normal[2], c.Metallic, c.Roughness,
                         c.Specular, c.Ior, c.Opacity, c.Transmission};
    for (int i = 0; i < 15; ++i) out[i] = v[i];
}

// Inverse of a 4x4 double matrix (the GfMatrix4d::GetInverse stand-in).
void oracle_inverse4(const double *m, double *out) {
    double a[4][4], o[4][4];
    for (int i = 0; i < 16; ++i) a[i / 4][i % 4] = m[i];
    Inverse4(a, o);
    for (int i = 0; i < 16; ++i) out[i] = o[i / 4][i % 4];
}

}  // extern "C"
