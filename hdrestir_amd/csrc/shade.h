// shade.h -- spectra, UsdPreviewSurface / GGX, lights and the direct-light estimators on the device.
//
// Every function restates one reference function (cited) with the same floating-point
// operation order, so the RNG stream of a pixel (how many draws, in which order) follows the
// CPU path. Two changes of form leave every value unchanged:
//   * the CIE fits and the three uplift Gaussians are evaluated once per pixel for its 4 hero
//     wavelengths (SpecTab) instead of ~50 times per candidate (source/core/spectrum.cpp:4-109);
//   * RIS keeps only the candidate weights (in LDS); the winning candidate is rebuilt afterwards
//     by replaying the generation draws from a copy of the RNG, instead of storing every
//     candidate (source/restir/ris_direct_light_integrator.cpp:84-215).
#pragma once

#include "dscene.h"
#include "mt64.h"

namespace hdr {

constexpr int kSpec = 4;

struct Spec {
    float v[kSpec];
};
HD Spec spec(float a) { return Spec{{a, a, a, a}}; }
HD Spec operator+(const Spec &a, const Spec &b) {
    return Spec{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}};
}
HD Spec operator*(const Spec &a, const Spec &b) {
    return Spec{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}};
}
HD Spec operator*(const Spec &a, float s) { return Spec{{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }
HD float spec_max(const Spec &a) { return smax(smax(smax(a.v[0], a.v[1]), a.v[2]), a.v[3]); }

// Per-pixel spectral table for the pixel's 4 wavelengths.
struct SpecTab {
    float l0;
    float xf[kSpec], yf[kSpec], zf[kSpec];
    float gr[kSpec], gg[kSpec], gb[kSpec];
};

__device__ __forceinline__ float lobe(float w, float mu, float s1, float s2) {
    const float t = (w - mu) * ((w < mu) ? s1 : s2);
    return exp_f(-0.5f * t * t);
}
// Wyman 2013 fits (spectrum.cpp:4-29) and Gaussian (spectrum.cpp:71-75)
__device__ __forceinline__ float xfit(float w) {
    return 0.362f * lobe(w, 442.0f, 0.0624f, 0.0374f) + 1.056f * lobe(w, 599.8f, 0.0264f, 0.0323f) -
           0.065f * lobe(w, 501.1f, 0.0490f, 0.0382f);
}
__device__ __forceinline__ float yfit(float w) {
    return 0.821f * lobe(w, 568.8f, 0.0213f, 0.0247f) + 0.286f * lobe(w, 530.9f, 0.0613f, 0.0322f);
}
__device__ __forceinline__ float zfit(float w) {
    return 1.217f * lobe(w, 437.0f, 0.0845f, 0.0278f) + 0.681f * lobe(w, 459.0f, 0.0385f, 0.0725f);
}
__device__ __forceinline__ float gauss(float x, float c) {
    const float t = (x - c) / 20.0f;
    return exp_f(-0.5f * t * t);
}

// SampledWavelengths::SampleUniform (spectrum.h:24-37) + the table.
__device__ __forceinline__ SpecTab make_tab(float u) {
    SpecTab t;
    float l[kSpec];
    l[0] = 360.0f + u * (830.0f - 360.0f);
#pragma unroll
    for (int i = 1; i < kSpec; ++i) {
        l[i] = l[0] + (float(i) * (830.0f - 360.0f) / float(kSpec));
        if (l[i] > 830.0f) l[i] -= (830.0f - 360.0f);
    }
    t.l0 = l[0];
#pragma unroll
    for (int i = 0; i < kSpec; ++i) {
        t.xf[i] = xfit(l[i]);
        t.yf[i] = yfit(l[i]);
        t.zf[i] = zfit(l[i]);
        t.gr[i] = gauss(l[i], 615.0f);
        t.gg[i] = gauss(l[i], 540.0f);
        t.gb[i] = gauss(l[i], 460.0f);
    }
    return t;
}

// RGBToSpectrum, spectrum.cpp:77-103
__device__ __forceinline__ Spec rgb2spec(const SpecTab &t, V3 rgb) {
    const float white = smin(rgb.x, smin(rgb.y, rgb.z));
    const float r = rgb.x - white, g = rgb.y - white, b = rgb.z - white;
    Spec s;
#pragma unroll
    for (int i = 0; i < kSpec; ++i) {
        float val = white;
        val += r * t.gr[i] * 1.295f;
        val += g * t.gg[i] * 1.556f;
        val += b * t.gb[i] * 1.382f;
        s.v[i] = smax(0.0f, val);
    }
    return s;
}

// SpectrumToRGB, spectrum.cpp:31-68
__device__ __forceinline__ V3 spec2rgb(const SpecTab &t, const Spec &s) {
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    float weight = (830.0f - 360.0f) / float(kSpec);
    weight /= 106.856f;
#pragma unroll
    for (int i = 0; i < kSpec; ++i) {
        X += s.v[i] * t.xf[i];
        Y += s.v[i] * t.yf[i];
        Z += s.v[i] * t.zf[i];
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
    return v3(smax(0.0f, r), smax(0.0f, g), smax(0.0f, b));
}

__device__ __forceinline__ float spec_lum(const SpecTab &t, const Spec &s) {  // spectrum.cpp:105-109
    const V3 rgb = spec2rgb(t, s);
    return smax(0.0f, 0.2126f * rgb.x + 0.7152f * rgb.y + 0.0722f * rgb.z);
}

// ------------------------------------------------------------ closure / BSDF

struct Clo {
    const GMaterial *m;
    V3 base;        // BaseColor (hit albedo for the default material)
    V3 n;           // BSDFClosure::Normal
    float subsurface;
};

__device__ __forceinline__ V3 mv3(const float *p) { return v3(p[0], p[1], p[2]); }

// GetMaterial + GetClosure (scene.cpp:273-281, preview_surface.cpp:61-124, default_material.cpp:55-63)
__device__ __forceinline__ Clo get_closure(const DeviceScene &s, const HitRec &h, bool enableSubsurface) {
    const GInstance &inst = s.instances[h.inst];
    Clo c;
    c.m = &s.materials[inst.material];
    if (c.m->is_default) {
        c.base = h.albedo;
        c.n = h.ng;
    } else {
        c.base = mv3(c.m->base);
        c.n = h.ns;
    }
    c.subsurface = enableSubsurface ? c.m->subsurface : 0.0f;
    return c;
}

__device__ __forceinline__ float fresnel_dielectric(float cosThetaI, float ior) {  // shading_helpers.h:51-65
    cosThetaI = sclamp(cosThetaI, -1.0f, 1.0f);
    float etaI = 1.0f, etaT = ior;
    if (cosThetaI > 0) {
        etaI = ior;
        etaT = 1.0f;
    }
    const float sinThetaT = etaI / etaT * sqrtf(smax(0.0f, 1.0f - cosThetaI * cosThetaI));
    if (sinThetaT >= 1.0f) return 1.0f;
    const float cosThetaT = sqrtf(smax(0.0f, 1.0f - sinThetaT * sinThetaT));
    cosThetaI = fabsf(cosThetaI);
    const float rParl = ((etaT * cosThetaI) - (etaI * cosThetaT)) / ((etaT * cosThetaI) + (etaI * cosThetaT));
    const float rPerp = ((etaI * cosThetaI) - (etaT * cosThetaT)) / ((etaI * cosThetaI) + (etaT * cosThetaT));
    return (rParl * rParl + rPerp * rPerp) / 2.0f;
}

// GGXBsdf::Eval, ggx.cpp:20-45
__device__ __forceinline__ V3 ggx_eval(const Clo &c, V3 n, V3 wo, V3 wi) {
    const GMaterial &m = *c.m;
    const V3 h = normalized(wo + wi);
    const float nDotL = smax(0.001f, dot(n, wi));
    const float nDotV = smax(0.001f, dot(n, wo));
    const float nDotH = smax(0.001f, dot(n, h));
    const float lDotH = smax(0.001f, dot(wi, h));
    const float rough = m.is_default ? 1.0f : m.roughness;
    const float metallic = m.metallic;
    const float alpha = smax(0.001f, rough * rough);
    const float alpha2 = alpha * alpha;
    const float denom = nDotH * nDotH * (alpha2 - 1.0f) + 1.0f;
    const float Dval = alpha2 / (kPiF * denom * denom);
    const float k_g = alpha / 2.0f;
    const float G_l = nDotL / (nDotL * (1.0f - k_g) + k_g);
    const float G_v = nDotV / (nDotV * (1.0f - k_g) + k_g);
    const float Gval = G_l * G_v;
    const V3 F0 = mv3(m.spec_color) * (1.0f - metallic) * 0.04f + c.base * metallic;
    const V3 Fval = F0 + (v3s(1.0f) - F0) * pow5_f(1.0f - lDotH);
    const V3 specBsdf = divd(Fval * Dval * Gval, double(4.0f * nDotL * nDotV));
    const V3 finalDiffuse = c.base * (1.0f - c.subsurface) + mv3(m.ss_color) * c.subsurface;
    const V3 diffBsdf = divd(finalDiffuse * (1.0f - metallic) * (1.0f - m.transmission), double(kPiF));
    return diffBsdf + specBsdf;
}

// GGXBsdf::Pdf, ggx.cpp:47-67
__device__ __forceinline__ float ggx_pdf(const Clo &c, V3 n, V3 wo, V3 wi) {
    const GMaterial &m = *c.m;
    const V3 h = normalized(wo + wi);
    const float nDotL = smax(0.001f, dot(n, wi));
    const float nDotV = smax(0.001f, dot(n, wo));
    const float nDotH = smax(0.001f, dot(n, h));
    const float lDotH = smax(0.001f, dot(wi, h));
    const float rough = m.is_default ? 1.0f : m.roughness;
    const float alpha = smax(0.001f, rough * rough);
    const float alpha2 = alpha * alpha;
    const float denom = nDotH * nDotH * (alpha2 - 1.0f) + 1.0f;
    const float Dval = alpha2 / (kPiF * denom * denom);
    float reflectProb = fresnel_dielectric(nDotV, m.ior) * m.specular;
    if (m.metallic > 0.0f) reflectProb = smax(reflectProb, m.metallic);
    const float diffPdf = nDotL / kPiF;
    const float specPdf = (Dval * nDotH) / (4.0f * lDotH);
    return reflectProb * specPdf + (1.0f - reflectProb) * (1.0f - m.transmission) * diffPdf;
}

// ------------------------------------------------------------ bounce sampling

struct Bounce {
    V3 o, d;              // NextRay
    Spec tmul;            // ThroughputMul
    Spec imul;            // ThroughputIntegrandMul
    float pdf;            // BsdfPdf.value (solid angle)
    bool impossible;      // ImpossibleNEEConnection
    bool skip_roulette;
};

__device__ __forceinline__ V3 sample_ggx(float u1, float u2, float roughness) {  // shading_helpers.h:33-41
    const float alpha = smax(0.001f, roughness * roughness);
    const float alpha2 = alpha * alpha;
    const float phi = 2.0f * kPiF * u1;
    const float cosTheta = sqrtf((1.0f - u2) / (1.0f + (alpha2 - 1.0f) * u2));
    const float sinTheta = sqrtf(smax(0.0f, 1.0f - cosTheta * cosTheta));
    return v3(sinTheta * cos_f(phi), cosTheta, sinTheta * sin_f(phi));
}
__device__ __forceinline__ V3 sample_cos_hemi(float u1, float u2) {  // shading_helpers.h:26-31
    const float r = sqrtf(u1);
    const float theta = 2.0f * kPiF * u2;
    return v3(r * cos_f(theta), sqrtf(1.0f - u1), r * sin_f(theta));
}
__device__ __forceinline__ V3 align_to_normal(V3 s, V3 n) {  // shading_helpers.h:43-49
    const V3 up = fabsf(n.y) < 0.999f ? v3(0, 1, 0) : v3(1, 0, 0);
    const V3 t = normalized(cross(up, n));
    const V3 b = cross(n, t);
    return s.x * t + s.y * n + s.z * b;
}
__device__ __forceinline__ float ggx_d_pdf(V3 h, V3 n, float alpha) {  // shading_helpers.h:68-74
    const float cosH = smax(0.f, dot(h, n));
    const float a2 = alpha * alpha;
    const float denom = cosH * cosH * (a2 - 1.f) + 1.f;
    return a2 / (kPiF * denom * denom) * cosH;
}
__device__ __forceinline__ float ggx_refl_weight(float nDotV, float nDotL, float vDotH, float nDotH,
                                                 float alpha) {  // shading_helpers.h:78-93
    nDotV = smax(0.001f, nDotV);
    nDotL = smax(0.001f, nDotL);
    vDotH = smax(0.001f, vDotH);
    nDotH = smax(0.001f, nDotH);
    alpha = smax(0.001f, alpha);
    const float alpha2 = alpha * alpha;
    const float g1V = 2.0f * nDotV / (nDotV + sqrtf(alpha2 + (1.0f - alpha2) * nDotV * nDotV));
    const float g1L = 2.0f * nDotL / (nDotL + sqrtf(alpha2 + (1.0f - alpha2) * nDotL * nDotL));
    const float weight = (g1V * g1L * vDotH) / (nDotV * nDotH);
    return smin(weight, 10.0f);
}
__device__ __forceinline__ V3 mirror(V3 d, V3 n) { return normalized(d - 2.0f * dot(d, n) * n); }

struct RoughRefl {
    V3 dir, h;
    float cosH, nDotV, nDotL, nDotH;
};
template <class R>
__device__ __forceinline__ RoughRefl rough_reflect(V3 rayDir, V3 n, float roughness, R &rng) {
    const float u1 = rng.next_float();
    const float u2 = rng.next_float();
    RoughRefl r;
    r.h = align_to_normal(sample_ggx(u1, u2, roughness), n);
    r.dir = normalized(rayDir - 2.0f * dot(rayDir, r.h) * r.h);
    if (dot(r.dir, n) < 0) r.dir = mirror(rayDir, n);
    r.cosH = smax(0.f, dot(r.dir, r.h));
    r.nDotV = smax(0.f, dot(n, -rayDir));
    r.nDotL = smax(0.f, dot(n, r.dir));
    r.nDotH = smax(0.f, dot(n, r.h));
    return r;
}

// Coat / sheen layer of PreviewSurfaceMaterial::SampleBounce (preview_surface.cpp:136-236).
// Returns 0 = layer not chosen, 1 = sample produced, 2 = bounce-limit error.
template <class R>
__device__ __forceinline__ int layer_bounce(const SpecTab &tab, V3 n, V3 hitPos, V3 rayDir, int maxRefl, int &refl,
                                            R &rng, float prob, float roughness, V3 tint, Bounce &out) {
    if (!(rng.next_float() < prob)) return 0;
    if (refl >= maxRefl) return 2;
    ++refl;
    if (roughness > 0.0f) {
        const RoughRefl r = rough_reflect(rayDir, n, roughness, rng);
        const float alpha = roughness * roughness;
        const float pdf = r.cosH > 0.f ? ggx_d_pdf(r.h, n, alpha) / (4.f * r.cosH) : 0.f;
        const float wgt = ggx_refl_weight(r.nDotV, r.nDotL, r.cosH, r.nDotH, alpha);
        const Spec tm = rgb2spec(tab, tint) * wgt;
        out.o = hitPos + n * 1e-4f;
        out.d = r.dir;
        out.tmul = tm;
        out.imul = tm * pdf;
        out.pdf = pdf;
        out.impossible = false;
    } else {
        out.o = hitPos + n * 1e-4f;
        out.d = mirror(rayDir, n);
        out.tmul = rgb2spec(tab, tint);
        out.imul = spec(0.0f);
        out.pdf = 0.f;
        out.impossible = true;
    }
    out.skip_roulette = true;
    return 1;
}

// IMaterial::SampleBounce: PreviewSurfaceMaterial (preview_surface.cpp:126-381) or
// DefaultMaterial (default_material.cpp:65-86). Returns false on a bounce-limit error.
template <class R>
__device__ __forceinline__ bool sample_bounce_inl(const SpecTab &tab, const Clo &c, V3 n, bool isInside, V3 hitPos,
                                                  V3 rayDir, int maxRefl, int maxRefr, int &refl, int &refr, R &rng,
                                                  Bounce &out) {
    const GMaterial &m = *c.m;
    out.skip_roulette = false;
    if (m.is_default) {
        const float sign = (n.z >= 0.0f) ? 1.0f : -1.0f;
        const float a = -1.0f / (sign + n.z);
        const float bb = n.x * n.y * a;
        const V3 t = v3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
        const V3 b = v3(bb, sign + n.y * n.y * a, -n.y);
        const float u1 = rng.next_float();
        const float u2 = rng.next_float();
        const float sinTheta = sqrtf(u1);
        const float phi = 2.0f * kPiF * u2;
        const V3 wiL = v3(sinTheta * cos_f(phi), sinTheta * sin_f(phi), sqrtf(smax(0.0f, 1.0f - u1)));
        const V3 wi = wiL.x * t + wiL.y * b + wiL.z * n;
        const float nDotL = wiL.z;
        const float rawPdf = nDotL / kPiF;
        const float safePdf = smax(rawPdf, 1e-6f);
        const float scale = rawPdf / safePdf;
        const Spec bs = rgb2spec(tab, c.base);
        out.o = hitPos + n * 1e-4f;
        out.d = wi;
        out.tmul = bs * scale;
        out.imul = bs * (nDotL / kPiF);
        out.pdf = safePdf;
        out.impossible = false;
        return true;
    }
    if (m.coat > 0.0f && !isInside) {
        const float coatFresnel = m.coat * fresnel_dielectric(dot(rayDir, n), m.coat_ior);
        const int r = layer_bounce(tab, n, hitPos, rayDir, maxRefl, refl, rng, coatFresnel, m.coat_roughness,
                                   mv3(m.coat_color), out);
        if (r) return r == 1;
    }
    if (m.sheen > 0.0f && !isInside) {
        const float cosTheta = smax(0.0f, dot(-rayDir, n));
        const float sheenFresnel = m.sheen * pow5_f(1.0f - cosTheta);
        const int r = layer_bounce(tab, n, hitPos, rayDir, maxRefl, refl, rng, sheenFresnel, m.sheen_roughness,
                                   mv3(m.sheen_color), out);
        if (r) return r == 1;
    }
    float localIor = m.ior;
    {
        const float C = 10000.0f;
        const float B = localIor - C / (589.3f * 589.3f);
        localIor = B + C / (tab.l0 * tab.l0);
    }
    const float fresnel = fresnel_dielectric(dot(rayDir, n), localIor);
    float reflectProb = fresnel * m.specular;
    if (m.metallic > 0.0f) reflectProb = smax(reflectProb, m.metallic);
    const float randVal = rng.next_float();
    if (randVal < reflectProb) {
        if (refl >= maxRefl) return false;
        ++refl;
        const V3 reflTint = mv3(m.spec_color) * (1.0f - m.metallic) + c.base * m.metallic;
        if (m.roughness > 0.0f) {
            const RoughRefl r = rough_reflect(rayDir, n, m.roughness, rng);
            const float baseAlpha = m.roughness * m.roughness;
            const float condPdf = r.cosH > 0.f ? ggx_d_pdf(r.h, n, baseAlpha) / (4.f * r.cosH) : 0.f;
            const float wgt = ggx_refl_weight(r.nDotV, r.nDotL, r.cosH, r.nDotH, baseAlpha);
            const Spec tm = rgb2spec(tab, reflTint) * wgt;
            out.o = hitPos + n * 1e-4f;
            out.d = r.dir;
            out.tmul = tm;
            out.imul = tm * condPdf;
            out.pdf = ggx_pdf(c, n, -rayDir, r.dir);
            out.impossible = false;
        } else {
            out.o = hitPos + n * 1e-4f;
            out.d = mirror(rayDir, n);
            out.tmul = rgb2spec(tab, reflTint);
            out.imul = spec(0.0f);
            out.pdf = 0.f;
            out.impossible = true;
        }
        return true;
    }
    const float remainingProb = (randVal - reflectProb) / (1.0f - reflectProb);
    if (m.transmission > 1e-6f && remainingProb < m.transmission) {
        float etaI = 1.0f, etaT = localIor;
        if (isInside) {
            etaI = localIor;
            etaT = 1.0f;
        }
        const float eta = etaI / etaT;
        const float cosThetaI = dot(rayDir, n);
        const float k = 1.0f - eta * eta * (1.0f - cosThetaI * cosThetaI);
        if (k >= 0) {
            if (refr >= maxRefr) return false;
            ++refr;
            V3 refractDir = normalized(eta * rayDir - (eta * cosThetaI + sqrtf(k)) * n);
            if (m.roughness > 0.0f) {
                const float u1 = rng.next_float();
                const float u2 = rng.next_float();
                const V3 h = align_to_normal(sample_ggx(u1, u2, m.roughness), n);
                const float cosH = dot(rayDir, h);
                const float kh = 1.0f - eta * eta * (1.0f - cosH * cosH);
                if (kh >= 0.0f) refractDir = normalized(eta * rayDir - (eta * cosH + sqrtf(kh)) * h);
                else refractDir = normalized(rayDir - 2.0f * cosH * h);
            }
            out.o = hitPos - n * 1e-4f;
            out.d = refractDir;
            out.tmul = rgb2spec(tab, mv3(m.trans_color));
            out.imul = spec(0.0f);
            out.pdf = 0.f;
            out.impossible = true;
            return true;
        }
        if (refl >= maxRefl) return false;
        ++refl;
        out.o = hitPos + n * 1e-4f;
        out.d = normalized(rayDir - 2.0f * cosThetaI * n);
        out.tmul = spec(1.0f);
        out.imul = spec(0.0f);
        out.pdf = 0.f;
        out.impossible = true;
        return true;
    }
    const float u1 = rng.next_float();
    const float u2 = rng.next_float();
    const V3 dir = align_to_normal(sample_cos_hemi(u1, u2), n);
    const float nDotL = smax(0.0f, dot(n, dir));
    const float rawPdf = nDotL / kPiF;
    const float safePdf = smax(rawPdf, 1e-6f);
    const float scale = rawPdf / safePdf;
    const V3 finalDiffuse = c.base * (1.0f - c.subsurface) + mv3(m.ss_color) * c.subsurface;
    const Spec fd = rgb2spec(tab, finalDiffuse);
    out.o = hitPos + n * 1e-4f;
    out.d = dir;
    out.tmul = fd * scale;
    out.imul = fd * (nDotL / kPiF);
    out.pdf = safePdf;
    out.impossible = false;
    return true;
}

// Out-of-line form for the integration megakernel (keeps its code size in check).
__device__ bool sample_bounce(const SpecTab &tab, const Clo &c, V3 n, bool isInside, V3 hitPos, V3 rayDir,
                              int maxRefl, int maxRefr, int &refl, int &refr, Mt64 &rng, Bounce &out) {
    return sample_bounce_inl(tab, c, n, isInside, hitPos, rayDir, maxRefl, maxRefr, refl, refr, rng, out);
}

// Consumes exactly the RNG draws sample_bounce() would for a fresh bounce state (limits 1 / 1, as
// for RIS BSDF candidates, ris_direct_light_integrator.cpp:298-302) without building the sample:
// the lobe decisions use the same float operations, the directions are never needed.
template <class R>
__device__ __forceinline__ void skip_bounce_draws_inl(const SpecTab &tab, const Clo &c, V3 n, bool isInside,
                                                      V3 rayDir, R &rng) {
    const GMaterial &m = *c.m;
    if (m.is_default) {
        rng.discard();
        rng.discard();
        return;
    }
    if (m.coat > 0.0f && !isInside) {
        const float coatFresnel = m.coat * fresnel_dielectric(dot(rayDir, n), m.coat_ior);
        if (rng.next_float() < coatFresnel) {
            if (m.coat_roughness > 0.0f) {
                rng.discard();
                rng.discard();
            }
            return;
        }
    }
    if (m.sheen > 0.0f && !isInside) {
        const float cosTheta = smax(0.0f, dot(-rayDir, n));
        const float sheenFresnel = m.sheen * pow5_f(1.0f - cosTheta);
        if (rng.next_float() < sheenFresnel) {
            if (m.sheen_roughness > 0.0f) {
                rng.discard();
                rng.discard();
            }
            return;
        }
    }
    float localIor = m.ior;
    {
        const float C = 10000.0f;
        const float B = localIor - C / (589.3f * 589.3f);
        localIor = B + C / (tab.l0 * tab.l0);
    }
    const float fresnel = fresnel_dielectric(dot(rayDir, n), localIor);
    float reflectProb = fresnel * m.specular;
    if (m.metallic > 0.0f) reflectProb = smax(reflectProb, m.metallic);
    const float randVal = rng.next_float();
    if (randVal < reflectProb) {
        if (m.roughness > 0.0f) {
            rng.discard();
            rng.discard();
        }
        return;
    }
    const float remainingProb = (randVal - reflectProb) / (1.0f - reflectProb);
    if (m.transmission > 1e-6f && remainingProb < m.transmission) {
        float etaI = 1.0f, etaT = localIor;
        if (isInside) {
            etaI = localIor;
            etaT = 1.0f;
        }
        const float eta = etaI / etaT;
        const float cosThetaI = dot(rayDir, n);
        const float k = 1.0f - eta * eta * (1.0f - cosThetaI * cosThetaI);
        if (k >= 0 && m.roughness > 0.0f) {
            rng.discard();
            rng.discard();
        }
        return;
    }
    rng.discard();
    rng.discard();
}

__device__ void skip_bounce_draws(const SpecTab &tab, const Clo &c, V3 n, bool isInside, V3 rayDir, Mt64 &rng) {
    skip_bounce_draws_inl(tab, c, n, isInside, rayDir, rng);
}

// BeerAbsorption, shading_helpers.h:133-145
__device__ __forceinline__ void beer(const SpecTab &tab, Spec &tp, const GMaterial &m, float t, bool isInside) {
    if (isInside && m.trans_depth > 0.0f && !m.thin_walled) {
        const Spec ts = rgb2spec(tab, mv3(m.trans_color));
#pragma unroll
        for (int i = 0; i < kSpec; ++i) {
            const float sigma_a = -log_f(smax(ts.v[i], 1e-4f)) / m.trans_depth;
            tp.v[i] *= exp_f(-sigma_a * t);
        }
    }
}

// ------------------------------------------------------------ lights

struct LSample {
    V3 dir, color, lnormal;
    float dist;
    float pdf;
    bool solid_angle;
};

// DomeLight texel lookup (dome_light.cpp:123-128)
__device__ __forceinline__ void dome_texel(const GDome &d, V3 dir, int &x, int &y) {
    const float theta = acos_f(sclamp(dir.y, -1.0f, 1.0f));
    float phi = atan2_f(dir.z, dir.x);
    if (phi < 0) phi = float(double(phi) + 2.0 * kPiD);
    x = iclamp(int(double(phi) / (2.0 * kPiD) * double(d.w)), 0, d.w - 1);
    y = iclamp(int(double(theta) / kPiD * double(d.h)), 0, d.h - 1);
}
__device__ __forceinline__ float dome_eval_pdf(const GDome &d, const GLight &L, V3 dir) {  // :119-133
    if (d.total_lum <= 0.f) return 0.f;
    int x, y;
    dome_texel(d, dir, x, y);
    const float *p = d.pixels + size_t(y * d.w + x) * 3;
    const float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
    return smax(lum / d.pdf_scale, 1e-6f);
}
__device__ __forceinline__ int lower_bound_f(const float *a, int n, float v) {
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// ILight::SampleLight for each light type. Returns false for std::nullopt.
template <class R>
__device__ __forceinline__ bool sample_light_inl(const DeviceScene &s, const GLight &L, V3 hitPos, R &rng,
                                                 LSample &ls) {
    switch (L.type) {
        case 2: {  // RectLight, rect_light.cpp:14-37
            const float u = rng.next_float() - 0.5f;
            const float v = rng.next_float() - 0.5f;
            const V3 lWorld = xform_point(L.m, v3(u * L.width, v * L.height, 0.0f));
            const V3 toLight = lWorld - hitPos;
            const float dist = length(toLight);
            const V3 lDir = divd(toLight, double(dist));
            if (L.area <= 0.0f) return false;
            const V3 lNormal = mv3(L.normal);
            const float cosL = smax(0.0f, dot(lNormal, -lDir));
            if (cosL <= 0.0f) return false;
            if (L.color[0] <= 0.0f && L.color[1] <= 0.0f && L.color[2] <= 0.0f) return false;
            ls.dir = lDir;
            ls.color = mv3(L.color);
            ls.lnormal = lNormal;
            ls.dist = dist;
            ls.pdf = L.inv_area;
            ls.solid_angle = false;
            return true;
        }
        case 3: {  // PointLight, point_light.cpp:16-55
            const V3 toLight = mv3(L.pos) - hitPos;
            const float dist = length(toLight);
            const V3 lDir = divd(toLight, double(dist));
            V3 col = mv3(L.color);
            if (L.cone_angle < 180.0f) {
                const float cosTheta = dot(mv3(L.normal), -lDir);
                if (cosTheta <= L.cos_cone) {
                    col = v3s(0.0f);
                } else if (L.softness > 0.0f) {
                    if (cosTheta < L.cos_inner) {
                        float f = (cosTheta - L.cos_cone) / (L.cos_inner - L.cos_cone);
                        f = f * f * (3.0f - 2.0f * f);
                        col = col * f;
                    }
                }
            }
            if (col.x <= 0.0f && col.y <= 0.0f && col.z <= 0.0f) return false;
            ls.dir = lDir;
            ls.color = col;
            ls.lnormal = -lDir;
            ls.dist = dist;
            ls.pdf = 1.0f;
            ls.solid_angle = true;
            return true;
        }
        case 0: {  // DistantLight, distant_light.cpp:10-17
            if (L.color[0] <= 0.0f && L.color[1] <= 0.0f && L.color[2] <= 0.0f) return false;
            ls.dir = mv3(L.normal);
            ls.color = mv3(L.color);
            ls.lnormal = -ls.dir;
            ls.dist = 1e30f;
            ls.pdf = 1.0f;
            ls.solid_angle = true;
            return true;
        }
        case 1: {  // DomeLight, dome_light.cpp:79-111
            if (L.dome < 0) return false;
            const GDome &d = s.domes[L.dome];
            const float u1 = rng.next_float();
            const float u2 = rng.next_float();
            const int iy = lower_bound_f(d.row_cdf, d.h + 1, u1);
            const int y = iclamp(iy - 1, 0, d.h - 1);
            const float *cc = d.col_cdf + size_t(y) * (d.w + 1);
            const int ix = lower_bound_f(cc, d.w + 1, u2);
            const int x = iclamp(ix - 1, 0, d.w - 1);
            const float theta = float(kPiD * double(float(y + 0.5f)) / double(float(d.h)));
            const float phi = float(2.0 * kPiD * double(float(x + 0.5f)) / double(float(d.w)));
            const V3 lDir = v3(sin_f(theta) * cos_f(phi), cos_f(theta), sin_f(theta) * sin_f(phi));
            const float *p = d.pixels + size_t(y * d.w + x) * 3;
            const V3 tex = v3(p[0], p[1], p[2]);
            const V3 col = cmul(tex, mv3(L.raw_color)) * L.eff_int;
            float pdf = 1.0f;
            if (d.total_lum > 0) {
                const float lum = 0.2126f * tex.x + 0.7152f * tex.y + 0.0722f * tex.z;
                pdf = smax(lum / d.pdf_scale, 1e-6f);
            }
            if (col.x <= 0.0f && col.y <= 0.0f && col.z <= 0.0f) return false;
            ls.dir = lDir;
            ls.color = col;
            ls.lnormal = -lDir;
            ls.dist = 1e30f;
            ls.pdf = pdf;
            ls.solid_angle = true;
            return true;
        }
    }
    return false;
}

__device__ bool sample_light(const DeviceScene &s, const GLight &L, V3 hitPos, Mt64 &rng, LSample &ls) {
    return sample_light_inl(s, L, hitPos, rng, ls);
}

// Number of RNG draws SampleLight consumes for a light type (rect / dome: 2; point / distant: 0).
__device__ __forceinline__ int light_draws(const DeviceScene &s, const GLight &L) {
    return (L.type == 2 || (L.type == 1 && L.dome >= 0)) ? 2 : 0;
}

// ILight::EvalPdf(hitPos, dir, dist, normal) value and space (rect: 1/area Area; dome: solid angle).
__device__ __forceinline__ float light_eval_pdf(const DeviceScene &s, const GLight &L, V3 dir, bool &solid) {
    if (L.type == 2) {
        solid = false;
        return L.area <= 0.f ? 0.f : L.inv_area;
    }
    if (L.type == 1 && L.dome >= 0) {
        solid = true;
        return dome_eval_pdf(s.domes[L.dome], L, dir);
    }
    solid = false;
    return 0.f;
}

// Pdf::ConvertTo(SolidAngle, dist2, cosY), pdf.h:20-27
__device__ __forceinline__ float to_solid_angle(float pdf, bool solid, float dist2, float cosY) {
    return solid ? pdf : pdf * dist2 / cosY;
}

// Environment radiance for a miss direction (IEnvironment::Sample)
__device__ __forceinline__ V3 env_radiance(const DeviceScene &s, V3 dir) {
    const GLight &L = s.lights[s.environment];
    if (s.env_dome < 0) return v3s(0.0f);
    const GDome &d = s.domes[s.env_dome];
    int x, y;
    dome_texel(d, dir, x, y);
    const float *p = d.pixels + size_t(y * d.w + x) * 3;
    return cmul(v3(p[0], p[1], p[2]), mv3(L.raw_color)) * L.eff_int;
}
__device__ __forceinline__ float env_pdf(const DeviceScene &s, V3 dir) {
    if (s.env_dome < 0) return 0.f;
    return dome_eval_pdf(s.domes[s.env_dome], s.lights[s.environment], dir);
}

// PowerHeuristic(f, nf, g, ng, h, nh), shading_helpers.h:121-130
__device__ __forceinline__ float power3(float f, int nf, float g, int ng, float h, int nh) {
    const float f2 = f * f, g2 = g * g;
    const float rest = float(nh) * (h * h) + 0.0f;
    const float denom = float(nf) * f2 + float(ng) * g2 + rest;
    return denom > 0.0f ? f2 / denom : 0.0f;
}
__device__ __forceinline__ float power2(float f, float g) {  // shading_helpers.h:95-100
    const float f2 = f * f, g2 = g * g;
    return f2 / (f2 + g2);
}

}  // namespace hdr
