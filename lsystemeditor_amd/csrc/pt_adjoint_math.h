// pt_adjoint_math.h -- derivative cores shared by the adjoint kernels: which material parameters a
// vertex is differentiable in, the contracted derivative of the opaque Principled BSDF value (RGB
// or per hero wavelength), the vector-Jacobian products of the firefly clamps, and the FP64 LDS
// accumulation of per-vertex gradients.
#pragma once

#include "pt_device.h"

namespace pt {

constexpr int kMaxGradMaterials = 64;
constexpr int kGradParams = 7;

struct Live {
    bool rgb, rough, metal, spec, emis, opaque, onRough;
};

PT_DEV bool untextured(uint32_t ti, const MeshAccelSceneGpu& s)
{
    return ti == 0u || s.textures == nullptr || ti >= s.textureCount;
}

// Parameters where the forward is differentiable: untextured, strictly inside (0, 1) for the
// clamped scalars, emission > 0.
PT_DEV Live live_of(const MaterialGpu& raw, const Mat& m, const MeshAccelSceneGpu& s)
{
    Live l;
    l.rgb = untextured(raw.albedoTex, s);
    l.rough = untextured(raw.roughnessTex, s) && raw.roughness > 0.0f && raw.roughness < 1.0f;
    l.metal = untextured(raw.metallicTex, s) && raw.metallic > 0.0f && raw.metallic < 1.0f;
    l.spec = raw.specular > 0.0f && raw.specular < 1.0f;
    l.emis = untextured(raw.emissionTex, s) && raw.emission > 0.0f;
    l.opaque = clamp01(m.transmission) == 0.0f && clamp01(m.subsurface) == 0.0f;
    l.onRough = raw.diffuseRoughness < 0.0f;
    return l;
}

// Contracted derivative of the opaque Principled BSDF f_k = base_k wd ON / pi + F_k ws ec D G /
// max(4 cosO cosI, 1e-8), F_k = (f0 + (base_k - f0) m)(1 - t^5) + t^5, with the normalised lobe
// weights wd = (1 - m) / (2 - m), ws = 1 / (2 - m), for a colour of n components (RGB, or the
// base colour at the four hero wavelengths): outBase[k] += a_k df_k/dbase_k and
// outScalar[0..2] += sum_k a_k df_k/d(roughness, metallic, specular). lm.rgb says whether the
// base colour is a parameter. The RGB instantiation's arithmetic is brdf_grad's.
template <class C>
PT_DEV void brdf_grad_t(const Ctx& c, V3 wi, const C& base, const C& a, const Live& lm, float* outBase,
                        float* outScalar)
{
    constexpr int n = ColN<C>::n;
    if (!lm.opaque) {
        return;
    }
    const float cosWi = max2(0.0f, dot(c.n, wi));
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return;
    }
    const Mat& m = c.m;
    const float mm = clamp01(m.metallic);
    const float inv2m = 1.0f / (2.0f - mm);
    const float wd = (1.0f - mm) * inv2m;
    const float ws = inv2m;
    // diffuse lobe
    const float dr = max2(0.0f, min2(1.0f, m.diffuseRoughness < 0.0f ? m.roughness : m.diffuseRoughness));
    const float sigma = dr * 0.5f * kPi;
    const float on = oren_nayar(sigma, cosWi, cosWo, dot(wi, c.wo));
    if (lm.rgb) {
        const float k = wd * on * kInvPi;
#pragma unroll
        for (int q = 0; q < n; ++q) {
            outBase[q] += cget(a, q) * k;
        }
    }
    float aBase = cget(a, 0) * cget(base, 0);
#pragma unroll
    for (int q = 1; q < n; ++q) {
        aBase = aBase + cget(a, q) * cget(base, q);
    }
    if (lm.metal) {
        outScalar[1] += aBase * on * kInvPi * (-inv2m * inv2m);
    }
    if (lm.rough && lm.onRough && sigma > 1.0e-4f) {
        const float s2 = sigma * sigma;
        const float sinI = sqrtf(max2(0.0f, 1.0f - cosWi * cosWi));
        const float sinO = sqrtf(max2(0.0f, 1.0f - cosWo * cosWo));
        float sa = cosWi > cosWo ? sinI : sinO;
        float sb = cosWi > cosWo ? sinO : sinI;
        float cosPhi = 1.0f;
        const float d = dot(wi, c.wo) - cosWi * cosWo;
        if (d <= 0.0f) {
            cosPhi = max2(-1.0f, min2(1.0f, d / max2(sa * sb, 1.0e-8f)));
        } else {
            sa = 0.0f;
            sb = 0.0f;
        }
        const float geo = max2(0.0f, cosPhi) * sa * (sb / max2(cosWi, cosWo));
        const float dA = -0.5f * 0.33f / ((s2 + 0.33f) * (s2 + 0.33f));
        const float dB = 0.45f * 0.09f / ((s2 + 0.09f) * (s2 + 0.09f));
        const float dOn = (dA + dB * geo) * 2.0f * sigma * 0.5f * kPi;
        outScalar[0] += aBase * wd * dOn * kInvPi;
    }
    // specular lobe
    const V3 h = normalize(add(wi, c.wo));
    const float cosH = max2(0.0f, dot(c.n, h));
    const float cosHo = max2(0.0f, dot(h, c.wo));
    if (cosHo <= 0.0f) {
        return;
    }
    const float rough = m.roughness;
    const float rc = max2(rough, 0.04f);
    const float alpha = rc * rc;
    const float a2 = alpha * alpha;
    const float c2 = cosH * cosH;
    const float dd = c2 * (a2 - 1.0f) + 1.0f;
    const float D = a2 / (kPi * dd * dd);
    float g1[2], dg1[2];
    const float cs[2] = {cosWo, cosWi};
    for (int q = 0; q < 2; ++q) {
        const float cc = max2(cs[q] * cs[q], 1.0e-8f);
        const float tan2 = (1.0f - cc) / cc;
        const float sq = sqrtf(1.0f + a2 * tan2);
        g1[q] = 2.0f / (1.0f + sq);
        dg1[q] = -tan2 / ((1.0f + sq) * (1.0f + sq) * sq);
    }
    const float G = g1[0] * g1[1];
    const float ior = max2(m.ior, 1.0e-3f);
    const float f0s = (ior - 1.0f) / (ior + 1.0f);
    const float f0v = f0s * f0s * clamp01(clamp01(m.specular));
    const float t = max2(0.0f, min2(1.0f, 1.0f - cosHo));
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    const float cw = max2(cosWo, 1.0e-4f);
    const float ess = 1.0f - alpha * (0.3316f + 0.3237f * alpha);
    const float ems = (1.0f - ess) * (1.0f - cw);
    const float ec = 1.0f + ems / max2(ess, 1.0e-4f);
    const float denom = max2(4.0f * cosWo * cosWi, 1.0e-8f);
    const float S = ws * ec * D * G / denom;
    float aF = cget(a, 0) * ((f0v + (cget(base, 0) - f0v) * mm) * (1.0f - t5) + t5);
    float aSum = cget(a, 0);
    float dF = cget(a, 0) * (cget(base, 0) - f0v);
#pragma unroll
    for (int q = 1; q < n; ++q) {
        aF = aF + cget(a, q) * ((f0v + (cget(base, q) - f0v) * mm) * (1.0f - t5) + t5);
        aSum = aSum + cget(a, q);
        dF = dF + cget(a, q) * (cget(base, q) - f0v);
    }
    if (lm.rgb) {
        const float k = mm * (1.0f - t5) * S;
#pragma unroll
        for (int q = 0; q < n; ++q) {
            outBase[q] += cget(a, q) * k;
        }
    }
    if (lm.metal) {
        outScalar[1] += dF * (1.0f - t5) * S + aF * S * inv2m;
    }
    if (lm.spec) {
        outScalar[2] += aSum * (1.0f - mm) * (1.0f - t5) * f0s * f0s * S;
    }
    if (lm.rough && rough > 0.04f) {
        const float dD = (dd - 2.0f * a2 * c2) / (kPi * dd * dd * dd) * 2.0f * alpha;
        const float dG = (dg1[0] * g1[1] + g1[0] * dg1[1]) * 2.0f * alpha;
        const float dEss = -(0.3316f + 2.0f * 0.3237f * alpha);
        const float dEms = -dEss * (1.0f - cw);
        const float dEc = ess > 1.0e-4f ? (dEms * ess - ems * dEss) / (ess * ess) : dEms / 1.0e-4f;
        const float dS = ws / denom * (dEc * D * G + ec * dD * G + ec * D * dG);
        outScalar[0] += aF * dS * 2.0f * rough;
    }
}

// out[k] += sum_c a_c d f_c / d theta_k for theta = (r, g, b, roughness, metallic, specular) of
// the opaque Principled BSDF (brdf_grad_t on the material's RGB base colour).
PT_DEV void brdf_grad(const Ctx& c, V3 wi, V3 a, const Live& lm, float out[6])
{
    brdf_grad_t(c, wi, mk(c.m.r, c.m.g, c.m.b), a, lm, out, out + 3);
}

// d(clamped radiance)/d(radiance)^T g for clamp_firefly (PathTracer.cu:49-67).
PT_DEV V3 clamp_vjp(V3 C, V3 g)
{
    if (!isfinite(C.x) || !isfinite(C.y) || !isfinite(C.z)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const V3 r = mk(max2(0.0f, C.x), max2(0.0f, C.y), max2(0.0f, C.z));
    const float lum = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    V3 gr = g;
    if (lum > 100.0f) {
        const float gdr = g.x * r.x + g.y * r.y + g.z * r.z;
        const float k = 100.0f / lum;
        gr = mk(k * (g.x - 0.2126f * gdr / lum), k * (g.y - 0.7152f * gdr / lum), k * (g.z - 0.0722f * gdr / lum));
    }
    return mk(C.x > 0.0f ? gr.x : 0.0f, C.y > 0.0f ? gr.y : 0.0f, C.z > 0.0f ? gr.z : 0.0f);
}

// d(clamped radiance)/d(radiance)^T g for clamp_firefly_spectral: negative channels pass through.
PT_DEV V3 clamp_vjp_spectral(V3 C, V3 g)
{
    if (!isfinite(C.x) || !isfinite(C.y) || !isfinite(C.z)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const float lum = C.x * 0.2126f + C.y * 0.7152f + C.z * 0.0722f;
    if (lum <= 100.0f) {
        return g;
    }
    const float gdr = g.x * C.x + g.y * C.y + g.z * C.z;
    const float k = 100.0f / lum;
    return mk(k * (g.x - 0.2126f * gdr / lum), k * (g.y - 0.7152f * gdr / lum), k * (g.z - 0.0722f * gdr / lum));
}

// Parameters of the spectral adjoint per material: the kPtSpectrumBins reflectance bins, the
// kPtSpectrumBins emission bins, roughness, metallic, specular and the emission scale.
constexpr int kSpecGradParams = 2 * kPtSpectrumBins + 4;
static_assert(kSpecGradParams == kPtSpectrumGradParams, "spectral gradient layout");
constexpr int kSpecRough = 2 * kPtSpectrumBins;
constexpr int kSpecEmission = 2 * kPtSpectrumBins + 3;

// lg points at this lane's copy of the block's per-material FP64 sums (lanes of a wave spread over
// several copies, so one LDS atomic instruction rarely has many lanes on the same address).
PT_DEV void flush(double* lg, int mi, const float acc[kGradParams])
{
    for (int k = 0; k < kGradParams; ++k) {
        if (acc[k] != 0.0f) {
            atomicAdd(&lg[mi * kGradParams + k], static_cast<double>(acc[k]));
        }
    }
}

// A spectral vertex record (derivatives with respect to the base colour at the four hero
// wavelengths, then roughness, metallic, specular) into row mi: wavelength k's term goes to its
// reflectance bin.
PT_DEV void flush_spectral(double* lg, int mi, const int bin[4], const float acc[kGradParams])
{
    double* row = lg + mi * kSpecGradParams;
    for (int k = 0; k < 4; ++k) {
        if (acc[k] != 0.0f) {
            atomicAdd(&row[bin[k]], static_cast<double>(acc[k]));
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (acc[4 + k] != 0.0f) {
            atomicAdd(&row[kSpecRough + k], static_cast<double>(acc[4 + k]));
        }
    }
}

} // namespace pt
