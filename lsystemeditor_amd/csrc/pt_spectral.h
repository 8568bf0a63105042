// pt_spectral.h -- spectral cores: CIE response, rgb2spec upsampling and hero wavelengths.
//
// The reference's spectral model (Spectral/Rgb2Spec.h, Spectral/SpectralCore.h,
// Spectral/SpectralState.cpp): an RGB reflectance becomes the smooth spectrum
// sigmoid(c0 lambda^2 + c1 lambda + c2), its coefficients interpolated trilinearly from a table
// indexed by the dominant channel, a scale (z) axis and the two other channels relative to it
// (Jakob and Hanika 2019); a spectral radiance sample becomes linear sRGB through the CIE 1931
// matching functions under D65, divided by the response of the table's white and by the
// wavelength pdf. Functions are host+device so the library computes the white normalisation
// with the same code the kernels use.
//
// Hero wavelengths: a spectral sample carries four wavelengths lambda_k = lambdaMin +
// frac(u + k/4) * range, one uniform draw u per sample. Sampling decisions (lobe choice,
// directions, Russian roulette) do not depend on the wavelength, so the four share one path
// and every colour quantity is evaluated per wavelength (S4).
#pragma once

#include "pt_cie.h"
#include "pt_device.h"

namespace pt {

constexpr float kLambdaMinNm = 360.0f;  // SpectralCore.h:34-36
constexpr float kLambdaMaxNm = 830.0f;
constexpr float kLambdaRangeNm = kLambdaMaxNm - kLambdaMinNm;

// What the kernels read in spectral mode: the coefficient table and white normalisation
// (SpectralStateGpu) and the hero-wavelength range.
struct SpecModel {
    const float* scale;
    const float* data;
    int res;
    float whiteR, whiteG, whiteB;
    float lambdaMin, lambdaRange;
};

PT_HD float sp_max(float a, float b) { return a > b ? a : b; }
PT_HD float sp_min(float a, float b) { return a < b ? a : b; }
PT_HD float sp_clamp01(float v) { return sp_max(0.0f, sp_min(1.0f, v)); }

// SpectralCore.h:41-48: linear interpolation in a 5-nm table starting at 360 nm
PT_HD float cie_interp(const float* table, float lambdaNm)
{
    float x = lambdaNm - kLambdaMinNm;
    x *= static_cast<float>(kCieSamples - 1) / kLambdaRangeNm;
    int o = static_cast<int>(x);
    o = o < 0 ? 0 : (o > kCieSamples - 2 ? kCieSamples - 2 : o);
    const float w = x - static_cast<float>(o);
    return table[o] * (1.0f - w) + table[o + 1] * w;
}

// SpectralCore.h:133-155: CIE matching functions times D65, then XYZ -> linear sRGB
PT_HD void cie_rgb_response(float lambdaNm, float& r, float& g, float& b)
{
    const float ill = cie_interp(kCieD65, lambdaNm);
    const float x = cie_interp(kCieXBar, lambdaNm) * ill;
    const float y = cie_interp(kCieYBar, lambdaNm) * ill;
    const float z = cie_interp(kCieZBar, lambdaNm) * ill;
    r = 3.240479f * x - 1.537150f * y - 0.498535f * z;
    g = -0.969256f * x + 1.875991f * y + 0.041556f * z;
    b = 0.055648f * x - 0.204043f * y + 1.057311f * z;
}

// Rgb2Spec.h:23-39: the last scale entry <= x (binary search), clamped to res - 2
PT_HD int rgb2spec_interval(const float* scale, int res, float x)
{
    int left = 0;
    const int last = res - 2;
    int size = last;
    while (size > 0) {
        const int half = size >> 1;
        const int middle = left + half + 1;
        if (scale[middle] <= x) {
            left = middle;
            size -= half + 1;
        } else {
            size = half;
        }
    }
    return left < last ? left : last;
}

// Rgb2Spec.h:65-136: trilinear coefficient fetch; zero coefficients (the constant 1/2
// spectrum) without a table or for black
PT_HD void rgb2spec_coeffs(const SpecModel& m, float r, float g, float b, float out[3])
{
    out[0] = 0.0f;
    out[1] = 0.0f;
    out[2] = 0.0f;
    if (m.scale == nullptr || m.data == nullptr || m.res <= 1) {
        return;
    }
    const int res = m.res;
    const float c[3] = {sp_clamp01(r), sp_clamp01(g), sp_clamp01(b)};
    int dom = 0;
    if (c[1] >= c[dom]) {
        dom = 1;
    }
    if (c[2] >= c[dom]) {
        dom = 2;
    }
    const float z = c[dom];
    if (z <= 1.0e-8f) {
        return;
    }
    const float sf = static_cast<float>(res - 1) / z;
    const float x = c[(dom + 1) % 3] * sf;
    const float y = c[(dom + 2) % 3] * sf;
    const uint32_t xi = static_cast<uint32_t>(sp_min(x, static_cast<float>(res - 2)));
    const uint32_t yi = static_cast<uint32_t>(sp_min(y, static_cast<float>(res - 2)));
    const int zi = rgb2spec_interval(m.scale, res, z);
    const int off = (((dom * res + zi) * res + static_cast<int>(yi)) * res + static_cast<int>(xi)) * 3;
    const int dx = 3;
    const int dy = 3 * res;
    const int dz = 3 * res * res;
    const float x1 = x - static_cast<float>(xi);
    const float x0 = 1.0f - x1;
    const float y1 = y - static_cast<float>(yi);
    const float y0 = 1.0f - y1;
    const float z1 = (z - m.scale[zi]) / sp_max(m.scale[zi + 1] - m.scale[zi], 1.0e-12f);
    const float z0 = 1.0f - z1;
    for (int k = 0; k < 3; ++k) {
        const float* d = m.data + off + k;
        out[k] = ((d[0] * x0 + d[dx] * x1) * y0 + (d[dy] * x0 + d[dy + dx] * x1) * y1) * z0 +
                 ((d[dz] * x0 + d[dz + dx] * x1) * y0 + (d[dz + dy] * x0 + d[dz + dy + dx] * x1) * y1) * z1;
    }
}

// Rgb2Spec.h:53-63: sigmoid(c0 l^2 + c1 l + c2) in the form 1/2 + x / (2 sqrt(1 + x^2))
PT_HD float rgb2spec_sigmoid(const float c[3], float lambdaNm)
{
    const float x = c[0] * lambdaNm * lambdaNm + c[1] * lambdaNm + c[2];
    const float y = 1.0f / sqrtf(x * x + 1.0f);
    return 0.5f * x * y + 0.5f;
}

// Rgb2Spec.h:138-148
PT_HD float rgb2spec_reflectance(const SpecModel& m, float r, float g, float b, float lambdaNm)
{
    float c[3];
    rgb2spec_coeffs(m, r, g, b, c);
    return rgb2spec_sigmoid(c, lambdaNm);
}

// Rgb2Spec.h:150-169: unbounded RGB scaled into the unit cube by its largest channel
PT_HD float rgb2spec_radiance_scale(float& r, float& g, float& b)
{
    const float mx = sp_max(r, sp_max(g, b));
    if (mx <= 1.0e-8f) {
        return 0.0f;
    }
    float s = 1.0f;
    if (mx > 1.0f) {
        s = mx;
        r /= s;
        g /= s;
        b /= s;
    }
    return s;
}

PT_HD float rgb2spec_radiance(const SpecModel& m, float r, float g, float b, float lambdaNm)
{
    const float s = rgb2spec_radiance_scale(r, g, b);
    if (s == 0.0f) {
        return 0.0f;
    }
    float c[3];
    rgb2spec_coeffs(m, r, g, b, c);
    return s * rgb2spec_sigmoid(c, lambdaNm);
}

// SpectralState.cpp:15-46: response of the table's white (1, 1, 1) summed over 360..830 nm
// in 1-nm steps, floored at 1e-8
inline void spectral_white_norm(const SpecModel& m, float out[3])
{
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (int l = static_cast<int>(kLambdaMinNm); l <= static_cast<int>(kLambdaMaxNm); ++l) {
        const float lam = static_cast<float>(l);
        const float s = rgb2spec_reflectance(m, 1.0f, 1.0f, 1.0f, lam);
        float r, g, b;
        cie_rgb_response(lam, r, g, b);
        sr += s * r;
        sg += s * g;
        sb += s * b;
    }
    out[0] = sp_max(sr, 1.0e-8f);
    out[1] = sp_max(sg, 1.0e-8f);
    out[2] = sp_max(sb, 1.0e-8f);
}

// SpectralCore.h:185-206 (spectralToRgb): linear sRGB of one wavelength's radiance estimate
PT_HD void spectral_to_rgb(const SpecModel& m, float L, float lambdaNm, float pdf, float& r, float& g, float& b)
{
    r = 0.0f;
    g = 0.0f;
    b = 0.0f;
    if (!__builtin_isfinite(L) || L <= 0.0f) {
        return;
    }
    const float s = L / sp_max(pdf, 1.0e-8f);
    float rr, rg, rb;
    cie_rgb_response(lambdaNm, rr, rg, rb);
    r = s * rr / sp_max(m.whiteR, 1.0e-8f);
    g = s * rg / sp_max(m.whiteG, 1.0e-8f);
    b = s * rb / sp_max(m.whiteB, 1.0e-8f);
}

// ----------------------------------------------------------------------------------------
// Colour policies: how a path's colour quantities are represented. RgbColors is the
// reference's RGB model (identity conversions); HeroColors evaluates every colour at the four
// hero wavelengths of the sample.
// ----------------------------------------------------------------------------------------
struct RgbColors {
    using Col = V3;
    PT_DEV V3 base(const Mat& m) const { return mk(m.r, m.g, m.b); }
    PT_DEV V3 radii(const Mat& m) const { return mk(max2(0.0f, m.srR), max2(0.0f, m.srG), max2(0.0f, m.srB)); }
    PT_DEV V3 emission(const Mat& m) const { return mk(m.r * m.emission, m.g * m.emission, m.b * m.emission); }
    // radiance of the emissive-list entry a next-event light sample chose (lr: its r, g, b x emission)
    PT_DEV V3 light(const MeshAccelSceneGpu&, int, V3 lr) const { return lr; }
    PT_DEV V3 radiance(V3 rgb) const { return rgb; }
    PT_DEV V3 reflectance(V3 rgb) const { return rgb; }
    // BrdfDispatch.h:29-33
    PT_DEV float lum(V3 t) const { return 0.2126f * t.x + 0.7152f * t.y + 0.0722f * t.z; }
};

// Bin of a hero wavelength in a binned spectrum: kPtSpectrumBins equal bins over the hero range.
PT_HD int spectrum_bin(const SpecModel& m, float lambdaNm)
{
    const int b = static_cast<int>((lambdaNm - m.lambdaMin) * (static_cast<float>(kPtSpectrumBins) / m.lambdaRange));
    return b < 0 ? 0 : (b > kPtSpectrumBins - 1 ? kPtSpectrumBins - 1 : b);
}

struct HeroColors {
    using Col = S4;
    SpecModel m;
    S4 lambda;
    int bin[4];  // spectrum_bin of each hero wavelength

    PT_DEV S4 binned(const float* v) const { return mk4(v[bin[0]], v[bin[1]], v[bin[2]], v[bin[3]]); }

    PT_DEV S4 sig4(const float c[3]) const
    {
        return mk4(rgb2spec_sigmoid(c, lambda.x), rgb2spec_sigmoid(c, lambda.y), rgb2spec_sigmoid(c, lambda.z),
                   rgb2spec_sigmoid(c, lambda.w));
    }
    PT_DEV S4 reflectance(V3 rgb) const
    {
        float c[3];
        rgb2spec_coeffs(m, rgb.x, rgb.y, rgb.z, c);
        return sig4(c);
    }
    PT_DEV S4 radiance(V3 rgb) const
    {
        float r = rgb.x, g = rgb.y, b = rgb.z;
        const float s = rgb2spec_radiance_scale(r, g, b);
        if (s == 0.0f) {
            return mk4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        float c[3];
        rgb2spec_coeffs(m, r, g, b, c);
        return scale(sig4(c), s);
    }
    // binned reflectance where the material has one, else the rgb2spec upsampling of its colour
    PT_DEV S4 base(const Mat& mt) const
    {
        if (mt.spec != nullptr && mt.spec[0] >= 0.0f) {
            return binned(mt.spec);
        }
        return reflectance(mk(mt.r, mt.g, mt.b));
    }
    PT_DEV S4 radii(const Mat& mt) const
    {
        return radiance(mk(max2(0.0f, mt.srR), max2(0.0f, mt.srG), max2(0.0f, mt.srB)));
    }
    // emission x binned emission where the material has one, else the upsampled r, g, b x emission
    PT_DEV S4 emission(const Mat& mt) const
    {
        if (mt.spec != nullptr && mt.spec[kPtSpectrumBins] >= 0.0f) {
            return scale(binned(mt.spec + kPtSpectrumBins), mt.emission);
        }
        return radiance(mk(mt.r * mt.emission, mt.g * mt.emission, mt.b * mt.emission));
    }
    // a next-event light sample's radiance: its material's binned emission (x the raw emission,
    // as the emissive list's r, g, b are) where it has one, else lr upsampled
    PT_DEV S4 light(const MeshAccelSceneGpu& s, int entry, V3 lr) const
    {
        if (s.materialSpectra != nullptr) {
            const uint32_t mi = tri_material(s, s.emissiveTriangleIndices[entry]);
            if (mi < s.materialCount) {
                const float* sp = s.materialSpectra + static_cast<size_t>(mi) * kPtSpectrumWords;
                if (sp[kPtSpectrumBins] >= 0.0f) {
                    return scale(binned(sp + kPtSpectrumBins), s.materials[mi].emission);
                }
            }
        }
        return radiance(lr);
    }
    PT_DEV float lum(S4 t) const { return 0.25f * (((t.x + t.y) + t.z) + t.w); }

    // Mean over the four wavelengths of their linear sRGB estimates.
    PT_DEV V3 to_rgb(S4 L) const
    {
        const float pdf = 1.0f / m.lambdaRange;
        float r0, g0, b0, r1, g1, b1, r2, g2, b2, r3, g3, b3;
        spectral_to_rgb(m, L.x, lambda.x, pdf, r0, g0, b0);
        spectral_to_rgb(m, L.y, lambda.y, pdf, r1, g1, b1);
        spectral_to_rgb(m, L.z, lambda.z, pdf, r2, g2, b2);
        spectral_to_rgb(m, L.w, lambda.w, pdf, r3, g3, b3);
        return mk(0.25f * (((r0 + r1) + r2) + r3), 0.25f * (((g0 + g1) + g2) + g3),
                  0.25f * (((b0 + b1) + b2) + b3));
    }
};

// lambda_k = lambdaMin + frac(u + k/4) * range
PT_DEV HeroColors hero_colors(const SpecModel& m, float u)
{
    HeroColors h;
    h.m = m;
    float v[4];
    for (int k = 0; k < 4; ++k) {
        float f = u + 0.25f * static_cast<float>(k);
        if (f >= 1.0f) {
            f -= 1.0f;
        }
        v[k] = m.lambdaMin + f * m.lambdaRange;
    }
    h.lambda = mk4(v[0], v[1], v[2], v[3]);
    for (int k = 0; k < 4; ++k) {
        h.bin[k] = spectrum_bin(m, v[k]);
    }
    return h;
}

// Firefly clamp of a spectral sample (PathTracer.cu:49-67 without the per-channel floor):
// non-finite -> 0, luminance above 100 scaled down to 100. Negative linear-sRGB channels are
// kept, because a wavelength sample's sRGB estimate is negative outside the gamut and
// flooring it would bias the colour.
PT_DEV V3 clamp_firefly_spectral(V3 r)
{
    if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const float lum = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    if (lum <= 100.0f) {
        return r;
    }
    const float k = 100.0f / lum;
    return mk(r.x * k, r.y * k, r.z * k);
}

} // namespace pt
