/*
 * orc_spectral.c -- spectral cores of the oracle: CIE response, rgb2spec upsampling, white
 * normalisation, spectral-to-RGB conversion and hero wavelengths.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows Spectral/Rgb2Spec.h:17-169, Spectral/SpectralCore.h:41-206 and
 * Spectral/SpectralState.cpp:15-46. The coefficient table (the rgb2spec project's srgb.coeff,
 * Jakob and Hanika 2019) is missing from the reference, so tables come from the caller.
 */
#include "orc_cie.h"
#include "orc_internal.h"

/* SpectralCore.h:41-48 */
float orc_cie_interp(int which, float lambdaNm)
{
    const float* t = which == 0 ? kOrcCieX : which == 1 ? kOrcCieY : which == 2 ? kOrcCieZ : kOrcD65;
    float x = lambdaNm - 360.0f;
    x *= (float)(ORC_CIE_SAMPLES - 1) / (830.0f - 360.0f);
    int o = (int)x;
    if (o < 0) {
        o = 0;
    }
    if (o > ORC_CIE_SAMPLES - 2) {
        o = ORC_CIE_SAMPLES - 2;
    }
    const float w = x - (float)o;
    return t[o] * (1.0f - w) + t[o + 1] * w;
}

/* SpectralCore.h:133-155 */
void orc_cie_rgb_response(float lambdaNm, float* out3)
{
    const float ill = orc_cie_interp(3, lambdaNm);
    const float x = orc_cie_interp(0, lambdaNm) * ill;
    const float y = orc_cie_interp(1, lambdaNm) * ill;
    const float z = orc_cie_interp(2, lambdaNm) * ill;
    out3[0] = 3.240479f * x - 1.537150f * y - 0.498535f * z;
    out3[1] = -0.969256f * x + 1.875991f * y + 0.041556f * z;
    out3[2] = 0.055648f * x - 0.204043f * y + 1.057311f * z;
}

/* Rgb2Spec.h:23-39 */
static int find_interval(const float* scale, int res, float x)
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

/* Rgb2Spec.h:65-136 */
void orc_rgb2spec_coeffs(const OSpecModel* m, float r, float g, float b, float* out3)
{
    out3[0] = out3[1] = out3[2] = 0.0f;
    if (m->scale == NULL || m->data == NULL || m->res <= 1) {
        return;
    }
    const int res = m->res;
    const float c[3] = {clamp01f(r), clamp01f(g), clamp01f(b)};
    int dom = 0;
    for (int k = 1; k < 3; ++k) {
        if (c[k] >= c[dom]) {
            dom = k;
        }
    }
    const float z = c[dom];
    if (z <= 1.0e-8f) {
        return;
    }
    const float sf = (float)(res - 1) / z;
    const float x = c[(dom + 1) % 3] * sf;
    const float y = c[(dom + 2) % 3] * sf;
    const uint32_t xi = (uint32_t)fmin2(x, (float)(res - 2));
    const uint32_t yi = (uint32_t)fmin2(y, (float)(res - 2));
    const int zi = find_interval(m->scale, res, z);
    const int off = (((dom * res + zi) * res + (int)yi) * res + (int)xi) * 3;
    const int dx = 3, dy = 3 * res, dz = 3 * res * res;
    const float x1 = x - (float)xi, x0 = 1.0f - x1;
    const float y1 = y - (float)yi, y0 = 1.0f - y1;
    const float z1 = (z - m->scale[zi]) / fmax2(m->scale[zi + 1] - m->scale[zi], 1.0e-12f);
    const float z0 = 1.0f - z1;
    for (int k = 0; k < 3; ++k) {
        const float* d = m->data + off + k;
        out3[k] = ((d[0] * x0 + d[dx] * x1) * y0 + (d[dy] * x0 + d[dy + dx] * x1) * y1) * z0 +
                  ((d[dz] * x0 + d[dz + dx] * x1) * y0 + (d[dz + dy] * x0 + d[dz + dy + dx] * x1) * y1) * z1;
    }
}

/* Rgb2Spec.h:53-63 */
float orc_rgb2spec_sigmoid(const float* c3, float lambdaNm)
{
    const float x = c3[0] * lambdaNm * lambdaNm + c3[1] * lambdaNm + c3[2];
    const float y = 1.0f / sqrtf(x * x + 1.0f);
    return 0.5f * x * y + 0.5f;
}

/* Rgb2Spec.h:138-148 */
float orc_rgb2spec_reflectance(const OSpecModel* m, float r, float g, float b, float lambdaNm)
{
    float c[3];
    orc_rgb2spec_coeffs(m, r, g, b, c);
    return orc_rgb2spec_sigmoid(c, lambdaNm);
}

/* Rgb2Spec.h:150-169 */
float orc_rgb2spec_radiance(const OSpecModel* m, float r, float g, float b, float lambdaNm)
{
    const float mx = fmax2(r, fmax2(g, b));
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
    float c[3];
    orc_rgb2spec_coeffs(m, r, g, b, c);
    return s * orc_rgb2spec_sigmoid(c, lambdaNm);
}

/* SpectralState.cpp:15-46 */
void orc_spectral_white_norm(const OSpecModel* m, float* out3)
{
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (int l = 360; l <= 830; ++l) {
        const float lam = (float)l;
        const float s = orc_rgb2spec_reflectance(m, 1.0f, 1.0f, 1.0f, lam);
        float resp[3];
        orc_cie_rgb_response(lam, resp);
        sr += s * resp[0];
        sg += s * resp[1];
        sb += s * resp[2];
    }
    out3[0] = fmax2(sr, 1.0e-8f);
    out3[1] = fmax2(sg, 1.0e-8f);
    out3[2] = fmax2(sb, 1.0e-8f);
}

/* SpectralCore.h:185-206 */
void orc_spectral_to_rgb(const OSpecModel* m, float L, float lambdaNm, float pdf, float* out3)
{
    out3[0] = out3[1] = out3[2] = 0.0f;
    if (!isfinite(L) || L <= 0.0f) {
        return;
    }
    const float s = L / fmax2(pdf, 1.0e-8f);
    float resp[3];
    orc_cie_rgb_response(lambdaNm, resp);
    out3[0] = s * resp[0] / fmax2(m->white[0], 1.0e-8f);
    out3[1] = s * resp[1] / fmax2(m->white[1], 1.0e-8f);
    out3[2] = s * resp[2] / fmax2(m->white[2], 1.0e-8f);
}

/* Hero wavelengths: lambda_k = lambdaMin + frac(u + k/4) * range. */
void orc_hero_lambdas(const OSpecModel* m, float u, float* out4)
{
    for (int k = 0; k < 4; ++k) {
        float f = u + 0.25f * (float)k;
        if (f >= 1.0f) {
            f -= 1.0f;
        }
        out4[k] = m->lambdaMin + f * m->lambdaRange;
    }
}

/* Mean of the four wavelengths' linear-sRGB estimates (pdf 1/range each). */
OVec3 orc_hero_to_rgb(const OSpecModel* m, const float* L4, const float* lambda4)
{
    const float pdf = 1.0f / m->lambdaRange;
    float c[4][3];
    for (int k = 0; k < 4; ++k) {
        orc_spectral_to_rgb(m, L4[k], lambda4[k], pdf, c[k]);
    }
    return v3(0.25f * (((c[0][0] + c[1][0]) + c[2][0]) + c[3][0]), 0.25f * (((c[0][1] + c[1][1]) + c[2][1]) + c[3][1]),
              0.25f * (((c[0][2] + c[1][2]) + c[2][2]) + c[3][2]));
}

/* The firefly clamp of a spectral sample: non-finite -> 0, luminance above 100 scaled down,
 * negative channels kept (PathTracer.cu:49-67 without the per-channel floor). */
OVec3 orc_clamp_firefly_spectral(OVec3 r)
{
    if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float lum = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    if (lum <= 100.0f) {
        return r;
    }
    const float k = 100.0f / lum;
    return v3(r.x * k, r.y * k, r.z * k);
}
