/*
 * orc_brdf.c -- oracle Principled BSDF (Brdf/BrdfBase.h, Brdf/PrincipledBrdf.h,
 * Brdf/SubsurfaceCore.h, Brdf/BrdfDispatch.h restated).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "orc_internal.h"

#include <string.h>

#define K_PI 3.14159265f
#define K_INV_PI (1.0f / K_PI)
#define K_MIN_PDF 1.0e-8f
#define K_MIN_RADIUS 1.0e-5f

enum {
    kForceTransmitLobeOnly = 1 << 0,
    kDisableRefract = 1 << 1,
    kDisableReflect = 1 << 2,
    kDisableTirFallback = 1 << 3,
};

/* BrdfBase.h:76-81 */
static void build_basis(OVec3 n, OVec3* t, OVec3* b)
{
    const OVec3 up = fabsf(n.y) < 0.999f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    *t = vnorm(vcross(up, n));
    *b = vnorm(vcross(n, *t));
}

/* BrdfBase.h:83-88 */
static OVec3 local_to_world(OVec3 l, OVec3 n, OVec3 t, OVec3 b)
{
    return vnorm(vadd(vadd(vscale(t, l.x), vscale(b, l.y)), vscale(n, l.z)));
}

/* BrdfBase.h:90-98 */
OVec3 orc_sample_cosine_local(float u1, float u2)
{
    const float phi = 2.0f * K_PI * u1;
    const float r = sqrtf(u2);
    const float x = r * orc_cosf(phi);
    const float y = r * orc_sinf(phi);
    const float z = sqrtf(fmax2(0.0f, 1.0f - u2));
    return v3(x, y, z);
}

/* BrdfBase.h:100-103 */
static float cos_hemi_pdf(float c) { return fmax2(0.0f, c) * K_INV_PI; }

/* BrdfBase.h:110-114 */
static float alpha_from_roughness(float r)
{
    const float c = fmax2(r, 0.04f);
    return c * c;
}

/* BrdfBase.h:121-127 */
float orc_ggx_d(float cosThetaH, float alpha)
{
    const float a2 = alpha * alpha;
    const float c2 = cosThetaH * cosThetaH;
    const float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (K_PI * d * d);
}

/* BrdfBase.h:129-135 */
float orc_smith_g1(float cosTheta, float alpha)
{
    const float a2 = alpha * alpha;
    const float c2 = fmax2(cosTheta * cosTheta, 1.0e-8f);
    const float tan2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * tan2));
}

/* BrdfBase.h:137-143 */
OVec3 orc_schlick_f(float cosTheta, OVec3 f0)
{
    const float t = fmax2(0.0f, fmin2(1.0f, 1.0f - cosTheta));
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    return vadd(vscale(f0, 1.0f - t5), v3(t5, t5, t5));
}

/* BrdfBase.h:145-155 */
OVec3 orc_sample_ggx_half_local(float alpha, float u1, float u2, float* pdf)
{
    const float phi = 2.0f * K_PI * u1;
    const float a2 = alpha * alpha;
    const float cosTheta = sqrtf((1.0f - u2) / (1.0f + (a2 - 1.0f) * u2));
    const float sinTheta = sqrtf(fmax2(0.0f, 1.0f - cosTheta * cosTheta));
    const float cosPhi = orc_cosf(phi);
    const float sinPhi = orc_sinf(phi);
    *pdf = (a2 * cosTheta) /
           (K_PI * (cosTheta * cosTheta * (a2 - 1.0f) + 1.0f) * (cosTheta * cosTheta * (a2 - 1.0f) + 1.0f));
    return v3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}

/* BrdfBase.h:157-175 */
float orc_dielectric_fresnel(float cosThetaI, float etaI, float etaT)
{
    float cosI = fmax2(-1.0f, fmin2(1.0f, cosThetaI));
    float eta = etaI / fmax2(etaT, 1.0e-8f);
    if (cosI < 0.0f) {
        cosI = -cosI;
        eta = 1.0f / eta;
    }
    const float sin2t = eta * eta * (1.0f - cosI * cosI);
    if (sin2t >= 1.0f) {
        return 1.0f;
    }
    const float cosT = sqrtf(fmax2(0.0f, 1.0f - sin2t));
    const float rPar = (eta * cosI - cosT) / fmax2(eta * cosI + cosT, 1.0e-8f);
    const float rPerp = (cosI - eta * cosT) / fmax2(cosI + eta * cosT, 1.0e-8f);
    return 0.5f * (rPar * rPar + rPerp * rPerp);
}

/* BrdfBase.h:177-198 */
static int refract_relative(OVec3 wo, OVec3 normal, float etaI, float etaT, OVec3* outWi)
{
    float cosO = vdot(normal, wo);
    OVec3 n = normal;
    float etaRel = etaI / fmax2(etaT, 1.0e-8f);
    if (cosO > 0.0f) {
        n = vscale(normal, -1.0f);
        etaRel = 1.0f / etaRel;
        cosO = -cosO;
    }
    const float sin2t = etaRel * etaRel * (1.0f - cosO * cosO);
    if (sin2t >= 1.0f) {
        return 0;
    }
    const float cosT = sqrtf(fmax2(0.0f, 1.0f - sin2t));
    *outWi = vnorm(vsub(vscale(wo, -etaRel), vscale(n, etaRel * cosO + cosT)));
    return 1;
}

/* BrdfBase.h:205-214 */
static float diffuse_roughness(const OMaterial* m)
{
    const float v = m->diffuseRoughness < 0.0f ? m->roughness : m->diffuseRoughness;
    return fmax2(0.0f, fmin2(1.0f, v));
}

/* BrdfBase.h:217-256 */
float orc_oren_nayar_factor(float sigma, float cosI, float cosO, float dotWoWi)
{
    if (sigma <= 1.0e-4f) {
        return 1.0f;
    }
    const float s2 = sigma * sigma;
    const float A = 1.0f - 0.5f * s2 / (s2 + 0.33f);
    const float B = 0.45f * s2 / (s2 + 0.09f);
    const float sinI = sqrtf(fmax2(0.0f, 1.0f - cosI * cosI));
    const float sinO = sqrtf(fmax2(0.0f, 1.0f - cosO * cosO));
    float sinAlpha, sinBeta;
    if (cosI > cosO) {
        sinAlpha = sinI;
        sinBeta = sinO;
    } else {
        sinAlpha = sinO;
        sinBeta = sinI;
    }
    float cosPhi = 1.0f;
    const float d = dotWoWi - cosI * cosO;
    if (d <= 0.0f) {
        const float denom = fmax2(sinAlpha * sinBeta, 1.0e-8f);
        cosPhi = fmax2(-1.0f, fmin2(1.0f, d / denom));
    } else {
        sinAlpha = 0.0f;
        sinBeta = 0.0f;
    }
    const float tanBeta = sinBeta / fmax2(cosI, cosO);
    return A + B * fmax2(0.0f, cosPhi) * sinAlpha * tanBeta;
}

/* BrdfBase.h:259-266 */
float orc_energy_compensation(float roughness, float cosTheta)
{
    const float alpha = alpha_from_roughness(roughness);
    const float c = fmax2(cosTheta, 1.0e-4f);
    const float Ess = 1.0f - alpha * (0.3316f + 0.3237f * alpha);
    const float Ems = (1.0f - Ess) * (1.0f - c);
    return 1.0f + Ems / fmax2(Ess, 1.0e-4f);
}

/* BrdfBase.h:268-280 */
static OVec3 scatter_radius(const OMaterial* m)
{
    return v3(fmax2(0.0f, m->scatterRadiusR), fmax2(0.0f, m->scatterRadiusG), fmax2(0.0f, m->scatterRadiusB));
}
static float max_scatter_radius(const OMaterial* m)
{
    const OVec3 r = scatter_radius(m);
    return fmax2(fmax2(r.x, r.y), r.z);
}

/* SubsurfaceCore.h:20-41 */
static float burley_sample_radius(float u, float d)
{
    const float sd = fmax2(d, K_MIN_RADIUS);
    return sd * sqrtf(-orc_logf(fmax2(1.0e-8f, 1.0f - u * (1.0f - orc_expf(-1.0f / sd)))));
}
static float burley_profile(float r, float d)
{
    const float sd = fmax2(d, K_MIN_RADIUS);
    return (orc_expf(-r / sd) + orc_expf(-r / (3.0f * sd))) / (4.0f * K_PI * sd);
}
static float burley_pdf_radius(float r, float d)
{
    const float sd = fmax2(d, K_MIN_RADIUS);
    const float t0 = orc_expf(-r / sd);
    const float t1 = orc_expf(-r / (3.0f * sd));
    const float norm = 1.0f - orc_expf(-1.0f / sd);
    return fmax2(0.0f, (t0 + t1) / fmax2(2.0f * sd * norm, 1.0e-8f));
}

/* Test hooks for the Burley cores (SubsurfaceCore.h:20-64). */
float orc_burley_sample_radius(float u, float d) { return burley_sample_radius(u, d); }
float orc_burley_profile(float r, float d) { return burley_profile(r, d); }
float orc_burley_pdf_radius(float r, float d) { return burley_pdf_radius(r, d); }

/* SubsurfaceCore.h:43-64 */
static OVec3 channel_weights(const OMaterial* m)
{
    const OVec3 r = scatter_radius(m);
    const float sum = fmax2(r.x + r.y + r.z, K_MIN_RADIUS);
    return v3(r.x / sum, r.y / sum, r.z / sum);
}
static float effective_radius(const OMaterial* m, float uc)
{
    const OVec3 r = scatter_radius(m);
    const float mr = max_scatter_radius(m);
    if (mr <= K_MIN_RADIUS) {
        return K_MIN_RADIUS;
    }
    if (uc < r.x / (r.x + r.y + r.z + K_MIN_RADIUS)) {
        return fmax2(r.x, K_MIN_RADIUS);
    }
    if (uc < (r.x + r.y) / (r.x + r.y + r.z + K_MIN_RADIUS)) {
        return fmax2(r.y, K_MIN_RADIUS);
    }
    return fmax2(r.z, K_MIN_RADIUS);
}

float orc_subsurface_effective_radius(const OMaterial* m, float uChannel) { return effective_radius(m, uChannel); }

/* SubsurfaceCore.h:66-82 */
static OVec3 eval_burley(const OMaterial* m, float dist, float cosI, float cosO)
{
    const OVec3 r = scatter_radius(m);
    const OVec3 w = channel_weights(m);
    const float pr = burley_profile(dist, fmax2(r.x, K_MIN_RADIUS));
    const float pg = burley_profile(dist, fmax2(r.y, K_MIN_RADIUS));
    const float pb = burley_profile(dist, fmax2(r.z, K_MIN_RADIUS));
    const float diffuse = fmax2(cosI, 0.0f) * fmax2(cosO, 0.0f);
    return v3(w.x * pr * diffuse, w.y * pg * diffuse, w.z * pb * diffuse);
}

/* SubsurfaceCore.h:84-106 */
static void sample_exit_offset(const OBrdfCtx* ctx, float u1, float u2, float u3, OVec3* off, float* dist,
                               float* pdf)
{
    OVec3 t, b;
    build_basis(ctx->normal, &t, &b);
    const float er = effective_radius(&ctx->material, u3);
    const float r = burley_sample_radius(u1, er);
    const float phi = 2.0f * K_PI * u2;
    *off = vadd(vscale(t, r * orc_cosf(phi)), vscale(b, r * orc_sinf(phi)));
    *dist = r;
    *pdf = burley_pdf_radius(r, er) / fmax2(2.0f * K_PI * r, 1.0e-8f);
}

/* PrincipledBrdf.h:21-24 + Spectral/SpectralCore.h:221-225 (lambda ignored) */
static float material_ior(const OBrdfCtx* ctx) { return fmax2(ctx->material.ior, 1.0e-3f); }

/* PrincipledBrdf.h:31-36 */
static OVec3 mirror_reflect(OVec3 wo, OVec3 n) { return vnorm(vsub(vscale(n, 2.0f * vdot(n, wo)), wo)); }

/* PrincipledBrdf.h:38-54 */
static OVec3 specular_f0(const OBrdfCtx* ctx)
{
    const OVec3 base = v3(ctx->material.r, ctx->material.g, ctx->material.b);
    const float metallic = clamp01f(ctx->material.metallic);
    const float ior = material_ior(ctx);
    const float f0 = (ior - 1.0f) / (ior + 1.0f);
    const float v = f0 * f0 * clamp01f(clamp01f(ctx->material.specular));
    return v3(v + (base.x - v) * metallic, v + (base.y - v) * metallic, v + (base.z - v) * metallic);
}

/* PrincipledBrdf.h:56-65 */
float orc_oren_nayar_ctx(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWi = fmax2(0.0f, vdot(ctx->normal, wi));
    const float cosWo = fmax2(0.0f, vdot(ctx->normal, ctx->wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return 0.0f;
    }
    const float sigma = diffuse_roughness(&ctx->material) * 0.5f * K_PI;
    return orc_oren_nayar_factor(sigma, cosWi, cosWo, vdot(wi, ctx->wo));
}

/* PrincipledBrdf.h:75-100 */
OLobeWeights orc_lobe_weights(const OBrdfCtx* ctx)
{
    OLobeWeights w = {0.0f, 0.0f, 0.0f, 0.0f};
    const float metallic = clamp01f(ctx->material.metallic);
    const float transmission = clamp01f(ctx->material.transmission);
    const float subsurface = clamp01f(ctx->material.subsurface);
    const float dielectric = 1.0f - metallic;
    w.transmit = dielectric * transmission;
    w.subsurface = dielectric * (1.0f - transmission) * subsurface;
    w.diffuse = dielectric * (1.0f - transmission) * (1.0f - subsurface);
    w.specular = metallic + dielectric * (1.0f - transmission);
    const float sum = w.diffuse + w.specular + w.transmit + w.subsurface;
    if (sum <= K_MIN_PDF) {
        w.diffuse = 1.0f;
        return w;
    }
    const float inv = 1.0f / sum;
    w.diffuse *= inv;
    w.specular *= inv;
    w.transmit *= inv;
    w.subsurface *= inv;
    return w;
}

/* PrincipledBrdf.h:102-116 */
static int pick_lobe(const OLobeWeights* w, float u)
{
    if (u < w->diffuse) {
        return 0;
    }
    u -= w->diffuse;
    if (u < w->specular) {
        return 1;
    }
    u -= w->specular;
    if (u < w->transmit) {
        return 2;
    }
    return 3;
}

/* PrincipledBrdf.h:133-145 */
static OVec3 eval_diffuse(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWi = fmax2(0.0f, vdot(ctx->normal, wi));
    const float cosWo = fmax2(0.0f, vdot(ctx->normal, ctx->wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const OVec3 base = v3(ctx->material.r, ctx->material.g, ctx->material.b);
    const float on = orc_oren_nayar_ctx(ctx, wi);
    return vscale(base, w.diffuse * on * K_INV_PI);
}

/* PrincipledBrdf.h:147-175 */
static OVec3 eval_subsurface(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWi = fmax2(0.0f, vdot(ctx->normal, wi));
    const float cosWo = fmax2(0.0f, vdot(ctx->normal, ctx->wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const OVec3 base = v3(ctx->material.r, ctx->material.g, ctx->material.b);
    const float maxR = max_scatter_radius(&ctx->material);
    if (maxR <= K_MIN_RADIUS) {
        const float on = orc_oren_nayar_ctx(ctx, wi);
        return vscale(base, w.subsurface * on * K_INV_PI);
    }
    OVec3 t, b;
    build_basis(ctx->normal, &t, &b);
    const OVec3 wt = vadd(vscale(t, vdot(wi, t)), vscale(b, vdot(wi, b)));
    const float dist = sqrtf(vdot(wt, wt));
    const OVec3 bur = eval_burley(&ctx->material, dist, cosWi, cosWo);
    return v3(base.x * w.subsurface * bur.x, base.y * w.subsurface * bur.y, base.z * w.subsurface * bur.z);
}

/* PrincipledBrdf.h:177-201 */
static OVec3 eval_specular(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWo = fmax2(0.0f, vdot(ctx->normal, ctx->wo));
    const float cosWi = fmax2(0.0f, vdot(ctx->normal, wi));
    if (cosWo <= 0.0f || cosWi <= 0.0f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OVec3 h = vnorm(vadd(wi, ctx->wo));
    const float cosH = fmax2(0.0f, vdot(ctx->normal, h));
    const float cosHo = fmax2(0.0f, vdot(h, ctx->wo));
    if (cosHo <= 0.0f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const float alpha = alpha_from_roughness(ctx->material.roughness);
    const float D = orc_ggx_d(cosH, alpha);
    const float G = orc_smith_g1(cosWo, alpha) * orc_smith_g1(cosWi, alpha);
    const OVec3 F = orc_schlick_f(cosHo, specular_f0(ctx));
    const float ec = orc_energy_compensation(ctx->material.roughness, cosWo);
    const float denom = fmax2(4.0f * cosWo * cosWi, 1.0e-8f);
    const float scale = w.specular * ec * (D * G) / denom;
    return v3(F.x * scale, F.y * scale, F.z * scale);
}

/* PrincipledBrdf.h:203-221 */
int orc_refract_direction(const OBrdfCtx* ctx, OVec3* outWi, float* etaI, float* etaT, float* nextEta)
{
    const float cosWo = vdot(ctx->normal, ctx->wo);
    const float ior = material_ior(ctx);
    *etaI = ctx->etaMedium;
    if (cosWo > 0.0f) {
        *etaT = ior;
        *nextEta = ior;
    } else {
        *etaT = 1.0f;
        *nextEta = 1.0f;
    }
    return refract_relative(ctx->wo, ctx->normal, *etaI, *etaT, outWi);
}

/* PrincipledBrdf.h:223-229 */
static int transmit_valid(const OBrdfCtx* ctx, float cosWoSigned)
{
    if (fabs1(cosWoSigned) <= K_MIN_PDF) {
        return 0;
    }
    return cosWoSigned > 0.0f || ctx->etaMedium > 1.0f + 1.0e-4f;
}

/* PrincipledBrdf.h:231-288 */
static OVec3 eval_transmit(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWoS = vdot(ctx->normal, ctx->wo);
    if (!transmit_valid(ctx, cosWoS)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const float thin = clamp01f(ctx->material.thin);
    const OVec3 base = v3(ctx->material.r, ctx->material.g, ctx->material.b);
    const float etaI = ctx->etaMedium;
    const float ior = material_ior(ctx);
    const float etaT = cosWoS > 0.0f ? ior : 1.0f;
    const float fres = orc_dielectric_fresnel(cosWoS, etaI, etaT);
    const float refl = fres;
    const float trans = 1.0f - refl;

    const OVec3 reflected = mirror_reflect(ctx->wo, ctx->normal);
    if (vdot(wi, reflected) > 0.999f) {
        const float cosWi = fabs1(vdot(ctx->normal, wi));
        return vscale(base, w.transmit * refl / fmax2(cosWi, 1.0e-8f));
    }
    OVec3 transmitted = v3(0.0f, 0.0f, 0.0f);
    float eI = 0.0f, eT = 0.0f, next = 1.0f;
    if (thin > 0.5f) {
        transmitted = vnorm(vscale(ctx->wo, -1.0f));
    } else if (!orc_refract_direction(ctx, &transmitted, &eI, &eT, &next)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    if (vdot(wi, transmitted) > 0.999f) {
        const float cosWi = fabs1(vdot(ctx->normal, transmitted));
        const OVec3 tint = thin > 0.5f ? base : v3(1.0f, 1.0f, 1.0f);
        return v3(tint.x * w.transmit * trans / fmax2(cosWi, 1.0e-8f),
                  tint.y * w.transmit * trans / fmax2(cosWi, 1.0e-8f),
                  tint.z * w.transmit * trans / fmax2(cosWi, 1.0e-8f));
    }
    const float alpha = alpha_from_roughness(ctx->material.roughness);
    if (alpha > 0.04f) {
        const OVec3 h = vnorm(vadd(wi, ctx->wo));
        const float cosH = fmax2(0.0f, fabs1(vdot(ctx->normal, h)));
        const float D = orc_ggx_d(cosH, alpha);
        const float roughPdf = D * cosH;
        if (roughPdf > K_MIN_PDF) {
            const OVec3 tint = thin > 0.5f ? base : v3(1.0f, 1.0f, 1.0f);
            const float scale = w.transmit * trans * roughPdf;
            return v3(tint.x * scale, tint.y * scale, tint.z * scale);
        }
    }
    return v3(0.0f, 0.0f, 0.0f);
}

/* PrincipledBrdf.h:290-294 */
static float pdf_diffuse(const OBrdfCtx* ctx, OVec3 wi)
{
    const OLobeWeights w = orc_lobe_weights(ctx);
    return w.diffuse * cos_hemi_pdf(fmax2(0.0f, vdot(ctx->normal, wi)));
}

/* PrincipledBrdf.h:296-315 */
static float pdf_subsurface(const OBrdfCtx* ctx, OVec3 wi)
{
    const OLobeWeights w = orc_lobe_weights(ctx);
    const float cosPdf = cos_hemi_pdf(fmax2(0.0f, vdot(ctx->normal, wi)));
    const float maxR = max_scatter_radius(&ctx->material);
    if (maxR <= K_MIN_RADIUS) {
        return w.subsurface * cosPdf;
    }
    OVec3 t, b;
    build_basis(ctx->normal, &t, &b);
    const OVec3 wt = vadd(vscale(t, vdot(wi, t)), vscale(b, vdot(wi, b)));
    const float dist = sqrtf(vdot(wt, wt));
    const float er = max_scatter_radius(&ctx->material);
    const float radial = burley_pdf_radius(dist, er);
    return w.subsurface * fmax2(cosPdf, radial * 0.5f);
}

/* PrincipledBrdf.h:317-336 */
static float pdf_specular(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWo = fmax2(0.0f, vdot(ctx->normal, ctx->wo));
    const float cosWi = fmax2(0.0f, vdot(ctx->normal, wi));
    if (cosWo <= 0.0f || cosWi <= 0.0f) {
        return 0.0f;
    }
    const OVec3 h = vnorm(vadd(wi, ctx->wo));
    const float cosH = fmax2(0.0f, vdot(ctx->normal, h));
    const float cosHo = fmax2(0.0f, vdot(h, ctx->wo));
    if (cosHo <= 0.0f) {
        return 0.0f;
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const float alpha = alpha_from_roughness(ctx->material.roughness);
    const float hPdf = orc_ggx_d(cosH, alpha) * cosH;
    return w.specular * hPdf / fmax2(4.0f * cosHo, 1.0e-8f);
}

/* PrincipledBrdf.h:338-379 */
static float pdf_transmit(const OBrdfCtx* ctx, OVec3 wi)
{
    const float cosWoS = vdot(ctx->normal, ctx->wo);
    if (!transmit_valid(ctx, cosWoS)) {
        return 0.0f;
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const float thin = clamp01f(ctx->material.thin);
    const float etaI = ctx->etaMedium;
    const float ior = material_ior(ctx);
    const float etaT = cosWoS > 0.0f ? ior : 1.0f;
    const float refl = orc_dielectric_fresnel(cosWoS, etaI, etaT);
    const float trans = 1.0f - refl;
    float pdf = 0.0f;
    const OVec3 reflected = mirror_reflect(ctx->wo, ctx->normal);
    if (vdot(wi, reflected) > 0.999f) {
        const float cosWi = fabs1(vdot(ctx->normal, wi));
        pdf += w.transmit * refl / fmax2(cosWi, 1.0e-8f);
    }
    OVec3 transmitted = v3(0.0f, 0.0f, 0.0f);
    if (thin > 0.5f) {
        transmitted = vnorm(vscale(ctx->wo, -1.0f));
    } else {
        float eI = 0.0f, eT = 0.0f, next = 1.0f;
        if (!orc_refract_direction(ctx, &transmitted, &eI, &eT, &next)) {
            return pdf;
        }
    }
    if (vdot(wi, transmitted) > 0.999f) {
        const float cosWi = fabs1(vdot(ctx->normal, transmitted));
        pdf += w.transmit * trans / fmax2(cosWi, 1.0e-8f);
    }
    return pdf;
}

/* PrincipledBrdf.h:385-392 */
OVec3 orc_brdf_eval(const OBrdfCtx* ctx, OVec3 wi)
{
    const OVec3 d = eval_diffuse(ctx, wi);
    const OVec3 s = eval_specular(ctx, wi);
    const OVec3 t = eval_transmit(ctx, wi);
    const OVec3 ss = eval_subsurface(ctx, wi);
    return vadd(vadd(d, s), vadd(t, ss));
}

/* PrincipledBrdf.h:581-587 */
float orc_brdf_pdf(const OBrdfCtx* ctx, OVec3 wi)
{
    return pdf_diffuse(ctx, wi) + pdf_specular(ctx, wi) + pdf_transmit(ctx, wi) + pdf_subsurface(ctx, wi);
}

static OSample empty_sample(void)
{
    OSample s;
    memset(&s, 0, sizeof(s));
    s.nextMediumEta = 1.0f;
    return s;
}

/* PrincipledBrdf.h:394-408 */
static OSample sample_diffuse(const OBrdfCtx* ctx, float u1, float u2)
{
    OSample r = empty_sample();
    OVec3 t, b;
    build_basis(ctx->normal, &t, &b);
    const OVec3 l = orc_sample_cosine_local(u1, u2);
    r.direction = local_to_world(l, ctx->normal, t, b);
    r.pdf = orc_brdf_pdf(ctx, r.direction);
    r.valid = vdot(ctx->normal, r.direction) > 0.0f && r.pdf > K_MIN_PDF;
    r.transmitted = 0;
    r.nextMediumEta = ctx->etaMedium;
    return r;
}

/* PrincipledBrdf.h:410-440 */
static OSample sample_subsurface(const OBrdfCtx* ctx, float u1, float u2)
{
    OSample r = empty_sample();
    const float maxR = max_scatter_radius(&ctx->material);
    if (maxR <= K_MIN_RADIUS) {
        return sample_diffuse(ctx, u1, u2);
    }
    OVec3 off;
    float dist = 0.0f, opdf = 0.0f;
    const float u3 = u1 * 0.318f + u2 * 0.682f;
    sample_exit_offset(ctx, u1, u2, u3, &off, &dist, &opdf);
    OVec3 t, b;
    build_basis(ctx->normal, &t, &b);
    const OVec3 l = orc_sample_cosine_local(fmax2(K_MIN_PDF, u1 * 0.5f + 0.25f), fmax2(K_MIN_PDF, u2 * 0.5f + 0.25f));
    r.direction = local_to_world(l, ctx->normal, t, b);
    r.subsurfaceScatter = 1;
    r.subsurfaceExitOffset = off;
    r.subsurfaceInternalSteps = clamp01f(ctx->material.thin) > 0.5f ? 0 : 2;
    r.pdf = orc_brdf_pdf(ctx, r.direction);
    r.valid = vdot(ctx->normal, r.direction) > 0.0f && r.pdf > K_MIN_PDF;
    r.transmitted = 0;
    r.nextMediumEta = ctx->etaMedium;
    return r;
}

/* PrincipledBrdf.h:442-474 */
static OSample sample_specular(const OBrdfCtx* ctx, float u1, float u2)
{
    OSample r = empty_sample();
    const float cosWo = fmax2(0.0f, vdot(ctx->normal, ctx->wo));
    if (cosWo <= 0.0f) {
        return r;
    }
    const float alpha = alpha_from_roughness(ctx->material.roughness);
    OVec3 t, b;
    build_basis(ctx->normal, &t, &b);
    float hPdf = 0.0f;
    const OVec3 hl = orc_sample_ggx_half_local(alpha, u1, u2, &hPdf);
    const OVec3 h = local_to_world(hl, ctx->normal, t, b);
    const float cosHo = fmax2(0.0f, vdot(h, ctx->wo));
    if (cosHo <= 0.0f || hPdf <= K_MIN_PDF) {
        return r;
    }
    r.direction = vnorm(vsub(vscale(h, 2.0f * cosHo), ctx->wo));
    const float cosWi = fmax2(0.0f, vdot(ctx->normal, r.direction));
    if (cosWi <= 0.0f) {
        return r;
    }
    r.pdf = orc_brdf_pdf(ctx, r.direction);
    r.valid = r.pdf > K_MIN_PDF;
    r.transmitted = 0;
    r.nextMediumEta = ctx->etaMedium;
    return r;
}

/* PrincipledBrdf.h:476-551 */
OSample orc_brdf_sample_transmit(const OBrdfCtx* ctx, float u1, float u2)
{
    OSample r = empty_sample();
    const float cosWoS = vdot(ctx->normal, ctx->wo);
    if (!transmit_valid(ctx, cosWoS)) {
        return r;
    }
    const float thin = clamp01f(ctx->material.thin);
    const float etaI = ctx->etaMedium;
    const float ior = material_ior(ctx);
    const float etaT = cosWoS > 0.0f ? ior : 1.0f;
    const float refl = orc_dielectric_fresnel(cosWoS, etaI, etaT);
    const float trans = 1.0f - refl;

    int refractBranch = u1 < trans;
    if ((ctx->debugFlags & kDisableRefract) != 0) {
        refractBranch = 0;
    }
    if ((ctx->debugFlags & kDisableReflect) != 0) {
        refractBranch = 1;
    }

    OBrdfCtx rctx = *ctx;
    const float alpha = alpha_from_roughness(ctx->material.roughness);
    if (refractBranch && thin < 0.5f && alpha > 0.04f) {
        OVec3 t, b;
        build_basis(ctx->normal, &t, &b);
        float hPdf = 0.0f;
        const OVec3 hl = orc_sample_ggx_half_local(alpha, u2, u1 * 0.413f + u2 * 0.587f, &hPdf);
        rctx.normal = local_to_world(hl, ctx->normal, t, b);
    }

    if (refractBranch) {
        if (thin > 0.5f) {
            r.direction = vnorm(vscale(ctx->wo, -1.0f));
            r.nextMediumEta = ctx->etaMedium;
        } else {
            float eI = 0.0f, eT = 0.0f;
            if (!orc_refract_direction(&rctx, &r.direction, &eI, &eT, &r.nextMediumEta)) {
                if ((ctx->debugFlags & kDisableTirFallback) != 0) {
                    return r;
                }
                r.direction = mirror_reflect(ctx->wo, ctx->normal);
                r.transmitted = 0;
                r.nextMediumEta = ctx->etaMedium;
            } else {
                r.transmitted = 1;
            }
        }
        if (thin > 0.5f) {
            r.transmitted = 1;
        }
        const float cosWi = fabs1(vdot(ctx->normal, r.direction));
        if (cosWi <= 0.0f) {
            return r;
        }
    } else {
        r.direction = mirror_reflect(ctx->wo, ctx->normal);
        const float cosWi = fabs1(vdot(ctx->normal, r.direction));
        if (cosWi <= 0.0f) {
            return r;
        }
        r.transmitted = 0;
        r.nextMediumEta = ctx->etaMedium;
    }
    r.pdf = orc_brdf_pdf(ctx, r.direction);
    r.valid = r.pdf > K_MIN_PDF;
    return r;
}

/* PrincipledBrdf.h:553-579 */
OSample orc_brdf_sample(const OBrdfCtx* ctx, float u1, float u2)
{
    if ((ctx->debugFlags & kForceTransmitLobeOnly) != 0) {
        return orc_brdf_sample_transmit(ctx, u1, u2);
    }
    const OLobeWeights w = orc_lobe_weights(ctx);
    const int lobe = pick_lobe(&w, u1);
    float lu = 0.0f;
    switch (lobe) {
    case 0:
        lu = u1 / fmax2(w.diffuse, K_MIN_PDF);
        return sample_diffuse(ctx, lu, u2);
    case 1:
        lu = (u1 - w.diffuse) / fmax2(w.specular, K_MIN_PDF);
        return sample_specular(ctx, lu, u2);
    case 2:
        lu = (u1 - w.diffuse - w.specular) / fmax2(w.transmit, K_MIN_PDF);
        return orc_brdf_sample_transmit(ctx, lu, u2);
    default:
        lu = (u1 - w.diffuse - w.specular - w.transmit) / fmax2(w.subsurface, K_MIN_PDF);
        return sample_subsurface(ctx, lu, u2);
    }
}

/* BrdfDispatch.h:35-40 */
int orc_skips_env_nee(const OMaterial* m)
{
    const float tr = fmax2(0.0f, fmin2(1.0f, m->transmission));
    const float me = fmax2(0.0f, fmin2(1.0f, m->metallic));
    return tr > 0.99f && me < 0.01f;
}

/* BrdfDispatch.h:42-76 */
OVec3 orc_apply_throughput(OVec3 T, const OBrdfCtx* ctx, const OSample* s, OVec3 bsdf)
{
    if (!s->valid || s->pdf <= K_MIN_PDF) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float tr = fmax2(0.0f, fmin2(1.0f, ctx->material.transmission));
    const float thin = fmax2(0.0f, fmin2(1.0f, ctx->material.thin));
    const int pure = tr > 0.99f && ctx->material.metallic < 0.01f;
    OVec3 scale;
    if (pure && thin < 0.5f) {
        scale = vscale(bsdf, 1.0f / s->pdf);
    } else if (s->transmitted) {
        const float c = fabs1(vdot(ctx->normal, s->direction));
        scale = vscale(bsdf, c / s->pdf);
        if (thin > 0.5f) {
            scale = v3(scale.x * ctx->material.r, scale.y * ctx->material.g, scale.z * ctx->material.b);
        }
    } else if (s->subsurfaceScatter) {
        const float c = fmax2(0.0f, vdot(ctx->normal, s->direction));
        scale = vscale(bsdf, c / s->pdf);
    } else {
        const float c = ctx->etaMedium > 1.0f + 1.0e-4f ? fabs1(vdot(ctx->normal, s->direction))
                                                         : fmax2(0.0f, vdot(ctx->normal, s->direction));
        scale = vscale(bsdf, c / s->pdf);
    }
    return v3(T.x * scale.x, T.y * scale.y, T.z * scale.z);
}

/* SubsurfaceCore.h:108-118 (the direction is computed and discarded by the caller) */
OVec3 orc_subsurface_step_direction(OVec3 n, float u1, float u2)
{
    OVec3 t, b;
    build_basis(n, &t, &b);
    const float ct = 1.0f - 2.0f * u1;
    const float st = sqrtf(fmax2(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * K_PI * u2;
    return local_to_world(v3(st * orc_cosf(phi), st * orc_sinf(phi), ct), n, t, b);
}

float orc_max_scatter_radius(const OMaterial* m) { return max_scatter_radius(m); }
