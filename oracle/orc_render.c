/*
 * orc_render.c -- oracle path integrator, per-sample kernel body, adaptive
 * accumulation, active-list compaction, render-loop schedule and display mapping.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "orc_internal.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define K_RAY_TMAX 1.0e6f
#define K_SHADOW_BIAS 1.0e-4f
#define K_REL_SHADOW_BIAS 1.0e-3f
#define K_MIN_PDF 1.0e-8f

OVec3 orc_subsurface_step_direction(OVec3 n, float u1, float u2);
float orc_max_scatter_radius(const OMaterial* m);

OVec3 orc_vec_normalize(OVec3 v) { return vnorm(v); }

/* Optional per-thread traversal statistics (algorithmic byte model). */
static __thread OTraceStats* tlsStats = NULL;

void orc_stats_begin(OTraceStats* stats)
{
    memset(stats, 0, sizeof(*stats));
    tlsStats = stats;
}

void orc_stats_end(void) { tlsStats = NULL; }

static OHit trace(OVec3 ro, OVec3 rd, const OScene* s, float tMin, float tMax, int closest)
{
    if (tlsStats == NULL) {
        return orc_trace_ray(ro, rd, s, tMin, tMax);
    }
    uint64_t nodes = 0, tests = 0;
    const OHit h = orc_trace_ray_counted(ro, rd, s, tMin, tMax, &nodes, &tests);
    tlsStats->rays += 1;
    tlsStats->closestRays += closest ? 1 : 0;
    tlsStats->shadowRays += closest ? 0 : 1;
    tlsStats->nodeFetches += nodes;
    tlsStats->triTests += tests;
    tlsStats->hits += h.hit ? 1 : 0;
    return h;
}

/* LightSamplingCore.h:187-197 */
static int occluded(OVec3 p, OVec3 n, OVec3 wi, const OScene* s)
{
    const OVec3 o = vadd(p, vscale(n, K_SHADOW_BIAS));
    return trace(o, wi, s, K_SHADOW_BIAS, K_RAY_TMAX, 0).hit;
}

/* LightSamplingCore.h:360-369 */
static int occluded_from(OVec3 p, OVec3 wi, const OScene* s)
{
    const OVec3 o = vadd(p, vscale(wi, K_SHADOW_BIAS));
    return trace(o, wi, s, K_SHADOW_BIAS, K_RAY_TMAX, 0).hit;
}

/* Light sampling (no reference counterpart; mode 0 is the reference's). Mode 1, multiple
 * importance sampling on area lights: the emissive shadow ray runs from p + wi * bias towards the
 * light point and stops at 0.9999 of that distance, and emission found by a BSDF-sampled ray takes
 * the balance weight pdf_bsdf / (pdf_bsdf + pdf_nee). */
static int gLightMode = 0;

void orc_set_light_sampling(int mode) { gLightMode = mode; }

static int occluded_toward(OVec3 p, OVec3 wi, OVec3 lp, const OScene* s)
{
    const OVec3 o = vadd(p, vscale(wi, K_SHADOW_BIAS));
    const OVec3 toP = vsub(lp, o);
    return trace(o, toP, s, K_SHADOW_BIAS, vlen(toP) * 0.9999f, 0).hit;
}

/* Solid-angle density with which emissive NEE picks the point at distance t along dir on original
 * triangle tri (sample_emissive's areaPdf * d2 / cosL); 0 where it never does. */
static float emissive_pdf_hit(const OScene* s, uint32_t tri, OVec3 dir, float t)
{
    if (s == NULL || s->emissiveCount == 0 || s->emissiveCdf == NULL || s->emissiveIndices == NULL) {
        return 0.0f;
    }
    const float total = s->emissiveCdf[s->emissiveCount];
    if (total <= K_MIN_PDF) {
        return 0.0f;
    }
    int lo = 0, hi = (int)s->emissiveCount - 1, li = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        const uint32_t v = s->emissiveIndices[mid];
        if (v == tri) {
            li = mid;
            break;
        }
        if (v < tri) {
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    if (li < 0) {
        return 0.0f;
    }
    const OTriangle* T = &s->triangles[s->emissiveIndices[li]];
    const OVec3 cr = vcross(vsub(T->v1, T->v0), vsub(T->v2, T->v0));
    const OVec3 ln = vnorm(cr);
    const float area = 0.5f * vlen(cr);
    if (area <= K_MIN_PDF) {
        return 0.0f;
    }
    const float cosL = fmax2(0.0f, vdot(ln, vscale(dir, -1.0f)));
    if (cosL <= 0.0f) {
        return 0.0f;
    }
    const float sel = (s->emissiveCdf[li + 1] - s->emissiveCdf[li]) / total;
    const float areaPdf = sel / area;
    return areaPdf * (t * t) / cosL;
}

float orc_emissive_pdf_hit(const OScene* s, uint32_t tri, OVec3 dir, float t) { return emissive_pdf_hit(s, tri, dir, t); }

/* Balance weight of emission found by a BSDF-sampled ray in mode 1: prevPdf is the pdf of the
 * sampled direction, reach whether it lies above the previous vertex's normal (the directions its
 * emissive NEE samples). */
static float emission_weight(const OScene* s, uint32_t tri, OVec3 dir, float t, float prevPdf, int reach)
{
    if (!reach) {
        return 1.0f;
    }
    const float lightPdf = emissive_pdf_hit(s, tri, dir, t);
    if (lightPdf <= 0.0f) {
        return 1.0f;
    }
    return orc_mis_balance(prevPdf, lightPdf);
}

static OBrdfCtx make_ctx(OVec3 n, OVec3 wo, const OMaterial* m, float eta, int flags)
{
    OBrdfCtx c;
    c.normal = n;
    c.wo = wo;
    c.material = *m;
    c.etaMedium = eta;
    c.wavelengthNm = 550.0f;
    c.debugFlags = flags;
    return c;
}

/* PathIntegratorRandCore.h:65-108 */
static OVec3 env_nee(OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta, const OScene* s, const OEnv* env,
                     const OParams* params, ORng* rng)
{
    if (orc_skips_env_nee(m)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    float lightPdf = 0.0f;
    const OVec3 wi = orc_env_sample(env, params, u1, u2, &lightPdf);
    if (lightPdf <= K_MIN_PDF) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    if (occluded(p, n, wi, s)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float cosT = fmax2(0.0f, vdot(n, wi));
    if (cosT <= 0.0f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OBrdfCtx ctx = make_ctx(n, wo, m, eta, 0);
    const OVec3 L = orc_env_eval(env, params, wi);
    const OVec3 f = orc_brdf_eval(&ctx, wi);
    const float bpdf = orc_brdf_pdf(&ctx, wi);
    const float mis = orc_mis_balance(lightPdf, bpdf);
    const float scale = mis * cosT / lightPdf;
    return vmul(f, vscale(L, scale));
}

/* PathIntegratorRandCore.h:110-167. In light mode 1 the path's last vertex (last) takes the draws
 * and adds nothing: its light lies one segment past the path length at which the reference
 * gathers emission, so both modes estimate the same image. */
static OVec3 emissive_nee(OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta, const OScene* s, ORng* rng,
                          int last)
{
    if (s == NULL || s->emissiveCount == 0) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float uTri = orc_rand01(rng);
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    if (gLightMode == 1 && last) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    OVec3 lp, ln, lr;
    float areaPdf = 0.0f;
    if (!orc_light_sample_emissive(s, uTri, u1, u2, &lp, &ln, &lr, &areaPdf)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OVec3 toL = vsub(lp, p);
    const float d2 = vdot(toL, toL);
    if (d2 <= K_MIN_PDF) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OVec3 wi = vnorm(toL);
    const float cosT = fmax2(0.0f, vdot(n, wi));
    const float cosL = fmax2(0.0f, vdot(ln, vscale(wi, -1.0f)));
    if (cosT <= 0.0f || cosL <= 0.0f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    if (gLightMode == 1 ? occluded_toward(p, wi, lp, s) : occluded_from(p, wi, s)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float lightPdf = areaPdf * d2 / cosL;
    if (lightPdf <= K_MIN_PDF) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OBrdfCtx ctx = make_ctx(n, wo, m, eta, 0);
    const OVec3 f = orc_brdf_eval(&ctx, wi);
    const float bpdf = orc_brdf_pdf(&ctx, wi);
    const float mis = orc_mis_balance(lightPdf, bpdf);
    const float scale = mis * cosT / lightPdf;
    return vmul(f, vscale(lr, scale));
}

/* PathIntegratorRandCore.h:169-204 */
static OVec3 subsurface_tier_b(OVec3 T, const OMaterial* m, OVec3 n, const OSample* s, ORng* rng)
{
    if (!s->subsurfaceScatter || s->subsurfaceInternalSteps <= 0) {
        return T;
    }
    const float maxR = orc_max_scatter_radius(m);
    if (maxR <= 1.0e-5f) {
        return T;
    }
    const OVec3 albedo = v3(m->r, m->g, m->b);
    for (int step = 0; step < s->subsurfaceInternalSteps; ++step) {
        const float a = orc_rand01(rng);
        const float b = orc_rand01(rng);
        const float c = orc_rand01(rng);
        (void)orc_subsurface_step_direction(n, a, b);
        const float mfp = maxR * (-orc_logf(fmax2(1.0e-8f, c)));
        const float surv = orc_expf(-mfp / maxR);
        T = vmul(T, vscale(albedo, surv));
    }
    const float ed = vlen(s->subsurfaceExitOffset);
    if (ed > 0.0f) {
        T = vscale(T, orc_expf(-ed / maxR));
    }
    return T;
}

/* BrdfDispatch.h:29-33 */
static float lum3(OVec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

/* PathIntegratorRandCore.h:206-363 */
OVec3 orc_trace_path(OVec3 ro, OVec3 rd, const OScene* s, const OParams* params, const OEnv* env, ORng* rng)
{
    const OHit first = trace(ro, rd, s, 0.0f, K_RAY_TMAX, 1);
    if (!first.hit) {
        return orc_env_eval(env, params, rd);
    }
    int maxDepth = 32;
    int rrMin = 3;
    int flags = 0;
    if (params != NULL) {
        if (params->maxPathDepth > 0) {
            maxDepth = params->maxPathDepth;
        } else if (params->maxPathDepth == 0) {
            maxDepth = 256;
        }
        if (params->russianRouletteMinDepth >= 0) {
            rrMin = params->russianRouletteMinDepth;
        }
        flags = params->brdfDebugFlags;
    }
    OVec3 L = v3(0.0f, 0.0f, 0.0f);
    OVec3 T = v3(1.0f, 1.0f, 1.0f);
    OVec3 origin = ro;
    OVec3 dir = rd;
    OHit hit = first;
    float eta = 1.0f;
    float prevPdf = 0.0f;
    int reach = 0;

    for (int depth = 0; depth < maxDepth; ++depth) {
        const OVec3 p = vray(origin, dir, hit.t);
        const OVec3 n = hit.normal;
        const OVec3 wo = vscale(dir, -1.0f);
        const OMaterial src = orc_fetch_material(s, hit.triangleIndex);
        const OMaterial m = orc_resolve_material(&src, hit.uv, s);

        if (gLightMode == 1 && depth > 0 && m.emission > 0.0f) {
            const float w = emission_weight(s, hit.triangleIndex, dir, hit.t, prevPdf, reach);
            L = vadd(L, vmul(T, vscale(v3(m.r * m.emission, m.g * m.emission, m.b * m.emission), w)));
        } else {
            L = vadd(L, vmul(T, v3(m.r * m.emission, m.g * m.emission, m.b * m.emission)));
        }
        L = vadd(L, vmul(T, env_nee(p, n, wo, &m, eta, s, env, params, rng)));
        L = vadd(L, vmul(T, emissive_nee(p, n, wo, &m, eta, s, rng, depth + 1 == maxDepth)));

        const OBrdfCtx ctx = make_ctx(n, wo, &m, eta, flags);
        const float u1 = orc_rand01(rng);
        const float u2 = orc_rand01(rng);
        const OSample smp = orc_brdf_sample(&ctx, u1, u2);
        if (!smp.valid || smp.pdf <= K_MIN_PDF) {
            break;
        }
        const OVec3 f = orc_brdf_eval(&ctx, smp.direction);
        T = orc_apply_throughput(T, &ctx, &smp, f);
        T = subsurface_tier_b(T, &m, n, &smp, rng);
        if ((flags & (1 << 4)) != 0) {
            T = smp.transmitted ? vmul(T, v3(0.2f, 1.0f, 0.2f)) : vmul(T, v3(1.0f, 0.2f, 0.2f));
        }
        const float tl = lum3(T);
        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
            break;
        }
        if (depth >= rrMin) {
            const float surv = fmin2(0.95f, fmax2(tl, 0.05f));
            if (orc_rand01(rng) > surv) {
                break;
            }
            T = vscale(T, 1.0f / surv);
        }
        eta = smp.nextMediumEta;
        prevPdf = smp.pdf;
        reach = vdot(n, smp.direction) > 0.0f;
        const float bias = fmax2(K_SHADOW_BIAS, K_REL_SHADOW_BIAS * hit.t);
        const float sign = vdot(n, smp.direction) >= 0.0f ? 1.0f : -1.0f;
        origin = vadd(p, vscale(n, sign * bias));
        dir = smp.direction;
        hit = trace(origin, dir, s, bias, K_RAY_TMAX, 1);
        if (!hit.hit) {
            const OVec3 envL = orc_env_eval(env, params, dir);
            const float lpdf = orc_env_pdf(env, params, dir);
            const float mis = orc_mis_balance(smp.pdf, lpdf);
            L = vadd(L, vmul(T, vscale(envL, mis)));
            break;
        }
    }
    return L;
}

/* ---- hero-wavelength integrator -----------------------------------------------------------
 * Same path as orc_trace_path (sampling, pdfs and Russian roulette use the RGB material); every
 * value is evaluated per wavelength k as one channel of the RGB model applied to a grey material
 * whose colour is the rgb2spec reflectance of the base colour at lambda_k and whose scatter
 * radius is the radiance-upsampled scatter radius at lambda_k. Light values are upsampled as
 * radiance. The tier-B walk keeps the RGB radius and multiplies each wavelength by its albedo. */
static OMaterial grey_material(const OMaterial* m, float base, float radius)
{
    OMaterial g = *m;
    g.r = base;
    g.g = base;
    g.b = base;
    g.scatterRadiusR = radius;
    g.scatterRadiusG = radius;
    g.scatterRadiusB = radius;
    return g;
}

typedef struct {
    const OSpecModel* model;
    float lambda[4];
    int bin[4]; /* binned-spectrum bin of each wavelength */
    OMaterial grey[4];
    OBrdfCtx ctx[4];
} SpecVertex;

/* Bin of a hero wavelength: ORC_SPECTRUM_BINS equal bins over the hero range. */
static int spectrum_bin(const OSpecModel* m, float lambdaNm)
{
    const int b = (int)((lambdaNm - m->lambdaMin) * ((float)ORC_SPECTRUM_BINS / m->lambdaRange));
    return b < 0 ? 0 : (b > ORC_SPECTRUM_BINS - 1 ? ORC_SPECTRUM_BINS - 1 : b);
}

/* The binned spectra record of a triangle's material, or NULL. */
static const float* material_spectrum(const OScene* s, uint32_t tri)
{
    if (s == NULL || s->spectra == NULL || s->materials == NULL || s->triangles == NULL || tri >= s->triangleCount) {
        return NULL;
    }
    const uint32_t mi = s->triangles[tri].materialIndex;
    return mi < s->materialCount ? s->spectra + (size_t)mi * ORC_SPECTRUM_WORDS : NULL;
}

static void spec_vertex(SpecVertex* v, const OMaterial* m, const float* sp, OVec3 n, OVec3 wo, float eta, int flags)
{
    for (int k = 0; k < 4; ++k) {
        const float b = sp != NULL && sp[0] >= 0.0f ? sp[v->bin[k]]
                                                    : orc_rgb2spec_reflectance(v->model, m->r, m->g, m->b, v->lambda[k]);
        const float r = orc_rgb2spec_radiance(v->model, fmax2(0.0f, m->scatterRadiusR), fmax2(0.0f, m->scatterRadiusG),
                                              fmax2(0.0f, m->scatterRadiusB), v->lambda[k]);
        v->grey[k] = grey_material(m, b, r);
        v->ctx[k] = make_ctx(n, wo, &v->grey[k], eta, flags);
    }
}

/* f_k * (light_k * scale) for a light sample of radiance Le4 at the four wavelengths */
static void spec_light_term4(const SpecVertex* v, OVec3 wi, const float* Le4, float scale, float* out4)
{
    for (int k = 0; k < 4; ++k) {
        const float f = orc_brdf_eval(&v->ctx[k], wi).x;
        out4[k] = f * (Le4[k] * scale);
    }
}

/* the same for a light sample of RGB radiance rgb (upsampled as radiance) */
static void spec_light_term(const SpecVertex* v, OVec3 wi, OVec3 rgb, float scale, float* out4)
{
    float Le[4];
    for (int k = 0; k < 4; ++k) {
        Le[k] = orc_rgb2spec_radiance(v->model, rgb.x, rgb.y, rgb.z, v->lambda[k]);
    }
    spec_light_term4(v, wi, Le, scale, out4);
}

/* Emitted radiance of a material at the four wavelengths: emission x its binned emission where
 * it has one, else the upsampled r, g, b x emission. */
static void spec_emission(const SpecVertex* v, const OMaterial* m, const float* sp, float* Le4)
{
    for (int k = 0; k < 4; ++k) {
        Le4[k] = sp != NULL && sp[ORC_SPECTRUM_BINS] >= 0.0f
                     ? sp[ORC_SPECTRUM_BINS + v->bin[k]] * m->emission
                     : orc_rgb2spec_radiance(v->model, m->r * m->emission, m->g * m->emission, m->b * m->emission,
                                             v->lambda[k]);
    }
}

static int spec_env_nee(const SpecVertex* v, OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta,
                        const OScene* s, const OEnv* env, const OParams* params, ORng* rng, float* out4)
{
    out4[0] = out4[1] = out4[2] = out4[3] = 0.0f;
    if (orc_skips_env_nee(m)) {
        return 0;
    }
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    float lightPdf = 0.0f;
    const OVec3 wi = orc_env_sample(env, params, u1, u2, &lightPdf);
    if (lightPdf <= K_MIN_PDF || occluded(p, n, wi, s)) {
        return 0;
    }
    const float cosT = fmax2(0.0f, vdot(n, wi));
    if (cosT <= 0.0f) {
        return 0;
    }
    const OBrdfCtx rgb = make_ctx(n, wo, m, eta, 0);
    const float mis = orc_mis_balance(lightPdf, orc_brdf_pdf(&rgb, wi));
    spec_light_term(v, wi, orc_env_eval(env, params, wi), mis * cosT / lightPdf, out4);
    return 1;
}

static int spec_emissive_nee(const SpecVertex* v, OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta,
                             const OScene* s, ORng* rng, float* out4, int last)
{
    out4[0] = out4[1] = out4[2] = out4[3] = 0.0f;
    if (s == NULL || s->emissiveCount == 0) {
        return 0;
    }
    const float uTri = orc_rand01(rng);
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    if (gLightMode == 1 && last) {
        return 0;
    }
    OVec3 lp, ln, lr;
    float areaPdf = 0.0f;
    uint32_t lightTri = 0u;
    if (!orc_light_sample_emissive_tri(s, uTri, u1, u2, &lp, &ln, &lr, &areaPdf, &lightTri)) {
        return 0;
    }
    const OVec3 toL = vsub(lp, p);
    const float d2 = vdot(toL, toL);
    if (d2 <= K_MIN_PDF) {
        return 0;
    }
    const OVec3 wi = vnorm(toL);
    const float cosT = fmax2(0.0f, vdot(n, wi));
    const float cosL = fmax2(0.0f, vdot(ln, vscale(wi, -1.0f)));
    if (cosT <= 0.0f || cosL <= 0.0f ||
        (gLightMode == 1 ? occluded_toward(p, wi, lp, s) : occluded_from(p, wi, s))) {
        return 0;
    }
    const float lightPdf = areaPdf * d2 / cosL;
    if (lightPdf <= K_MIN_PDF) {
        return 0;
    }
    const OBrdfCtx rgb = make_ctx(n, wo, m, eta, 0);
    const float mis = orc_mis_balance(lightPdf, orc_brdf_pdf(&rgb, wi));
    /* a light with binned emission: its raw emission x the bins (as the emissive list's r, g, b) */
    const float* lsp = material_spectrum(s, lightTri);
    if (lsp != NULL && lsp[ORC_SPECTRUM_BINS] >= 0.0f) {
        const OMaterial lm = orc_fetch_material(s, lightTri);
        float Le[4];
        for (int k = 0; k < 4; ++k) {
            Le[k] = lsp[ORC_SPECTRUM_BINS + v->bin[k]] * lm.emission;
        }
        spec_light_term4(v, wi, Le, mis * cosT / lightPdf, out4);
    } else {
        spec_light_term(v, wi, lr, mis * cosT / lightPdf, out4);
    }
    return 1;
}

void orc_trace_path_spectral(OVec3 ro, OVec3 rd, const OScene* s, const OParams* params, const OEnv* env,
                             ORng* rng, const OSpecModel* model, const float* lambda4, float* L4)
{
    for (int k = 0; k < 4; ++k) {
        L4[k] = 0.0f;
    }
    const OHit first = trace(ro, rd, s, 0.0f, K_RAY_TMAX, 1);
    if (!first.hit) {
        const OVec3 bg = orc_env_eval(env, params, rd);
        for (int k = 0; k < 4; ++k) {
            L4[k] = orc_rgb2spec_radiance(model, bg.x, bg.y, bg.z, lambda4[k]);
        }
        return;
    }
    int maxDepth = 32;
    int rrMin = 3;
    int flags = 0;
    if (params != NULL) {
        if (params->maxPathDepth > 0) {
            maxDepth = params->maxPathDepth;
        } else if (params->maxPathDepth == 0) {
            maxDepth = 256;
        }
        if (params->russianRouletteMinDepth >= 0) {
            rrMin = params->russianRouletteMinDepth;
        }
        flags = params->brdfDebugFlags;
    }
    float T[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    OVec3 origin = ro;
    OVec3 dir = rd;
    OHit hit = first;
    float eta = 1.0f;
    float prevPdf = 0.0f;
    int reach = 0;
    SpecVertex v;
    v.model = model;
    for (int k = 0; k < 4; ++k) {
        v.lambda[k] = lambda4[k];
        v.bin[k] = spectrum_bin(model, lambda4[k]);
    }

    for (int depth = 0; depth < maxDepth; ++depth) {
        const OVec3 p = vray(origin, dir, hit.t);
        const OVec3 n = hit.normal;
        const OVec3 wo = vscale(dir, -1.0f);
        const OMaterial src = orc_fetch_material(s, hit.triangleIndex);
        const OMaterial m = orc_resolve_material(&src, hit.uv, s);
        const float* sp = material_spectrum(s, hit.triangleIndex);
        spec_vertex(&v, &m, sp, n, wo, eta, 0);

        const int weighted = gLightMode == 1 && depth > 0 && m.emission > 0.0f;
        const float w = weighted ? emission_weight(s, hit.triangleIndex, dir, hit.t, prevPdf, reach) : 1.0f;
        float Le4[4];
        spec_emission(&v, &m, sp, Le4);
        for (int k = 0; k < 4; ++k) {
            L4[k] = L4[k] + T[k] * (weighted ? Le4[k] * w : Le4[k]);
        }
        float c[4];
        spec_env_nee(&v, p, n, wo, &m, eta, s, env, params, rng, c);
        for (int k = 0; k < 4; ++k) {
            L4[k] = L4[k] + T[k] * c[k];
        }
        spec_emissive_nee(&v, p, n, wo, &m, eta, s, rng, c, depth + 1 == maxDepth);
        for (int k = 0; k < 4; ++k) {
            L4[k] = L4[k] + T[k] * c[k];
        }

        const OBrdfCtx ctx = make_ctx(n, wo, &m, eta, flags);
        const float u1 = orc_rand01(rng);
        const float u2 = orc_rand01(rng);
        const OSample smp = orc_brdf_sample(&ctx, u1, u2);
        if (!smp.valid || smp.pdf <= K_MIN_PDF) {
            break;
        }
        for (int k = 0; k < 4; ++k) {
            OBrdfCtx ck = v.ctx[k];
            ck.debugFlags = flags;
            const float f = orc_brdf_eval(&ck, smp.direction).x;
            T[k] = orc_apply_throughput(v3(T[k], T[k], T[k]), &ck, &smp, v3(f, f, f)).x;
        }
        /* tier-B walk (PathIntegratorRandCore.h:169-204) with the RGB radius */
        if (smp.subsurfaceScatter && smp.subsurfaceInternalSteps > 0) {
            const float maxR = orc_max_scatter_radius(&m);
            if (maxR > 1.0e-5f) {
                for (int step = 0; step < smp.subsurfaceInternalSteps; ++step) {
                    (void)orc_rand01(rng);
                    (void)orc_rand01(rng);
                    const float u3 = orc_rand01(rng);
                    const float mfp = maxR * (-orc_logf(fmax2(1.0e-8f, u3)));
                    const float surv = orc_expf(-mfp / maxR);
                    for (int k = 0; k < 4; ++k) {
                        T[k] = T[k] * (v.grey[k].r * surv);
                    }
                }
                const float ed = vlen(smp.subsurfaceExitOffset);
                if (ed > 0.0f) {
                    const float a = orc_expf(-ed / maxR);
                    for (int k = 0; k < 4; ++k) {
                        T[k] = T[k] * a;
                    }
                }
            }
        }
        if ((flags & (1 << 4)) != 0) {
            const OVec3 tint = smp.transmitted ? v3(0.2f, 1.0f, 0.2f) : v3(1.0f, 0.2f, 0.2f);
            for (int k = 0; k < 4; ++k) {
                T[k] = T[k] * orc_rgb2spec_reflectance(model, tint.x, tint.y, tint.z, v.lambda[k]);
            }
        }
        const float tl = 0.25f * (((T[0] + T[1]) + T[2]) + T[3]);
        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
            break;
        }
        if (depth >= rrMin) {
            const float surv = fmin2(0.95f, fmax2(tl, 0.05f));
            if (orc_rand01(rng) > surv) {
                break;
            }
            for (int k = 0; k < 4; ++k) {
                T[k] = T[k] * (1.0f / surv);
            }
        }
        eta = smp.nextMediumEta;
        prevPdf = smp.pdf;
        reach = vdot(n, smp.direction) > 0.0f;
        const float bias = fmax2(K_SHADOW_BIAS, K_REL_SHADOW_BIAS * hit.t);
        const float sign = vdot(n, smp.direction) >= 0.0f ? 1.0f : -1.0f;
        origin = vadd(p, vscale(n, sign * bias));
        dir = smp.direction;
        hit = trace(origin, dir, s, bias, K_RAY_TMAX, 1);
        if (!hit.hit) {
            const OVec3 bg = orc_env_eval(env, params, dir);
            const float lpdf = orc_env_pdf(env, params, dir);
            const float mis = orc_mis_balance(smp.pdf, lpdf);
            for (int k = 0; k < 4; ++k) {
                const float envL = orc_rgb2spec_radiance(model, bg.x, bg.y, bg.z, v.lambda[k]);
                L4[k] = L4[k] + T[k] * (envL * mis);
            }
            break;
        }
    }
}

/* Spectral model used by orc_pixel_radiance (NULL: RGB). */
static const OSpecModel* gSpecModel = NULL;

void orc_set_spectral(const OSpecModel* model) { gSpecModel = model; }

/* PathTracer.cu:49-67 */
OVec3 orc_clamp_firefly(OVec3 r)
{
    if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    r = v3(fmax2(0.0f, r.x), fmax2(0.0f, r.y), fmax2(0.0f, r.z));
    const float lum = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    if (lum <= 100.0f) {
        return r;
    }
    const float k = 100.0f / lum;
    return v3(r.x * k, r.y * k, r.z * k);
}

/* PathTracer.cu:69-105 */
OVec3 orc_pixel_radiance(int idx, int x, int y, int width, int height, int jitter, int sampleIndex,
                         const OCamera* cam, const OScene* s, const OEnv* env, const OParams* params,
                         uint32_t globalSeed)
{
    ORng rng;
    orc_rng_init(&rng, idx, sampleIndex, globalSeed);
    const float ju = jitter ? orc_rand01(&rng) : 0.5f;
    const float jv = jitter ? orc_rand01(&rng) : 0.5f;
    const float u = ((float)x + ju) / (float)width;
    const float v = ((float)y + jv) / (float)height;
    const float l1 = orc_rand01(&rng);
    const float l2 = orc_rand01(&rng);
    OVec3 ro, rd;
    orc_camera_ray(cam, u, v, l1, l2, &ro, &rd);
    if (gSpecModel != NULL) {
        float lambda[4], L4[4];
        orc_hero_lambdas(gSpecModel, orc_rand01(&rng), lambda);
        orc_trace_path_spectral(ro, rd, s, params, env, &rng, gSpecModel, lambda, L4);
        return orc_clamp_firefly_spectral(orc_hero_to_rgb(gSpecModel, L4, lambda));
    }
    return orc_clamp_firefly(orc_trace_path(ro, rd, s, params, env, &rng));
}

/* PathTracer.cu:107-130 */
int orc_is_converged(uint32_t n, float mean, float m2, int minSamples, float thr)
{
    if (n < (uint32_t)minSamples) {
        return 0;
    }
    if (n == 0) {
        return 0;
    }
    if (mean < 1.0e-5f) {
        return 0;
    }
    const float var = m2 / (float)n;
    const float sd = sqrtf(var);
    const float denom = fmaxf(mean, 1.0e-4f);
    return (sd / denom) < thr;
}

/* PathTracer.cu:275-303 */
void orc_clear_accumulator(float* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, int* active,
                           int* activeCount, int width, int height)
{
    const int n = width * height;
    for (int i = 0; i < n; ++i) {
        acc[4 * i + 0] = 0.0f;
        acc[4 * i + 1] = 0.0f;
        acc[4 * i + 2] = 0.0f;
        acc[4 * i + 3] = 1.0f;
        counts[i] = 0;
        mean[i] = 0.0f;
        m2[i] = 0.0f;
        conv[i] = 0;
        active[i] = i;
    }
    if (activeCount) {
        *activeCount = n;
    }
}

/* PathTracer.cu:186-273: the body run for one active pixel. */
static void adaptive_one(float* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, int idx, int width,
                         int height, const OCamera* cam, const OScene* s, const OEnv* env, const OParams* params,
                         uint32_t seed)
{
    if (idx < 0 || conv[idx] != 0) {
        return;
    }
    const int x = idx % width;
    const int y = idx / width;
    const uint32_t n = counts[idx];
    const OVec3 r = orc_pixel_radiance(idx, x, y, width, height, 1, (int)n, cam, s, env, params, seed);
    const float sl = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    const float pm = mean[idx];
    const float pM2 = m2[idx];
    const float delta = sl - pm;
    const float nm = pm + delta / (float)(n + 1);
    const float nM2 = pM2 + delta * (sl - nm);
    mean[idx] = nm;
    m2[idx] = nM2;
    const float inv = 1.0f / (float)(n + 1);
    float* a = &acc[4 * (size_t)idx];
    a[0] = (a[0] * (float)n + r.x) * inv;
    a[1] = (a[1] * (float)n + r.y) * inv;
    a[2] = (a[2] * (float)n + r.z) * inv;
    a[3] = 1.0f;
    const uint32_t nc = n + 1;
    counts[idx] = nc;
    const int maxS = params->maxSamplesPerPixel;
    if (maxS > 0 && nc >= (uint32_t)maxS) {
        conv[idx] = 1;
        return;
    }
    if ((nc % 4u) != 0) {
        return;
    }
    if (orc_is_converged(nc, nm, nM2, params->minSamples, params->relativeErrorThreshold)) {
        conv[idx] = 1;
    }
}

typedef struct {
    float* acc;
    uint32_t* counts;
    float* mean;
    float* m2;
    uint8_t* conv;
    const int* active;
    int begin, end;
    int width, height;
    const OCamera* cam;
    const OScene* s;
    const OEnv* env;
    const OParams* params;
    uint32_t seed;
    int preview;
    OTraceStats* stats;
} Job;

static void* job_run(void* arg)
{
    Job* j = (Job*)arg;
    if (j->stats) {
        orc_stats_begin(j->stats);
    }
    for (int t = j->begin; t < j->end; ++t) {
        if (j->preview) {
            const int x = t % j->width;
            const int y = t / j->width;
            const OVec3 r = orc_pixel_radiance(t, x, y, j->width, j->height, 0, 0, j->cam, j->s, j->env, j->params,
                                               j->seed);
            j->acc[4 * (size_t)t + 0] = r.x;
            j->acc[4 * (size_t)t + 1] = r.y;
            j->acc[4 * (size_t)t + 2] = r.z;
            j->acc[4 * (size_t)t + 3] = 1.0f;
        } else {
            adaptive_one(j->acc, j->counts, j->mean, j->m2, j->conv, j->active[t], j->width, j->height, j->cam, j->s,
                         j->env, j->params, j->seed);
        }
    }
    if (j->stats) {
        orc_stats_end();
    }
    return NULL;
}

static void run_jobs(Job proto, int count, int threads)
{
    if (threads < 1) {
        threads = 1;
    }
    if (threads > 256) {
        threads = 256;
    }
    if (threads == 1 || count < 64) {
        proto.begin = 0;
        proto.end = count;
        job_run(&proto);
        return;
    }
    pthread_t tid[256];
    Job jobs[256];
    /* interleaved chunks for load balance */
    const int chunk = 64;
    int next = 0;
    (void)next;
    for (int t = 0; t < threads; ++t) {
        jobs[t] = proto;
        const long long b = (long long)count * t / threads;
        const long long e = (long long)count * (t + 1) / threads;
        jobs[t].begin = (int)b;
        jobs[t].end = (int)e;
    }
    (void)chunk;
    for (int t = 0; t < threads; ++t) {
        pthread_create(&tid[t], NULL, job_run, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
    }
}

void orc_adaptive_sample(float* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, const int* active,
                         int activeCount, int width, int height, const OCamera* cam, const OScene* s, const OEnv* env,
                         const OParams* params, uint32_t seed, int threads)
{
    Job j;
    memset(&j, 0, sizeof(j));
    j.acc = acc;
    j.counts = counts;
    j.mean = mean;
    j.m2 = m2;
    j.conv = conv;
    j.active = active;
    j.width = width;
    j.height = height;
    j.cam = cam;
    j.s = s;
    j.env = env;
    j.params = params;
    j.seed = seed;
    run_jobs(j, activeCount, threads);
}

/* PathTracer.cu:132-184 (preview pass: no jitter, sampleIndex 0, overwrite) */
void orc_preview_sample(float* acc, int width, int height, const OCamera* cam, const OScene* s, const OEnv* env,
                        const OParams* params, uint32_t seed, int threads)
{
    Job j;
    memset(&j, 0, sizeof(j));
    j.acc = acc;
    j.width = width;
    j.height = height;
    j.cam = cam;
    j.s = s;
    j.env = env;
    j.params = params;
    j.seed = seed;
    j.preview = 1;
    run_jobs(j, width * height, threads);
}

/* PathTracer.cu:372-411, written in input order (the reference's atomicAdd order is
 * unspecified; per-pixel results do not depend on it). Returns the new count. */
int orc_compact_active(const int* in, int* out, const uint8_t* conv, const uint32_t* counts, int maxSpp, int inCount,
                       int width, int useRegion, int rx0, int ry0, int rx1, int ry1)
{
    int n = 0;
    for (int t = 0; t < inCount; ++t) {
        const int p = in[t];
        if (conv[p] != 0) {
            continue;
        }
        if (maxSpp > 0 && counts[p] >= (uint32_t)maxSpp) {
            continue;
        }
        if (useRegion) {
            const int x = p % width;
            const int y = p / width;
            if (x < rx0 || x > rx1 || y < ry0 || y > ry1) {
                continue;
            }
        }
        out[n++] = p;
    }
    return n;
}

/* PathTracer.cpp:2220-2543 schedule without preview levels: globalSeed = iter + 1;
 * compaction when (iter - firstIter + 1) % 4 == 0; stop when no pixel is active. */
int orc_render_adaptive(float* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, int width, int height,
                        const OCamera* cam, const OScene* s, const OEnv* env, const OParams* params, int firstIter,
                        int maxPasses, int threads)
{
    const int n = width * height;
    int* active = (int*)malloc(sizeof(int) * (size_t)n);
    int* scratch = (int*)malloc(sizeof(int) * (size_t)n);
    int activeCount = 0;
    orc_clear_accumulator(acc, counts, mean, m2, conv, active, &activeCount, width, height);
    int passes = 0;
    for (int iter = firstIter; activeCount > 0 && passes < maxPasses; ++iter) {
        orc_adaptive_sample(acc, counts, mean, m2, conv, active, activeCount, width, height, cam, s, env, params,
                            (uint32_t)iter + 1u, threads);
        ++passes;
        const int wave = iter - firstIter + 1;
        if (wave % 4 == 0) {
            const int nc = orc_compact_active(active, scratch, conv, counts, params->maxSamplesPerPixel, activeCount,
                                              width, 0, 0, 0, 0, 0);
            int* tmp = active;
            active = scratch;
            scratch = tmp;
            activeCount = nc;
        }
    }
    free(active);
    free(scratch);
    return passes;
}

/* PathTracer.cu:421-512: exposure, Reinhard x/(1+x), clamp, x255 truncation. */
void orc_copy_to_pbo(const float* acc, const uint32_t* counts, const uint8_t* conv, uint8_t* rgba, int outW, int outH,
                     int srcW, int srcH, int downscale, int useAcc, const OParams* params, float exposure)
{
    for (int y = 0; y < outH; ++y) {
        for (int x = 0; x < outW; ++x) {
            int si = y * outW + x;
            if (downscale > 1 && srcW > 0 && srcH > 0) {
                const int ds = downscale < 1 ? 1 : downscale;
                const int px = x / ds;
                const int py = y / ds;
                const int cx = px < srcW ? px : srcW - 1;
                const int cy = py < srcH ? py : srcH - 1;
                si = cy * srcW + cx;
            }
            float r, g, b;
            uint8_t* o = &rgba[4 * ((size_t)y * outW + x)];
            if (params != NULL && params->debugOverlayMode == 2) {
                if (useAcc && conv != NULL) {
                    if (conv[si] != 0) {
                        r = 0.0f;
                        g = 0.35f;
                        b = 0.0f;
                    } else {
                        r = 1.0f;
                        g = 0.0f;
                        b = 0.0f;
                    }
                } else {
                    r = 1.0f;
                    g = 0.0f;
                    b = 0.0f;
                }
            } else {
                if (params == NULL) {
                    r = g = b = 10.0f / 255.0f;
                } else {
                    r = params->backgroundR;
                    g = params->backgroundG;
                    b = params->backgroundB;
                }
                if (acc != NULL) {
                    if (useAcc) {
                        if (counts != NULL && counts[si] > 0) {
                            r = acc[4 * (size_t)si];
                            g = acc[4 * (size_t)si + 1];
                            b = acc[4 * (size_t)si + 2];
                        }
                    } else {
                        r = acc[4 * (size_t)si];
                        g = acc[4 * (size_t)si + 1];
                        b = acc[4 * (size_t)si + 2];
                    }
                }
                r = r * exposure;
                g = g * exposure;
                b = b * exposure;
                r = r / (1.0f + r);
                g = g / (1.0f + g);
                b = b / (1.0f + b);
            }
            o[0] = (uint8_t)(fminf(fmaxf(r, 0.0f), 1.0f) * 255.0f);
            o[1] = (uint8_t)(fminf(fmaxf(g, 0.0f), 1.0f) * 255.0f);
            o[2] = (uint8_t)(fminf(fmaxf(b, 0.0f), 1.0f) * 255.0f);
            o[3] = 255;
        }
    }
}

/* PathTracer.cu:514-548: pinhole primary ray (CameraDevice.cuh:82-86), closest hit in
 * [0.001, 1e30), fractional UV as red/green, background (black without params) on a miss,
 * same clamp and x255 truncation as the display path. */
void orc_uv_debug(const OCamera* cam, const OScene* scene, uint8_t* rgba, int width, int height,
                  const OParams* params)
{
    OCamera pin = *cam;
    pin.focusValid = 0;
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            const float u = ((float)x + 0.5f) / (float)width;
            const float v = ((float)y + 0.5f) / (float)height;
            OVec3 ro, rd;
            orc_camera_ray(&pin, u, v, 0.0f, 0.0f, &ro, &rd);
            float r = 0.0f, g = 0.0f, b = 0.0f;
            if (params != NULL) {
                r = params->backgroundR;
                g = params->backgroundG;
                b = params->backgroundB;
            }
            const OHit h = orc_trace_ray(ro, rd, scene, 0.001f, 1.0e30f);
            if (h.hit) {
                r = h.uv.x - floorf(h.uv.x);
                g = h.uv.y - floorf(h.uv.y);
                b = 0.0f;
            }
            uint8_t* o = &rgba[4 * ((size_t)y * width + x)];
            o[0] = (uint8_t)(fminf(fmaxf(r, 0.0f), 1.0f) * 255.0f);
            o[1] = (uint8_t)(fminf(fmaxf(g, 0.0f), 1.0f) * 255.0f);
            o[2] = (uint8_t)(fminf(fmaxf(b, 0.0f), 1.0f) * 255.0f);
            o[3] = 255;
        }
    }
}

/* PathTracer.cu:305-344: reset the pixels of [minX,maxX]x[minY,maxY] (acc w = 1) and list them
 * row-major as the active set. Returns the region's pixel count (0 for an empty region). */
int orc_clear_region(float* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, int* active, int width,
                     int minX, int minY, int maxX, int maxY)
{
    const int rw = maxX - minX + 1;
    const int rh = maxY - minY + 1;
    if (rw <= 0 || rh <= 0) {
        return 0;
    }
    for (int t = 0; t < rw * rh; ++t) {
        const int idx = (minY + t / rw) * width + (minX + t % rw);
        acc[4 * (size_t)idx + 0] = 0.0f;
        acc[4 * (size_t)idx + 1] = 0.0f;
        acc[4 * (size_t)idx + 2] = 0.0f;
        acc[4 * (size_t)idx + 3] = 1.0f;
        counts[idx] = 0;
        mean[idx] = 0.0f;
        m2[idx] = 0.0f;
        conv[idx] = 0;
        active[t] = idx;
    }
    return rw * rh;
}

/* PathTracer.cu:346-370: every pixel neither converged nor at maxSamplesPerPixel, ascending. */
int orc_rebuild_active(const uint8_t* conv, const uint32_t* counts, int* out, int pixelCount, int maxSpp)
{
    int n = 0;
    for (int p = 0; p < pixelCount; ++p) {
        if (conv[p] != 0) {
            continue;
        }
        if (maxSpp > 0 && counts[p] >= (uint32_t)maxSpp) {
            continue;
        }
        out[n++] = p;
    }
    return n;
}

/* ---- differentiable rendering: detached estimator and adjoint path replay ------------------
 *
 * Parameters per material k: grad[7k + 0..6] = base r, g, b, roughness, metallic, specular,
 * emission. Derivatives use "detached" sampling: directions, pdfs, MIS weights and Russian
 * roulette are those of the primal parameters, only the evaluated values (emission, BSDF values
 * in throughputs and next-event terms) depend on the parameters. This is an unbiased estimator
 * of the derivative of the expected image (sampling with any fixed density is unbiased), and it
 * is exactly the derivative of orc_trace_path_ad's radiance with respect to evalMats at
 * evalMats = the scene's materials -- which is what the finite-difference tests use.
 * BSDF derivatives cover the opaque Principled lobes (diffuse + specular); a material with
 * transmission or subsurface weight contributes emission derivatives only. Parameters are live
 * where the forward is differentiable: untextured, inside the open clamp interval (0, 1) for
 * roughness/metallic/specular, emission > 0. */

typedef struct {
    int rgb, rough, metal, spec, emis, opaque;
    int onRough; /* Oren-Nayar sigma follows roughness (unresolved diffuseRoughness < 0) */
} LiveMask;

static int untextured(uint32_t ti, const OScene* s)
{
    return ti == 0u || s == NULL || s->textures == NULL || ti >= s->textureCount;
}

static LiveMask live_mask(const OMaterial* raw, const OMaterial* resolved, const OScene* s)
{
    LiveMask l;
    l.rgb = untextured(raw->albedoTex, s);
    l.rough = untextured(raw->roughnessTex, s) && raw->roughness > 0.0f && raw->roughness < 1.0f;
    l.metal = untextured(raw->metallicTex, s) && raw->metallic > 0.0f && raw->metallic < 1.0f;
    l.spec = raw->specular > 0.0f && raw->specular < 1.0f;
    l.emis = untextured(raw->emissionTex, s) && raw->emission > 0.0f;
    l.opaque = clamp01f(resolved->transmission) == 0.0f && clamp01f(resolved->subsurface) == 0.0f;
    l.onRough = raw->diffuseRoughness < 0.0f;
    return l;
}

/* out[k] += sum_c a_c * d f_c / d theta_k, theta = (r, g, b, roughness, metallic, specular), for
 * the opaque Principled BSDF f = base wd ON / pi + F ws ec D G / max(4 cosO cosI, 1e-8) with
 * wd = (1 - m) / (2 - m), ws = 1 / (2 - m) (normalised lobe weights at transmission =
 * subsurface = 0). The geometry is computed as the forward does, the algebra in double. */
static void brdf_grad(const OBrdfCtx* c, OVec3 wi, const double a[3], LiveMask lm, double out[6])
{
    if (!lm.opaque) {
        return;
    }
    const OMaterial* m = &c->material;
    const float cosWi = fmax2(0.0f, vdot(c->normal, wi));
    const float cosWo = fmax2(0.0f, vdot(c->normal, c->wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return;
    }
    const double pi = (double)3.14159265f;
    const double base[3] = {m->r, m->g, m->b};
    const double mm = clamp01f(m->metallic);
    const double wd = (1.0 - mm) / (2.0 - mm);
    const double ws = 1.0 / (2.0 - mm);
    /* diffuse lobe */
    const float dr = fmax2(0.0f, fmin2(1.0f, m->diffuseRoughness < 0.0f ? m->roughness : m->diffuseRoughness));
    const float sigma = dr * 0.5f * 3.14159265f;
    const double on = orc_oren_nayar_factor(sigma, cosWi, cosWo, vdot(wi, c->wo));
    for (int k = 0; k < 3; ++k) {
        if (lm.rgb) {
            out[k] += a[k] * wd * on / pi;
        }
    }
    const double aBase = a[0] * base[0] + a[1] * base[1] + a[2] * base[2];
    if (lm.metal) {
        out[4] += aBase * on / pi * (-1.0 / ((2.0 - mm) * (2.0 - mm)));
    }
    if (lm.rough && lm.onRough && sigma > 1.0e-4f) {
        /* ON = A(s2) + B(s2) * geo with geo = max(0, cosPhi) sinAlpha tanBeta (independent of sigma) */
        const double s2 = (double)sigma * sigma;
        const float sinI = sqrtf(fmax2(0.0f, 1.0f - cosWi * cosWi));
        const float sinO = sqrtf(fmax2(0.0f, 1.0f - cosWo * cosWo));
        float sa = cosWi > cosWo ? sinI : sinO;
        float sb = cosWi > cosWo ? sinO : sinI;
        float cosPhi = 1.0f;
        const float d = vdot(wi, c->wo) - cosWi * cosWo;
        if (d <= 0.0f) {
            cosPhi = fmax2(-1.0f, fmin2(1.0f, d / fmax2(sa * sb, 1.0e-8f)));
        } else {
            sa = 0.0f;
            sb = 0.0f;
        }
        const double geo = (double)fmax2(0.0f, cosPhi) * sa * (sb / fmax2(cosWi, cosWo));
        const double dA = -0.5 * 0.33 / ((s2 + 0.33) * (s2 + 0.33));
        const double dB = 0.45 * 0.09 / ((s2 + 0.09) * (s2 + 0.09));
        const double dOn = (dA + dB * geo) * 2.0 * sigma * 0.5 * pi;
        out[3] += aBase * wd * dOn / pi;
    }
    /* specular lobe */
    const OVec3 h = vnorm(vadd(wi, c->wo));
    const float cosH = fmax2(0.0f, vdot(c->normal, h));
    const float cosHo = fmax2(0.0f, vdot(h, c->wo));
    if (cosHo <= 0.0f) {
        return;
    }
    const float rough = m->roughness;
    const double rc = fmax2(rough, 0.04f);
    const double alpha = rc * rc;
    const double a2 = alpha * alpha;
    const double c2 = (double)cosH * cosH;
    const double dd = c2 * (a2 - 1.0) + 1.0;
    const double D = a2 / (pi * dd * dd);
    double g1[2], dg1[2];
    const float cs[2] = {cosWo, cosWi};
    for (int q = 0; q < 2; ++q) {
        const double cc = fmax((double)cs[q] * cs[q], 1.0e-8);
        const double tan2 = (1.0 - cc) / cc;
        const double sq = sqrt(1.0 + a2 * tan2);
        g1[q] = 2.0 / (1.0 + sq);
        dg1[q] = -tan2 / ((1.0 + sq) * (1.0 + sq) * sq); /* d G1 / d a2 */
    }
    const double G = g1[0] * g1[1];
    const double ior = fmax2(m->ior, 1.0e-3f);
    const double f0s = (ior - 1.0) / (ior + 1.0);
    const double specv = clamp01f(clamp01f(m->specular));
    const double f0v = f0s * f0s * specv;
    const double t = fmax(0.0, fmin(1.0, 1.0 - (double)cosHo));
    const double t5 = t * t * t * t * t;
    double F[3];
    for (int k = 0; k < 3; ++k) {
        F[k] = (f0v + (base[k] - f0v) * mm) * (1.0 - t5) + t5;
    }
    const double cw = fmax(cosWo, 1.0e-4);
    const double ess = 1.0 - alpha * (0.3316 + 0.3237 * alpha);
    const double ems = (1.0 - ess) * (1.0 - cw);
    const double essC = fmax(ess, 1.0e-4);
    const double ec = 1.0 + ems / essC;
    const double denom = fmax(4.0 * cosWo * cosWi, 1.0e-8);
    const double S = ws * ec * D * G / denom;
    const double aF = a[0] * F[0] + a[1] * F[1] + a[2] * F[2];
    for (int k = 0; k < 3; ++k) {
        if (lm.rgb) {
            out[k] += a[k] * mm * (1.0 - t5) * S;
        }
    }
    if (lm.metal) {
        double dF = 0.0;
        for (int k = 0; k < 3; ++k) {
            dF += a[k] * (base[k] - f0v) * (1.0 - t5);
        }
        out[4] += dF * S + aF * S / (2.0 - mm);
    }
    if (lm.spec) {
        out[5] += (a[0] + a[1] + a[2]) * (1.0 - mm) * (1.0 - t5) * f0s * f0s * S;
    }
    if (lm.rough && rough > 0.04f) {
        const double dD = (dd - 2.0 * a2 * c2) / (pi * dd * dd * dd) * 2.0 * alpha;   /* dD/dalpha */
        const double dG = (dg1[0] * g1[1] + g1[0] * dg1[1]) * 2.0 * alpha;          /* dG/dalpha */
        const double dEss = -(0.3316 + 2.0 * 0.3237 * alpha);
        const double dEms = -dEss * (1.0 - cw);
        const double dEc = ess > 1.0e-4 ? (dEms * ess - ems * dEss) / (ess * ess) : dEms / 1.0e-4;
        const double dS = ws / denom * (dEc * D * G + ec * dD * G + ec * D * dG);
        out[3] += aF * dS * 2.0 * rough;
    }
}

/* The value of the BSDF brdf_grad differentiates, in double: the same float geometry (cosines,
 * half vector, Oren-Nayar angles) with the parameters' algebra in double, so central differences
 * of it in a parameter carry no float rounding. Materials with transmission or subsurface weight
 * (whose BSDF values brdf_grad leaves undifferentiated) use the float evaluation. */
static void brdf_value_f64(const OBrdfCtx* c, OVec3 wi, LiveMask lm, double f[3])
{
    f[0] = f[1] = f[2] = 0.0;
    if (!lm.opaque) {
        const OVec3 v = orc_brdf_eval(c, wi);
        f[0] = v.x;
        f[1] = v.y;
        f[2] = v.z;
        return;
    }
    const OMaterial* m = &c->material;
    const float cosWi = fmax2(0.0f, vdot(c->normal, wi));
    const float cosWo = fmax2(0.0f, vdot(c->normal, c->wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return;
    }
    const double pi = (double)3.14159265f;
    const double base[3] = {m->r, m->g, m->b};
    const double mm = clamp01f(m->metallic);
    const double wd = (1.0 - mm) / (2.0 - mm);
    const double ws = 1.0 / (2.0 - mm);
    /* diffuse lobe: Oren-Nayar A(s2) + B(s2) geo, sigma = diffuse roughness * pi / 2 */
    const float dr = fmax2(0.0f, fmin2(1.0f, m->diffuseRoughness < 0.0f ? m->roughness : m->diffuseRoughness));
    double on = 1.0;
    if (dr * 0.5f * 3.14159265f > 1.0e-4f) {
        const double sigma = (double)dr * 0.5 * pi;
        const double s2 = sigma * sigma;
        const float sinI = sqrtf(fmax2(0.0f, 1.0f - cosWi * cosWi));
        const float sinO = sqrtf(fmax2(0.0f, 1.0f - cosWo * cosWo));
        float sa = cosWi > cosWo ? sinI : sinO;
        float sb = cosWi > cosWo ? sinO : sinI;
        float cosPhi = 1.0f;
        const float d = vdot(wi, c->wo) - cosWi * cosWo;
        if (d <= 0.0f) {
            cosPhi = fmax2(-1.0f, fmin2(1.0f, d / fmax2(sa * sb, 1.0e-8f)));
        } else {
            sa = 0.0f;
            sb = 0.0f;
        }
        const double geo = (double)fmax2(0.0f, cosPhi) * sa * (sb / fmax2(cosWi, cosWo));
        on = 1.0 - 0.5 * s2 / (s2 + 0.33) + 0.45 * s2 / (s2 + 0.09) * geo;
    }
    for (int k = 0; k < 3; ++k) {
        f[k] = base[k] * wd * on / pi;
    }
    /* specular lobe: F ws ec D G / max(4 cosO cosI, 1e-8) */
    const OVec3 h = vnorm(vadd(wi, c->wo));
    const float cosH = fmax2(0.0f, vdot(c->normal, h));
    const float cosHo = fmax2(0.0f, vdot(h, c->wo));
    if (cosHo <= 0.0f) {
        return;
    }
    const double rc = fmax((double)m->roughness, 0.04);
    const double alpha = rc * rc;
    const double a2 = alpha * alpha;
    const double c2 = (double)cosH * cosH;
    const double dd = c2 * (a2 - 1.0) + 1.0;
    const double D = a2 / (pi * dd * dd);
    double G = 1.0;
    const float cs[2] = {cosWo, cosWi};
    for (int q = 0; q < 2; ++q) {
        const double cc = fmax((double)cs[q] * cs[q], 1.0e-8);
        G *= 2.0 / (1.0 + sqrt(1.0 + a2 * (1.0 - cc) / cc));
    }
    const double ior = fmax2(m->ior, 1.0e-3f);
    const double f0s = (ior - 1.0) / (ior + 1.0);
    const double f0v = f0s * f0s * clamp01f(clamp01f(m->specular));
    const double t = fmax(0.0, fmin(1.0, 1.0 - (double)cosHo));
    const double t5 = t * t * t * t * t;
    const double cw = fmax(cosWo, 1.0e-4);
    const double ess = 1.0 - alpha * (0.3316 + 0.3237 * alpha);
    const double ec = 1.0 + (1.0 - ess) * (1.0 - cw) / fmax(ess, 1.0e-4);
    const double S = ws * ec * D * G / fmax(4.0 * cosWo * cosWi, 1.0e-8);
    for (int k = 0; k < 3; ++k) {
        f[k] += ((f0v + (base[k] - f0v) * mm) * (1.0 - t5) + t5) * S;
    }
}

void orc_brdf_eval_f64(const OBrdfCtx* c, OVec3 wi, const OMaterial* raw, double* f3)
{
    OMaterial none = *raw;
    none.albedoTex = none.roughnessTex = none.metallicTex = none.emissionTex = 0u;
    brdf_value_f64(c, wi, live_mask(&none, &c->material, NULL), f3);
}

void orc_brdf_grad(const OBrdfCtx* c, OVec3 wi, const double* a3, const OMaterial* raw, double* out6)
{
    OMaterial none = *raw;
    none.albedoTex = none.roughnessTex = none.metallicTex = none.emissionTex = 0u;
    const LiveMask lm = live_mask(&none, &c->material, NULL);
    brdf_grad(c, wi, a3, lm, out6);
}

/* Next-event terms of the sampling material (pdfs, MIS) as f-weights: contribution = f(wi) * w. */
static int env_nee_terms(OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta, const OScene* s, const OEnv* env,
                         const OParams* params, ORng* rng, OVec3* wiOut, OVec3* wOut)
{
    if (orc_skips_env_nee(m)) {
        return 0;
    }
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    float lightPdf = 0.0f;
    const OVec3 wi = orc_env_sample(env, params, u1, u2, &lightPdf);
    if (lightPdf <= K_MIN_PDF || occluded(p, n, wi, s)) {
        return 0;
    }
    const float cosT = fmax2(0.0f, vdot(n, wi));
    if (cosT <= 0.0f) {
        return 0;
    }
    const OBrdfCtx ctx = make_ctx(n, wo, m, eta, 0);
    const OVec3 L = orc_env_eval(env, params, wi);
    const float bpdf = orc_brdf_pdf(&ctx, wi);
    const float mis = orc_mis_balance(lightPdf, bpdf);
    *wiOut = wi;
    *wOut = vscale(L, mis * cosT / lightPdf);
    return 1;
}

static int emissive_nee_terms(OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta, const OScene* s, ORng* rng,
                              OVec3* wiOut, OVec3* wOut)
{
    if (s == NULL || s->emissiveCount == 0) {
        return 0;
    }
    const float uTri = orc_rand01(rng);
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    OVec3 lp, ln, lr;
    float areaPdf = 0.0f;
    if (!orc_light_sample_emissive(s, uTri, u1, u2, &lp, &ln, &lr, &areaPdf)) {
        return 0;
    }
    const OVec3 toL = vsub(lp, p);
    const float d2 = vdot(toL, toL);
    if (d2 <= K_MIN_PDF) {
        return 0;
    }
    const OVec3 wi = vnorm(toL);
    const float cosT = fmax2(0.0f, vdot(n, wi));
    const float cosL = fmax2(0.0f, vdot(ln, vscale(wi, -1.0f)));
    if (cosT <= 0.0f || cosL <= 0.0f || occluded_from(p, wi, s)) {
        return 0;
    }
    const float lightPdf = areaPdf * d2 / cosL;
    if (lightPdf <= K_MIN_PDF) {
        return 0;
    }
    const OBrdfCtx ctx = make_ctx(n, wo, m, eta, 0);
    const float bpdf = orc_brdf_pdf(&ctx, wi);
    const float mis = orc_mis_balance(lightPdf, bpdf);
    *wiOut = wi;
    *wOut = vscale(lr, mis * cosT / lightPdf);
    return 1;
}

/* Fetches material `tri`'s record from an alternative table (same index rules as
 * orc_fetch_material). */
static OMaterial fetch_from(const OScene* s, const OMaterial* table, uint32_t tri, int* index)
{
    *index = -1;
    if (table != NULL && s != NULL && s->triangles != NULL && tri < s->triangleCount &&
        s->triangles[tri].materialIndex < s->materialCount) {
        *index = (int)s->triangles[tri].materialIndex;
        return table[*index];
    }
    return orc_fetch_material(s, tri);
}

/* Ld += Td * f(wi) * w with the BSDF value in double. */
static void add_term_f64(double* Ld, const double Td[3], const OBrdfCtx* ctxE, OVec3 wi, LiveMask lm, OVec3 w)
{
    double fd[3];
    brdf_value_f64(ctxE, wi, lm, fd);
    Ld[0] += Td[0] * fd[0] * w.x;
    Ld[1] += Td[1] * fd[1] * w.y;
    Ld[2] += Td[2] * fd[2] * w.z;
}

/* With Ld != NULL the radiance is also evaluated in double (Ld, 3 doubles): the same sampling,
 * pdfs, MIS weights and roulette factors (float, parameter-free), with emission, BSDF values and
 * the throughput they build in double. */
static OVec3 trace_path_ad(OVec3 ro, OVec3 rd, const OScene* s, const OMaterial* evalMats, const OParams* params,
                           const OEnv* env, ORng* rng, const double* gw, OVec3 C, double* grad, double* Ld)
{
    const OHit first = trace(ro, rd, s, 0.0f, K_RAY_TMAX, 1);
    if (!first.hit) {
        const OVec3 e = orc_env_eval(env, params, rd);
        if (Ld != NULL) {
            Ld[0] = e.x;
            Ld[1] = e.y;
            Ld[2] = e.z;
        }
        return e;
    }
    int maxDepth = 32, rrMin = 3, flags = 0;
    if (params != NULL) {
        maxDepth = params->maxPathDepth > 0 ? params->maxPathDepth : (params->maxPathDepth == 0 ? 256 : 32);
        rrMin = params->russianRouletteMinDepth >= 0 ? params->russianRouletteMinDepth : 3;
        flags = params->brdfDebugFlags;
    }
    const OMaterial* evalTable = evalMats != NULL ? evalMats : (s != NULL ? s->materials : NULL);
    OVec3 L = v3(0.0f, 0.0f, 0.0f);
    OVec3 T = v3(1.0f, 1.0f, 1.0f);  /* evaluated throughput */
    OVec3 T0 = T;                    /* primal throughput: Russian roulette and termination */
    OVec3 origin = ro, dir = rd;
    OHit hit = first;
    float eta = 1.0f;
    double Td[3] = {1.0, 1.0, 1.0};
    if (Ld != NULL) {
        Ld[0] = Ld[1] = Ld[2] = 0.0;
    }
    for (int depth = 0; depth < maxDepth; ++depth) {
        const OVec3 p = vray(origin, dir, hit.t);
        const OVec3 n = hit.normal;
        const OVec3 wo = vscale(dir, -1.0f);
        const OMaterial src = orc_fetch_material(s, hit.triangleIndex);
        const OMaterial m = orc_resolve_material(&src, hit.uv, s);
        int mi = -1;
        const OMaterial srcE = fetch_from(s, evalTable, hit.triangleIndex, &mi);
        const OMaterial mE = orc_resolve_material(&srcE, hit.uv, s);
        const LiveMask lm = live_mask(&srcE, &mE, s);
        const OBrdfCtx ctxE = make_ctx(n, wo, &mE, eta, 0);
        double* g = (grad != NULL && gw != NULL && mi >= 0) ? grad + 7 * mi : NULL;

        L = vadd(L, vmul(T, v3(mE.r * mE.emission, mE.g * mE.emission, mE.b * mE.emission)));
        if (Ld != NULL) {
            const double base[3] = {mE.r, mE.g, mE.b};
            for (int k = 0; k < 3; ++k) {
                Ld[k] += Td[k] * base[k] * (double)mE.emission;
            }
        }
        if (g != NULL) {
            const double tb[3] = {(double)T.x * mE.r, (double)T.y * mE.g, (double)T.z * mE.b};
            if (lm.emis) {
                g[6] += gw[0] * tb[0] + gw[1] * tb[1] + gw[2] * tb[2];
            }
            if (lm.rgb) {
                g[0] += gw[0] * T.x * mE.emission;
                g[1] += gw[1] * T.y * mE.emission;
                g[2] += gw[2] * T.z * mE.emission;
            }
        }
        OVec3 wi, w;
        if (env_nee_terms(p, n, wo, &m, eta, s, env, params, rng, &wi, &w)) {
            L = vadd(L, vmul(T, vmul(orc_brdf_eval(&ctxE, wi), w)));
            if (Ld != NULL) {
                add_term_f64(Ld, Td, &ctxE, wi, lm, w);
            }
            if (g != NULL) {
                const double a[3] = {gw[0] * T.x * w.x, gw[1] * T.y * w.y, gw[2] * T.z * w.z};
                brdf_grad(&ctxE, wi, a, lm, g);
            }
        }
        if (emissive_nee_terms(p, n, wo, &m, eta, s, rng, &wi, &w)) {
            L = vadd(L, vmul(T, vmul(orc_brdf_eval(&ctxE, wi), w)));
            if (Ld != NULL) {
                add_term_f64(Ld, Td, &ctxE, wi, lm, w);
            }
            if (g != NULL) {
                const double a[3] = {gw[0] * T.x * w.x, gw[1] * T.y * w.y, gw[2] * T.z * w.z};
                brdf_grad(&ctxE, wi, a, lm, g);
            }
        }
        const OBrdfCtx ctx = make_ctx(n, wo, &m, eta, flags);
        const float u1 = orc_rand01(rng);
        const float u2 = orc_rand01(rng);
        const OSample smp = orc_brdf_sample(&ctx, u1, u2);
        if (!smp.valid || smp.pdf <= K_MIN_PDF) {
            break;
        }
        const OVec3 f0 = orc_brdf_eval(&ctx, smp.direction);
        const OVec3 fE = orc_brdf_eval(&ctxE, smp.direction);
        if (g != NULL) {
            /* radiance still to come after this vertex: linear in fE */
            const double R[3] = {(double)C.x - L.x, (double)C.y - L.y, (double)C.z - L.z};
            const double fe[3] = {fE.x, fE.y, fE.z};
            double a[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = fe[k] != 0.0 ? gw[k] * R[k] / fe[k] : 0.0;
            }
            brdf_grad(&ctxE, smp.direction, a, lm, g);
        }
        const OVec3 T0b = T0;
        T0 = orc_apply_throughput(T0, &ctx, &smp, f0);
        const OVec3 T0mid = T0;
        if (Ld != NULL) {
            /* the factor the throughput update multiplies the BSDF value by (cosine / pdf) */
            double fd[3];
            brdf_value_f64(&ctxE, smp.direction, lm, fd);
            const OVec3 sc = orc_apply_throughput(v3(1.0f, 1.0f, 1.0f), &ctx, &smp, v3(1.0f, 1.0f, 1.0f));
            Td[0] *= fd[0] * sc.x;
            Td[1] *= fd[1] * sc.y;
            Td[2] *= fd[2] * sc.z;
        }
        if (evalMats == NULL) {
            T = T0;
            T0 = subsurface_tier_b(T0, &m, n, &smp, rng);
            T = T0;
        } else {
            T = orc_apply_throughput(T, &ctx, &smp, fE);
            const OVec3 before = T0;
            T0 = subsurface_tier_b(T0, &m, n, &smp, rng);
            const OVec3 k3 = v3(before.x != 0.0f ? T0.x / before.x : 0.0f, before.y != 0.0f ? T0.y / before.y : 0.0f,
                                before.z != 0.0f ? T0.z / before.z : 0.0f);
            T = vmul(T, k3);
        }
        (void)T0b;
        if (Ld != NULL) {
            /* the tier-B walk's factor on the primal throughput */
            Td[0] *= T0mid.x != 0.0f ? (double)T0.x / T0mid.x : 0.0;
            Td[1] *= T0mid.y != 0.0f ? (double)T0.y / T0mid.y : 0.0;
            Td[2] *= T0mid.z != 0.0f ? (double)T0.z / T0mid.z : 0.0;
        }
        if ((flags & (1 << 4)) != 0) {
            const OVec3 tint = smp.transmitted ? v3(0.2f, 1.0f, 0.2f) : v3(1.0f, 0.2f, 0.2f);
            T0 = vmul(T0, tint);
            T = vmul(T, tint);
            if (Ld != NULL) {
                Td[0] *= tint.x;
                Td[1] *= tint.y;
                Td[2] *= tint.z;
            }
        }
        const float tl = lum3(T0);
        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
            break;
        }
        if (depth >= rrMin) {
            const float surv = fmin2(0.95f, fmax2(tl, 0.05f));
            if (orc_rand01(rng) > surv) {
                break;
            }
            T0 = vscale(T0, 1.0f / surv);
            T = vscale(T, 1.0f / surv);
            if (Ld != NULL) {
                for (int k = 0; k < 3; ++k) {
                    Td[k] *= (double)(1.0f / surv);
                }
            }
        }
        eta = smp.nextMediumEta;
        const float bias = fmax2(K_SHADOW_BIAS, K_REL_SHADOW_BIAS * hit.t);
        const float sign = vdot(n, smp.direction) >= 0.0f ? 1.0f : -1.0f;
        origin = vadd(p, vscale(n, sign * bias));
        dir = smp.direction;
        hit = trace(origin, dir, s, bias, K_RAY_TMAX, 1);
        if (!hit.hit) {
            const OVec3 envL = orc_env_eval(env, params, dir);
            const float lpdf = orc_env_pdf(env, params, dir);
            const float mis = orc_mis_balance(smp.pdf, lpdf);
            L = vadd(L, vmul(T, vscale(envL, mis)));
            if (Ld != NULL) {
                const OVec3 e = vscale(envL, mis);
                Ld[0] += Td[0] * e.x;
                Ld[1] += Td[1] * e.y;
                Ld[2] += Td[2] * e.z;
            }
            break;
        }
    }
    return L;
}

OVec3 orc_trace_path_ad(OVec3 ro, OVec3 rd, const OScene* s, const OMaterial* evalMats, const OParams* params,
                        const OEnv* env, ORng* rng, const double* gw, OVec3 C, double* grad)
{
    return trace_path_ad(ro, rd, s, evalMats, params, env, rng, gw, C, grad, NULL);
}

/* The firefly clamp (PathTracer.cu:49-67) in double. */
static void clamp_firefly_f64(const double r[3], double out[3])
{
    if (!isfinite(r[0]) || !isfinite(r[1]) || !isfinite(r[2])) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    for (int k = 0; k < 3; ++k) {
        out[k] = fmax(0.0, r[k]);
    }
    const double lum = out[0] * 0.2126 + out[1] * 0.7152 + out[2] * 0.0722;
    if (lum > 100.0) {
        for (int k = 0; k < 3; ++k) {
            out[k] *= 100.0 / lum;
        }
    }
}

/* d(clamped radiance)/d(radiance)^T g for the firefly clamp (PathTracer.cu:49-67). */
void orc_clamp_vjp(OVec3 C, const double* g, double* out)
{
    out[0] = out[1] = out[2] = 0.0;
    if (!isfinite(C.x) || !isfinite(C.y) || !isfinite(C.z)) {
        return;
    }
    const double r[3] = {fmax2(0.0f, C.x), fmax2(0.0f, C.y), fmax2(0.0f, C.z)};
    const int pos[3] = {C.x > 0.0f, C.y > 0.0f, C.z > 0.0f};
    const double wl[3] = {0.2126, 0.7152, 0.0722};
    const float lum = fmax2(0.0f, C.x) * 0.2126f + fmax2(0.0f, C.y) * 0.7152f + fmax2(0.0f, C.z) * 0.0722f;
    double gr[3] = {g[0], g[1], g[2]};
    if (lum > 100.0f) {
        const double l = lum;
        const double gdotr = g[0] * r[0] + g[1] * r[1] + g[2] * r[2];
        for (int k = 0; k < 3; ++k) {
            gr[k] = 100.0 / l * (g[k] - wl[k] * gdotr / l);
        }
    }
    for (int k = 0; k < 3; ++k) {
        out[k] = pos[k] ? gr[k] : 0.0;
    }
}

typedef struct {
    float* img;
    double* imgD;        /* FP64 detached image (orc_render_detached_f64_seeded) */
    const double* mask;  /* with imgD: pixels whose three mask values are 0 are skipped */
    const double* gradImg;
    double* grad;
    int w, h, spp, y0, y1, gradCount;
    uint32_t firstSeed;
    const OCamera* cam;
    const OScene* s;
    const OMaterial* evalMats;
    const OEnv* env;
    const OParams* params;
} AdJob;

static void ad_rows(const AdJob* j)
{
    for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < j->w; ++x) {
            const int idx = y * j->w + x;
            float acc[3] = {0.0f, 0.0f, 0.0f};
            for (int k = 0; k < j->spp; ++k) {
                ORng rng;
                orc_rng_init(&rng, idx, k, j->firstSeed + (uint32_t)k);
                const float ju = orc_rand01(&rng);
                const float jv = orc_rand01(&rng);
                const float u = ((float)x + ju) / (float)j->w;
                const float v = ((float)y + jv) / (float)j->h;
                const float l1 = orc_rand01(&rng);
                const float l2 = orc_rand01(&rng);
                OVec3 ro, rd;
                orc_camera_ray(j->cam, u, v, l1, l2, &ro, &rd);
                if (j->imgD != NULL) {
                    if (j->mask != NULL && j->mask[3 * idx] == 0.0 && j->mask[3 * idx + 1] == 0.0 &&
                        j->mask[3 * idx + 2] == 0.0) {
                        break;
                    }
                    double Ld[3], c[3];
                    (void)trace_path_ad(ro, rd, j->s, j->evalMats, j->params, j->env, &rng, NULL, v3(0, 0, 0), NULL,
                                        Ld);
                    clamp_firefly_f64(Ld, c);
                    for (int q = 0; q < 3; ++q) {
                        j->imgD[3 * idx + q] += c[q] / j->spp;
                    }
                } else if (j->gradImg == NULL) {
                    const OVec3 c = orc_clamp_firefly(
                        orc_trace_path_ad(ro, rd, j->s, j->evalMats, j->params, j->env, &rng, NULL, v3(0, 0, 0), NULL));
                    const float inv = 1.0f / (float)(k + 1);
                    acc[0] = (acc[0] * (float)k + c.x) * inv;
                    acc[1] = (acc[1] * (float)k + c.y) * inv;
                    acc[2] = (acc[2] * (float)k + c.z) * inv;
                } else if (j->gradImg[3 * idx] == 0.0 && j->gradImg[3 * idx + 1] == 0.0 &&
                           j->gradImg[3 * idx + 2] == 0.0) {
                    break; /* pixel outside the loss */
                } else {
                    ORng replay = rng;
                    const OVec3 C = orc_trace_path_ad(ro, rd, j->s, NULL, j->params, j->env, &rng, NULL, v3(0, 0, 0),
                                                      NULL);
                    double gs[3] = {j->gradImg[3 * idx] / j->spp, j->gradImg[3 * idx + 1] / j->spp,
                                    j->gradImg[3 * idx + 2] / j->spp};
                    double gw[3];
                    orc_clamp_vjp(C, gs, gw);
                    if (gw[0] != 0.0 || gw[1] != 0.0 || gw[2] != 0.0) {
                        (void)orc_trace_path_ad(ro, rd, j->s, NULL, j->params, j->env, &replay, gw, C, j->grad);
                    }
                }
            }
            if (j->img != NULL) {
                j->img[3 * idx] = acc[0];
                j->img[3 * idx + 1] = acc[1];
                j->img[3 * idx + 2] = acc[2];
            }
        }
    }
}

static void* ad_worker(void* arg)
{
    ad_rows((const AdJob*)arg);
    return NULL;
}

static void ad_run(AdJob base, int threads, double* gradOut)
{
    if (threads < 1) {
        threads = 1;
    }
    if (threads > 64) {
        threads = 64;
    }
    pthread_t tid[64];
    AdJob jobs[64];
    double* partial = NULL;
    if (gradOut != NULL) {
        partial = (double*)calloc((size_t)threads * (size_t)base.gradCount, sizeof(double));
    }
    for (int t = 0; t < threads; ++t) {
        jobs[t] = base;
        jobs[t].y0 = base.h * t / threads;
        jobs[t].y1 = base.h * (t + 1) / threads;
        jobs[t].grad = partial != NULL ? partial + (size_t)t * base.gradCount : NULL;
        pthread_create(&tid[t], NULL, ad_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
    }
    if (partial != NULL) {
        for (int t = 0; t < threads; ++t) {
            for (int k = 0; k < base.gradCount; ++k) {
                gradOut[k] += partial[(size_t)t * base.gradCount + k];
            }
        }
        free(partial);
    }
}

/* ---- differentiable hero-wavelength rendering -----------------------------------------------
 * orc_trace_path_spectral with detached sampling: directions, pdfs, MIS weights and Russian
 * roulette from the primal parameters (the scene's materials and spectra), per-wavelength values
 * (base colour, emission, BSDF values) from evalMats / evalSpectra. Parameters per material
 * (ORC_SPECTRAL_GRAD_PARAMS): the reflectance bins, the emission bins, roughness, metallic,
 * specular and the emission scale; the bins are parameters where the material has binned spectra
 * (the base colour at lambda_k is bin(lambda_k), the emitted radiance emission x bin(lambda_k)).
 * Wavelength k is the RGB model's first channel on a grey material (spec_vertex), so the RGB
 * derivative cores apply with the weight vector (a_k, 0, 0): their r-derivative is the
 * derivative with respect to the base colour at lambda_k. Light mode 0 only. */

/* wi and the per-wavelength light weights Le_k * scale of the environment NEE (0: none) */
static int spec_env_weights(const SpecVertex* v, OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta,
                            const OScene* s, const OEnv* env, const OParams* params, ORng* rng, OVec3* wiOut,
                            float* w4)
{
    if (orc_skips_env_nee(m)) {
        return 0;
    }
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    float lightPdf = 0.0f;
    const OVec3 wi = orc_env_sample(env, params, u1, u2, &lightPdf);
    if (lightPdf <= K_MIN_PDF || occluded(p, n, wi, s)) {
        return 0;
    }
    const float cosT = fmax2(0.0f, vdot(n, wi));
    if (cosT <= 0.0f) {
        return 0;
    }
    const OBrdfCtx rgb = make_ctx(n, wo, m, eta, 0);
    const float mis = orc_mis_balance(lightPdf, orc_brdf_pdf(&rgb, wi));
    const float sc = mis * cosT / lightPdf;
    const OVec3 bg = orc_env_eval(env, params, wi);
    for (int k = 0; k < 4; ++k) {
        w4[k] = orc_rgb2spec_radiance(v->model, bg.x, bg.y, bg.z, v->lambda[k]) * sc;
    }
    *wiOut = wi;
    return 1;
}

/* the same for the emissive NEE; the light's radiance comes from `values` (its materials and
 * spectra), the light choice from s */
static int spec_emissive_weights(const SpecVertex* v, OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta,
                                 const OScene* s, const OScene* values, ORng* rng, OVec3* wiOut, float* w4)
{
    if (s == NULL || s->emissiveCount == 0) {
        return 0;
    }
    const float uTri = orc_rand01(rng);
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
    OVec3 lp, ln, lr;
    float areaPdf = 0.0f;
    uint32_t lightTri = 0u;
    if (!orc_light_sample_emissive_tri(s, uTri, u1, u2, &lp, &ln, &lr, &areaPdf, &lightTri)) {
        return 0;
    }
    const OVec3 toL = vsub(lp, p);
    const float d2 = vdot(toL, toL);
    if (d2 <= K_MIN_PDF) {
        return 0;
    }
    const OVec3 wi = vnorm(toL);
    const float cosT = fmax2(0.0f, vdot(n, wi));
    const float cosL = fmax2(0.0f, vdot(ln, vscale(wi, -1.0f)));
    if (cosT <= 0.0f || cosL <= 0.0f || occluded_from(p, wi, s)) {
        return 0;
    }
    const float lightPdf = areaPdf * d2 / cosL;
    if (lightPdf <= K_MIN_PDF) {
        return 0;
    }
    const OBrdfCtx rgb = make_ctx(n, wo, m, eta, 0);
    const float mis = orc_mis_balance(lightPdf, orc_brdf_pdf(&rgb, wi));
    const float sc = mis * cosT / lightPdf;
    const OMaterial lm = orc_fetch_material(values, lightTri);
    const float* lsp = material_spectrum(values, lightTri);
    for (int k = 0; k < 4; ++k) {
        const float Le = lsp != NULL && lsp[ORC_SPECTRUM_BINS] >= 0.0f
                             ? lsp[ORC_SPECTRUM_BINS + v->bin[k]] * lm.emission
                             : orc_rgb2spec_radiance(v->model, lm.r * lm.emission, lm.g * lm.emission,
                                                     lm.b * lm.emission, v->lambda[k]);
        w4[k] = Le * sc;
    }
    *wiOut = wi;
    return 1;
}

/* g += the per-wavelength BSDF derivative with weights a4 at the grey contexts of ve */
static void spec_brdf_grad(const SpecVertex* ve, OVec3 wi, const double* a4, LiveMask lm, double* g)
{
    for (int k = 0; k < 4; ++k) {
        if (a4[k] == 0.0) {
            continue;
        }
        const double a[3] = {a4[k], 0.0, 0.0};
        double out[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        brdf_grad(&ve->ctx[k], wi, a, lm, out);
        g[ve->bin[k]] += out[0];
        g[2 * ORC_SPECTRUM_BINS + 0] += out[3];
        g[2 * ORC_SPECTRUM_BINS + 1] += out[4];
        g[2 * ORC_SPECTRUM_BINS + 2] += out[5];
    }
}

/* L4 (float) and, with Ld4, the same radiance in double; with gw4 (dLoss/dL_k) and C4 (the primal
 * radiance at the four wavelengths) the replay accumulates grad. */
static void trace_path_ad_hero(OVec3 ro, OVec3 rd, const OScene* s, const OScene* values, const OParams* params,
                               const OEnv* env, ORng* rng, const OSpecModel* model, const float* lambda4,
                               const double* gw4, const float* C4, double* grad, float* L4, double* Ld4)
{
    for (int k = 0; k < 4; ++k) {
        L4[k] = 0.0f;
        if (Ld4 != NULL) {
            Ld4[k] = 0.0;
        }
    }
    const OHit first = trace(ro, rd, s, 0.0f, K_RAY_TMAX, 1);
    if (!first.hit) {
        const OVec3 bg = orc_env_eval(env, params, rd);
        for (int k = 0; k < 4; ++k) {
            L4[k] = orc_rgb2spec_radiance(model, bg.x, bg.y, bg.z, lambda4[k]);
            if (Ld4 != NULL) {
                Ld4[k] = L4[k];
            }
        }
        return;
    }
    int maxDepth = 32, rrMin = 3, flags = 0;
    if (params != NULL) {
        maxDepth = params->maxPathDepth > 0 ? params->maxPathDepth : (params->maxPathDepth == 0 ? 256 : 32);
        rrMin = params->russianRouletteMinDepth >= 0 ? params->russianRouletteMinDepth : 3;
        flags = params->brdfDebugFlags;
    }
    float T[4] = {1.0f, 1.0f, 1.0f, 1.0f};   /* evaluated throughput */
    float T0[4] = {1.0f, 1.0f, 1.0f, 1.0f};  /* primal throughput: Russian roulette and termination */
    double Td[4] = {1.0, 1.0, 1.0, 1.0};
    OVec3 origin = ro, dir = rd;
    OHit hit = first;
    float eta = 1.0f;
    SpecVertex v, ve;
    v.model = ve.model = model;
    for (int k = 0; k < 4; ++k) {
        v.lambda[k] = ve.lambda[k] = lambda4[k];
        v.bin[k] = ve.bin[k] = spectrum_bin(model, lambda4[k]);
    }
    for (int depth = 0; depth < maxDepth; ++depth) {
        const OVec3 p = vray(origin, dir, hit.t);
        const OVec3 n = hit.normal;
        const OVec3 wo = vscale(dir, -1.0f);
        const OMaterial src = orc_fetch_material(s, hit.triangleIndex);
        const OMaterial m = orc_resolve_material(&src, hit.uv, s);
        int mi = -1;
        const OMaterial srcE = fetch_from(s, values->materials, hit.triangleIndex, &mi);
        const OMaterial mE = orc_resolve_material(&srcE, hit.uv, s);
        const float* sp = material_spectrum(s, hit.triangleIndex);
        const float* spE = material_spectrum(values, hit.triangleIndex);
        spec_vertex(&v, &m, sp, n, wo, eta, 0);
        spec_vertex(&ve, &mE, spE, n, wo, eta, 0);
        LiveMask lm = live_mask(&srcE, &mE, s);
        lm.rgb = spE != NULL && spE[0] >= 0.0f;
        double* g = (grad != NULL && gw4 != NULL && mi >= 0) ? grad + ORC_SPECTRAL_GRAD_PARAMS * mi : NULL;

        /* emission */
        float Le[4];
        spec_emission(&ve, &mE, spE, Le);
        const int binnedE = spE != NULL && spE[ORC_SPECTRUM_BINS] >= 0.0f;
        for (int k = 0; k < 4; ++k) {
            L4[k] = L4[k] + T[k] * Le[k];
            if (Ld4 != NULL) {
                Ld4[k] += Td[k] * (binnedE ? (double)spE[ORC_SPECTRUM_BINS + ve.bin[k]] * mE.emission : (double)Le[k]);
            }
            if (g != NULL && binnedE) {
                g[ORC_SPECTRUM_BINS + ve.bin[k]] += gw4[k] * T[k] * mE.emission;
                if (lm.emis) {
                    g[ORC_SPECTRAL_GRAD_PARAMS - 1] += gw4[k] * T[k] * spE[ORC_SPECTRUM_BINS + ve.bin[k]];
                }
            }
        }
        /* next-event terms: sampling with the primal vertex, values with the evaluated one */
        OVec3 wi;
        float w4[4];
        for (int q = 0; q < 2; ++q) {
            const int ok = q == 0 ? spec_env_weights(&v, p, n, wo, &m, eta, s, env, params, rng, &wi, w4)
                                  : spec_emissive_weights(&v, p, n, wo, &m, eta, s, values, rng, &wi, w4);
            if (!ok) {
                continue;
            }
            double a4[4];
            for (int k = 0; k < 4; ++k) {
                const float f = orc_brdf_eval(&ve.ctx[k], wi).x;
                L4[k] = L4[k] + T[k] * (f * w4[k]);
                if (Ld4 != NULL) {
                    double fd[3];
                    brdf_value_f64(&ve.ctx[k], wi, lm, fd);
                    Ld4[k] += Td[k] * fd[0] * w4[k];
                }
                a4[k] = g != NULL ? gw4[k] * T[k] * w4[k] : 0.0;
            }
            if (g != NULL) {
                spec_brdf_grad(&ve, wi, a4, lm, g);
            }
        }
        /* BSDF sampling with the primal RGB context */
        const OBrdfCtx ctx = make_ctx(n, wo, &m, eta, flags);
        const float u1 = orc_rand01(rng);
        const float u2 = orc_rand01(rng);
        const OSample smp = orc_brdf_sample(&ctx, u1, u2);
        if (!smp.valid || smp.pdf <= K_MIN_PDF) {
            break;
        }
        double a4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < 4; ++k) {
            OBrdfCtx ck = v.ctx[k];
            ck.debugFlags = flags;
            OBrdfCtx cke = ve.ctx[k];
            cke.debugFlags = flags;
            const float f0 = orc_brdf_eval(&ck, smp.direction).x;
            const float fE = orc_brdf_eval(&cke, smp.direction).x;
            if (g != NULL && fE != 0.0f) {
                /* radiance still to come after this vertex: linear in fE */
                a4[k] = gw4[k] * ((double)C4[k] - L4[k]) / fE;
            }
            if (Ld4 != NULL) {
                double fd[3];
                brdf_value_f64(&ve.ctx[k], smp.direction, lm, fd);
                Td[k] *= fd[0] * orc_apply_throughput(v3(1.0f, 1.0f, 1.0f), &cke, &smp, v3(1.0f, 1.0f, 1.0f)).x;
            }
            T0[k] = orc_apply_throughput(v3(T0[k], T0[k], T0[k]), &ck, &smp, v3(f0, f0, f0)).x;
            T[k] = orc_apply_throughput(v3(T[k], T[k], T[k]), &cke, &smp, v3(fE, fE, fE)).x;
        }
        if (g != NULL) {
            spec_brdf_grad(&ve, smp.direction, a4, lm, g);
        }
        /* tier-B walk (PathIntegratorRandCore.h:169-204) with the RGB radius */
        if (smp.subsurfaceScatter && smp.subsurfaceInternalSteps > 0) {
            const float maxR = orc_max_scatter_radius(&m);
            if (maxR > 1.0e-5f) {
                for (int step = 0; step < smp.subsurfaceInternalSteps; ++step) {
                    (void)orc_rand01(rng);
                    (void)orc_rand01(rng);
                    const float u3 = orc_rand01(rng);
                    const float mfp = maxR * (-orc_logf(fmax2(1.0e-8f, u3)));
                    const float surv = orc_expf(-mfp / maxR);
                    for (int k = 0; k < 4; ++k) {
                        T0[k] = T0[k] * (v.grey[k].r * surv);
                        T[k] = T[k] * (ve.grey[k].r * surv);
                        Td[k] *= (double)(ve.grey[k].r * surv);
                    }
                }
                const float ed = vlen(smp.subsurfaceExitOffset);
                if (ed > 0.0f) {
                    const float a = orc_expf(-ed / maxR);
                    for (int k = 0; k < 4; ++k) {
                        T0[k] = T0[k] * a;
                        T[k] = T[k] * a;
                        Td[k] *= a;
                    }
                }
            }
        }
        if ((flags & (1 << 4)) != 0) {
            const OVec3 tint = smp.transmitted ? v3(0.2f, 1.0f, 0.2f) : v3(1.0f, 0.2f, 0.2f);
            for (int k = 0; k < 4; ++k) {
                const float r = orc_rgb2spec_reflectance(model, tint.x, tint.y, tint.z, v.lambda[k]);
                T0[k] = T0[k] * r;
                T[k] = T[k] * r;
                Td[k] *= r;
            }
        }
        const float tl = 0.25f * (((T0[0] + T0[1]) + T0[2]) + T0[3]);
        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
            break;
        }
        if (depth >= rrMin) {
            const float surv = fmin2(0.95f, fmax2(tl, 0.05f));
            if (orc_rand01(rng) > surv) {
                break;
            }
            for (int k = 0; k < 4; ++k) {
                T0[k] = T0[k] * (1.0f / surv);
                T[k] = T[k] * (1.0f / surv);
                Td[k] *= (double)(1.0f / surv);
            }
        }
        eta = smp.nextMediumEta;
        const float bias = fmax2(K_SHADOW_BIAS, K_REL_SHADOW_BIAS * hit.t);
        const float sign = vdot(n, smp.direction) >= 0.0f ? 1.0f : -1.0f;
        origin = vadd(p, vscale(n, sign * bias));
        dir = smp.direction;
        hit = trace(origin, dir, s, bias, K_RAY_TMAX, 1);
        if (!hit.hit) {
            const OVec3 bg = orc_env_eval(env, params, dir);
            const float lpdf = orc_env_pdf(env, params, dir);
            const float mis = orc_mis_balance(smp.pdf, lpdf);
            for (int k = 0; k < 4; ++k) {
                const float envL = orc_rgb2spec_radiance(model, bg.x, bg.y, bg.z, v.lambda[k]);
                L4[k] = L4[k] + T[k] * (envL * mis);
                if (Ld4 != NULL) {
                    Ld4[k] += Td[k] * (double)(envL * mis);
                }
            }
            break;
        }
    }
}

/* The spectral firefly clamp in double (non-finite -> 0, luminance above 100 scaled down,
 * negative channels kept). */
static void clamp_firefly_spectral_f64(const double r[3], double out[3])
{
    if (!isfinite(r[0]) || !isfinite(r[1]) || !isfinite(r[2])) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double lum = r[0] * 0.2126 + r[1] * 0.7152 + r[2] * 0.0722;
    const double k = lum > 100.0 ? 100.0 / lum : 1.0;
    for (int q = 0; q < 3; ++q) {
        out[q] = r[q] * k;
    }
}

typedef struct {
    float* img;
    double* imgD;
    const double* mask;
    const double* gradImg;
    double* grad;
    int w, h, spp, y0, y1, gradCount;
    uint32_t firstSeed;
    const OCamera* cam;
    const OScene* s;
    const OScene* values;
    const OEnv* env;
    const OParams* params;
    const OSpecModel* model;
} HeroJob;

static void hero_rows(const HeroJob* j)
{
    const float pdf = 1.0f / j->model->lambdaRange;
    for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < j->w; ++x) {
            const int idx = y * j->w + x;
            const double* gp = j->gradImg != NULL ? j->gradImg + 3 * idx : (j->mask != NULL ? j->mask + 3 * idx : NULL);
            if (gp != NULL && gp[0] == 0.0 && gp[1] == 0.0 && gp[2] == 0.0) {
                continue; /* pixel outside the loss */
            }
            float acc[3] = {0.0f, 0.0f, 0.0f};
            for (int k = 0; k < j->spp; ++k) {
                ORng rng;
                orc_rng_init(&rng, idx, k, j->firstSeed + (uint32_t)k);
                const float ju = orc_rand01(&rng);
                const float jv = orc_rand01(&rng);
                const float u = ((float)x + ju) / (float)j->w;
                const float v = ((float)y + jv) / (float)j->h;
                const float l1 = orc_rand01(&rng);
                const float l2 = orc_rand01(&rng);
                OVec3 ro, rd;
                orc_camera_ray(j->cam, u, v, l1, l2, &ro, &rd);
                float lambda[4];
                orc_hero_lambdas(j->model, orc_rand01(&rng), lambda);
                float L4[4];
                if (j->gradImg == NULL) {
                    double Ld[4];
                    trace_path_ad_hero(ro, rd, j->s, j->values, j->params, j->env, &rng, j->model, lambda, NULL, NULL,
                                       NULL, L4, j->imgD != NULL ? Ld : NULL);
                    if (j->imgD != NULL) {
                        double rgb[3] = {0.0, 0.0, 0.0}, c[3];
                        for (int q = 0; q < 4; ++q) {
                            if (!isfinite(Ld[q]) || Ld[q] <= 0.0) {
                                continue;
                            }
                            float resp[3];
                            orc_cie_rgb_response(lambda[q], resp);
                            for (int ch = 0; ch < 3; ++ch) {
                                rgb[ch] += 0.25 * (Ld[q] / fmax2(pdf, 1.0e-8f)) * resp[ch] /
                                           fmax2(j->model->white[ch], 1.0e-8f);
                            }
                        }
                        clamp_firefly_spectral_f64(rgb, c);
                        for (int ch = 0; ch < 3; ++ch) {
                            j->imgD[3 * idx + ch] += c[ch] / j->spp;
                        }
                    } else {
                        const OVec3 c = orc_clamp_firefly_spectral(orc_hero_to_rgb(j->model, L4, lambda));
                        const float inv = 1.0f / (float)(k + 1);
                        acc[0] = (acc[0] * (float)k + c.x) * inv;
                        acc[1] = (acc[1] * (float)k + c.y) * inv;
                        acc[2] = (acc[2] * (float)k + c.z) * inv;
                    }
                } else {
                    ORng replay = rng;
                    float C4[4];
                    trace_path_ad_hero(ro, rd, j->s, j->s, j->params, j->env, &rng, j->model, lambda, NULL, NULL, NULL,
                                       C4, NULL);
                    /* clamp VJP of the sample's linear sRGB, then each wavelength's response */
                    const OVec3 Crgb = orc_hero_to_rgb(j->model, C4, lambda);
                    const double gs[3] = {j->gradImg[3 * idx] / j->spp, j->gradImg[3 * idx + 1] / j->spp,
                                          j->gradImg[3 * idx + 2] / j->spp};
                    double gw[3] = {gs[0], gs[1], gs[2]};
                    if (!isfinite(Crgb.x) || !isfinite(Crgb.y) || !isfinite(Crgb.z)) {
                        gw[0] = gw[1] = gw[2] = 0.0;
                    } else {
                        const double cr[3] = {Crgb.x, Crgb.y, Crgb.z};
                        const double lum = cr[0] * 0.2126 + cr[1] * 0.7152 + cr[2] * 0.0722;
                        if (lum > 100.0) {
                            const double wl[3] = {0.2126, 0.7152, 0.0722};
                            const double gdr = gs[0] * cr[0] + gs[1] * cr[1] + gs[2] * cr[2];
                            for (int ch = 0; ch < 3; ++ch) {
                                gw[ch] = 100.0 / lum * (gs[ch] - wl[ch] * gdr / lum);
                            }
                        }
                    }
                    double gw4[4];
                    int any = 0;
                    for (int q = 0; q < 4; ++q) {
                        gw4[q] = 0.0;
                        if (isfinite(C4[q]) && C4[q] > 0.0f) {
                            float resp[3];
                            orc_cie_rgb_response(lambda[q], resp);
                            for (int ch = 0; ch < 3; ++ch) {
                                gw4[q] += 0.25 / fmax2(pdf, 1.0e-8f) * gw[ch] * resp[ch] /
                                          fmax2(j->model->white[ch], 1.0e-8f);
                            }
                        }
                        any = any || gw4[q] != 0.0;
                    }
                    if (any) {
                        trace_path_ad_hero(ro, rd, j->s, j->s, j->params, j->env, &replay, j->model, lambda, gw4, C4,
                                           j->grad, L4, NULL);
                    }
                }
            }
            if (j->img != NULL) {
                j->img[3 * idx] = acc[0];
                j->img[3 * idx + 1] = acc[1];
                j->img[3 * idx + 2] = acc[2];
            }
        }
    }
}

static void* hero_worker(void* arg)
{
    hero_rows((const HeroJob*)arg);
    return NULL;
}

static void hero_run(HeroJob base, int threads, double* gradOut)
{
    threads = threads < 1 ? 1 : (threads > 64 ? 64 : threads);
    pthread_t tid[64];
    HeroJob jobs[64];
    double* partial = NULL;
    if (gradOut != NULL) {
        partial = (double*)calloc((size_t)threads * (size_t)base.gradCount, sizeof(double));
    }
    for (int t = 0; t < threads; ++t) {
        jobs[t] = base;
        jobs[t].y0 = base.h * t / threads;
        jobs[t].y1 = base.h * (t + 1) / threads;
        jobs[t].grad = partial != NULL ? partial + (size_t)t * base.gradCount : NULL;
        pthread_create(&tid[t], NULL, hero_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
    }
    if (partial != NULL) {
        for (int t = 0; t < threads; ++t) {
            for (int k = 0; k < base.gradCount; ++k) {
                gradOut[k] += partial[(size_t)t * base.gradCount + k];
            }
        }
        free(partial);
    }
}

/* The spectral detached estimator's image: paths sampled with the scene (its materials and
 * spectra), values from evalMats / evalSpectra (NULL: the scene's). img3 (float, running mean of
 * the clamped linear sRGB, as the forward) or imgD (double, added into, mask3 selecting pixels as
 * orc_render_detached_f64_seeded) -- one of them. */
void orc_render_detached_spectral(float* img3, double* imgD, const double* mask3, int width, int height,
                                  const OCamera* cam, const OScene* s, const OMaterial* evalMats,
                                  const float* evalSpectra, const OEnv* env, const OParams* params,
                                  const OSpecModel* model, int spp, uint32_t firstSeed, int threads)
{
    OScene values = *s;
    if (evalMats != NULL) {
        values.materials = evalMats;
    }
    if (evalSpectra != NULL) {
        values.spectra = evalSpectra;
    }
    HeroJob j;
    memset(&j, 0, sizeof(j));
    j.img = imgD != NULL ? NULL : img3;
    j.imgD = imgD;
    j.mask = imgD != NULL ? mask3 : NULL;
    j.w = width;
    j.h = height;
    j.spp = spp;
    j.firstSeed = firstSeed;
    j.cam = cam;
    j.s = s;
    j.values = &values;
    j.env = env;
    j.params = params;
    j.model = model;
    hero_run(j, threads, NULL);
}

/* The FP64 adjoint of the hero-wavelength render: grad[ORC_SPECTRAL_GRAD_PARAMS * materialCount]
 * += sum over pixels and samples of gradImg . d(pixel)/d(theta) (gradImg = dLoss/dImage, 3
 * doubles per pixel; sample k uses seed firstSeed + k). */
void orc_render_adjoint_spectral(const double* gradImg, int width, int height, const OCamera* cam, const OScene* s,
                                 const OEnv* env, const OParams* params, const OSpecModel* model, int spp,
                                 uint32_t firstSeed, double* grad, int threads)
{
    HeroJob j;
    memset(&j, 0, sizeof(j));
    j.gradImg = gradImg;
    j.w = width;
    j.h = height;
    j.spp = spp;
    j.firstSeed = firstSeed;
    j.cam = cam;
    j.s = s;
    j.values = s;
    j.env = env;
    j.params = params;
    j.model = model;
    j.gradCount = ORC_SPECTRAL_GRAD_PARAMS * (int)s->materialCount;
    hero_run(j, threads, grad);
}

/* Fixed-spp image of the detached estimator: sample k of every pixel uses sample index k and
 * seed k + 1 (the adaptive schedule with minSamples = maxSamplesPerPixel = spp); img holds the
 * running mean of the clamped radiance (3 floats per pixel). */
void orc_render_detached(float* img3, int width, int height, const OCamera* cam, const OScene* s,
                         const OMaterial* evalMats, const OEnv* env, const OParams* params, int spp, int threads)
{
    orc_render_detached_seeded(img3, width, height, cam, s, evalMats, env, params, spp, 1u, threads);
}

/* The same image over the samples of seeds firstSeed .. firstSeed + spp - 1. */
void orc_render_detached_seeded(float* img3, int width, int height, const OCamera* cam, const OScene* s,
                                const OMaterial* evalMats, const OEnv* env, const OParams* params, int spp,
                                uint32_t firstSeed, int threads)
{
    AdJob j;
    memset(&j, 0, sizeof(j));
    j.firstSeed = firstSeed;
    j.img = img3;
    j.w = width;
    j.h = height;
    j.spp = spp;
    j.cam = cam;
    j.s = s;
    j.evalMats = evalMats;
    j.env = env;
    j.params = params;
    ad_run(j, threads, NULL);
}

/* The same estimator evaluated in double (trace_path_ad with Ld): the per-pixel mean of the
 * clamped radiance over seeds firstSeed .. firstSeed + spp - 1, added into img3 (3 doubles per
 * pixel, zero it first). mask3 != NULL restricts it to the pixels with a nonzero mask value. Its
 * central differences in a parameter are those of the detached estimator without float rounding,
 * so they pin orc_render_adjoint_seeded's derivative directly. */
void orc_render_detached_f64_seeded(double* img3, int width, int height, const OCamera* cam, const OScene* s,
                                    const OMaterial* evalMats, const OEnv* env, const OParams* params, int spp,
                                    uint32_t firstSeed, const double* mask3, int threads)
{
    AdJob j;
    memset(&j, 0, sizeof(j));
    j.firstSeed = firstSeed;
    j.imgD = img3;
    j.mask = mask3;
    j.w = width;
    j.h = height;
    j.spp = spp;
    j.cam = cam;
    j.s = s;
    j.evalMats = evalMats;
    j.env = env;
    j.params = params;
    ad_run(j, threads, NULL);
}

/* Adjoint of the same render: grad[7 * materialCount] += sum over pixels and samples of
 * gradImg . d(pixel)/d(theta), with gradImg = dLoss/dImage (3 doubles per pixel). Sample k of a
 * pixel uses sample index k and seed firstSeed + k (orc_render_adjoint: firstSeed = 1). */
void orc_render_adjoint_seeded(const double* gradImg, int width, int height, const OCamera* cam, const OScene* s,
                               const OEnv* env, const OParams* params, int spp, uint32_t firstSeed, double* grad,
                               int threads)
{
    AdJob j;
    memset(&j, 0, sizeof(j));
    j.firstSeed = firstSeed;
    j.gradImg = gradImg;
    j.w = width;
    j.h = height;
    j.spp = spp;
    j.cam = cam;
    j.s = s;
    j.env = env;
    j.params = params;
    j.gradCount = 7 * (int)s->materialCount;
    ad_run(j, threads, grad);
}

void orc_render_adjoint(const double* gradImg, int width, int height, const OCamera* cam, const OScene* s,
                        const OEnv* env, const OParams* params, int spp, double* grad, int threads)
{
    orc_render_adjoint_seeded(gradImg, width, height, cam, s, env, params, spp, 1u, grad, threads);
}
