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

/* PathIntegratorRandCore.h:110-167 */
static OVec3 emissive_nee(OVec3 p, OVec3 n, OVec3 wo, const OMaterial* m, float eta, const OScene* s, ORng* rng)
{
    if (s == NULL || s->emissiveCount == 0) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const float uTri = orc_rand01(rng);
    const float u1 = orc_rand01(rng);
    const float u2 = orc_rand01(rng);
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
    if (occluded_from(p, wi, s)) {
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

    for (int depth = 0; depth < maxDepth; ++depth) {
        const OVec3 p = vray(origin, dir, hit.t);
        const OVec3 n = hit.normal;
        const OVec3 wo = vscale(dir, -1.0f);
        const OMaterial src = orc_fetch_material(s, hit.triangleIndex);
        const OMaterial m = orc_resolve_material(&src, hit.uv, s);

        L = vadd(L, vmul(T, v3(m.r * m.emission, m.g * m.emission, m.b * m.emission)));
        L = vadd(L, vmul(T, env_nee(p, n, wo, &m, eta, s, env, params, rng)));
        L = vadd(L, vmul(T, emissive_nee(p, n, wo, &m, eta, s, rng)));

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
