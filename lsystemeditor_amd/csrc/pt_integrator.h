// pt_integrator.h -- the per-sample path integrator shared by the megakernel launchers and the
// adjoint (PathIntegratorRandCore.h:65-363).
#pragma once

#include "pt_device.h"

namespace pt {

// PathIntegratorRandCore.h:65-108. Returns the NEE term (before multiplication by throughput).
PT_DEV V3 env_nee(const MeshAccelSceneGpu& s, const RenderParamsGpu& p, const EnvironmentMapGpu* env, V3 pos,
                  const Ctx& c, Rng& rng, TraceStats* st)
{
    if (skips_env_nee(c.m)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const float u1 = rand01(rng);
    const float u2 = rand01(rng);
    float lightPdf = 0.0f;
    const V3 wi = env_sample(env, u1, u2, lightPdf);
    if (lightPdf <= kMinPdf) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    // LightSamplingCore.h:187-197: origin offset along the normal, closest-hit -> any hit
    const V3 o = add(pos, scale(c.n, kShadowBias));
    if (trace_any(s, o, wi, kShadowBias, kRayTMax, st)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const float cosT = max2(0.0f, dot(c.n, wi));
    if (cosT <= 0.0f) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const V3 L = env_eval(env, p, wi);
    const V3 f = brdf_eval(c, wi);
    const float bpdf = brdf_pdf(c, wi);
    const float mis = mis_balance(lightPdf, bpdf);
    const float sc = mis * cosT / lightPdf;
    return mul(f, scale(L, sc));
}

// PathIntegratorRandCore.h:110-167
PT_DEV V3 emissive_nee(const MeshAccelSceneGpu& s, V3 pos, const Ctx& c, Rng& rng, TraceStats* st)
{
    if (s.emissiveTriangleCount == 0) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const float uTri = rand01(rng);
    const float u1 = rand01(rng);
    const float u2 = rand01(rng);
    V3 lp, ln, lr;
    float areaPdf = 0.0f;
    if (!sample_emissive(s, uTri, u1, u2, lp, ln, lr, areaPdf)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const V3 toL = sub(lp, pos);
    const float d2 = dot(toL, toL);
    if (d2 <= kMinPdf) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const V3 wi = normalize(toL);
    const float cosT = max2(0.0f, dot(c.n, wi));
    const float cosL = max2(0.0f, dot(ln, scale(wi, -1.0f)));
    if (cosT <= 0.0f || cosL <= 0.0f) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const float lightPdf = areaPdf * d2 / cosL;
    // The result is zero whenever lightPdf <= kMinPdf, occluded or not, so the shadow ray
    // (LightSamplingCore.h:360-369) is only traced when it can matter.
    if (lightPdf <= kMinPdf) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const V3 o = add(pos, scale(wi, kShadowBias));
    if (trace_any(s, o, wi, kShadowBias, kRayTMax, st)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    const V3 f = brdf_eval(c, wi);
    const float bpdf = brdf_pdf(c, wi);
    const float mis = mis_balance(lightPdf, bpdf);
    const float sc = mis * cosT / lightPdf;
    return mul(f, scale(lr, sc));
}

// PathIntegratorRandCore.h:169-204 (the step direction is computed and discarded there;
// it has no effect on the result and is not computed here)
PT_DEV V3 subsurface_tier_b(V3 T, const Mat& m, const Sample& smp, Rng& rng)
{
    if (!smp.sss || smp.steps <= 0) {
        return T;
    }
    const float mr = max_radius(m);
    if (mr <= kMinRadius) {
        return T;
    }
    const V3 albedo = mk(m.r, m.g, m.b);
    for (int step = 0; step < smp.steps; ++step) {
        (void)rand01(rng);
        (void)rand01(rng);
        const float u3 = rand01(rng);
        const float mfp = mr * (-logf_pt(max2(1.0e-8f, u3)));
        const float surv = expf_pt(-mfp / mr);
        T = mul(T, scale(albedo, surv));
    }
    const float ed = length(smp.exitOffset);
    if (ed > 0.0f) {
        T = scale(T, expf_pt(-ed / mr));
    }
    return T;
}

// PathIntegratorRandCore.h:206-363
PT_DEV V3 trace_path(const MeshAccelSceneGpu& s, const RenderParamsGpu& p, const EnvironmentMapGpu* env, V3 ro,
                     V3 rd, Rng& rng, TraceStats* st)
{
    Hit hit = trace_closest(s, ro, rd, 0.0f, kRayTMax, st);
    if (!hit.hit) {
        return env_eval(env, p, rd);
    }
    int maxDepth = 32;
    if (p.maxPathDepth > 0) {
        maxDepth = p.maxPathDepth;
    } else if (p.maxPathDepth == 0) {
        maxDepth = 256;
    }
    const int rrMin = p.russianRouletteMinDepth >= 0 ? p.russianRouletteMinDepth : 3;
    const int flags = p.brdfDebugFlags;

    V3 L = mk(0.0f, 0.0f, 0.0f);
    V3 T = mk(1.0f, 1.0f, 1.0f);
    V3 origin = ro;
    V3 dir = rd;
    float eta = 1.0f;
    for (int depth = 0; depth < maxDepth; ++depth) {
        const V3 pos = add(origin, scale(dir, hit.t));
        Ctx c;
        c.n = hit.normal;
        c.wo = scale(dir, -1.0f);
        c.m = resolve_material(s, hit.tri, hit.uvx, hit.uvy);
        c.w = lobe_weights(c.m);
        c.eta = eta;
        c.flags = 0;

        L = add(L, mul(T, mk(c.m.r * c.m.emission, c.m.g * c.m.emission, c.m.b * c.m.emission)));
        L = add(L, mul(T, env_nee(s, p, env, pos, c, rng, st)));
        L = add(L, mul(T, emissive_nee(s, pos, c, rng, st)));

        c.flags = flags;
        const float u1 = rand01(rng);
        const float u2 = rand01(rng);
        const Sample smp = brdf_sample(c, u1, u2);
        if (!smp.valid || smp.pdf <= kMinPdf) {
            break;
        }
        const V3 f = brdf_eval(c, smp.dir);
        T = apply_throughput(T, c, smp, f);
        T = subsurface_tier_b(T, c.m, smp, rng);
        if ((flags & 16) != 0) {
            T = smp.transmitted ? mul(T, mk(0.2f, 1.0f, 0.2f)) : mul(T, mk(1.0f, 0.2f, 0.2f));
        }
        const float tl = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
            break;
        }
        if (depth >= rrMin) {
            const float surv = min2(0.95f, max2(tl, 0.05f));
            if (rand01(rng) > surv) {
                break;
            }
            T = scale(T, 1.0f / surv);
        }
        eta = smp.nextEta;
        const float bias = max2(kShadowBias, kRelShadowBias * hit.t);
        const float sgn = dot(c.n, smp.dir) >= 0.0f ? 1.0f : -1.0f;
        origin = add(pos, scale(c.n, sgn * bias));
        dir = smp.dir;
        hit = trace_closest(s, origin, dir, bias, kRayTMax, st);
        if (!hit.hit) {
            const V3 envL = env_eval(env, p, dir);
            const float lpdf = env_pdf(env, dir);
            const float mis = mis_balance(smp.pdf, lpdf);
            L = add(L, mul(T, scale(envL, mis)));
            break;
        }
    }
    return L;
}

} // namespace pt
