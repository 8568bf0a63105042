// pt_integrator.h -- the per-sample path integrator shared by the megakernel launchers and the
// adjoint (PathIntegratorRandCore.h:65-363). Generic over the colour policy (pt_spectral.h):
// RgbColors is the reference's RGB integrator, HeroColors the hero-wavelength spectral one.
#pragma once

#include "pt_device.h"
#include "pt_spectral.h"

namespace pt {

// PathIntegratorRandCore.h:65-108. Returns the NEE term (before multiplication by throughput).
template <class Pol>
PT_DEV typename Pol::Col env_nee_t(const MeshAccelSceneGpu& s, const RenderParamsGpu& p, const EnvironmentMapGpu* env,
                                   V3 pos, const Ctx& c, const typename Pol::Col& base,
                                   const typename Pol::Col& rad, const Pol& pol, Rng& rng, TraceStats* st)
{
    using C = typename Pol::Col;
    const C zero = ColN<C>::splat(0.0f);
    if (skips_env_nee(c.m)) {
        return zero;
    }
    const float u1 = rand01(rng);
    const float u2 = rand01(rng);
    float lightPdf = 0.0f;
    const V3 wi = env_sample(env, u1, u2, lightPdf);
    if (lightPdf <= kMinPdf) {
        return zero;
    }
    // LightSamplingCore.h:187-197: origin offset along the normal, closest-hit -> any hit
    const V3 o = add(pos, scale(c.n, kShadowBias));
    if (trace_any(s, o, wi, kShadowBias, kRayTMax, st)) {
        return zero;
    }
    const float cosT = max2(0.0f, dot(c.n, wi));
    if (cosT <= 0.0f) {
        return zero;
    }
    const C L = pol.radiance(env_eval(env, p, wi));
    const C f = brdf_eval_t(c, wi, base, rad);
    const float bpdf = brdf_pdf(c, wi);
    const float mis = mis_balance(lightPdf, bpdf);
    const float sc = mis * cosT / lightPdf;
    return mul(f, scale(L, sc));
}

// PathIntegratorRandCore.h:110-167; lightMode kLightMis bounds the shadow ray short of the light
// point (pt_device.h). Under kLightMis the last vertex of a path (last) adds nothing: the light its
// estimate finds lies one segment past the path's length, where the reference never gathers
// emission, so both modes estimate the same image. Its draws are taken all the same.
template <class Pol>
PT_DEV typename Pol::Col emissive_nee_t(const MeshAccelSceneGpu& s, V3 pos, const Ctx& c,
                                        const typename Pol::Col& base, const typename Pol::Col& rad,
                                        const Pol& pol, Rng& rng, TraceStats* st, int lightMode, bool last)
{
    using C = typename Pol::Col;
    const C zero = ColN<C>::splat(0.0f);
    if (s.emissiveTriangleCount == 0) {
        return zero;
    }
    const float uTri = rand01(rng);
    const float u1 = rand01(rng);
    const float u2 = rand01(rng);
    if (lightMode == kLightMis && last) {
        return zero;
    }
    V3 lp, ln, lr;
    float areaPdf = 0.0f;
    int entry = 0;
    if (!sample_emissive(s, uTri, u1, u2, lp, ln, lr, areaPdf, &entry)) {
        return zero;
    }
    const V3 toL = sub(lp, pos);
    const float d2 = dot(toL, toL);
    if (d2 <= kMinPdf) {
        return zero;
    }
    const V3 wi = normalize(toL);
    const float cosT = max2(0.0f, dot(c.n, wi));
    const float cosL = max2(0.0f, dot(ln, scale(wi, -1.0f)));
    if (cosT <= 0.0f || cosL <= 0.0f) {
        return zero;
    }
    const float lightPdf = areaPdf * d2 / cosL;
    // The result is zero whenever lightPdf <= kMinPdf, occluded or not, so the shadow ray
    // (LightSamplingCore.h:360-369) is only traced when it can matter.
    if (lightPdf <= kMinPdf) {
        return zero;
    }
    const V3 o = add(pos, scale(wi, kShadowBias));
    if (lightMode == kLightMis) {
        const V3 toP = sub(lp, o);
        if (trace_any(s, o, toP, kShadowBias, length(toP) * kBoundedShadowScale, st)) {
            return zero;
        }
    } else if (emissive_blocks(s, entry, make_ray(o, wi), kShadowBias, kRayTMax) ||
               trace_any(s, o, wi, kShadowBias, kRayTMax, st)) {
        return zero;
    }
    const C f = brdf_eval_t(c, wi, base, rad);
    const float bpdf = brdf_pdf(c, wi);
    const float mis = mis_balance(lightPdf, bpdf);
    const float sc = mis * cosT / lightPdf;
    return mul(f, scale(pol.light(s, entry, lr), sc));
}

// PathIntegratorRandCore.h:169-204 (the step direction is computed and discarded there;
// it has no effect on the result and is not computed here)
template <class C>
PT_DEV C subsurface_tier_b_t(C T, const Mat& m, const C& albedo, const Sample& smp, Rng& rng)
{
    if (!smp.sss || smp.steps <= 0) {
        return T;
    }
    const float mr = max_radius(m);
    if (mr <= kMinRadius) {
        return T;
    }
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

// PathIntegratorRandCore.h:206-363. lightMode kLightMis weights emission found by BSDF sampling
// against the emissive next-event estimate of the vertex before (pt_device.h).
template <class Pol>
PT_DEV typename Pol::Col trace_path_t(const MeshAccelSceneGpu& s, const RenderParamsGpu& p,
                                      const EnvironmentMapGpu* env, V3 ro, V3 rd, const Pol& pol, Rng& rng,
                                      TraceStats* st, int lightMode = kLightReference)
{
    using C = typename Pol::Col;
    Hit hit = trace_closest(s, ro, rd, 0.0f, kRayTMax, st);
    if (!hit.hit) {
        return pol.radiance(env_eval(env, p, rd));
    }
    int maxDepth = 32;
    if (p.maxPathDepth > 0) {
        maxDepth = p.maxPathDepth;
    } else if (p.maxPathDepth == 0) {
        maxDepth = 256;
    }
    const int rrMin = p.russianRouletteMinDepth >= 0 ? p.russianRouletteMinDepth : 3;
    const int flags = p.brdfDebugFlags;

    C L = ColN<C>::splat(0.0f);
    C T = ColN<C>::splat(1.0f);
    V3 origin = ro;
    V3 dir = rd;
    float eta = 1.0f;
    float prevPdf = 0.0f;   // pdf of the BSDF sample that found the current hit
    bool reach = false;     // whether that direction lies where the emissive NEE before samples
    for (int depth = 0; depth < maxDepth; ++depth) {
        const V3 pos = add(origin, scale(dir, hit.t));
        Ctx c;
        c.n = hit.normal;
        c.wo = scale(dir, -1.0f);
        c.m = resolve_material(s, hit.tri, hit.uvx, hit.uvy);
        c.w = lobe_weights(c.m);
        c.eta = eta;
        c.flags = 0;

        const C base = pol.base(c.m);
        const C rad = c.w.subsurface != 0.0f ? pol.radii(c.m) : ColN<C>::splat(0.0f);

        if (lightMode == kLightMis && depth > 0 && c.m.emission > 0.0f) {
            const float w = emission_mis_weight(s, hit.tri, dir, hit.t, prevPdf, reach);
            L = add(L, mul(T, scale(pol.emission(c.m), w)));
        } else {
            L = add(L, mul(T, pol.emission(c.m)));
        }
        L = add(L, mul(T, env_nee_t(s, p, env, pos, c, base, rad, pol, rng, st)));
        L = add(L, mul(T, emissive_nee_t(s, pos, c, base, rad, pol, rng, st, lightMode, depth + 1 == maxDepth)));

        c.flags = flags;
        const float u1 = rand01(rng);
        const float u2 = rand01(rng);
        const Sample smp = brdf_sample(c, u1, u2);
        if (!smp.valid || smp.pdf <= kMinPdf) {
            break;
        }
        const C f = brdf_eval_t(c, smp.dir, base, rad);
        T = apply_throughput_t(T, c, smp, f, base);
        T = subsurface_tier_b_t(T, c.m, base, smp, rng);
        if ((flags & 16) != 0) {
            T = smp.transmitted ? mul(T, pol.reflectance(mk(0.2f, 1.0f, 0.2f)))
                                : mul(T, pol.reflectance(mk(1.0f, 0.2f, 0.2f)));
        }
        const float tl = pol.lum(T);
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
        prevPdf = smp.pdf;
        reach = dot(c.n, smp.dir) > 0.0f;
        const float bias = max2(kShadowBias, kRelShadowBias * hit.t);
        const float sgn = dot(c.n, smp.dir) >= 0.0f ? 1.0f : -1.0f;
        origin = add(pos, scale(c.n, sgn * bias));
        dir = smp.dir;
        hit = trace_closest(s, origin, dir, bias, kRayTMax, st);
        if (!hit.hit) {
            const C envL = pol.radiance(env_eval(env, p, dir));
            const float lpdf = env_pdf(env, dir);
            const float mis = mis_balance(smp.pdf, lpdf);
            L = add(L, mul(T, scale(envL, mis)));
            break;
        }
    }
    return L;
}

PT_DEV V3 env_nee(const MeshAccelSceneGpu& s, const RenderParamsGpu& p, const EnvironmentMapGpu* env, V3 pos,
                  const Ctx& c, Rng& rng, TraceStats* st)
{
    return env_nee_t(s, p, env, pos, c, mat_base(c.m), mat_radii(c.m), RgbColors{}, rng, st);
}

PT_DEV V3 emissive_nee(const MeshAccelSceneGpu& s, V3 pos, const Ctx& c, Rng& rng, TraceStats* st)
{
    return emissive_nee_t(s, pos, c, mat_base(c.m), mat_radii(c.m), RgbColors{}, rng, st, kLightReference, false);
}

PT_DEV V3 subsurface_tier_b(V3 T, const Mat& m, const Sample& smp, Rng& rng)
{
    return subsurface_tier_b_t(T, m, mat_base(m), smp, rng);
}

PT_DEV V3 trace_path(const MeshAccelSceneGpu& s, const RenderParamsGpu& p, const EnvironmentMapGpu* env, V3 ro,
                     V3 rd, Rng& rng, TraceStats* st, int lightMode = kLightReference)
{
    return trace_path_t(s, p, env, ro, rd, RgbColors{}, rng, st, lightMode);
}

} // namespace pt
