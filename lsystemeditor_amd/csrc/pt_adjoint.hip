// pt_adjoint.hip -- adjoint (path replay) gradients of an image loss with respect to material
// parameters, for the fixed-spp render (sample k of every pixel: sample index k, seed k + 1).
//
// Each thread takes one (pixel, sample): it traces the primal path (radiance C), maps the pixel's
// loss gradient through the firefly clamp, and replays the same path from the same RNG state,
// accumulating at every vertex the derivative of that vertex's contributions with respect to its
// material: emission, next-event BSDF values, and the BSDF value of the sampled continuation,
// which scales everything the path gathers afterwards (C minus what was gathered so far).
// Sampling is detached: directions, pdfs, MIS weights and Russian roulette are those of the
// primal parameters, so the estimator stays unbiased and nothing is stored per path.
// Derivatives are FP32 per vertex; the per-parameter reduction is FP64 (LDS per block, then one
// global atomic per block and parameter). Parameters per material: base r, g, b, roughness,
// metallic, specular, emission; the BSDF derivative covers the opaque Principled lobes.
#include "pt_device.h"
#include "pt_integrator.h"
#include "pt_internal.h"

#include "../../include/pt_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>

using namespace pt;

namespace {

constexpr int kAdBlock = 256;
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

// out[k] += sum_c a_c d f_c / d theta_k for theta = (r, g, b, roughness, metallic, specular) of
// the opaque Principled BSDF f = base wd ON / pi + F ws ec D G / max(4 cosO cosI, 1e-8), with the
// normalised lobe weights wd = (1 - m) / (2 - m), ws = 1 / (2 - m).
PT_DEV void brdf_grad(const Ctx& c, V3 wi, V3 a, const Live& lm, float out[6])
{
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
        out[0] += a.x * k;
        out[1] += a.y * k;
        out[2] += a.z * k;
    }
    const float aBase = a.x * m.r + a.y * m.g + a.z * m.b;
    if (lm.metal) {
        out[4] += aBase * on * kInvPi * (-inv2m * inv2m);
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
        out[3] += aBase * wd * dOn * kInvPi;
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
    const V3 F = mk((f0v + (m.r - f0v) * mm) * (1.0f - t5) + t5, (f0v + (m.g - f0v) * mm) * (1.0f - t5) + t5,
                    (f0v + (m.b - f0v) * mm) * (1.0f - t5) + t5);
    const float cw = max2(cosWo, 1.0e-4f);
    const float ess = 1.0f - alpha * (0.3316f + 0.3237f * alpha);
    const float ems = (1.0f - ess) * (1.0f - cw);
    const float ec = 1.0f + ems / max2(ess, 1.0e-4f);
    const float denom = max2(4.0f * cosWo * cosWi, 1.0e-8f);
    const float S = ws * ec * D * G / denom;
    const float aF = a.x * F.x + a.y * F.y + a.z * F.z;
    if (lm.rgb) {
        const float k = mm * (1.0f - t5) * S;
        out[0] += a.x * k;
        out[1] += a.y * k;
        out[2] += a.z * k;
    }
    if (lm.metal) {
        const float dF = (a.x * (m.r - f0v) + a.y * (m.g - f0v) + a.z * (m.b - f0v)) * (1.0f - t5);
        out[4] += dF * S + aF * S * inv2m;
    }
    if (lm.spec) {
        out[5] += (a.x + a.y + a.z) * (1.0f - mm) * (1.0f - t5) * f0s * f0s * S;
    }
    if (lm.rough && rough > 0.04f) {
        const float dD = (dd - 2.0f * a2 * c2) / (kPi * dd * dd * dd) * 2.0f * alpha;
        const float dG = (dg1[0] * g1[1] + g1[0] * dg1[1]) * 2.0f * alpha;
        const float dEss = -(0.3316f + 2.0f * 0.3237f * alpha);
        const float dEms = -dEss * (1.0f - cw);
        const float dEc = ess > 1.0e-4f ? (dEms * ess - ems * dEss) / (ess * ess) : dEms / 1.0e-4f;
        const float dS = ws / denom * (dEc * D * G + ec * dD * G + ec * D * dG);
        out[3] += aF * dS * 2.0f * rough;
    }
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

// The replay of trace_path (pt_integrator.h) with gradient accumulation; same RNG use and
// decisions as the primal walk.
PT_DEV void adjoint_walk(const MeshAccelSceneGpu& s, const RenderParamsGpu& p, const EnvironmentMapGpu* env, V3 ro,
                         V3 rd, Rng& rng, V3 C, V3 gw, double* lg)
{
    Hit hit = trace_closest(s, ro, rd, 0.0f, kRayTMax, nullptr);
    if (!hit.hit) {
        return;
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
        int mi = -1;
        Live lm{false, false, false, false, false, false, false};
        if (s.materials != nullptr && hit.tri < s.triangleCount) {
            const uint32_t idx = tri_material(s, hit.tri);
            if (idx < s.materialCount && idx < static_cast<uint32_t>(kMaxGradMaterials)) {
                mi = static_cast<int>(idx);
                lm = live_of(s.materials[idx], c.m, s);
            }
        }
        float acc[kGradParams] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};

        // emission
        L = add(L, mul(T, mk(c.m.r * c.m.emission, c.m.g * c.m.emission, c.m.b * c.m.emission)));
        if (lm.emis) {
            acc[6] += gw.x * T.x * c.m.r + gw.y * T.y * c.m.g + gw.z * T.z * c.m.b;
        }
        if (lm.rgb) {
            acc[0] += gw.x * T.x * c.m.emission;
            acc[1] += gw.y * T.y * c.m.emission;
            acc[2] += gw.z * T.z * c.m.emission;
        }
        // environment NEE (env_nee)
        if (!skips_env_nee(c.m)) {
            const float u1 = rand01(rng);
            const float u2 = rand01(rng);
            float lightPdf = 0.0f;
            const V3 wi = env_sample(env, u1, u2, lightPdf);
            if (lightPdf > kMinPdf && !trace_any(s, add(pos, scale(c.n, kShadowBias)), wi, kShadowBias, kRayTMax,
                                                  nullptr)) {
                const float cosT = max2(0.0f, dot(c.n, wi));
                if (cosT > 0.0f) {
                    const V3 Le = env_eval(env, p, wi);
                    const V3 f = brdf_eval(c, wi);
                    const float mis = mis_balance(lightPdf, brdf_pdf(c, wi));
                    const V3 w = scale(Le, mis * cosT / lightPdf);
                    L = add(L, mul(T, mul(f, w)));
                    if (mi >= 0) {
                        brdf_grad(c, wi, mul(gw, mul(T, w)), lm, acc);
                    }
                }
            }
        }
        // emissive-triangle NEE (emissive_nee)
        if (s.emissiveTriangleCount != 0) {
            const float uTri = rand01(rng);
            const float u1 = rand01(rng);
            const float u2 = rand01(rng);
            V3 lp, ln, lr;
            float areaPdf = 0.0f;
            if (sample_emissive(s, uTri, u1, u2, lp, ln, lr, areaPdf)) {
                const V3 toL = sub(lp, pos);
                const float d2 = dot(toL, toL);
                if (d2 > kMinPdf) {
                    const V3 wi = normalize(toL);
                    const float cosT = max2(0.0f, dot(c.n, wi));
                    const float cosL = max2(0.0f, dot(ln, scale(wi, -1.0f)));
                    if (cosT > 0.0f && cosL > 0.0f) {
                        const float lightPdf = areaPdf * d2 / cosL;
                        if (lightPdf > kMinPdf &&
                            !trace_any(s, add(pos, scale(wi, kShadowBias)), wi, kShadowBias, kRayTMax, nullptr)) {
                            const V3 f = brdf_eval(c, wi);
                            const float mis = mis_balance(lightPdf, brdf_pdf(c, wi));
                            const V3 w = scale(lr, mis * cosT / lightPdf);
                            L = add(L, mul(T, mul(f, w)));
                            if (mi >= 0) {
                                brdf_grad(c, wi, mul(gw, mul(T, w)), lm, acc);
                            }
                        }
                    }
                }
            }
        }
        // BSDF sampling; everything gathered later is linear in this vertex's BSDF value
        c.flags = flags;
        const float u1 = rand01(rng);
        const float u2 = rand01(rng);
        const Sample smp = brdf_sample(c, u1, u2);
        if (!smp.valid || smp.pdf <= kMinPdf) {
            if (mi >= 0) {
                flush(lg, mi, acc);
            }
            break;
        }
        const V3 f = brdf_eval(c, smp.dir);
        if (mi >= 0) {
            const V3 R = sub(C, L);
            const V3 a = mk(f.x != 0.0f ? gw.x * R.x / f.x : 0.0f, f.y != 0.0f ? gw.y * R.y / f.y : 0.0f,
                            f.z != 0.0f ? gw.z * R.z / f.z : 0.0f);
            brdf_grad(c, smp.dir, a, lm, acc);
            flush(lg, mi, acc);
        }
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
        hit = trace_closest(s, origin, dir, bias, kRayTMax, nullptr);
        if (!hit.hit) {
            break;
        }
    }
}

__global__ __launch_bounds__(kAdBlock) void adjoint_kernel(const float* __restrict__ gradImg, int width, int height,
                                                           int spp, const CameraGpu* camera,
                                                           const MeshAccelSceneGpu* scene,
                                                           const EnvironmentMapGpu* env,
                                                           const RenderParamsGpu* params, int materialCount,
                                                           const float* __restrict__ sampleRadiance,
                                                           double* __restrict__ grad, int copies)
{
    // `copies` replicas of the per-material sums in dynamic LDS; lane l adds to replica l % copies
    extern __shared__ double lgAll[];
    const int perCopy = materialCount * kGradParams;
    for (int k = static_cast<int>(threadIdx.x); k < copies * perCopy; k += kAdBlock) {
        lgAll[k] = 0.0;
    }
    double* lg = lgAll + (static_cast<int>(threadIdx.x) % copies) * perCopy;
    __syncthreads();
    const MeshAccelSceneGpu s = *scene;
    const RenderParamsGpu p = *params;
    const CameraGpu cam = *camera;
    const long long total = static_cast<long long>(width) * height * spp;
    const long long stride = static_cast<long long>(gridDim.x) * kAdBlock;
    const float invSpp = 1.0f / static_cast<float>(spp);
    for (long long t = static_cast<long long>(blockIdx.x) * kAdBlock + threadIdx.x; t < total; t += stride) {
        const int idx = static_cast<int>(t / spp);
        const int k = static_cast<int>(t - static_cast<long long>(idx) * spp);
        const V3 gs = mk(gradImg[3 * idx] * invSpp, gradImg[3 * idx + 1] * invSpp, gradImg[3 * idx + 2] * invSpp);
        if (gs.x == 0.0f && gs.y == 0.0f && gs.z == 0.0f) {
            continue;
        }
        const int x = idx % width;
        const int y = idx / width;
        Rng rng;
        rng_init(rng, idx, k, static_cast<uint32_t>(k) + 1u);
        const float ju = rand01(rng);
        const float jv = rand01(rng);
        const float u = (static_cast<float>(x) + ju) / static_cast<float>(width);
        const float v = (static_cast<float>(y) + jv) / static_cast<float>(height);
        const float l1 = rand01(rng);
        const float l2 = rand01(rng);
        V3 ro, rd;
        camera_ray(cam, u, v, l1, l2, ro, rd);
        Rng replay = rng;
        V3 C;
        if (sampleRadiance != nullptr) {
            const size_t q = 3 * static_cast<size_t>(t);
            C = mk(sampleRadiance[q], sampleRadiance[q + 1], sampleRadiance[q + 2]);
        } else {
            C = trace_path(s, p, env, ro, rd, rng, nullptr);
        }
        const V3 gw = clamp_vjp(C, gs);
        if (gw.x != 0.0f || gw.y != 0.0f || gw.z != 0.0f) {
            adjoint_walk(s, p, env, ro, rd, replay, C, gw, lg);
        }
    }
    __syncthreads();
    for (int k = static_cast<int>(threadIdx.x); k < perCopy; k += kAdBlock) {
        double sum = 0.0;
        for (int c = 0; c < copies; ++c) {
            sum += lgAll[c * perCopy + k];
        }
        if (sum != 0.0) {
            atomicAdd(&grad[k], sum);
        }
    }
}

// Radiance of the detached estimator: sampling (directions, pdfs, MIS, Russian roulette from the
// primal throughput T0) with the scene's materials, values (emission, BSDF values) with evalMats.
// Finite differences of this function in evalMats are what the adjoint computes.
PT_DEV V3 trace_path_detached(const MeshAccelSceneGpu& s, const MaterialGpu* evalMats, const RenderParamsGpu& p,
                              const EnvironmentMapGpu* env, V3 ro, V3 rd, Rng& rng)
{
    Hit hit = trace_closest(s, ro, rd, 0.0f, kRayTMax, nullptr);
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
    MeshAccelSceneGpu se = s;
    se.materials = evalMats;
    V3 L = mk(0.0f, 0.0f, 0.0f);
    V3 T = mk(1.0f, 1.0f, 1.0f);
    V3 T0 = T;
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
        Ctx ce = c;
        ce.m = resolve_material(se, hit.tri, hit.uvx, hit.uvy);
        ce.w = lobe_weights(ce.m);
        L = add(L, mul(T, mk(ce.m.r * ce.m.emission, ce.m.g * ce.m.emission, ce.m.b * ce.m.emission)));
        if (!skips_env_nee(c.m)) {
            const float u1 = rand01(rng);
            const float u2 = rand01(rng);
            float lightPdf = 0.0f;
            const V3 wi = env_sample(env, u1, u2, lightPdf);
            if (lightPdf > kMinPdf && !trace_any(s, add(pos, scale(c.n, kShadowBias)), wi, kShadowBias, kRayTMax,
                                                  nullptr)) {
                const float cosT = max2(0.0f, dot(c.n, wi));
                if (cosT > 0.0f) {
                    const float mis = mis_balance(lightPdf, brdf_pdf(c, wi));
                    L = add(L, mul(T, mul(brdf_eval(ce, wi), scale(env_eval(env, p, wi), mis * cosT / lightPdf))));
                }
            }
        }
        if (s.emissiveTriangleCount != 0) {
            const float uTri = rand01(rng);
            const float u1 = rand01(rng);
            const float u2 = rand01(rng);
            V3 lp, ln, lr;
            float areaPdf = 0.0f;
            if (sample_emissive(s, uTri, u1, u2, lp, ln, lr, areaPdf)) {
                const V3 toL = sub(lp, pos);
                const float d2 = dot(toL, toL);
                if (d2 > kMinPdf) {
                    const V3 wi = normalize(toL);
                    const float cosT = max2(0.0f, dot(c.n, wi));
                    const float cosL = max2(0.0f, dot(ln, scale(wi, -1.0f)));
                    if (cosT > 0.0f && cosL > 0.0f) {
                        const float lightPdf = areaPdf * d2 / cosL;
                        if (lightPdf > kMinPdf &&
                            !trace_any(s, add(pos, scale(wi, kShadowBias)), wi, kShadowBias, kRayTMax, nullptr)) {
                            const float mis = mis_balance(lightPdf, brdf_pdf(c, wi));
                            L = add(L, mul(T, mul(brdf_eval(ce, wi), scale(lr, mis * cosT / lightPdf))));
                        }
                    }
                }
            }
        }
        c.flags = flags;
        const float u1 = rand01(rng);
        const float u2 = rand01(rng);
        const Sample smp = brdf_sample(c, u1, u2);
        if (!smp.valid || smp.pdf <= kMinPdf) {
            break;
        }
        const V3 before = T0;
        T0 = apply_throughput(T0, c, smp, brdf_eval(c, smp.dir));
        T = apply_throughput(T, c, smp, brdf_eval(ce, smp.dir));
        const V3 mid = T0;
        T0 = subsurface_tier_b(T0, c.m, smp, rng);
        T = mk(mid.x != 0.0f ? T.x * (T0.x / mid.x) : 0.0f, mid.y != 0.0f ? T.y * (T0.y / mid.y) : 0.0f,
               mid.z != 0.0f ? T.z * (T0.z / mid.z) : 0.0f);
        (void)before;
        if ((flags & 16) != 0) {
            const V3 tint = smp.transmitted ? mk(0.2f, 1.0f, 0.2f) : mk(1.0f, 0.2f, 0.2f);
            T0 = mul(T0, tint);
            T = mul(T, tint);
        }
        const float tl = 0.2126f * T0.x + 0.7152f * T0.y + 0.0722f * T0.z;
        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
            break;
        }
        if (depth >= rrMin) {
            const float surv = min2(0.95f, max2(tl, 0.05f));
            if (rand01(rng) > surv) {
                break;
            }
            T0 = scale(T0, 1.0f / surv);
            T = scale(T, 1.0f / surv);
        }
        eta = smp.nextEta;
        const float bias = max2(kShadowBias, kRelShadowBias * hit.t);
        const float sgn = dot(c.n, smp.dir) >= 0.0f ? 1.0f : -1.0f;
        origin = add(pos, scale(c.n, sgn * bias));
        dir = smp.dir;
        hit = trace_closest(s, origin, dir, bias, kRayTMax, nullptr);
        if (!hit.hit) {
            const float mis = mis_balance(smp.pdf, env_pdf(env, dir));
            L = add(L, mul(T, scale(env_eval(env, p, dir), mis)));
            break;
        }
    }
    return L;
}

// One thread per pixel, samples in order, running mean of the clamped radiance (as accumulate).
__global__ __launch_bounds__(kAdBlock) void detached_kernel(float* __restrict__ img, int width, int height, int spp,
                                                            const CameraGpu* camera, const MeshAccelSceneGpu* scene,
                                                            const MaterialGpu* evalMats,
                                                            const EnvironmentMapGpu* env,
                                                            const RenderParamsGpu* params)
{
    const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (idx >= width * height) {
        return;
    }
    const MeshAccelSceneGpu s = *scene;
    const RenderParamsGpu p = *params;
    const CameraGpu cam = *camera;
    const int x = idx % width;
    const int y = idx / width;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int k = 0; k < spp; ++k) {
        Rng rng;
        rng_init(rng, idx, k, static_cast<uint32_t>(k) + 1u);
        const float ju = rand01(rng);
        const float jv = rand01(rng);
        const float u = (static_cast<float>(x) + ju) / static_cast<float>(width);
        const float v = (static_cast<float>(y) + jv) / static_cast<float>(height);
        const float l1 = rand01(rng);
        const float l2 = rand01(rng);
        V3 ro, rd;
        camera_ray(cam, u, v, l1, l2, ro, rd);
        const V3 r = clamp_firefly(trace_path_detached(s, evalMats, p, env, ro, rd, rng));
        const float inv = 1.0f / static_cast<float>(k + 1);
        ax = (ax * static_cast<float>(k) + r.x) * inv;
        ay = (ay * static_cast<float>(k) + r.y) * inv;
        az = (az * static_cast<float>(k) + r.z) * inv;
    }
    img[3 * idx] = ax;
    img[3 * idx + 1] = ay;
    img[3 * idx + 2] = az;
}

} // namespace

extern "C" bool pathTracerDetachedRender(float* d_image, int width, int height, int samplesPerPixel,
                                         const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                         const MaterialGpu* d_evalMaterials, const EnvironmentMapGpu* d_env,
                                         const RenderParamsGpu* d_params, hipStream_t stream)
{
    if (d_image == nullptr || d_camera == nullptr || d_scene == nullptr || d_evalMaterials == nullptr ||
        d_params == nullptr || width <= 0 || height <= 0 || samplesPerPixel <= 0) {
        ptSetError("pathTracerDetachedRender: invalid arguments");
        return false;
    }
    if (pt_spectral_enabled()) {
        ptSetError("pathTracerDetachedRender: RGB rendering only (hero-wavelength sampling is on)");
        return false;
    }
    const long long n = static_cast<long long>(width) * height;
    hipLaunchKernelGGL(detached_kernel, dim3(static_cast<unsigned>((n + kAdBlock - 1) / kAdBlock)), dim3(kAdBlock), 0,
                       stream, d_image, width, height, samplesPerPixel, d_camera, d_scene, d_evalMaterials, d_env,
                       d_params);
    if (hipGetLastError() != hipSuccess) {
        ptSetError("pathTracerDetachedRender: launch failed");
        return false;
    }
    return true;
}

extern "C" bool pathTracerAdjointRender(const float* d_gradImage, int width, int height, int samplesPerPixel,
                                        const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                        const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                        int materialCount, const float* d_sampleRadiance, double* d_grad,
                                        hipStream_t stream)
{
    if (d_gradImage == nullptr || d_camera == nullptr || d_scene == nullptr || d_params == nullptr ||
        d_grad == nullptr || width <= 0 || height <= 0 || samplesPerPixel <= 0 || materialCount <= 0 ||
        materialCount > kMaxGradMaterials) {
        ptSetError("pathTracerAdjointRender: invalid arguments (materialCount must be 1..%d)", kMaxGradMaterials);
        return false;
    }
    if (pt_spectral_enabled()) {
        ptSetError("pathTracerAdjointRender: RGB rendering only (hero-wavelength sampling is on)");
        return false;
    }
    int device = 0, cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    const long long total = static_cast<long long>(width) * height * samplesPerPixel;
    const long long blocks = std::min<long long>((total + kAdBlock - 1) / kAdBlock, static_cast<long long>(cus) * 16);
    // LDS replicas of the per-material sums: up to 16, within 24 KiB per block
    int copies = 16;
    while (copies > 1 && static_cast<size_t>(copies) * materialCount * kGradParams * sizeof(double) > 24576) {
        copies >>= 1;
    }
    const size_t lds = static_cast<size_t>(copies) * materialCount * kGradParams * sizeof(double);
    hipLaunchKernelGGL(adjoint_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kAdBlock), lds, stream, d_gradImage,
                       width, height, samplesPerPixel, d_camera, d_scene, d_env, d_params, materialCount,
                       d_sampleRadiance, d_grad, copies);
    if (hipGetLastError() != hipSuccess) {
        ptSetError("pathTracerAdjointRender: launch failed");
        return false;
    }
    return true;
}
