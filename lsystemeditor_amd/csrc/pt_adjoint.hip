// pt_adjoint.hip -- the differentiable-rendering entry points: the adjoint (path replay) gradients
// of an image loss with respect to material parameters (argument checks here; the replay runs on
// the wavefront scheduler in pt_wavefront.hip), and the detached estimator whose finite
// differences the adjoint computes.
#include "pt_device.h"
#include "pt_adjoint_math.h"
#include "pt_integrator.h"
#include "pt_internal.h"

#include "../../include/pt_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>

using namespace pt;

namespace {

constexpr int kAdBlock = 256;

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
    if (pt_light_mode() != 0) {
        ptSetError("pathTracerDetachedRender: the reference light sampling only (pathTracerSetLightSampling(0))");
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

extern "C" bool pathTracerAdjointRenderSeeded(const float* d_gradImage, int width, int height, int samplesPerPixel,
                                              const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                              const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                              int materialCount, const float* d_sampleRadiance, double* d_grad,
                                              unsigned int firstSeed, hipStream_t stream)
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
    if (pt_light_mode() != 0) {
        ptSetError("pathTracerAdjointRender: the reference light sampling only (pathTracerSetLightSampling(0))");
        return false;
    }
    return pt_adjoint_wavefront(d_gradImage, width, height, samplesPerPixel, d_camera, d_scene, d_env, d_params,
                                materialCount, d_sampleRadiance, d_grad, firstSeed, stream);
}

extern "C" bool pathTracerAdjointRenderSpectral(const float* d_gradImage, int width, int height, int samplesPerPixel,
                                                const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                                const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                                int materialCount, const float* d_sampleRadiance, double* d_grad,
                                                unsigned int firstSeed, hipStream_t stream)
{
    if (d_gradImage == nullptr || d_camera == nullptr || d_scene == nullptr || d_params == nullptr ||
        d_grad == nullptr || width <= 0 || height <= 0 || samplesPerPixel <= 0 || materialCount <= 0 ||
        materialCount > kMaxGradMaterials) {
        ptSetError("pathTracerAdjointRenderSpectral: invalid arguments (materialCount must be 1..%d)",
                   kMaxGradMaterials);
        return false;
    }
    if (!pt_spectral_enabled()) {
        ptSetError("pathTracerAdjointRenderSpectral: hero-wavelength rendering only (pathTracerSetSpectralSampling)");
        return false;
    }
    if (pt_light_mode() != 0) {
        ptSetError("pathTracerAdjointRenderSpectral: the reference light sampling only (pathTracerSetLightSampling(0))");
        return false;
    }
    return pt_adjoint_wavefront(d_gradImage, width, height, samplesPerPixel, d_camera, d_scene, d_env, d_params,
                                materialCount, d_sampleRadiance, d_grad, firstSeed, stream);
}

extern "C" bool pathTracerAdjointRender(const float* d_gradImage, int width, int height, int samplesPerPixel,
                                        const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                        const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                        int materialCount, const float* d_sampleRadiance, double* d_grad,
                                        hipStream_t stream)
{
    return pathTracerAdjointRenderSeeded(d_gradImage, width, height, samplesPerPixel, d_camera, d_scene, d_env,
                                         d_params, materialCount, d_sampleRadiance, d_grad, 1u, stream);
}
