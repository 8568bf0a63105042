/*
 * oracle.h -- CPU restatement of PathTracerLib's FP32 path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls
 * this code. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker / the timed CPU baseline.
 *
 * Every function restates the reference's arithmetic operation for operation
 * (unfused multiply-add, same evaluation order) and cites the file:line it follows
 * (paths relative to the reference repository root). Two deliberate replacements:
 *   - RNG: cuRAND XORWOW (PathTracerLib/src/Sampling/RandCore.h:7-23) is replaced by
 *     Philox-4x32-10 (Salmon et al., SC'11) keyed by the reference's seed hash.
 *   - sinf/cosf/tanf/expf/logf/atan2f/acosf are the project's double-precision
 *     evaluations rounded to float (Cody-Waite reduction + Taylor series), so the
 *     CPU and GPU produce identical bits. Their accuracy against libm is tested.
 *
 * Parity status of each function is documented beside it (pinned by golden
 * vectors, closed forms or brute force in tests/test_oracle_*.py).
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POD types, byte-identical to the reference layouts ------------------- */
/* Geometry/GeometryTypes.h:3-14 */
typedef struct { float x, y, z; } OVec3;
typedef struct { float x, y; } OVec2;
typedef struct { _Alignas(16) float x; float y, z, w; } OFloat4;

/* MeshAccel/MeshAccelTypes.h:49-86 (96 B) */
typedef struct {
    _Alignas(16) float r;
    float g, b, roughness, metallic, transmission, thin, ior, subsurface, emission,
        diffuseRoughness, scatterRadiusR, scatterRadiusG, scatterRadiusB, specular;
    uint32_t albedoTex, roughnessTex, metallicTex, transmissionTex, thinTex, iorTex,
        subsurfaceTex, emissionTex;
} OMaterial;

/* MeshAccel/MeshAccelTypes.h:22-28 (64 B) */
typedef struct {
    _Alignas(16) uint32_t kind;
    OFloat4 p0, p1, p2;
} OTexDesc;

/* MeshAccel/MeshAccelTypes.h:90-102 (112 B) */
typedef struct {
    _Alignas(16) OVec3 v0;
    OVec3 v1, v2, n0, n1, n2;
    OVec2 uv0, uv1, uv2;
    uint32_t materialIndex;
} OTriangle;

/* MeshAccel/MeshAccelTypes.h:106-120 (48 B) */
typedef struct {
    _Alignas(16) OVec3 boundsMin;
    float pad0;
    OVec3 boundsMax;
    float pad1;
    uint32_t leftIndex, rightIndex, triStart, triCount;
} ONode;

/* MeshAccel/MeshAccelTypes.h:125-140 (80 B) */
typedef struct {
    _Alignas(16) const ONode* nodes;
    const OTriangle* triangles;
    const OMaterial* materials;
    const uint32_t* emissiveIndices;
    const float* emissiveCdf;
    const OTexDesc* textures;
    uint32_t nodeCount, triangleCount, rootIndex, materialCount, emissiveCount, textureCount;
    /* binned spectra for the hero-wavelength integrator (NULL: none): ORC_SPECTRUM_WORDS floats per
     * material, ORC_SPECTRUM_BINS reflectance values then ORC_SPECTRUM_BINS emission values
     * (MeshAccelSceneGpu::materialSpectra; no counterpart in the reference) */
    const float* spectra;
} OScene;
#define ORC_SPECTRUM_BINS 16
#define ORC_SPECTRUM_WORDS 32

/* CameraGpu.h:5-16 (80 B; float4 member forces 16-byte alignment) */
typedef struct {
    OVec3 position;
    OFloat4 orientation; /* (w, x, y, z) in (.x, .y, .z, .w) */
    float fovY, aspect, nearPlane, farPlane, apertureRadius, focusDistance;
    OVec3 focusPoint;
    int focusValid;
} OCamera;

/* RenderTypes.h:15-29 (44 B) */
typedef struct {
    float backgroundR, backgroundG, backgroundB, environmentIntensity;
    int maxPathDepth, russianRouletteMinDepth, minSamples, maxSamplesPerPixel,
        debugOverlayMode, brdfDebugFlags;
    float relativeErrorThreshold;
} OParams;

/* RenderTypes.h:5-13 */
typedef struct {
    const float* pixels;
    const float* marginalCdf;
    const float* rowCdf;
    int width, height, valid;
} OEnv;

/* MeshAccel/MeshAccelTypes.h:144-151 */
typedef struct {
    int hit;
    float t;
    OVec3 normal;
    OVec2 uv;
    uint32_t triangleIndex;
} OHit;

/* Brdf/BrdfBase.h:39-50 */
typedef struct {
    OVec3 normal, wo;
    OMaterial material;
    float etaMedium, wavelengthNm;
    int debugFlags;
} OBrdfCtx;

/* Brdf/BrdfBase.h:22-37 */
typedef struct {
    OVec3 direction;
    float pdf;
    int valid, transmitted;
    float nextMediumEta;
    int subsurfaceScatter;
    OVec3 subsurfaceExitOffset;
    int subsurfaceInternalSteps;
} OSample;

typedef struct { float diffuse, specular, transmit, subsurface; } OLobeWeights;

/* Counter-based RNG state (replaces curandState). */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int avail;
} ORng;

/* ---- project math (bit-identical on CPU and GPU) -------------------------- */
float orc_sinf(float x);
float orc_cosf(float x);
float orc_tanf(float x);
float orc_expf(float x);
float orc_logf(float x);
float orc_atan2f(float y, float x);
float orc_acosf(float x);
OVec3 orc_vec_normalize(OVec3 v);

/* ---- RNG ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_seed_hash(int pixelIndex, int sampleIndex, uint32_t globalSeed);
void orc_rng_init(ORng* rng, int pixelIndex, int sampleIndex, uint32_t globalSeed);
float orc_rand01(ORng* rng);
float orc_u32_to_unit(uint32_t x);

/* ---- BVH + traversal ------------------------------------------------------ */
int orc_bvh_build(const OTriangle* tris, uint32_t count, ONode* outNodes, uint32_t* outRoot);
int orc_ray_aabb(OVec3 ro, OVec3 rd, OVec3 invRd, OVec3 bmin, OVec3 bmax, float tMin, float tMax,
                 float* outEnter, float* outExit);
int orc_ray_triangle(OVec3 ro, OVec3 rd, const OTriangle* tri, float tMin, float tMax, float* outT,
                     OVec3* outNormal, OVec2* outUv);
OHit orc_trace_ray(OVec3 ro, OVec3 rd, const OScene* scene, float tMin, float tMax);
/* Same traversal, also counting node records fetched and triangles tested. */
OHit orc_trace_ray_counted(OVec3 ro, OVec3 rd, const OScene* scene, float tMin, float tMax,
                           uint64_t* nodeFetches, uint64_t* triTests);
uint32_t orc_build_emissive(const OTriangle* tris, uint32_t count, const OMaterial* mats,
                            uint32_t matCount, uint32_t* outIndices, float* outCdf);

/* ---- BRDF ----------------------------------------------------------------- */
OLobeWeights orc_lobe_weights(const OBrdfCtx* ctx);
OVec3 orc_brdf_eval(const OBrdfCtx* ctx, OVec3 wi);
float orc_brdf_pdf(const OBrdfCtx* ctx, OVec3 wi);
OSample orc_brdf_sample(const OBrdfCtx* ctx, float u1, float u2);
OSample orc_brdf_sample_transmit(const OBrdfCtx* ctx, float u1, float u2);
float orc_oren_nayar_ctx(const OBrdfCtx* ctx, OVec3 wi);
int orc_refract_direction(const OBrdfCtx* ctx, OVec3* outWi, float* etaI, float* etaT, float* nextEta);
int orc_skips_env_nee(const OMaterial* m);
OVec3 orc_apply_throughput(OVec3 throughput, const OBrdfCtx* ctx, const OSample* s, OVec3 bsdf);
float orc_ggx_d(float cosThetaH, float alpha);
float orc_smith_g1(float cosTheta, float alpha);
float orc_dielectric_fresnel(float cosThetaI, float etaI, float etaT);
OVec3 orc_schlick_f(float cosTheta, OVec3 f0);
float orc_energy_compensation(float roughness, float cosTheta);
float orc_oren_nayar_factor(float sigma, float cosI, float cosO, float dotWoWi);
/* Burley diffusion cores (SubsurfaceCore.h:20-64). */
float orc_burley_sample_radius(float u, float d);
float orc_burley_profile(float r, float d);
float orc_burley_pdf_radius(float r, float d);
float orc_subsurface_effective_radius(const OMaterial* m, float uChannel);
OVec3 orc_sample_cosine_local(float u1, float u2);
OVec3 orc_sample_ggx_half_local(float alpha, float u1, float u2, float* pdf);

/* ---- lights, MIS, materials, textures ------------------------------------- */
float orc_mis_balance(float pdfA, float pdfB);
OVec3 orc_lat_long_to_dir(float u, float v);
OVec3 orc_dir_to_lat_long(OVec3 d);
OVec3 orc_env_eval(const OEnv* env, const OParams* params, OVec3 dir);
float orc_env_pdf(const OEnv* env, const OParams* params, OVec3 dir);
OVec3 orc_env_sample(const OEnv* env, const OParams* params, float u1, float u2, float* pdf);
int orc_light_sample_emissive(const OScene* scene, float uTri, float u1, float u2, OVec3* pos,
                              OVec3* normal, OVec3* radiance, float* pdf);
int orc_light_sample_emissive_tri(const OScene* scene, float uTri, float u1, float u2, OVec3* pos, OVec3* normal,
                                  OVec3* radiance, float* pdf, uint32_t* triOut);
OMaterial orc_fetch_material(const OScene* scene, uint32_t tri);
OMaterial orc_resolve_material(const OMaterial* m, OVec2 uv, const OScene* scene);
int orc_env_build_tables(const float* rgb, int width, int height, float* marginalCdf, float* rowCdf);

/* ---- camera --------------------------------------------------------------- */
void orc_camera_ray(const OCamera* cam, float u, float v, float lensU1, float lensU2, OVec3* ro, OVec3* rd);

/* ---- integrator ----------------------------------------------------------- */
OVec3 orc_trace_path(OVec3 ro, OVec3 rd, const OScene* scene, const OParams* params, const OEnv* env, ORng* rng);
OVec3 orc_clamp_firefly(OVec3 radiance);
OVec3 orc_pixel_radiance(int idx, int x, int y, int width, int height, int jitter, int sampleIndex,
                         const OCamera* cam, const OScene* scene, const OEnv* env, const OParams* params,
                         uint32_t globalSeed);
/* Per-sample traversal statistics of orc_pixel_radiance, for the algorithmic byte model. */
typedef struct {
    uint64_t rays, closestRays, shadowRays, nodeFetches, triTests, hits;
} OTraceStats;
void orc_stats_begin(OTraceStats* stats);
void orc_stats_end(void);

/* ---- adaptive accumulation (PathTracer.cu:186-411) ------------------------ */
int orc_is_converged(uint32_t n, float lumMean, float m2, int minSamples, float threshold);
void orc_clear_accumulator(float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* converged,
                           int* activeIdx, int* activeCount, int width, int height);
void orc_adaptive_sample(float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* converged,
                         const int* activeIdx, int activeCount, int width, int height, const OCamera* cam,
                         const OScene* scene, const OEnv* env, const OParams* params, uint32_t globalSeed,
                         int threads);
int orc_compact_active(const int* in, int* out, const uint8_t* converged, const uint32_t* counts,
                       int maxSpp, int inCount, int width, int useRegion, int rx0, int ry0, int rx1, int ry1);
void orc_preview_sample(float* acc4, int width, int height, const OCamera* cam, const OScene* scene,
                        const OEnv* env, const OParams* params, uint32_t globalSeed, int threads);
/* Runs the renderLoop schedule (PathTracer.cpp:2220-2543) without previews:
 * passes until no pixel is active or maxPasses is reached. Returns passes run. */
int orc_render_adaptive(float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* converged,
                        int width, int height, const OCamera* cam, const OScene* scene, const OEnv* env,
                        const OParams* params, int firstIter, int maxPasses, int threads);
void orc_copy_to_pbo(const float* acc4, const uint32_t* counts, const uint8_t* converged, uint8_t* rgba,
                     int outW, int outH, int srcW, int srcH, int downscale, int useAcc, const OParams* params,
                     float exposure);
/* ---- differentiable rendering (detached sampling, adjoint path replay) ----- */
/* Radiance of one path whose sampling uses the scene's materials and whose values use evalMats
 * (NULL: the scene's, giving orc_trace_path's result). With gw != NULL the path is replayed as
 * the adjoint: C is its primal radiance, gw = dLoss/dRadiance (3 doubles), and
 * grad[7 * material + (r, g, b, roughness, metallic, specular, emission)] accumulates. */
OVec3 orc_trace_path_ad(OVec3 ro, OVec3 rd, const OScene* scene, const OMaterial* evalMats, const OParams* params,
                        const OEnv* env, ORng* rng, const double* gw, OVec3 C, double* grad);
/* Contracted derivative of the opaque Principled BSDF value: out6[k] += sum_c a3[c] df_c/dtheta_k,
 * theta = (r, g, b, roughness, metallic, specular); raw = the unresolved material (live mask). */
void orc_brdf_grad(const OBrdfCtx* ctx, OVec3 wi, const double* a3, const OMaterial* raw, double* out6);
/* The value orc_brdf_grad differentiates, in double (f3 = 3 doubles). */
void orc_brdf_eval_f64(const OBrdfCtx* ctx, OVec3 wi, const OMaterial* raw, double* f3);
void orc_clamp_vjp(OVec3 C, const double* g3, double* out3);
void orc_render_detached(float* img3, int width, int height, const OCamera* cam, const OScene* scene,
                         const OMaterial* evalMats, const OEnv* env, const OParams* params, int spp, int threads);
void orc_render_adjoint(const double* gradImg3, int width, int height, const OCamera* cam, const OScene* scene,
                        const OEnv* env, const OParams* params, int spp, double* grad, int threads);
/* The detached image and the adjoint over the samples of seeds firstSeed .. firstSeed + spp - 1. */
void orc_render_detached_seeded(float* img3, int width, int height, const OCamera* cam, const OScene* scene,
                                const OMaterial* evalMats, const OEnv* env, const OParams* params, int spp,
                                uint32_t firstSeed, int threads);
void orc_render_adjoint_seeded(const double* gradImg3, int width, int height, const OCamera* cam,
                               const OScene* scene, const OEnv* env, const OParams* params, int spp,
                               uint32_t firstSeed, double* grad, int threads);
/* The detached estimator evaluated in double, added into img3; mask3 (or NULL) selects pixels. */
void orc_render_detached_f64_seeded(double* img3, int width, int height, const OCamera* cam, const OScene* scene,
                                    const OMaterial* evalMats, const OEnv* env, const OParams* params, int spp,
                                    uint32_t firstSeed, const double* mask3, int threads);

/* ---- spectral (Spectral/Rgb2Spec.h, SpectralCore.h, SpectralState.cpp) ----------------------- */
typedef struct {
    const float* scale;  /* res entries */
    const float* data;   /* 3 * 3 * res^3 coefficients */
    int res;
    float white[3];      /* white normalisation */
    float lambdaMin, lambdaRange; /* hero-wavelength range */
} OSpecModel;
/* Differentiable hero-wavelength rendering (light mode 0): the spectral detached estimator's image
 * (img3 float, or imgD double added into with mask3 selecting pixels) and its FP64 adjoint,
 * ORC_SPECTRAL_GRAD_PARAMS per material: reflectance bins, emission bins, roughness, metallic,
 * specular, emission scale. */
#define ORC_SPECTRAL_GRAD_PARAMS 36
void orc_render_detached_spectral(float* img3, double* imgD, const double* mask3, int width, int height,
                                  const OCamera* cam, const OScene* scene, const OMaterial* evalMats,
                                  const float* evalSpectra, const OEnv* env, const OParams* params,
                                  const OSpecModel* model, int spp, uint32_t firstSeed, int threads);
void orc_render_adjoint_spectral(const double* gradImg, int width, int height, const OCamera* cam,
                                 const OScene* scene, const OEnv* env, const OParams* params,
                                 const OSpecModel* model, int spp, uint32_t firstSeed, double* grad, int threads);
float orc_cie_interp(int which, float lambdaNm); /* 0 x, 1 y, 2 z, 3 D65 */
void orc_cie_rgb_response(float lambdaNm, float* out3);
void orc_rgb2spec_coeffs(const OSpecModel* m, float r, float g, float b, float* out3);
float orc_rgb2spec_sigmoid(const float* c3, float lambdaNm);
float orc_rgb2spec_reflectance(const OSpecModel* m, float r, float g, float b, float lambdaNm);
float orc_rgb2spec_radiance(const OSpecModel* m, float r, float g, float b, float lambdaNm);
void orc_spectral_white_norm(const OSpecModel* m, float* out3);
void orc_spectral_to_rgb(const OSpecModel* m, float L, float lambdaNm, float pdf, float* out3);
void orc_hero_lambdas(const OSpecModel* m, float u, float* out4);
OVec3 orc_hero_to_rgb(const OSpecModel* m, const float* L4, const float* lambda4);
OVec3 orc_clamp_firefly_spectral(OVec3 r);
void orc_trace_path_spectral(OVec3 ro, OVec3 rd, const OScene* scene, const OParams* params, const OEnv* env,
                             ORng* rng, const OSpecModel* model, const float* lambda4, float* L4);
/* Hero-wavelength sampling for orc_pixel_radiance and everything built on it (NULL: RGB). The
 * hero draw follows the lens draws. Process-wide; not for concurrent use with other settings. */
/* Light sampling of orc_trace_path / orc_trace_path_spectral: 0 the reference's, 1 MIS on area
 * lights (bounded emissive shadow rays, balance-weighted emission at BSDF hits). */
void orc_set_light_sampling(int mode);
/* Test hook: the solid-angle pdf with which emissive NEE picks the point at distance t along dir on
 * triangle tri (0 where it never does). */
float orc_emissive_pdf_hit(const OScene* scene, uint32_t tri, OVec3 dir, float t);
void orc_set_spectral(const OSpecModel* model);

/* UV debug view: pinhole primary ray, fractional hit UV as red/green. */
/* Region reset (acc w = 1) listing the region row-major as the active set; returns its size. */
int orc_clear_region(float* acc4, uint32_t* counts, float* mean, float* m2, uint8_t* converged, int* active,
                     int width, int minX, int minY, int maxX, int maxY);
/* Active pixels (not converged, below maxSamplesPerPixel when > 0) in ascending order. */
int orc_rebuild_active(const uint8_t* converged, const uint32_t* counts, int* out, int pixelCount, int maxSpp);
void orc_uv_debug(const OCamera* cam, const OScene* scene, uint8_t* rgba, int width, int height,
                  const OParams* params);

/* Procedural mesh (orc_mesh.c): the reference's lofted tube of a frame list (13 floats per frame:
 * position, tangent, normal, binormal, radius) as 25-word HostTriangles, and its root transform. */
int orc_loft_triangles(const float* frames, int frameCount, int circularSegments, float* out);
void orc_root_transform(float* tris, int count, const float translation[3], const float rotationDeg[3]);

#ifdef __cplusplus
}
#endif
#endif
