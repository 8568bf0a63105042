/*
 * pt_hip.h -- C ABI of the MI355X (gfx950) path tracer.
 *
 * Drop-in for PathTracerLib's CUDA launcher API (PathTracerLib/src/PathTracerCuda.h:12-126)
 * and its data-entry boundary (MeshAccel/MeshAccelScene.cpp:141-194). Every launcher keeps
 * the reference's name, argument order, argument meaning, validation order and `bool`
 * result; CUDA types are replaced by plain PODs of identical size and `hipStream_t`.
 * All pointers named d_* are device pointers owned by the caller; calls are asynchronous
 * on the given stream. On failure a launcher returns false before launching anything and
 * pathTracerLastError() describes why.
 *
 * Layout note: MeshAccelSceneGpu is produced by ptSceneBuild() and consumed by the
 * launchers. It keeps the reference's role (one device struct describing the scene) but
 * not its field list: the BVH is stored as 64-byte internal nodes holding both child boxes,
 * triangle positions are packed in BVH leaf order, and shading attributes are separate.
 */
#ifndef PT_HIP_H
#define PT_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__cplusplus)
#define PT_ALIGN16 alignas(16)
#else
#define PT_ALIGN16 _Alignas(16)
#endif

/* ---- POD types (byte-identical to the reference where the reference defines them) ---- */

/* Geometry/GeometryTypes.h:3-14 */
typedef struct { float x, y, z; } Vec3;
typedef struct { float x, y; } Vec2;
/* Replaces CUDA's float4 / uchar4 (same size and alignment). */
typedef struct { PT_ALIGN16 float x; float y, z, w; } PtFloat4;
typedef struct { uint8_t x, y, z, w; } PtUchar4;

/* MeshAccel/MeshAccelTypes.h:49-86 (96 bytes) */
typedef struct {
    PT_ALIGN16 float r;
    float g, b, roughness, metallic, transmission, thin, ior, subsurface, emission, diffuseRoughness,
        scatterRadiusR, scatterRadiusG, scatterRadiusB, specular;
    uint32_t albedoTex, roughnessTex, metallicTex, transmissionTex, thinTex, iorTex, subsurfaceTex,
        emissionTex;
} MaterialGpu;

/* MeshAccel/MeshAccelTypes.h:22-28 (64 bytes); kind: 0 ConstantScalar, 1 ConstantRgb, 2 Grid2D, 3 Stripe1D */
typedef struct {
    PT_ALIGN16 uint32_t kind;
    PtFloat4 p0, p1, p2;
} TextureDescGpu;

/* Application/src/MeshBuilder/HostMesh.h:9-21 (100 bytes, packed floats) */
typedef struct {
    Vec3 v0, v1, v2, n0, n1, n2;
    Vec2 uv0, uv1, uv2;
    uint32_t materialIndex;
} HostTriangle;

/* MeshAccel/MeshAccelTypes.h:106-120 (48 bytes) -- reference BVH node layout, used only to
 * export the tree for inspection and parity checks (ptSceneReadBvh). */
typedef struct {
    PT_ALIGN16 Vec3 boundsMin;
    float pad0;
    Vec3 boundsMax;
    float pad1;
    uint32_t leftIndex, rightIndex, triStart, triCount;
} MeshBvhNode;

/* CameraGpu.h:5-16 (80 bytes) */
typedef struct {
    Vec3 position;
    PtFloat4 orientation; /* quaternion (w, x, y, z) stored in (.x, .y, .z, .w) */
    float fovY, aspect, nearPlane, farPlane, apertureRadius, focusDistance;
    Vec3 focusPoint;
    int focusValid;
} CameraGpu;

/* RenderTypes.h:15-29 (44 bytes); reference defaults: background 10/255, intensity 1,
 * depth 32, RR min depth 3, minSamples 16, maxSamplesPerPixel 1024, threshold 0.02 */
typedef struct {
    float backgroundR, backgroundG, backgroundB, environmentIntensity;
    int maxPathDepth, russianRouletteMinDepth, minSamples, maxSamplesPerPixel, debugOverlayMode,
        brdfDebugFlags;
    float relativeErrorThreshold;
} RenderParamsGpu;

/* RenderTypes.h:5-13: lat-long RGB float image with importance CDFs
 * (marginalCdf: height+1 floats, rowCdf: height*(width+1) floats). */
typedef struct {
    const float* pixels;
    const float* marginalCdf;
    const float* rowCdf;
    int width, height, valid;
} EnvironmentMapGpu;

/* Spectral/SpectralState.h:11-19 */
typedef struct {
    const float* rgb2specScale;
    const float* rgb2specData;
    int rgb2specRes;
    float whiteNormR, whiteNormG, whiteNormB;
} SpectralStateGpu;

/* Device scene description (role of MeshAccel/MeshAccelTypes.h:125-140). */
typedef struct {
    PT_ALIGN16 const PtFloat4* nodes;   /* internal nodes, 4 x float4 each (64 B):
                                           (lmin.xyz, lmax.x) (lmax.yz, rmin.xy) (rmin.z, rmax.xyz)
                                           (leftRef, rightRef, -, -) as int bits; ref < 0 is leaf slot ~ref */
    const PtFloat4* triPos;              /* 3 x float4 per leaf slot: (v0.xyz, original index bits),
                                           (v1-v0, material index bits), (v2-v0, -) */
    const PtFloat4* triShade;            /* 4 x float4 per original triangle:
                                           (n0.xyz, n1.x) (n1.yz, n2.xy) (n2.z, uv0.xy, uv1.x)
                                           (uv1.y, uv2.xy, materialIndex bits) */
    const MaterialGpu* materials;
    const uint32_t* emissiveTriangleIndices;
    const float* emissiveTriangleCdf;    /* emissiveTriangleCount + 1 floats */
    const PtFloat4* emissiveTris;        /* 3 x float4 per emissive entry: (v0, r*e) (v1, g*e) (v2, b*e) */
    const TextureDescGpu* textures;
    uint32_t nodeCount, triangleCount, materialCount, emissiveTriangleCount, textureCount;
    int rootRef;                          /* 0 = internal node 0; < 0 = single leaf ~rootRef */
    float rootMin[3];
    float rootMax[3];
    const PtFloat4* wideNodes;           /* 4-wide tree of the SAH build (NULL otherwise): 8 x float4 per
                                           node, min x, min y, min z, max x, max y, max z of the four
                                           children, then their references (int bits: >= 0 wide node,
                                           < 0 leaf slot ~ref, INT_MIN none), one unused */
    uint32_t wideNodeCount;
    int wideRootRef;                      /* as rootRef, in the wide tree */
    const float* materialSpectra;        /* binned spectra for hero-wavelength rendering (NULL: none):
                                           kPtSpectrumWords floats per material, see ptSceneSetSpectra */
} MeshAccelSceneGpu;

/* ---- error reporting ------------------------------------------------------------------ */
const char* pathTracerLastError(void);

/* ---- launchers: PathTracerLib/src/PathTracerCuda.h:12-126 ------------------------------ */

/* PathTracerCuda.h:12-23 (PathTracer.cu:614-647): acc = (0,0,0,1), counts/mean/m2/converged = 0,
 * activeIndices[i] = i, *activeCount = width*height. d_activeScratch is unused, as in the reference. */
bool pathTracerClearAccumulator(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                uint8_t* d_converged, int* d_activeIndices, int* d_activeScratch,
                                int* d_activeCount, int width, int height, hipStream_t stream);

/* PathTracerCuda.h:25-39 (PathTracer.cu:649-695) */
bool pathTracerClearRegionAccumulator(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean,
                                      float* d_m2, uint8_t* d_converged, int* d_activeIndices,
                                      int* d_activeCount, int width, int height, int minX, int minY,
                                      int maxX, int maxY, hipStream_t stream);

/* PathTracerCuda.h:41-49 (PathTracer.cu:697-730). Unlike the reference's atomicAdd order,
 * the rebuilt list is in ascending pixel order (per-pixel results do not depend on it). */
bool pathTracerRebuildActiveList(const uint8_t* d_converged, const uint32_t* d_counts, int* d_activeIndices,
                                 int* d_activeCount, int width, int height, int maxSamplesPerPixel,
                                 hipStream_t stream);

/* PathTracerCuda.h:51-66 (PathTracer.cu:732-780): one path sample for every active,
 * non-converged pixel; Welford update of luminance, running mean of RGB, convergence. */
bool pathTracerAdaptiveSample(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                              uint8_t* d_converged, const int* d_activeIndices, int activeCount, int width,
                              int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                              const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                              unsigned int globalSeed, hipStream_t stream);

/* PathTracerCuda.h:68-83 (PathTracer.cu:782-822). Output order is the input order. */
bool pathTracerCompactActiveList(const int* d_activeIndicesIn, int* d_activeIndicesOut, int* d_activeCount,
                                 const uint8_t* d_converged, const uint32_t* d_counts, int maxSamplesPerPixel,
                                 int inCount, int width, int height, bool useRegionFilter, int regionMinX,
                                 int regionMinY, int regionMaxX, int regionMaxY, hipStream_t stream);

/* No counterpart in the reference: pathTracerCompactActiveList for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and result; its tile counts live in a fixed 256 KiB device
 * array per GPU instead of the library's grow-only allocation, so inCount is at most 64M
 * (67,108,864) entries and larger counts return false with an error. Calls that overlap on one
 * GPU (other streams or threads) share that array and race. */
bool pathTracerCompactActiveListNoAlloc(const int* d_activeIndicesIn, int* d_activeIndicesOut, int* d_activeCount,
                                        const uint8_t* d_converged, const uint32_t* d_counts,
                                        int maxSamplesPerPixel, int inCount, int width, int height,
                                        bool useRegionFilter, int regionMinX, int regionMinY, int regionMaxX,
                                        int regionMaxY, hipStream_t stream);

/* PathTracerCuda.h:85-94 (PathTracer.cu:824-846): no jitter, sampleIndex 0, overwrite. */
bool pathTracerPreviewSample(PtFloat4* d_buffer, int width, int height, const CameraGpu* d_camera,
                             const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                             const RenderParamsGpu* d_params, unsigned int globalSeed, hipStream_t stream);

/* PathTracerCuda.h:96-105 (PathTracer.cu:848-881): exposure, x/(1+x), clamp, x255 truncation. */
bool pathTracerCopyToPbo(const PtFloat4* acc, const uint32_t* counts, const uint8_t* converged, PtUchar4* pbo,
                         int width, int height, const RenderParamsGpu* d_params, float exposure,
                         hipStream_t stream);

/* PathTracerCuda.h:107-117 (PathTracer.cu:883-916) */
bool pathTracerUpsamplePreviewToPbo(const PtFloat4* preview, int previewWidth, int previewHeight, int downscale,
                                    PtUchar4* pbo, int fullWidth, int fullHeight,
                                    const RenderParamsGpu* d_params, float exposure, hipStream_t stream);

/* PathTracerCuda.h:119-126 (PathTracer.cu:918-936) */
bool pathTracerWriteUvDebugToPbo(const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene, PtUchar4* pbo,
                                 int width, int height, const RenderParamsGpu* d_params, hipStream_t stream);

/* PathTracer.cu:27-33. Stores the spectral model (device tables + white normalisation) that
 * the launchers pass to their kernels when hero-wavelength sampling is on. The tables must stay
 * allocated while it is stored. */
hipError_t spectralUploadStateToSymbol(const SpectralStateGpu* hostState);

/* ---- light sampling (no reference counterpart) ------------------------------------------------
 * 0 (default): the reference's light sampling. Its emissive next-event shadow ray runs to 1e6 from
 * p + wi * 1e-4 and meets the sampled light itself (LightSamplingCore.h:360-369), so area lights are
 * lit by BSDF sampling alone, with unweighted emission (PathIntegratorRandCore.h:110-167, 248).
 * 1: multiple importance sampling on area lights. The shadow ray stops at 0.9999 of the distance
 * to the light point, and emission found by BSDF sampling takes the balance weight against the
 * next-event estimate of the vertex before. The expected image is the same as in mode 0, with less
 * noise where area lights dominate. Applies to every forward launcher and wavefront render. The
 * adjoint and detached renders accept mode 0 only. Returns false for other modes. */
bool pathTracerSetLightSampling(int mode);
int pathTracerLightSampling(void);

/* ---- spectral state (role of SpectralStateHost, Spectral/SpectralState.cpp:15-141) --------- */

typedef struct PtSpectral PtSpectral;

/* White normalisation of an rgb2spec table: the linear-sRGB response of the table's (1,1,1)
 * spectrum summed over 360..830 nm in 1-nm steps (SpectralState.cpp:15-46). Host only. */
bool ptSpectralWhiteNorm(const float* scale, const float* data, int res, float* out3);
/* Host tables (scale[res], data[3*3*res^3]) -> device copies + white normalisation. */
PtSpectral* ptSpectralCreate(const float* scale, const float* data, int res);
/* SpectralState.cpp:48-103: reads a "SPEC" coefficient file, then as ptSpectralCreate. */
PtSpectral* ptSpectralLoadCoeffFile(const char* path);
/* The SpectralStateGpu to pass to spectralUploadStateToSymbol. */
bool ptSpectralGetState(const PtSpectral* spectral, SpectralStateGpu* out);
/* Frees the device tables; if they are the stored state, the library returns to RGB rendering. */
void ptSpectralDestroy(PtSpectral* spectral);

/* Hero-wavelength rendering (extension; the reference's integrator is RGB): with
 * wavelengthsPerSample = 4 every launcher and wavefront render samples four wavelengths per
 * path sample, lambda_k = lambdaMin + frac(u + k/4) (lambdaMax - lambdaMin), from one extra
 * uniform draw after the lens draws; colours are upsampled with the stored rgb2spec table and
 * each sample is converted to linear sRGB (CIE 1931 x D65, divided by the white response and
 * the wavelength pdf) before the firefly clamp, which keeps negative channels in this mode.
 * 0 returns to RGB. Needs 360 <= lambdaMin < lambdaMax <= 830 and a stored table. Process-wide
 * setting; change it only between renders. The RGB adjoint entry points reject spectral mode;
 * pathTracerAdjointRenderSpectral differentiates it. */
bool pathTracerSetSpectralSampling(int wavelengthsPerSample, float lambdaMinNm, float lambdaMaxNm);
int pathTracerSpectralSampling(void);

/* ---- extensions of the launcher API (no reference counterpart) -------------------------- */

/* Runs `passCount` consecutive adaptive passes in one launch. Pass p uses globalSeed
 * firstGlobalSeed + p and samples every pixel of the active list that is not converged at
 * the start of that pass -- the per-pixel result equals calling pathTracerAdaptiveSample
 * passCount times (PathTracer.cpp:2379-2465), without host round trips. Pixels are handed
 * out through an internal work counter, so the list order does not matter. */
bool pathTracerAdaptiveRenderPasses(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                    uint8_t* d_converged, const int* d_activeIndices, int activeCount,
                                    int width, int height, const CameraGpu* d_camera,
                                    const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                                    const RenderParamsGpu* d_params, unsigned int firstGlobalSeed,
                                    int passCount, hipStream_t stream);

/* pathTracerAdaptiveRenderPasses with traversal counters: adds, over all samples taken, the
 * internal nodes fetched, triangles tested, rays traced and closest hits into d_totals[0..3]
 * and the number of samples into d_totals[4] (5 x uint64, accumulated, not cleared). */
bool pathTracerAdaptiveRenderPassesCounted(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                           uint8_t* d_converged, const int* d_activeIndices, int activeCount,
                                           int width, int height, const CameraGpu* d_camera,
                                           const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                                           const RenderParamsGpu* d_params, unsigned int firstGlobalSeed,
                                           int passCount, uint64_t* d_totals, hipStream_t stream);

/* Wavefront-scheduled equivalent of pathTracerAdaptiveRenderPasses: the same per-pixel results
 * (pass p uses globalSeed firstGlobalSeed + p; every active, unconverged pixel is sampled once
 * per pass; convergence is tested at counts that are multiples of 4). Groups of consecutive
 * samples of all active pixels are traced together, one bounce per launch: closest-hit and
 * any-hit traversal kernels with an LDS stack, and a shading kernel that appends the next rays
 * to queues. maxPathDepth is the depth the params resolve to (maxPathDepth > 0 ? that : 0 -> 256;
 * PathIntegratorRandCore.h:219-227). maxPathsInFlight bounds the paths of one group (0 = 64M;
 * at least 4 per active pixel are always allowed). The first group takes every sample that no
 * pixel can skip (up to minSamples rounded up to 4); later groups take 4 samples per active
 * pixel, widened speculatively up to 256 (samples drawn after a pixel converged are discarded,
 * so results do not depend on the grouping).
 * Uses a library-owned device workspace (grown on demand); calls on one device must not overlap.
 * The caller's active list is read, not modified. Synchronises the stream once per group. */
bool pathTracerAdaptiveRenderWavefront(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                       uint8_t* d_converged, const int* d_activeIndices, int activeCount, int width,
                                       int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                       const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                       unsigned int firstGlobalSeed, int passCount, int maxPathDepth,
                                       int maxPathsInFlight, hipStream_t stream);

/* Fixed-spp wavefront render with no host synchronisation, so it can be captured into a HIP
 * graph: every pixel takes samplesPerPixel samples (sample index k, seed k + 1, as the adaptive
 * schedule with minSamples = maxSamplesPerPixel = samplesPerPixel), in groups of up to
 * maxPathsInFlight paths (0 = 64M), accumulated into the cleared accumulator as the reference
 * does. While a stream is capturing, the workspace is not allocated: reserve it first with
 * pathTracerWavefrontReserve. Calls on one device must not overlap (shared workspace).
 * No reference counterpart. */
bool pathTracerRenderFixedWavefront(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                    uint8_t* d_converged, int width, int height, const CameraGpu* d_camera,
                                    const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                                    const RenderParamsGpu* d_params, int samplesPerPixel, int maxPathDepth,
                                    int maxPathsInFlight, hipStream_t stream);

/* Grows the wavefront workspace for groups of maxPathsInFlight paths over `pixels` pixels at depth
 * maxPathDepth (grow-only; a single allocation). */
bool pathTracerWavefrontReserve(int maxPathsInFlight, int pixels, int maxPathDepth);

/* Workspace pins (no reference counterpart). A fixed-spp or batch render enqueued while its stream
 * is capturing a graph takes one pin: the graph keeps the workspace's addresses, so while any pin
 * is held the workspace does not move, and a call that needs a larger one fails with an error
 * (reserve enough first). pathTracerWavefrontPin takes a pin for a caller's own capture,
 * pathTracerWavefrontPins reports the pins held, and pathTracerWavefrontUnpin releases `count` of
 * them once the graphs that took them are destroyed (false if fewer are held). */
bool pathTracerWavefrontPin(void);
int pathTracerWavefrontPins(void);
bool pathTracerWavefrontUnpin(int count);

/* Measurement hooks of pathTracerAdaptiveRenderWavefront (no reference counterpart).
 * timing != 0: every launch of the render is bracketed by hipEvents on its stream and the device
 * time is accumulated per kernel class; counted != 0: the trace kernels also count internal
 * nodes fetched, triangles tested and rays. Both reset the accumulated values. */
bool pathTracerWavefrontProfile(int timing, int counted);

/* out[0..3]: ms of trace-extension, trace-shadow, shade and other launches (generation,
 * accumulation); out[4..7]: their launch counts; out[8..13]: extension nodes, triangles, rays,
 * shadow nodes, triangles, rays; out[14] (when n >= 15): samples started, including speculative
 * ones that were discarded. */
bool pathTracerWavefrontProfileRead(double* out, int n);

/* Scheduling knob of pathTracerAdaptiveRenderWavefront: once pixels may converge mid-group,
 * groups are widened (in steps of 4 samples per pixel, at most 64) while the group's paths stay
 * within speculativePaths (0 = the workspace capacity, the default). Results do not depend on
 * it; larger values trade discarded speculative samples for fewer, fuller launches. */
bool pathTracerWavefrontSchedule(int speculativePaths);

/* Per-bounce breakdown of the profiled renders (pathTracerWavefrontProfile with timing on):
 * for bounce b < n (n <= 258), out[5b..5b+4] = ms of the extension-trace, shadow-trace and
 * shade launches, extension rays, shadow rays. */
bool pathTracerWavefrontProfileBounces(double* out, int n);

/* Traces one ray per entry (origins/directions as 3 floats each) through the same closest-hit
 * traversal the integrator uses; writes hit flag, t, normal (3), uv (2), triangle index as
 * 8 floats per ray (int fields bit-cast). tMin/tMax as in MeshAccelIntersectCore.h:110-115. */
bool pathTracerTraceRays(const MeshAccelSceneGpu* d_scene, const float* d_origins, const float* d_dirs,
                         int rayCount, float tMin, float tMax, float* d_out, hipStream_t stream);

/* Per-sample traversal statistics for the algorithmic byte model: for each active pixel of
 * the given pass, writes 4 uint32 (internal nodes fetched, triangles tested, rays traced,
 * closest hits). Does not modify the accumulator. */
bool pathTracerCountTraversal(const uint32_t* d_samples, const int* d_activeIndices, int activeCount, int width,
                              int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                              const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                              unsigned int globalSeed, uint32_t* d_stats, hipStream_t stream);

/* Adjoint (path replay) of the fixed-spp render -- sample k of every pixel uses sample index k
 * and seed k + 1, i.e. pathTracerAdaptiveRenderWavefront or the per-pass loop with
 * minSamples = maxSamplesPerPixel = samplesPerPixel from a cleared accumulator. d_gradImage
 * holds dLoss/dPixel (3 floats per pixel, row-major); d_grad (7 doubles per material:
 * base r, g, b, roughness, metallic, specular, emission) accumulates dLoss/dParameter.
 * Sampling decisions are detached (those of the current parameters); BSDF derivatives cover the
 * opaque Principled lobes; parameters are differentiated where the forward is differentiable
 * (untextured; roughness/metallic/specular strictly inside (0, 1); emission > 0).
 * materialCount <= 64. d_sampleRadiance (optional) holds the primal, unclamped radiance of every
 * (pixel, sample), 3 floats at (pixel * samplesPerPixel + sample), as
 * pathTracerWavefrontCaptureRadiance records it during the forward; without it the fixed-spp
 * render of the same samples runs first. Reads the render parameters from the device (one stream
 * synchronisation) and uses the wavefront workspace: calls on one device must not overlap. Paths
 * run in groups of up to 16M, or of the reserved workspace's size when that is larger
 * (pathTracerWavefrontReserve); the grouping changes only the order of the FP64 sums.
 * No reference counterpart. */
bool pathTracerAdjointRender(const float* d_gradImage, int width, int height, int samplesPerPixel,
                             const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                             const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params, int materialCount,
                             const float* d_sampleRadiance, double* d_grad, hipStream_t stream);

/* pathTracerAdjointRender over another sample set: sample k of every pixel uses seed
 * firstSeed + k (pathTracerAdjointRender is firstSeed = 1). Replaying independent samples each
 * optimisation step keeps the gradient's sampling error from acting as a fixed bias. No
 * reference counterpart. */
bool pathTracerAdjointRenderSeeded(const float* d_gradImage, int width, int height, int samplesPerPixel,
                                   const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                   const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                   int materialCount, const float* d_sampleRadiance, double* d_grad,
                                   unsigned int firstSeed, hipStream_t stream);

/* The adjoint of a hero-wavelength render (pathTracerSetSpectralSampling(4, ...) on, light
 * sampling 0): as pathTracerAdjointRenderSeeded, with d_grad accumulating kPtSpectrumGradParams
 * doubles per material: the kPtSpectrumBins reflectance bins, the kPtSpectrumBins emission bins,
 * roughness, metallic, specular and the emission scale of the material's binned spectra
 * (ptSceneSetSpectra). The bins of a material are parameters where it has binned spectra (its
 * base colour / emitted radiance at a hero wavelength is that wavelength's bin, x emission for
 * the latter); roughness, metallic and specular are parameters as in the RGB adjoint. Each
 * sample's loss weights pass through the firefly clamp and the conversion of its four
 * wavelengths to linear sRGB. d_sampleRadiance, if given, holds 4 floats per sample
 * (pathTracerWavefrontCaptureRadiance in spectral mode). No reference counterpart. */
#define kPtSpectrumGradParams 36
bool pathTracerAdjointRenderSpectral(const float* d_gradImage, int width, int height, int samplesPerPixel,
                                     const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                     const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                     int materialCount, const float* d_sampleRadiance, double* d_grad,
                                     unsigned int firstSeed, hipStream_t stream);

/* While d_out is set, pathTracerAdaptiveRenderWavefront writes the unclamped radiance of every
 * finished sample with index < samplesPerPixel to d_out[3 * (pixel * samplesPerPixel + sample)]
 * (width * height * samplesPerPixel * 3 floats; in hero-wavelength mode the radiance at the four
 * wavelengths, 4 floats per sample). NULL turns it off. No reference counterpart. */
bool pathTracerWavefrontCaptureRadiance(float* d_out, int samplesPerPixel);

/* ---- batches of scenes (no reference counterpart) --------------------------------------------
 * A batch renders many scenes in one wavefront: every bounce's trace and shade launches serve the
 * paths of all scenes, so the cost of a launch's slowest rays is paid once per bounce for the
 * whole batch instead of once per scene. Each scene's results equal those of its own render. */

/* One scene of a batch: the device pointers the single-scene launchers take (env may be NULL:
 * solid background) and the image size. */
typedef struct {
    const MeshAccelSceneGpu* scene;
    const CameraGpu* camera;
    const RenderParamsGpu* params;
    const EnvironmentMapGpu* env;
    int width, height;
} PtBatchSceneGpu;

typedef struct PtBatch PtBatch;

/* Batch of sceneCount (1..4096) scenes from a host array. The scenes' pixels are concatenated in
 * order (scene i starts at pixel offset sum_{j<i} width_j * height_j of the batch's per-pixel
 * buffers) and so are their materials' gradient rows (scene i's material m is row
 * sum_{j<i} materialCount_j + m). Reads every scene's description and parameters once
 * (synchronously) to size the batch; the renders read the records' current contents, so material,
 * camera and parameter updates need no new batch, while a rebuilt scene or a larger
 * maxPathDepth does. NULL on failure. */
PtBatch* ptBatchCreate(const PtBatchSceneGpu* scenes, int sceneCount);
void ptBatchDestroy(PtBatch* batch);
/* Total pixels, gradient rows, the largest resolved path depth, and (arrays of sceneCount ints,
 * each optional) the scenes' pixel offsets and first gradient rows. */
bool ptBatchInfo(const PtBatch* batch, int* totalPixels, int* gradRows, int* maxPathDepth, int* pixelOffsets,
                 int* gradRowOffsets);

/* pathTracerRenderFixedWavefront over a batch: per-pixel buffers hold the batch's totalPixels
 * pixels; every scene takes samplesPerPixel samples, sample k with seed firstSeed + k (1 gives the
 * plain fixed render), each with its own parameters. Capturable into a HIP graph once the
 * workspace is reserved (pathTracerWavefrontReserve with the batch's pixels as path bound).
 * Calls on one device must not overlap (shared workspace). */
bool pathTracerRenderFixedBatch(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                uint8_t* d_converged, const PtBatch* batch, int samplesPerPixel, int maxPathsInFlight,
                                unsigned int firstSeed, hipStream_t stream);

/* pathTracerAdjointRenderSeeded over a batch: d_gradImage holds dLoss/dPixel for the batch's
 * pixels (3 floats each), d_grad accumulates 7 doubles per gradient row (at most 512 rows); in
 * hero-wavelength mode, kPtSpectrumGradParams doubles per row as pathTracerAdjointRenderSpectral
 * (at most 224 rows). The samples (seeds firstSeed..) are rendered first, then replayed. Not
 * capturable. */
bool pathTracerAdjointRenderBatch(const float* d_gradImage, const PtBatch* batch, int samplesPerPixel,
                                  unsigned int firstSeed, double* d_grad, hipStream_t stream);

/* The detached estimator of the same render as an image (3 floats per pixel, running mean of the
 * clamped radiance): paths sampled with the scene's materials (directions, pdfs, MIS weights,
 * Russian roulette), values evaluated with d_evalMaterials (same count and layout). With
 * d_evalMaterials equal to the scene's table it is the fixed-spp render itself; central
 * differences in d_evalMaterials are the finite-difference reference of
 * pathTracerAdjointRender. No reference counterpart. */
bool pathTracerDetachedRender(float* d_image, int width, int height, int samplesPerPixel, const CameraGpu* d_camera,
                              const MeshAccelSceneGpu* d_scene, const MaterialGpu* d_evalMaterials,
                              const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params, hipStream_t stream);

/* ---- scene (replaces MeshAccelScene::build + upload, MeshAccelScene.cpp:141-341) ---------- */
typedef struct PtScene PtScene;

PtScene* ptSceneCreate(void);
void ptSceneDestroy(PtScene* scene);

/* Builds the scene from host arrays: BVH on the GPU (median split on the longest axis of the
 * triangle bounds, one triangle per leaf, MeshBvhBuilder.cpp:52-142, ties broken by triangle
 * index), emissive list + CDF (MeshAccelScene.cpp:62-95), default material when none is given
 * (MeshAccelScene.cpp:27-46), and the device layout described above. Synchronous. */
bool ptSceneBuild(PtScene* scene, const HostTriangle* triangles, uint32_t triangleCount,
                  const MaterialGpu* materials, uint32_t materialCount, const TextureDescGpu* textures,
                  uint32_t textureCount, hipStream_t stream);

/* Device pointer to the scene description, for the launchers. NULL before a build. */
const MeshAccelSceneGpu* ptSceneDevice(const PtScene* scene);

/* Host copy of the scene description (device pointers inside). */
bool ptSceneDescription(const PtScene* scene, MeshAccelSceneGpu* out);

/* Exports the tree in the reference's 48-byte pre-order layout (2N-1 nodes, root 0). */
bool ptSceneReadBvh(const PtScene* scene, MeshBvhNode* hostNodes, uint32_t capacity, uint32_t* outCount,
                    uint32_t* outRoot);

/* Test and inspection hook: copies the device emissive list built by ptSceneBuild
 * (MeshAccelScene.cpp:62-95): the emissive triangles' original indices (ascending), the
 * unnormalised prefix CDF of area x emission x luminance (count + 1 floats) and the packed
 * records the sampler reads (3 x float4 per entry: v0, v1, v2 with r, g, b x emission in .w).
 * outCount receives the entry count; NULL outputs are skipped; capacity is in entries. */
bool ptSceneReadEmissive(const PtScene* scene, uint32_t* indices, float* cdf, float* records, uint32_t capacity,
                         uint32_t* outCount);

/* Milliseconds of the last GPU BVH build (device time between events). */
float ptSceneLastBuildMs(const PtScene* scene);

/* Tree rebuilds that a caller can capture into a graph (no reference counterpart: the reference
 * builds its tree once on the CPU). A scene built in a binned-SAH mode (1 or 2) keeps its
 * triangles on the device; ptSceneRebuild enqueues the whole build again on them, in the mode the
 * scene was built with (with the 4-wide collapse in mode 1; a later ptSceneSetBuildMode applies
 * to the next ptSceneBuild only), into the scene's own buffers: a fixed sequence of launches on
 * `stream` with no allocation and no host synchronisation. The rebuilt tree is the one ptSceneBuild
 * made (the build is deterministic and the triangles are unchanged), so a rebuild enqueues just
 * the levels that first build needed. The workspace holds the build's temporaries for scenes of
 * up to maxTriangles triangles (about 1.1 KB per triangle); rebuilds that share one workspace
 * must be ordered on one stream. Every rebuild ends by copying its status into the scene's own
 * status record, so ptSceneRebuildCheck (which waits for `stream`) checks that scene's last
 * rebuild whatever was rebuilt after it, and returns false, with pathTracerLastError set, if that
 * rebuild failed, produced a tree of another shape than the scene's, or none has run since the
 * scene's build or its last check (each replay of a captured rebuild rewrites the record). The
 * workspace argument of the check is not used. */
typedef struct PtBuildWorkspace PtBuildWorkspace;
PtBuildWorkspace* ptBuildWorkspaceCreate(int maxTriangles);
void ptBuildWorkspaceDestroy(PtBuildWorkspace* workspace);
bool ptSceneRebuild(PtScene* scene, PtBuildWorkspace* workspace, hipStream_t stream);
/* As ptSceneRebuild, but enqueues every level the depth cap allows (49), as a build of triangles
 * whose tree depth is not known yet must: the cost of rebuilding changed geometry in a graph. */
bool ptSceneRebuildCold(PtScene* scene, PtBuildWorkspace* workspace, hipStream_t stream);
bool ptSceneRebuildCheck(PtScene* scene, PtBuildWorkspace* workspace, hipStream_t stream);

/* Replaces the material table of a built scene (no reference counterpart; for differentiable
 * rendering and optimisation loops): same count; a material may not start or stop emitting.
 * The emissive list keeps its triangles; their radiance and the selection CDF are recomputed
 * as the build does. Synchronous. */
bool ptSceneUpdateMaterials(PtScene* scene, const MaterialGpu* materials, uint32_t materialCount,
                            hipStream_t stream);

/* Binned spectra of the scene's materials for hero-wavelength rendering (no reference
 * counterpart): kPtSpectrumWords floats per material, kPtSpectrumBins reflectance values, then
 * kPtSpectrumBins emission values. Bin b covers [lambdaMin + b w, lambdaMin + (b + 1) w) of the
 * hero-wavelength range, w = (lambdaMax - lambdaMin) / kPtSpectrumBins, so a table is read
 * against the range pathTracerSetSpectralSampling sets. Where a material's first reflectance
 * value is >= 0, its base colour at a hero wavelength is that wavelength's bin (instead of the
 * rgb2spec upsampling of r, g, b); where its first emission value is >= 0, its emitted radiance
 * is emission x that wavelength's bin (on surface hits and for next-event light samples). RGB
 * rendering, light selection and the emissive list keep using r, g, b and emission. NULL (or
 * materialCount 0) removes the table. The table survives ptSceneBuild calls that keep the
 * material count and ptSceneUpdateMaterials; it is uploaded on `stream` and copied from the host
 * before the call returns. */
#define kPtSpectrumBins 16
#define kPtSpectrumWords 32
bool ptSceneSetSpectra(PtScene* scene, const float* spectra, uint32_t materialCount, hipStream_t stream);

/* BVH build used by the next ptSceneBuild (no reference counterpart): 0 = the reference's
 * median split (default; node arrays equal the reference's), 1 = binned SAH over Morton-ordered
 * triangles (same device layout, geometry-dependent topology, depth <= 48; ptSceneReadBvh then
 * reports 0 nodes) plus its 4-wide collapse, which the wavefront trace kernels traverse;
 * 2 = the binned SAH tree alone. Closest hits are the same except for exact ties between
 * triangles. */
bool ptSceneSetBuildMode(PtScene* scene, int mode);

/* The 4-wide tree of a scene built in mode 1: copies its nodes (8 x float4 each, layout of
 * MeshAccelSceneGpu::wideNodes) into a host buffer of *count entries (nodes may be NULL to ask
 * for the count) and reports the root reference. No reference counterpart. */
bool ptSceneReadWide(const PtScene* scene, float* nodes, int* count, int* rootRef);

/* Copies the device layout: internal nodes (nodeCount x 16 floats) and leaf slots
 * (triangleCount x 12 floats); NULL skips one. levels receives the SAH build's level count
 * (0 for the median split). */
bool ptSceneReadLayout(const PtScene* scene, float* nodes, float* triPos, int* levels);

/* ---- environment map (replaces EnvironmentMap::buildImportanceTables + upload,
 *      PathTracerLib/src/EnvironmentMap.cpp:76-127) --------------------------------------- */
typedef struct PtEnvironment PtEnvironment;

/* Builds the importance tables of a lat-long RGB float image (3 floats per texel, rows top to
 * bottom) and uploads image and tables. Synchronous. NULL on failure. */
PtEnvironment* ptEnvironmentCreate(const float* rgb, int width, int height, hipStream_t stream);

/* Device pointer for the launchers' d_env argument (valid = 1). */
const EnvironmentMapGpu* ptEnvironmentDevice(const PtEnvironment* env);

/* Copies the tables: marginal (height + 1 floats) and rows (height * (width + 1) floats). */
bool ptEnvironmentReadTables(const PtEnvironment* env, float* marginal, float* rows);

void ptEnvironmentDestroy(PtEnvironment* env);

/* Reads a Radiance RGBE picture (.hdr; the reference's stbi_loadf path, EnvironmentMap.cpp:191-210):
 * flat or run-length scanlines, orientation -Y H +X W, texel = (r, g, b) * 2^(e - 136) (0 when
 * e = 0). Returns width * height * 3 floats (rows top to bottom, as ptEnvironmentCreate takes
 * them) to be released with ptEnvironmentFreeImage, or NULL with the error set. */
float* ptEnvironmentLoadHdr(const char* path, int* width, int* height);
void ptEnvironmentFreeImage(float* rgb);

#ifdef __cplusplus
}
#endif

#endif
