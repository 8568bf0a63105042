/*
 * pt_tracer.h -- headless renderer facade over the launcher ABI of pt_hip.h.
 *
 * Mirrors `class PathTracer` (PathTracerLib/src/PathTracer.h:30-152, implemented in
 * PathTracer.cpp): the same method set, clamps, reset semantics and worker-thread render loop
 * (PathTracer.cpp:2220-2543: preview levels, adaptive passes with globalSeed = iteration + 1,
 * active-list compaction every 4 full-resolution passes, region rendering, sample budget).
 * Qt/glm/GL types are replaced by plain ones: colours are three floats, regions are four ints,
 * points are Vec3, the two GL PBOs are two library-owned device RGBA8 buffers that
 * ptTracerReadDisplay copies out, and the HDR file path is replaced by a decoded RGB float image.
 *
 * The C functions below wrap one PathTracer each (ptTracer* = PathTracer::*); C++ callers can
 * use the class directly.
 */
#ifndef PT_TRACER_H
#define PT_TRACER_H

#include "pt_hip.h"
#include "pt_procedural.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PtTracer PtTracer;

/* Frame-ready notification, invoked on the worker thread after every completed pass. */
typedef void (*PtFrameReadyFn)(void* user);

PtTracer* ptTracerCreate(void);
void ptTracerDestroy(PtTracer* t);
void ptTracerSetFrameReadyCallback(PtTracer* t, PtFrameReadyFn fn, void* user);

/* PathTracer::configure (PathTracer.cpp:1196-1395): accumulator, preview levels, display
 * buffers, camera/params, spectral state and the mesh scene. false on invalid sizes or when an
 * allocation or the scene build fails. */
bool ptTracerConfigure(PtTracer* t, int width, int height, const HostTriangle* triangles, uint32_t triangleCount,
                       const MaterialGpu* materials, uint32_t materialCount, const TextureDescGpu* textures,
                       uint32_t textureCount);
bool ptTracerRebuildMeshScene(PtTracer* t, const HostTriangle* triangles, uint32_t triangleCount,
                              const MaterialGpu* materials, uint32_t materialCount, const TextureDescGpu* textures,
                              uint32_t textureCount);
/* PathTracer::configure / rebuildMeshScene with the reference's scene description
 * (PathTracer.h:44-50, 137-139): the instances go through meshSceneBuild
 * (MeshSceneContent.cpp:5-52; ptProceduralBuildScene with circularSegments and creaseAngleDeg
 * from params, NULL = {64, 50}). An empty list builds an empty scene. false, with the mesher's
 * message ("Procedural instance <i>: ..."), when an instance fails. */
bool ptTracerConfigureProcedural(PtTracer* t, int width, int height, const PtProceduralInstance* instances,
                                 uint32_t count, const PtMeshSceneBuildParams* params);
bool ptTracerRebuildMeshSceneProcedural(PtTracer* t, const PtProceduralInstance* instances, uint32_t count,
                                        const PtMeshSceneBuildParams* params);

void ptTracerStart(PtTracer* t);
void ptTracerStop(PtTracer* t);
void ptTracerResetAccumulation(PtTracer* t);
void ptTracerResetRegionAccumulation(PtTracer* t);
bool ptTracerIsRunning(const PtTracer* t);
void ptTracerReleaseOutputSurfaces(PtTracer* t);

/* Display publish into library-owned slots 0/1 (PathTracer.cpp:1477-1708). */
bool ptTracerBeginDisplayPublish(PtTracer* t, int slot);
bool ptTracerIsDisplayPublishReady(const PtTracer* t);
bool ptTracerHasDisplayPublishInFlight(const PtTracer* t);
int ptTracerFinishDisplayPublish(PtTracer* t);
bool ptTracerPublishDisplayFrame(PtTracer* t, int slot);
/* Copies slot's RGBA8 frame (width * height texels) to host memory. */
bool ptTracerReadDisplay(PtTracer* t, int slot, PtUchar4* hostRgba);

void ptTracerSetMaxSamplesPerPixel(PtTracer* t, int max);
int ptTracerMaxSamplesPerPixel(const PtTracer* t);
void ptTracerSetMinSamples(PtTracer* t, int min);
int ptTracerMinSamples(const PtTracer* t);
void ptTracerSetRelativeErrorThreshold(PtTracer* t, float threshold);
float ptTracerRelativeErrorThreshold(const PtTracer* t);
void ptTracerSetDebugOverlayMode(PtTracer* t, int mode);
int ptTracerDebugOverlayMode(const PtTracer* t);
void ptTracerSetBrdfDebugFlags(PtTracer* t, int flags);
int ptTracerBrdfDebugFlags(const PtTracer* t);
void ptTracerSetPreviewStepsPerLevel(PtTracer* t, int steps);
int ptTracerPreviewStepsPerLevel(const PtTracer* t);
int ptTracerCurrentSampleCount(const PtTracer* t);
int ptTracerCurrentActivePixelCount(const PtTracer* t);
int ptTracerSampleBudgetTotalIterations(const PtTracer* t);
bool ptTracerIsSampleBudgetExhausted(const PtTracer* t);
void ptTracerSetCamera(PtTracer* t, const CameraGpu* camera);
void ptTracerLastSampleCamera(const PtTracer* t, CameraGpu* out);
void ptTracerSetClearColor(PtTracer* t, float r, float g, float b);
void ptTracerSetPhysicalCamera(PtTracer* t, float fStop, float shutterSpeedSeconds, float iso);
void ptTracerSetFocusPoint(PtTracer* t, Vec3 worldPoint);
void ptTracerClearFocusPoint(PtTracer* t);
bool ptTracerPickSurface(const PtTracer* t, Vec3 ro, Vec3 rd, Vec3* hitPoint);
/* Replaces setEnvironmentHdrPath: a decoded lat-long RGB float image (NULL = solid background). */
bool ptTracerSetEnvironmentImage(PtTracer* t, const float* rgb, int width, int height);
void ptTracerSetEnvironmentIntensity(PtTracer* t, float intensity);
float ptTracerEnvironmentIntensity(const PtTracer* t);
void ptTracerSetRussianRouletteMinDepth(PtTracer* t, int depth);
int ptTracerRussianRouletteMinDepth(const PtTracer* t);
void ptTracerSetRegionRenderEnabled(PtTracer* t, bool enabled);
bool ptTracerRegionRenderEnabled(const PtTracer* t);
void ptTracerSetRenderRegion(PtTracer* t, int minX, int minY, int maxX, int maxY);
void ptTracerRenderRegion(const PtTracer* t, int* minX, int* minY, int* maxX, int* maxY);

/* Extensions for headless callers (no reference counterpart). */
/* Blocks until the worker can take no further sample (budget exhausted) or timeoutMs passes;
 * true when idle. */
bool ptTracerWaitUntilIdle(PtTracer* t, int timeoutMs);
/* Copies the accumulator: acc (4 floats/px), counts, lumMean, m2, converged; NULL skips one. */
bool ptTracerReadAccumulator(PtTracer* t, float* acc4, uint32_t* counts, float* lumMean, float* m2,
                             uint8_t* converged);
/* Copies preview level `level` (its width * height float4 texels) and reports its size. */
bool ptTracerReadPreviewLevel(PtTracer* t, int level, float* acc4, int* width, int* height, int* downscale);
/* Exposure multiplier of the physical camera (PhysicalCamera.cpp:397-402). */
float ptTracerExposureMultiplier(const PtTracer* t);
/* Full-resolution maxPathDepth (the reference fixes it at the RenderTypes.h default, 32);
 * 0 means 256 and negative values 32, as the integrator resolves them. Resets accumulation. */
void ptTracerSetMaxPathDepth(PtTracer* t, int depth);
int ptTracerMaxPathDepth(const PtTracer* t);
/* Full-resolution passes per worker iteration (1..256, default 1 = the reference's cadence: one
 * pass, then a frame-ready callback). Above 1 the worker renders that many passes of the active
 * pixels in one wavefront call (pathTracerAdaptiveRenderWavefront: the same per-pixel results as
 * the passes one by one), compacts the active list, adds them to the sample count and calls the
 * frame-ready callback once. For offline rendering; no reference counterpart. */
void ptTracerSetPassesPerLaunch(PtTracer* t, int passes);
int ptTracerPassesPerLaunch(const PtTracer* t);
/* BVH build of the next configure / rebuildMeshScene (ptSceneSetBuildMode: 0 = the reference's
 * median split, the default; 1 = binned SAH with its 4-wide tree; 2 = binned SAH alone). No
 * reference counterpart. */
void ptTracerSetBvhBuildMode(PtTracer* t, int mode);
int ptTracerBvhBuildMode(const PtTracer* t);
/* The error text of the worker thread's last failed launch (pathTracerLastError is per thread);
 * false when none failed. No reference counterpart. */
bool ptTracerLastWorkerError(const PtTracer* t, char* buffer, int capacity);

#ifdef __cplusplus
} /* extern "C" */

#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace ptdetail {
struct TracerImpl;
}

/* Plain-type region (QRect in the reference). */
struct PtRegion {
    int minX, minY, maxX, maxY;
};

/* Mesh handed to configure/rebuildMeshScene (HostMesh, Application/src/MeshBuilder/HostMesh.h:23-28). */
struct PtHostMesh {
    const HostTriangle* triangles = nullptr;
    uint32_t triangleCount = 0;
    const MaterialGpu* materials = nullptr;
    uint32_t materialCount = 0;
    const TextureDescGpu* textures = nullptr;
    uint32_t textureCount = 0;
};

class PathTracer {
public:
    using FrameReadyCallback = std::function<void()>;

    PathTracer();
    ~PathTracer();
    PathTracer(const PathTracer&) = delete;
    PathTracer& operator=(const PathTracer&) = delete;

    void setFrameReadyCallback(FrameReadyCallback callback);
    bool configure(int width, int height, const PtHostMesh& mesh);
    bool configure(int width, int height, const std::vector<PtProceduralInstance>& proceduralInstances,
                   const PtMeshSceneBuildParams& meshParams = PtMeshSceneBuildParams{64, 50.0f});
    void start();
    void stop();
    void resetAccumulation();
    bool isRunning() const;
    void releaseOutputSurfaces();

    bool beginDisplayPublish(int slot);
    bool isDisplayPublishReady() const;
    bool hasDisplayPublishInFlight() const;
    int finishDisplayPublish();
    bool publishDisplayFrame(int slot);
    bool readDisplay(int slot, PtUchar4* hostRgba) const;

    void setMaxSamplesPerPixel(int max);
    int maxSamplesPerPixel() const;
    void setMinSamples(int min);
    int minSamples() const;
    void setRelativeErrorThreshold(float threshold);
    float relativeErrorThreshold() const;
    void setDebugOverlayMode(int mode);
    int debugOverlayMode() const;
    void setBrdfDebugFlags(int flags);
    int brdfDebugFlags() const;
    void setPreviewStepsPerLevel(int steps);
    int previewStepsPerLevel() const;
    int currentSampleCount() const;
    int currentActivePixelCount() const;
    int sampleBudgetTotalIterations() const;
    bool isSampleBudgetExhausted() const;
    void setCamera(const CameraGpu& camera);
    CameraGpu lastSampleCamera() const;
    void setClearColor(float r, float g, float b);
    void setPhysicalCamera(float fStop, float shutterSpeedSeconds, float iso);
    void setFocusPoint(const Vec3& worldPoint);
    void clearFocusPoint();
    bool pickSurface(const Vec3& ro, const Vec3& rd, Vec3* hitPoint) const;
    bool setEnvironmentImage(const float* rgb, int width, int height);
    void setEnvironmentIntensity(float intensity);
    float environmentIntensity() const;
    void setRussianRouletteMinDepth(int depth);
    int russianRouletteMinDepth() const;
    bool rebuildMeshScene(const PtHostMesh& mesh);
    bool rebuildMeshScene(const std::vector<PtProceduralInstance>& proceduralInstances,
                          const PtMeshSceneBuildParams& meshParams = PtMeshSceneBuildParams{64, 50.0f});
    void setRegionRenderEnabled(bool enabled);
    bool regionRenderEnabled() const;
    void setRenderRegion(int minX, int minY, int maxX, int maxY);
    PtRegion renderRegion() const;
    void resetRegionAccumulation();

    bool waitUntilIdle(int timeoutMs);
    bool readAccumulator(float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* converged) const;
    bool readPreviewLevel(int level, float* acc4, int* width, int* height, int* downscale) const;
    float exposureMultiplier() const;
    void setMaxPathDepth(int depth);
    int maxPathDepth() const;
    void setPassesPerLaunch(int passes);
    int passesPerLaunch() const;
    std::string lastWorkerError() const;
    void setBvhBuildMode(int mode);
    int bvhBuildMode() const;

private:
    void renderLoop();
    void notifyWorker();
    void invokeFrameReadyCallback();

    std::unique_ptr<ptdetail::TracerImpl> m_impl;
};
#endif

#endif /* PT_TRACER_H */
