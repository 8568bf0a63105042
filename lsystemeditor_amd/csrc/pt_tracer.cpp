// Headless renderer facade: class PathTracer over the launcher ABI (pt_hip.h).
//
// Follows PathTracerLib/src/PathTracer.cpp: configure (1196-1395), start/stop/reset (1397-1434),
// region handling (1436-1470), display publish (1477-1708), setters with their clamps (1710-1990),
// pickSurface (1991-2013) and the worker's renderLoop (2220-2543). Policies restated from
// PathTracerSampleBudget.h, PathTracerPreviewLevels.h and PhysicalCamera.cpp.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pt_tracer.h"
#include "pt_internal.h"

namespace {

// PathTracer.cpp:126-131
constexpr int kMaxSamplesUpperBound = 1000000;
constexpr int kMaxPreviewStepsPerLevel = 128;
constexpr int kMinMinSamples = 1;
constexpr int kMaxMinSamples = 10000;
constexpr float kMinRelativeErrorThreshold = 0.001f;
constexpr float kMaxRelativeErrorThreshold = 1.0f;
constexpr int kCompactActiveListInterval = 4; // PathTracerAdaptiveSampling.h:7
constexpr int kPreviewMaxPathDepth = 4;       // PathTracerPreviewLevels.h:13

// PhysicalCamera.h:36-51
constexpr float kDefaultFStop = 2.8f;
constexpr float kDefaultShutterSpeedSeconds = 1.0f / 125.0f;
constexpr float kDefaultIso = 100.0f;
constexpr float kMinIso = 100.0f;
constexpr float kMaxIso = 25600.0f;
constexpr float kMinFStop = 1.2f;
constexpr float kMaxFStop = 22.0f;
constexpr float kMinShutterSpeedSeconds = 1.0f / 8000.0f;
constexpr float kMaxShutterSpeedSeconds = 120.0f;
constexpr float kDefaultFarPlane = 1000000.0f;
constexpr float kDefaultFocusDistance = 1000.0f;
constexpr float kSensorHeightMm = 23.9f;
constexpr float kMinFocalLengthMm = 14.0f;
constexpr float kMaxFocalLengthMm = 1000.0f;
constexpr float kDefaultFocalLengthMm = 24.0f;

int clampMaxSamples(int v) { return v < 0 ? 0 : (v > kMaxSamplesUpperBound ? kMaxSamplesUpperBound : v); }
int clampMinSamples(int v) { return v < kMinMinSamples ? kMinMinSamples : (v > kMaxMinSamples ? kMaxMinSamples : v); }
float clampRelativeErrorThreshold(float v)
{
    return v < kMinRelativeErrorThreshold ? kMinRelativeErrorThreshold
                                          : (v > kMaxRelativeErrorThreshold ? kMaxRelativeErrorThreshold : v);
}
int clampPreviewStepsPerLevel(int v) { return v < 0 ? 0 : (v > kMaxPreviewStepsPerLevel ? kMaxPreviewStepsPerLevel : v); }

// PathTracerSampleBudget.h
int budgetTotalIterations(int previewSteps, int maxSpp) { return maxSpp <= 0 ? -1 : previewSteps + maxSpp; }
bool canTakeSampleAtIteration(int sampleCount, int previewSteps, int maxSpp)
{
    return maxSpp <= 0 || sampleCount < previewSteps + maxSpp;
}
bool budgetExhaustedAtIteration(int sampleCount, int previewSteps, int maxSpp)
{
    return maxSpp > 0 && sampleCount >= previewSteps + maxSpp;
}

// PathTracerPreviewLevels.h: level 0 is 1/2^N of full resolution, level N-1 is 1/2.
struct LevelDims {
    int width, height, downscale;
};
LevelDims previewLevelDims(int fullW, int fullH, int levels, int index)
{
    int ds = 1;
    if (levels > 0 && index >= 0 && index < levels) {
        ds = 1 << (levels - index);
    }
    auto axis = [ds](int full) { return full <= 0 || ds <= 0 ? 0 : std::max(1, full / ds); };
    return {axis(fullW), axis(fullH), ds};
}

// PhysicalCamera.cpp:397-402
float exposureMultiplierOf(float fStop, float shutter, float iso)
{
    const float f = fStop < 0.1f ? 0.1f : fStop;
    const float i = iso < 1.0f ? 1.0f : iso;
    return shutter * (i / 100.0f) / (f * f);
}

float clampFocal(float mm) { return std::max(kMinFocalLengthMm, std::min(mm, kMaxFocalLengthMm)); }
float fovYToFocalLengthMm(float fovY)
{
    const float tanHalf = std::tan(fovY * 0.5f);
    return tanHalf <= 0.0f ? kMaxFocalLengthMm : clampFocal(kSensorHeightMm / (2.0f * tanHalf));
}

// Rotates v by the unit quaternion stored as (w, x, y, z) in CameraGpu::orientation.
Vec3 rotateByQuat(const PtFloat4& q, Vec3 v)
{
    const float w = q.x, qx = q.y, qy = q.z, qz = q.w;
    const Vec3 t = {2.0f * (qy * v.z - qz * v.y), 2.0f * (qz * v.x - qx * v.z), 2.0f * (qx * v.y - qy * v.x)};
    return {v.x + w * t.x + (qy * t.z - qz * t.y), v.y + w * t.y + (qz * t.x - qx * t.z),
            v.z + w * t.z + (qx * t.y - qy * t.x)};
}

enum class WaitResult { Completed, ResetDuringWait, Stopped, Error };

} // namespace

namespace ptdetail {

struct Accumulator {
    PtFloat4* buffer = nullptr;
    uint32_t* samples = nullptr;
    float* lumMean = nullptr;
    float* m2 = nullptr;
    uint8_t* converged = nullptr;
    int* activeIndices = nullptr;
    int* activeScratch = nullptr;
    int* activeCount = nullptr;
    int width = 0;
    int height = 0;
};

struct PreviewLevel {
    PtFloat4* buffer = nullptr;
    int width = 0, height = 0, downscale = 1;
};

// PhysicalCamera state the renderer uses (geometry, exposure, focus).
struct HostCamera {
    CameraGpu geom{};
    float fStop = kDefaultFStop;
    float shutter = kDefaultShutterSpeedSeconds;
    float iso = kDefaultIso;
    float focalLengthMm = kDefaultFocalLengthMm;
    float focusDistance = kDefaultFocusDistance;
    bool focusValid = true;

    Vec3 forward() const { return rotateByQuat(geom.orientation, Vec3{0.0f, 0.0f, -1.0f}); }

    // PhysicalCamera::toGpu (PhysicalCamera.cpp:314-329) with the aperture of 291-298.
    CameraGpu toGpu() const
    {
        CameraGpu g = geom;
        g.apertureRadius = focusValid && fStop > 0.0f ? focalLengthMm / (2.0f * fStop) : 0.0f;
        g.focusDistance = focusValid ? focusDistance : 0.0f;
        const Vec3 f = forward();
        g.focusPoint = Vec3{geom.position.x + f.x * focusDistance, geom.position.y + f.y * focusDistance,
                            geom.position.z + f.z * focusDistance};
        g.focusValid = focusValid ? 1 : 0;
        return g;
    }

    // PhysicalCamera::applyGpuGeometry (PhysicalCamera.cpp:348-363)
    void applyGpuGeometry(const CameraGpu& c)
    {
        geom = c;
        focalLengthMm = fovYToFocalLengthMm(c.fovY);
        focusDistance = c.focusDistance;
        focusValid = c.focusValid != 0;
    }
};

struct TracerImpl {
    Accumulator acc;
    std::vector<PreviewLevel> previewLevels;
    PtUchar4* display[2] = {nullptr, nullptr};
    CameraGpu* dCamera = nullptr;
    RenderParamsGpu* dParams = nullptr;
    PtScene* scene = nullptr;
    PtEnvironment* env = nullptr;
    std::vector<float> envImage;
    int envWidth = 0, envHeight = 0;
    hipStream_t sampleStream = nullptr;
    hipStream_t displayStream = nullptr;
    hipEvent_t sampleCompleteEvent = nullptr;
    hipEvent_t displayReadyEvent = nullptr;

    std::thread worker;
    std::mutex workerMutex;
    std::condition_variable workerCv;
    mutable std::mutex cameraMutex;
    mutable std::mutex lastSampleCameraMutex;
    mutable std::mutex sceneMutex;
    std::mutex streamMutex;
    std::mutex errorMutex;
    std::string workerError;  // the last failed launch of the worker thread (its own error text)
    std::mutex callbackMutex;
    PathTracer::FrameReadyCallback callback;

    HostCamera hostCamera;
    CameraGpu lastSampleCamera{};
    RenderParamsGpu hostParams{};
    bool deviceParamsArePreview = false;

    std::atomic<bool> cameraDirty{true};
    std::atomic<bool> displayPublishInFlight{false};
    std::atomic<int> displayPublishSlot{-1};
    std::atomic<bool> running{false};
    std::atomic<bool> resetRequested{false};
    std::atomic<bool> regionResetRequested{false};
    std::atomic<bool> configured{false};
    std::atomic<int> maxSamplesPerPixel{8};
    std::atomic<int> minSamples{16};
    std::atomic<float> relativeErrorThreshold{0.02f};
    std::atomic<int> debugOverlayMode{0};
    std::atomic<int> brdfDebugFlags{0};
    std::atomic<int> activePixelCount{0};
    std::atomic<int> previewStepsPerLevel{0};
    std::atomic<int> passesPerLaunch{1};
    std::atomic<int> bvhBuildMode{0};
    std::atomic<int> sampleCount{0};
    std::atomic<int> finestCompletedPreview{-1};
    std::atomic<bool> displayBuffersValid{false};
    std::atomic<bool> environmentDirty{true};
    std::atomic<bool> regionEnabled{false};
    std::atomic<int> regionMinX{0}, regionMinY{0}, regionMaxX{0}, regionMaxY{0};
    std::atomic<bool> renderParamsDirty{true};
    std::atomic<bool> idle{false};

    TracerImpl()
    {
        hostParams.backgroundR = hostParams.backgroundG = hostParams.backgroundB = 10.0f / 255.0f;
        hostParams.environmentIntensity = 1.0f;
        hostParams.maxPathDepth = 32;
        hostParams.russianRouletteMinDepth = 3;
        hostParams.minSamples = 16;
        hostParams.maxSamplesPerPixel = 1024;
        hostParams.relativeErrorThreshold = 0.02f;
        CameraGpu& g = hostCamera.geom;
        g.orientation = PtFloat4{1.0f, 0.0f, 0.0f, 0.0f};
        g.fovY = 2.0f * std::atan(kSensorHeightMm / (2.0f * kDefaultFocalLengthMm));
        g.aspect = 1.0f;
        g.nearPlane = kDefaultFocalLengthMm;
        g.farPlane = kDefaultFarPlane;
    }
};

} // namespace ptdetail

using ptdetail::TracerImpl;

namespace {

template <typename T>
bool devAlloc(T** p, size_t count)
{
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(1, count) * sizeof(T)) == hipSuccess;
}

template <typename T>
void devFree(T*& p)
{
    if (p != nullptr) {
        (void)hipFree(p);
        p = nullptr;
    }
}

void freeAccumulator(ptdetail::Accumulator& a)
{
    devFree(a.buffer);
    devFree(a.samples);
    devFree(a.lumMean);
    devFree(a.m2);
    devFree(a.converged);
    devFree(a.activeIndices);
    devFree(a.activeScratch);
    devFree(a.activeCount);
    a.width = a.height = 0;
}

bool initAccumulator(ptdetail::Accumulator& a, int w, int h)
{
    freeAccumulator(a);
    const size_t n = static_cast<size_t>(w) * static_cast<size_t>(h);
    if (!devAlloc(&a.buffer, n) || !devAlloc(&a.samples, n) || !devAlloc(&a.lumMean, n) || !devAlloc(&a.m2, n) ||
        !devAlloc(&a.converged, n) || !devAlloc(&a.activeIndices, n) || !devAlloc(&a.activeScratch, n) ||
        !devAlloc(&a.activeCount, 1)) {
        freeAccumulator(a);
        return false;
    }
    a.width = w;
    a.height = h;
    return true;
}

void freePreviewLevels(std::vector<ptdetail::PreviewLevel>& levels)
{
    for (auto& l : levels) {
        devFree(l.buffer);
    }
    levels.clear();
}

bool initPreviewLevels(std::vector<ptdetail::PreviewLevel>& levels, int w, int h, int count)
{
    freePreviewLevels(levels);
    for (int i = 0; i < count; ++i) {
        const LevelDims d = previewLevelDims(w, h, count, i);
        ptdetail::PreviewLevel l;
        l.width = d.width;
        l.height = d.height;
        l.downscale = d.downscale;
        if (!devAlloc(&l.buffer, static_cast<size_t>(d.width) * static_cast<size_t>(d.height))) {
            freePreviewLevels(levels);
            return false;
        }
        levels.push_back(l);
    }
    return true;
}

bool clearPreviewLevels(TracerImpl* impl)
{
    for (const auto& l : impl->previewLevels) {
        if (hipMemsetAsync(l.buffer, 0, sizeof(PtFloat4) * static_cast<size_t>(l.width) * l.height,
                           impl->sampleStream) != hipSuccess) {
            return false;
        }
    }
    return true;
}

// PathTracer.cpp:511-542
bool clearAccumulator(TracerImpl* impl)
{
    auto& a = impl->acc;
    if (a.buffer == nullptr || a.width <= 0) {
        return true;
    }
    if (!pathTracerClearAccumulator(a.buffer, a.samples, a.lumMean, a.m2, a.converged, a.activeIndices,
                                    a.activeScratch, a.activeCount, a.width, a.height, impl->sampleStream)) {
        return false;
    }
    impl->activePixelCount.store(a.width * a.height);
    return true;
}

// PathTracer.cpp:544-583
bool clearRegionAccumulator(TracerImpl* impl)
{
    auto& a = impl->acc;
    if (a.buffer == nullptr || a.width <= 0) {
        return true;
    }
    const int minX = impl->regionMinX.load(), minY = impl->regionMinY.load();
    const int maxX = impl->regionMaxX.load(), maxY = impl->regionMaxY.load();
    if (!pathTracerClearRegionAccumulator(a.buffer, a.samples, a.lumMean, a.m2, a.converged, a.activeIndices,
                                          a.activeCount, a.width, a.height, minX, minY, maxX, maxY,
                                          impl->sampleStream)) {
        return false;
    }
    impl->activePixelCount.store((maxX - minX + 1) * (maxY - minY + 1));
    return true;
}

// PathTracer.cpp:585-632
bool rebuildFullActiveList(TracerImpl* impl)
{
    auto& a = impl->acc;
    if (a.buffer == nullptr || a.width <= 0) {
        return true;
    }
    if (!pathTracerRebuildActiveList(a.converged, a.samples, a.activeIndices, a.activeCount, a.width, a.height,
                                     impl->maxSamplesPerPixel.load(), impl->sampleStream)) {
        return false;
    }
    int count = 0;
    if (hipMemcpyAsync(&count, a.activeCount, sizeof(int), hipMemcpyDeviceToHost, impl->sampleStream) != hipSuccess ||
        hipStreamSynchronize(impl->sampleStream) != hipSuccess) {
        return false;
    }
    impl->activePixelCount.store(count);
    return true;
}

CameraGpu buildCameraGpu(const TracerImpl* impl)
{
    CameraGpu c;
    {
        std::lock_guard<std::mutex> lock(impl->cameraMutex);
        c = impl->hostCamera.toGpu();
    }
    if (impl->acc.width > 0 && impl->acc.height > 0) {
        c.aspect = static_cast<float>(impl->acc.width) / static_cast<float>(impl->acc.height);
    }
    return c;
}

// PathTracer.cpp:873-898
bool uploadCameraIfDirty(TracerImpl* impl, hipStream_t stream)
{
    if (impl->dCamera == nullptr) {
        return false;
    }
    if (!impl->cameraDirty.load()) {
        return true;
    }
    static thread_local CameraGpu staged;
    staged = buildCameraGpu(impl);
    if (hipMemcpyAsync(impl->dCamera, &staged, sizeof(CameraGpu), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        return false;
    }
    impl->cameraDirty.store(false);
    return true;
}

// PathTracer.cpp:727-775 (previewPass: depth kPreviewMaxPathDepth)
bool syncRenderParamsToDevice(TracerImpl* impl, hipStream_t stream, bool previewPass, bool force = false)
{
    if (impl->dParams == nullptr) {
        return false;
    }
    if (previewPass) {
        if (impl->deviceParamsArePreview) {
            return true;
        }
    } else if (!force && !impl->renderParamsDirty.load() && !impl->deviceParamsArePreview) {
        return true;
    }
    RenderParamsGpu p = impl->hostParams;
    p.minSamples = clampMinSamples(impl->minSamples.load());
    p.maxSamplesPerPixel = clampMaxSamples(impl->maxSamplesPerPixel.load());
    p.relativeErrorThreshold = clampRelativeErrorThreshold(impl->relativeErrorThreshold.load());
    p.debugOverlayMode = impl->debugOverlayMode.load();
    p.brdfDebugFlags = impl->brdfDebugFlags.load();
    if (previewPass) {
        p.maxPathDepth = kPreviewMaxPathDepth;
    }
    if (hipMemcpyAsync(impl->dParams, &p, sizeof(p), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        return false;
    }
    impl->deviceParamsArePreview = previewPass;
    if (!previewPass) {
        impl->renderParamsDirty.store(false);
    }
    return true;
}

bool uploadEnvironmentIfDirty(TracerImpl* impl, hipStream_t stream)
{
    if (!impl->environmentDirty.load()) {
        return true;
    }
    if (impl->env != nullptr) {
        ptEnvironmentDestroy(impl->env);
        impl->env = nullptr;
    }
    if (!impl->envImage.empty()) {
        impl->env = ptEnvironmentCreate(impl->envImage.data(), impl->envWidth, impl->envHeight, stream);
        if (impl->env == nullptr) {
            return false;
        }
    }
    impl->environmentDirty.store(false);
    return true;
}

bool canTakeSample(const TracerImpl* impl)
{
    const int previewSteps = impl->previewStepsPerLevel.load();
    const int count = impl->sampleCount.load();
    if (count < previewSteps) {
        return canTakeSampleAtIteration(count, previewSteps, impl->maxSamplesPerPixel.load());
    }
    return impl->activePixelCount.load() > 0;
}

bool hasDisplayableContent(const TracerImpl& impl)
{
    const int count = impl.sampleCount.load();
    const int previewCount = impl.previewStepsPerLevel.load();
    if (count <= 0) {
        return false;
    }
    if (previewCount > 0 && count <= previewCount) {
        return impl.finestCompletedPreview.load() >= 0;
    }
    return true;
}

WaitResult waitForSampleEvent(TracerImpl* impl)
{
    bool resetSeen = false;
    for (;;) {
        if (!impl->running.load()) {
            return WaitResult::Stopped;
        }
        if (impl->resetRequested.load() || impl->regionResetRequested.load()) {
            resetSeen = true;
        }
        const hipError_t q = hipEventQuery(impl->sampleCompleteEvent);
        if (q == hipSuccess) {
            return resetSeen ? WaitResult::ResetDuringWait : WaitResult::Completed;
        }
        if (q != hipErrorNotReady) {
            return WaitResult::Error;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
}

// PathTracer.cpp:1477-1617 with the two GL PBOs replaced by device buffers.
bool launchDisplayPublishLocked(TracerImpl* impl, int slot)
{
    if (slot < 0 || slot > 1 || impl->display[slot] == nullptr) {
        return false;
    }
    const int overlay = impl->debugOverlayMode.load();
    const bool uvMode = overlay == 3; // RenderViewOverlayMode::Uv
    if (!uvMode && (!impl->displayBuffersValid.load() || !hasDisplayableContent(*impl))) {
        return false;
    }
    if (!uvMode && hipStreamWaitEvent(impl->displayStream, impl->sampleCompleteEvent, 0) != hipSuccess) {
        return false;
    }
    if (!syncRenderParamsToDevice(impl, impl->displayStream, false, impl->renderParamsDirty.load()) ||
        !uploadCameraIfDirty(impl, impl->displayStream)) {
        return false;
    }
    const int previewCount = impl->previewStepsPerLevel.load();
    const int iter = impl->sampleCount.load();
    const int finest = impl->finestCompletedPreview.load();
    const bool showPreview = !uvMode && !impl->regionEnabled.load() && previewCount > 0 && finest >= 0 &&
                             finest < static_cast<int>(impl->previewLevels.size()) && iter <= previewCount;
    float exposure = 1.0f;
    {
        std::lock_guard<std::mutex> lock(impl->cameraMutex);
        exposure = exposureMultiplierOf(impl->hostCamera.fStop, impl->hostCamera.shutter, impl->hostCamera.iso);
    }
    bool ok = false;
    if (uvMode) {
        ok = pathTracerWriteUvDebugToPbo(impl->dCamera, ptSceneDevice(impl->scene), impl->display[slot],
                                         impl->acc.width, impl->acc.height, impl->dParams, impl->displayStream);
    } else if (showPreview) {
        const auto& l = impl->previewLevels[static_cast<size_t>(finest)];
        ok = pathTracerUpsamplePreviewToPbo(l.buffer, l.width, l.height, l.downscale, impl->display[slot],
                                            impl->acc.width, impl->acc.height, impl->dParams, exposure,
                                            impl->displayStream);
    } else {
        ok = pathTracerCopyToPbo(impl->acc.buffer, impl->acc.samples, impl->acc.converged, impl->display[slot],
                                 impl->acc.width, impl->acc.height, impl->dParams, exposure, impl->displayStream);
    }
    return ok && hipEventRecord(impl->displayReadyEvent, impl->displayStream) == hipSuccess;
}

void releaseAll(TracerImpl* impl)
{
    freeAccumulator(impl->acc);
    freePreviewLevels(impl->previewLevels);
    devFree(impl->display[0]);
    devFree(impl->display[1]);
    devFree(impl->dCamera);
    devFree(impl->dParams);
    if (impl->env != nullptr) {
        ptEnvironmentDestroy(impl->env);
        impl->env = nullptr;
    }
    impl->environmentDirty.store(true);
    if (impl->scene != nullptr) {
        ptSceneDestroy(impl->scene);
        impl->scene = nullptr;
    }
}

bool buildScene(TracerImpl* impl, const PtHostMesh& mesh)
{
    std::lock_guard<std::mutex> lock(impl->sceneMutex);
    if (impl->scene == nullptr) {
        impl->scene = ptSceneCreate();
        if (impl->scene == nullptr) {
            return false;
        }
    }
    if (!ptSceneSetBuildMode(impl->scene, impl->bvhBuildMode.load())) {
        return false;
    }
    return ptSceneBuild(impl->scene, mesh.triangles, mesh.triangleCount, mesh.materials, mesh.materialCount,
                        mesh.textures, mesh.textureCount, impl->sampleStream);
}

} // namespace

// ---- PathTracer -----------------------------------------------------------------------------

PathTracer::PathTracer() : m_impl(new TracerImpl()) {}

PathTracer::~PathTracer()
{
    stop();
    releaseAll(m_impl.get());
    if (m_impl->sampleCompleteEvent != nullptr) {
        (void)hipEventDestroy(m_impl->sampleCompleteEvent);
    }
    if (m_impl->displayReadyEvent != nullptr) {
        (void)hipEventDestroy(m_impl->displayReadyEvent);
    }
    if (m_impl->sampleStream != nullptr) {
        (void)hipStreamDestroy(m_impl->sampleStream);
    }
    if (m_impl->displayStream != nullptr) {
        (void)hipStreamDestroy(m_impl->displayStream);
    }
}

void PathTracer::setFrameReadyCallback(FrameReadyCallback callback)
{
    std::lock_guard<std::mutex> lock(m_impl->callbackMutex);
    m_impl->callback = std::move(callback);
}

bool PathTracer::configure(int width, int height, const PtHostMesh& mesh)
{
    TracerImpl* impl = m_impl.get();
    if (width <= 0 || height <= 0) {
        ptSetError("PathTracer::configure: invalid dimensions");
        return false;
    }
    if (impl->running.load()) {
        stop();
    }
    impl->configured.store(false);
    releaseAll(impl);
    if ((impl->sampleStream == nullptr && hipStreamCreate(&impl->sampleStream) != hipSuccess) ||
        (impl->displayStream == nullptr && hipStreamCreate(&impl->displayStream) != hipSuccess) ||
        (impl->sampleCompleteEvent == nullptr &&
         hipEventCreateWithFlags(&impl->sampleCompleteEvent, hipEventDisableTiming) != hipSuccess) ||
        (impl->displayReadyEvent == nullptr &&
         hipEventCreateWithFlags(&impl->displayReadyEvent, hipEventDisableTiming) != hipSuccess)) {
        ptSetError("PathTracer::configure: stream or event creation failed");
        return false;
    }
    impl->displayPublishInFlight.store(false);
    impl->displayPublishSlot.store(-1);
    const size_t px = static_cast<size_t>(width) * static_cast<size_t>(height);
    bool ok = initAccumulator(impl->acc, width, height) &&
              initPreviewLevels(impl->previewLevels, width, height, impl->previewStepsPerLevel.load()) &&
              devAlloc(&impl->display[0], px) && devAlloc(&impl->display[1], px) && devAlloc(&impl->dCamera, 1) &&
              devAlloc(&impl->dParams, 1);
    if (!ok) {
        ptSetError("PathTracer::configure: device allocation failed");
        releaseAll(impl);
        return false;
    }
    impl->cameraDirty.store(true);
    impl->renderParamsDirty.store(true);
    impl->deviceParamsArePreview = false;
    ok = buildScene(impl, mesh) && uploadEnvironmentIfDirty(impl, impl->sampleStream) &&
         uploadCameraIfDirty(impl, impl->sampleStream) && syncRenderParamsToDevice(impl, impl->sampleStream, false, true) &&
         clearAccumulator(impl) && clearPreviewLevels(impl) && hipStreamSynchronize(impl->sampleStream) == hipSuccess &&
         hipStreamSynchronize(impl->displayStream) == hipSuccess;
    if (!ok) {
        if (pathTracerLastError()[0] == '\0') {
            ptSetError("PathTracer::configure: scene build or upload failed");
        }
        releaseAll(impl);
        return false;
    }
    impl->resetRequested.store(false);
    impl->regionResetRequested.store(false);
    impl->sampleCount.store(0);
    impl->finestCompletedPreview.store(-1);
    impl->displayBuffersValid.store(false);
    impl->configured.store(true);
    return true;
}

bool PathTracer::rebuildMeshScene(const PtHostMesh& mesh)
{
    TracerImpl* impl = m_impl.get();
    if (!impl->configured.load()) {
        return false;
    }
    bool ok = false;
    {
        std::lock_guard<std::mutex> streamLock(impl->streamMutex);
        ok = hipStreamSynchronize(impl->sampleStream) == hipSuccess && buildScene(impl, mesh);
    }
    if (ok) {
        resetAccumulation();
    }
    return ok;
}

void PathTracer::start()
{
    if (m_impl->running.load() || !m_impl->configured.load()) {
        return;
    }
    m_impl->running.store(true);
    m_impl->worker = std::thread([this]() { renderLoop(); });
}

void PathTracer::stop()
{
    if (!m_impl->running.load()) {
        if (m_impl->worker.joinable()) {
            m_impl->worker.join();
        }
        return;
    }
    m_impl->running.store(false);
    notifyWorker();
    if (m_impl->worker.joinable()) {
        m_impl->worker.join();
    }
}

void PathTracer::resetAccumulation()
{
    m_impl->resetRequested.store(true);
    notifyWorker();
}

void PathTracer::resetRegionAccumulation()
{
    m_impl->regionResetRequested.store(true);
    notifyWorker();
}

bool PathTracer::isRunning() const { return m_impl->running.load(); }

void PathTracer::releaseOutputSurfaces()
{
    stop();
    devFree(m_impl->display[0]);
    devFree(m_impl->display[1]);
    m_impl->configured.store(false);
}

bool PathTracer::beginDisplayPublish(int slot)
{
    TracerImpl* impl = m_impl.get();
    if (slot < 0 || slot > 1 || !impl->configured.load() || impl->displayPublishInFlight.load()) {
        return false;
    }
    std::lock_guard<std::mutex> streamLock(impl->streamMutex);
    if (!launchDisplayPublishLocked(impl, slot)) {
        return false;
    }
    impl->displayPublishSlot.store(slot);
    impl->displayPublishInFlight.store(true);
    return true;
}

bool PathTracer::isDisplayPublishReady() const
{
    if (!m_impl->displayPublishInFlight.load() || m_impl->displayReadyEvent == nullptr) {
        return false;
    }
    return hipEventQuery(m_impl->displayReadyEvent) == hipSuccess;
}

bool PathTracer::hasDisplayPublishInFlight() const { return m_impl->displayPublishInFlight.load(); }

int PathTracer::finishDisplayPublish()
{
    if (!isDisplayPublishReady()) {
        return -1;
    }
    const int slot = m_impl->displayPublishSlot.exchange(-1);
    m_impl->displayPublishInFlight.store(false);
    return slot;
}

bool PathTracer::publishDisplayFrame(int slot)
{
    if (!beginDisplayPublish(slot)) {
        return false;
    }
    while (!isDisplayPublishReady()) {
        std::this_thread::yield();
    }
    return finishDisplayPublish() >= 0;
}

bool PathTracer::readDisplay(int slot, PtUchar4* hostRgba) const
{
    const TracerImpl* impl = m_impl.get();
    if (slot < 0 || slot > 1 || hostRgba == nullptr || impl->display[slot] == nullptr) {
        return false;
    }
    const size_t n = static_cast<size_t>(impl->acc.width) * static_cast<size_t>(impl->acc.height);
    return hipStreamSynchronize(impl->displayStream) == hipSuccess &&
           hipMemcpy(hostRgba, impl->display[slot], n * sizeof(PtUchar4), hipMemcpyDeviceToHost) == hipSuccess;
}

void PathTracer::setMaxSamplesPerPixel(int max)
{
    const int c = clampMaxSamples(max);
    m_impl->maxSamplesPerPixel.store(c);
    m_impl->hostParams.maxSamplesPerPixel = c;
    m_impl->renderParamsDirty.store(true);
    notifyWorker();
}
int PathTracer::maxSamplesPerPixel() const { return m_impl->maxSamplesPerPixel.load(); }

void PathTracer::setMinSamples(int min)
{
    const int c = clampMinSamples(min);
    m_impl->minSamples.store(c);
    m_impl->hostParams.minSamples = c;
    m_impl->renderParamsDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
    notifyWorker();
}
int PathTracer::minSamples() const { return m_impl->minSamples.load(); }

void PathTracer::setRelativeErrorThreshold(float threshold)
{
    const float c = clampRelativeErrorThreshold(threshold);
    m_impl->relativeErrorThreshold.store(c);
    m_impl->hostParams.relativeErrorThreshold = c;
    m_impl->renderParamsDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
    notifyWorker();
}
float PathTracer::relativeErrorThreshold() const { return m_impl->relativeErrorThreshold.load(); }

void PathTracer::setDebugOverlayMode(int mode)
{
    mode = mode < 0 ? 0 : (mode > 3 ? 3 : mode);
    m_impl->debugOverlayMode.store(mode);
    m_impl->hostParams.debugOverlayMode = mode;
    m_impl->renderParamsDirty.store(true);
    notifyWorker();
}
int PathTracer::debugOverlayMode() const { return m_impl->debugOverlayMode.load(); }

void PathTracer::setBrdfDebugFlags(int flags)
{
    m_impl->brdfDebugFlags.store(flags);
    m_impl->hostParams.brdfDebugFlags = flags;
    m_impl->renderParamsDirty.store(true);
    notifyWorker();
}
int PathTracer::brdfDebugFlags() const { return m_impl->brdfDebugFlags.load(); }

void PathTracer::setPreviewStepsPerLevel(int steps)
{
    TracerImpl* impl = m_impl.get();
    const int c = clampPreviewStepsPerLevel(steps);
    impl->previewStepsPerLevel.store(c);
    if (impl->configured.load() && impl->acc.width > 0) {
        std::lock_guard<std::mutex> streamLock(impl->streamMutex);
        if (hipStreamSynchronize(impl->sampleStream) == hipSuccess &&
            initPreviewLevels(impl->previewLevels, impl->acc.width, impl->acc.height, c) && clearPreviewLevels(impl)) {
            (void)hipStreamSynchronize(impl->sampleStream);
        }
    }
    notifyWorker();
}
int PathTracer::previewStepsPerLevel() const { return m_impl->previewStepsPerLevel.load(); }

int PathTracer::currentSampleCount() const { return m_impl->sampleCount.load(); }
int PathTracer::currentActivePixelCount() const { return m_impl->activePixelCount.load(); }

int PathTracer::sampleBudgetTotalIterations() const
{
    return budgetTotalIterations(m_impl->previewStepsPerLevel.load(), m_impl->maxSamplesPerPixel.load());
}

bool PathTracer::isSampleBudgetExhausted() const
{
    const int previewSteps = m_impl->previewStepsPerLevel.load();
    const int count = m_impl->sampleCount.load();
    if (count < previewSteps) {
        return budgetExhaustedAtIteration(count, previewSteps, m_impl->maxSamplesPerPixel.load());
    }
    return m_impl->activePixelCount.load() <= 0; // isAdaptiveSampleBudgetExhausted
}

void PathTracer::setCamera(const CameraGpu& camera)
{
    CameraGpu stored = camera;
    if (m_impl->acc.width > 0 && m_impl->acc.height > 0) {
        stored.aspect = static_cast<float>(m_impl->acc.width) / static_cast<float>(m_impl->acc.height);
    }
    {
        std::lock_guard<std::mutex> lock(m_impl->cameraMutex);
        m_impl->hostCamera.applyGpuGeometry(stored);
    }
    {
        std::lock_guard<std::mutex> lock(m_impl->lastSampleCameraMutex);
        m_impl->lastSampleCamera = stored;
    }
    m_impl->cameraDirty.store(true);
    notifyWorker();
}

CameraGpu PathTracer::lastSampleCamera() const
{
    std::lock_guard<std::mutex> lock(m_impl->lastSampleCameraMutex);
    return m_impl->lastSampleCamera;
}

void PathTracer::setClearColor(float r, float g, float b)
{
    m_impl->hostParams.backgroundR = r;
    m_impl->hostParams.backgroundG = g;
    m_impl->hostParams.backgroundB = b;
    m_impl->renderParamsDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
}

void PathTracer::setPhysicalCamera(float fStop, float shutterSpeedSeconds, float iso)
{
    std::lock_guard<std::mutex> lock(m_impl->cameraMutex);
    m_impl->hostCamera.fStop = std::max(kMinFStop, std::min(fStop, kMaxFStop));
    m_impl->hostCamera.shutter = std::max(kMinShutterSpeedSeconds, std::min(shutterSpeedSeconds, kMaxShutterSpeedSeconds));
    m_impl->hostCamera.iso = std::max(kMinIso, std::min(iso, kMaxIso));
    m_impl->cameraDirty.store(true);
}

void PathTracer::setFocusPoint(const Vec3& p)
{
    {
        std::lock_guard<std::mutex> lock(m_impl->cameraMutex);
        auto& c = m_impl->hostCamera;
        const Vec3 f = c.forward();
        const float d = (p.x - c.geom.position.x) * f.x + (p.y - c.geom.position.y) * f.y +
                        (p.z - c.geom.position.z) * f.z;
        c.focusDistance = d > 0.0f ? d : 0.0f;
        c.focusValid = c.focusDistance > 0.0f;
    }
    m_impl->cameraDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
    notifyWorker();
}

void PathTracer::clearFocusPoint()
{
    {
        std::lock_guard<std::mutex> lock(m_impl->cameraMutex);
        m_impl->hostCamera.focusValid = false;
    }
    m_impl->cameraDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
    notifyWorker();
}

// PathTracer.cpp:1991-2013: closest hit in [0.001, 1e30) along normalize(rd), through the GPU
// traversal (the reference walks its host copy of the BVH).
bool PathTracer::pickSurface(const Vec3& ro, const Vec3& rd, Vec3* hitPoint) const
{
    if (hitPoint == nullptr) {
        return false;
    }
    std::lock_guard<std::mutex> lock(m_impl->sceneMutex);
    const MeshAccelSceneGpu* d_scene = ptSceneDevice(m_impl->scene);
    if (d_scene == nullptr) {
        return false;
    }
    const float len = std::sqrt(rd.x * rd.x + rd.y * rd.y + rd.z * rd.z);
    if (!(len > 0.0f)) {
        return false;
    }
    const float o[3] = {ro.x, ro.y, ro.z};
    const float d[3] = {rd.x / len, rd.y / len, rd.z / len};
    float* dev = nullptr;
    if (!devAlloc(&dev, 14)) {
        return false;
    }
    float out[8] = {};
    bool ok = hipMemcpy(dev, o, sizeof(o), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(dev + 3, d, sizeof(d), hipMemcpyHostToDevice) == hipSuccess &&
              pathTracerTraceRays(d_scene, dev, dev + 3, 1, 0.001f, 1.0e30f, dev + 6, nullptr) &&
              hipMemcpy(out, dev + 6, sizeof(out), hipMemcpyDeviceToHost) == hipSuccess;
    devFree(dev);
    int hit = 0;
    std::memcpy(&hit, &out[0], sizeof(int));
    if (!ok || hit == 0) {
        return false;
    }
    *hitPoint = Vec3{ro.x + d[0] * out[1], ro.y + d[1] * out[1], ro.z + d[2] * out[1]};
    return true;
}

bool PathTracer::setEnvironmentImage(const float* rgb, int width, int height)
{
    if (rgb != nullptr && (width <= 0 || height <= 0)) {
        return false;
    }
    {
        std::lock_guard<std::mutex> streamLock(m_impl->streamMutex);
        if (rgb == nullptr) {
            m_impl->envImage.clear();
            m_impl->envWidth = m_impl->envHeight = 0;
        } else {
            m_impl->envImage.assign(rgb, rgb + 3 * static_cast<size_t>(width) * static_cast<size_t>(height));
            m_impl->envWidth = width;
            m_impl->envHeight = height;
        }
        m_impl->environmentDirty.store(true);
    }
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
    return true;
}

void PathTracer::setEnvironmentIntensity(float intensity)
{
    intensity = intensity < 0.0f ? 0.0f : (intensity > 100.0f ? 100.0f : intensity);
    if (m_impl->hostParams.environmentIntensity == intensity) {
        return;
    }
    m_impl->hostParams.environmentIntensity = intensity;
    m_impl->renderParamsDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
}
float PathTracer::environmentIntensity() const { return m_impl->hostParams.environmentIntensity; }

void PathTracer::setRussianRouletteMinDepth(int depth)
{
    depth = depth < 0 ? 0 : (depth > 256 ? 256 : depth);
    if (m_impl->hostParams.russianRouletteMinDepth == depth) {
        return;
    }
    m_impl->hostParams.russianRouletteMinDepth = depth;
    m_impl->renderParamsDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
}
int PathTracer::russianRouletteMinDepth() const { return m_impl->hostParams.russianRouletteMinDepth; }

void PathTracer::setRegionRenderEnabled(bool enabled)
{
    m_impl->regionEnabled.store(enabled);
    notifyWorker();
}
bool PathTracer::regionRenderEnabled() const { return m_impl->regionEnabled.load(); }

void PathTracer::setRenderRegion(int minX, int minY, int maxX, int maxY)
{
    const int mx = std::max(0, m_impl->acc.width - 1);
    const int my = std::max(0, m_impl->acc.height - 1);
    m_impl->regionMinX.store(std::max(0, std::min(minX, maxX)));
    m_impl->regionMaxX.store(std::min(mx, std::max(minX, maxX)));
    m_impl->regionMinY.store(std::max(0, std::min(minY, maxY)));
    m_impl->regionMaxY.store(std::min(my, std::max(minY, maxY)));
}

PtRegion PathTracer::renderRegion() const
{
    return PtRegion{m_impl->regionMinX.load(), m_impl->regionMinY.load(), m_impl->regionMaxX.load(),
                    m_impl->regionMaxY.load()};
}

bool PathTracer::waitUntilIdle(int timeoutMs)
{
    const auto end = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeoutMs);
    for (;;) {
        if (m_impl->idle.load() && !m_impl->resetRequested.load() && !m_impl->regionResetRequested.load()) {
            return true;
        }
        if (!m_impl->running.load()) {
            return !canTakeSample(m_impl.get());
        }
        if (std::chrono::steady_clock::now() >= end) {
            return false;
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
}

bool PathTracer::readAccumulator(float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* converged) const
{
    const auto& a = m_impl->acc;
    if (a.buffer == nullptr) {
        return false;
    }
    const size_t n = static_cast<size_t>(a.width) * static_cast<size_t>(a.height);
    std::lock_guard<std::mutex> streamLock(m_impl->streamMutex);
    if (hipStreamSynchronize(m_impl->sampleStream) != hipSuccess) {
        return false;
    }
    return (acc4 == nullptr || hipMemcpy(acc4, a.buffer, n * sizeof(PtFloat4), hipMemcpyDeviceToHost) == hipSuccess) &&
           (counts == nullptr || hipMemcpy(counts, a.samples, n * 4, hipMemcpyDeviceToHost) == hipSuccess) &&
           (lumMean == nullptr || hipMemcpy(lumMean, a.lumMean, n * 4, hipMemcpyDeviceToHost) == hipSuccess) &&
           (m2 == nullptr || hipMemcpy(m2, a.m2, n * 4, hipMemcpyDeviceToHost) == hipSuccess) &&
           (converged == nullptr || hipMemcpy(converged, a.converged, n, hipMemcpyDeviceToHost) == hipSuccess);
}

bool PathTracer::readPreviewLevel(int level, float* acc4, int* width, int* height, int* downscale) const
{
    std::lock_guard<std::mutex> streamLock(m_impl->streamMutex);
    if (level < 0 || level >= static_cast<int>(m_impl->previewLevels.size())) {
        return false;
    }
    const auto& l = m_impl->previewLevels[static_cast<size_t>(level)];
    if (width != nullptr) {
        *width = l.width;
    }
    if (height != nullptr) {
        *height = l.height;
    }
    if (downscale != nullptr) {
        *downscale = l.downscale;
    }
    return hipStreamSynchronize(m_impl->sampleStream) == hipSuccess &&
           (acc4 == nullptr || hipMemcpy(acc4, l.buffer, sizeof(PtFloat4) * static_cast<size_t>(l.width) * l.height,
                                         hipMemcpyDeviceToHost) == hipSuccess);
}

float PathTracer::exposureMultiplier() const
{
    std::lock_guard<std::mutex> lock(m_impl->cameraMutex);
    return exposureMultiplierOf(m_impl->hostCamera.fStop, m_impl->hostCamera.shutter, m_impl->hostCamera.iso);
}

void PathTracer::setMaxPathDepth(int depth)
{
    depth = depth > 256 ? 256 : depth;
    if (m_impl->hostParams.maxPathDepth == depth) {
        return;
    }
    m_impl->hostParams.maxPathDepth = depth;
    m_impl->renderParamsDirty.store(true);
    if (m_impl->configured.load()) {
        resetAccumulation();
    }
}
int PathTracer::maxPathDepth() const { return m_impl->hostParams.maxPathDepth; }

std::string PathTracer::lastWorkerError() const
{
    std::lock_guard<std::mutex> lock(m_impl->errorMutex);
    return m_impl->workerError;
}

void PathTracer::setPassesPerLaunch(int passes)
{
    m_impl->passesPerLaunch.store(passes < 1 ? 1 : (passes > 256 ? 256 : passes));
}

int PathTracer::passesPerLaunch() const { return m_impl->passesPerLaunch.load(); }

void PathTracer::setBvhBuildMode(int mode) { m_impl->bvhBuildMode.store(mode < 0 ? 0 : (mode > 2 ? 2 : mode)); }

int PathTracer::bvhBuildMode() const { return m_impl->bvhBuildMode.load(); }

void PathTracer::notifyWorker()
{
    {
        std::lock_guard<std::mutex> lock(m_impl->workerMutex);
    }
    m_impl->workerCv.notify_all();
}

void PathTracer::invokeFrameReadyCallback()
{
    FrameReadyCallback cb;
    {
        std::lock_guard<std::mutex> lock(m_impl->callbackMutex);
        cb = m_impl->callback;
    }
    if (cb) {
        cb();
    }
}

// PathTracer.cpp:2220-2543
void PathTracer::renderLoop()
{
    TracerImpl* impl = m_impl.get();
    while (impl->running.load()) {
        const bool fullReset = impl->resetRequested.exchange(false);
        const bool regionReset = impl->regionResetRequested.exchange(false);
        if (fullReset || regionReset) {
            impl->idle.store(false);
            bool invalidateDisplay = false;
            bool cleared = false;
            {
                std::lock_guard<std::mutex> streamLock(impl->streamMutex);
                if (impl->acc.buffer != nullptr) {
                    if (fullReset) {
                        if (impl->regionEnabled.load()) {
                            cleared = clearRegionAccumulator(impl);
                        } else if (clearAccumulator(impl) && clearPreviewLevels(impl)) {
                            cleared = true;
                            invalidateDisplay = true;
                        }
                    } else {
                        cleared = impl->regionEnabled.load() ? clearRegionAccumulator(impl) : rebuildFullActiveList(impl);
                    }
                }
            }
            if (cleared && hipStreamSynchronize(impl->sampleStream) == hipSuccess && !invalidateDisplay &&
                impl->sampleCount.load() > 0) {
                invokeFrameReadyCallback();
            }
            if (invalidateDisplay) {
                impl->displayBuffersValid.store(false);
            }
            if (fullReset) {
                impl->sampleCount.store(0);
                impl->finestCompletedPreview.store(-1);
            }
        }
        if (impl->acc.width <= 0 || impl->acc.height <= 0) {
            continue;
        }
        {
            std::unique_lock<std::mutex> lock(impl->workerMutex);
            if (!canTakeSample(impl) && !impl->resetRequested.load() && !impl->regionResetRequested.load()) {
                impl->idle.store(true);
            }
            impl->workerCv.wait(lock, [impl]() {
                return !impl->running.load() || impl->resetRequested.load() || impl->regionResetRequested.load() ||
                       canTakeSample(impl);
            });
            if (!impl->running.load()) {
                break;
            }
        }
        if (!canTakeSample(impl)) {
            continue;
        }
        impl->idle.store(false);

        const int previewCount = impl->previewStepsPerLevel.load();
        const int iter = impl->sampleCount.load();
        const bool previewPass = !impl->regionEnabled.load() && previewCount > 0 && iter < previewCount &&
                                 iter < static_cast<int>(impl->previewLevels.size());
        bool launched = false;
        bool pendingUpdate = false;
        int passesTaken = 1;
        static thread_local int pendingCount = 0;
        {
            std::lock_guard<std::mutex> streamLock(impl->streamMutex);
            if (impl->displayPublishInFlight.load() &&
                hipStreamWaitEvent(impl->sampleStream, impl->displayReadyEvent, 0) != hipSuccess) {
                continue;
            }
            if (!uploadCameraIfDirty(impl, impl->sampleStream) ||
                !syncRenderParamsToDevice(impl, impl->sampleStream, previewPass) ||
                !uploadEnvironmentIfDirty(impl, impl->sampleStream)) {
                continue;
            }
            const unsigned int globalSeed = static_cast<unsigned int>(iter) + 1u;
            const MeshAccelSceneGpu* d_scene = nullptr;
            {
                std::lock_guard<std::mutex> sceneLock(impl->sceneMutex);
                d_scene = ptSceneDevice(impl->scene);
            }
            const EnvironmentMapGpu* d_env = impl->env != nullptr ? ptEnvironmentDevice(impl->env) : nullptr;
            if (previewPass) {
                const auto& l = impl->previewLevels[static_cast<size_t>(iter)];
                launched = pathTracerPreviewSample(l.buffer, l.width, l.height, impl->dCamera, d_scene, d_env,
                                                   impl->dParams, globalSeed, impl->sampleStream);
            } else {
                const int active = impl->activePixelCount.load();
                const int fullResWave = iter - previewCount + 1;
                const int perLaunch = impl->passesPerLaunch.load();
                bool runCompact = active > 0 && (fullResWave % kCompactActiveListInterval) == 0;
                auto& a = impl->acc;
                if (active > 0 && perLaunch > 1) {
                    // several passes in one wavefront render (same per-pixel results), then a
                    // compaction of the caller-side active list
                    const int depth = impl->hostParams.maxPathDepth;
                    const int resolved = depth > 0 ? depth : (depth == 0 ? 256 : 32);
                    passesTaken = perLaunch;
                    launched = pathTracerAdaptiveRenderWavefront(
                        a.buffer, a.samples, a.lumMean, a.m2, a.converged, a.activeIndices, active, a.width, a.height,
                        impl->dCamera, d_scene, d_env, impl->dParams, globalSeed, perLaunch, resolved, 0,
                        impl->sampleStream);
                    runCompact = true;
                } else if (active > 0) {
                    launched = pathTracerAdaptiveSample(a.buffer, a.samples, a.lumMean, a.m2, a.converged,
                                                        a.activeIndices, active, a.width, a.height, impl->dCamera,
                                                        d_scene, d_env, impl->dParams, globalSeed, impl->sampleStream);
                }
                if (launched && runCompact) {
                    launched = pathTracerCompactActiveList(
                        a.activeIndices, a.activeScratch, a.activeCount, a.converged, a.samples,
                        impl->maxSamplesPerPixel.load(), active, a.width, a.height, impl->regionEnabled.load(),
                        impl->regionMinX.load(), impl->regionMinY.load(), impl->regionMaxX.load(),
                        impl->regionMaxY.load(), impl->sampleStream);
                    if (launched) {
                        pendingCount = 0;
                        launched = hipMemcpyAsync(&pendingCount, a.activeCount, sizeof(int), hipMemcpyDeviceToHost,
                                                  impl->sampleStream) == hipSuccess;
                        pendingUpdate = launched;
                    }
                }
            }
            if (launched && hipEventRecord(impl->sampleCompleteEvent, impl->sampleStream) != hipSuccess) {
                launched = false;
                pendingUpdate = false;
            }
        }
        if (!launched) {
            {
                std::lock_guard<std::mutex> errorLock(impl->errorMutex);
                impl->workerError = pathTracerLastError();
            }
            if (!previewPass && impl->activePixelCount.load() > 0) {
                std::this_thread::sleep_for(std::chrono::milliseconds(1));
            }
            continue;
        }
        const WaitResult wr = waitForSampleEvent(impl);
        if (wr == WaitResult::Error) {
            continue;
        }
        if (wr == WaitResult::Stopped) {
            break;
        }
        if (wr == WaitResult::ResetDuringWait) {
            continue;
        }
        if (pendingUpdate) {
            impl->activePixelCount.store(pendingCount);
            std::swap(impl->acc.activeIndices, impl->acc.activeScratch);
        }
        if (previewCount > 0 && iter < previewCount) {
            impl->finestCompletedPreview.store(iter);
        }
        impl->sampleCount.fetch_add(passesTaken);
        impl->displayBuffersValid.store(true);
        invokeFrameReadyCallback();
    }
    impl->idle.store(false);
}

// ---- C wrappers -----------------------------------------------------------------------------

struct PtTracer {
    PathTracer tracer;
};

namespace {
PtHostMesh meshOf(const HostTriangle* tris, uint32_t n, const MaterialGpu* mats, uint32_t m,
                  const TextureDescGpu* tex, uint32_t t)
{
    PtHostMesh mesh;
    mesh.triangles = tris;
    mesh.triangleCount = n;
    mesh.materials = mats;
    mesh.materialCount = m;
    mesh.textures = tex;
    mesh.textureCount = t;
    return mesh;
}
} // namespace

extern "C" {

PtTracer* ptTracerCreate(void) { return new PtTracer(); }
void ptTracerDestroy(PtTracer* t) { delete t; }
void ptTracerSetFrameReadyCallback(PtTracer* t, PtFrameReadyFn fn, void* user)
{
    if (t == nullptr) {
        return;
    }
    if (fn == nullptr) {
        t->tracer.setFrameReadyCallback(nullptr);
    } else {
        t->tracer.setFrameReadyCallback([fn, user]() { fn(user); });
    }
}
bool ptTracerConfigure(PtTracer* t, int width, int height, const HostTriangle* triangles, uint32_t triangleCount,
                       const MaterialGpu* materials, uint32_t materialCount, const TextureDescGpu* textures,
                       uint32_t textureCount)
{
    return t != nullptr && t->tracer.configure(width, height, meshOf(triangles, triangleCount, materials,
                                                                      materialCount, textures, textureCount));
}
bool ptTracerRebuildMeshScene(PtTracer* t, const HostTriangle* triangles, uint32_t triangleCount,
                              const MaterialGpu* materials, uint32_t materialCount, const TextureDescGpu* textures,
                              uint32_t textureCount)
{
    return t != nullptr &&
           t->tracer.rebuildMeshScene(meshOf(triangles, triangleCount, materials, materialCount, textures, textureCount));
}
void ptTracerStart(PtTracer* t) { if (t) t->tracer.start(); }
void ptTracerStop(PtTracer* t) { if (t) t->tracer.stop(); }
void ptTracerResetAccumulation(PtTracer* t) { if (t) t->tracer.resetAccumulation(); }
void ptTracerResetRegionAccumulation(PtTracer* t) { if (t) t->tracer.resetRegionAccumulation(); }
bool ptTracerIsRunning(const PtTracer* t) { return t != nullptr && t->tracer.isRunning(); }
void ptTracerReleaseOutputSurfaces(PtTracer* t) { if (t) t->tracer.releaseOutputSurfaces(); }
bool ptTracerBeginDisplayPublish(PtTracer* t, int slot) { return t != nullptr && t->tracer.beginDisplayPublish(slot); }
bool ptTracerIsDisplayPublishReady(const PtTracer* t) { return t != nullptr && t->tracer.isDisplayPublishReady(); }
bool ptTracerHasDisplayPublishInFlight(const PtTracer* t) { return t != nullptr && t->tracer.hasDisplayPublishInFlight(); }
int ptTracerFinishDisplayPublish(PtTracer* t) { return t != nullptr ? t->tracer.finishDisplayPublish() : -1; }
bool ptTracerPublishDisplayFrame(PtTracer* t, int slot) { return t != nullptr && t->tracer.publishDisplayFrame(slot); }
bool ptTracerReadDisplay(PtTracer* t, int slot, PtUchar4* hostRgba) { return t != nullptr && t->tracer.readDisplay(slot, hostRgba); }
void ptTracerSetMaxSamplesPerPixel(PtTracer* t, int max) { if (t) t->tracer.setMaxSamplesPerPixel(max); }
int ptTracerMaxSamplesPerPixel(const PtTracer* t) { return t ? t->tracer.maxSamplesPerPixel() : 0; }
void ptTracerSetMinSamples(PtTracer* t, int min) { if (t) t->tracer.setMinSamples(min); }
int ptTracerMinSamples(const PtTracer* t) { return t ? t->tracer.minSamples() : 0; }
void ptTracerSetRelativeErrorThreshold(PtTracer* t, float v) { if (t) t->tracer.setRelativeErrorThreshold(v); }
float ptTracerRelativeErrorThreshold(const PtTracer* t) { return t ? t->tracer.relativeErrorThreshold() : 0.0f; }
void ptTracerSetDebugOverlayMode(PtTracer* t, int mode) { if (t) t->tracer.setDebugOverlayMode(mode); }
int ptTracerDebugOverlayMode(const PtTracer* t) { return t ? t->tracer.debugOverlayMode() : 0; }
void ptTracerSetBrdfDebugFlags(PtTracer* t, int flags) { if (t) t->tracer.setBrdfDebugFlags(flags); }
int ptTracerBrdfDebugFlags(const PtTracer* t) { return t ? t->tracer.brdfDebugFlags() : 0; }
void ptTracerSetPreviewStepsPerLevel(PtTracer* t, int steps) { if (t) t->tracer.setPreviewStepsPerLevel(steps); }
int ptTracerPreviewStepsPerLevel(const PtTracer* t) { return t ? t->tracer.previewStepsPerLevel() : 0; }
int ptTracerCurrentSampleCount(const PtTracer* t) { return t ? t->tracer.currentSampleCount() : 0; }
int ptTracerCurrentActivePixelCount(const PtTracer* t) { return t ? t->tracer.currentActivePixelCount() : 0; }
int ptTracerSampleBudgetTotalIterations(const PtTracer* t) { return t ? t->tracer.sampleBudgetTotalIterations() : -1; }
bool ptTracerIsSampleBudgetExhausted(const PtTracer* t) { return t != nullptr && t->tracer.isSampleBudgetExhausted(); }
void ptTracerSetCamera(PtTracer* t, const CameraGpu* c) { if (t && c) t->tracer.setCamera(*c); }
void ptTracerLastSampleCamera(const PtTracer* t, CameraGpu* out) { if (t && out) *out = t->tracer.lastSampleCamera(); }
void ptTracerSetClearColor(PtTracer* t, float r, float g, float b) { if (t) t->tracer.setClearColor(r, g, b); }
void ptTracerSetPhysicalCamera(PtTracer* t, float f, float s, float iso) { if (t) t->tracer.setPhysicalCamera(f, s, iso); }
void ptTracerSetFocusPoint(PtTracer* t, Vec3 p) { if (t) t->tracer.setFocusPoint(p); }
void ptTracerClearFocusPoint(PtTracer* t) { if (t) t->tracer.clearFocusPoint(); }
bool ptTracerPickSurface(const PtTracer* t, Vec3 ro, Vec3 rd, Vec3* hit) { return t != nullptr && t->tracer.pickSurface(ro, rd, hit); }
bool ptTracerSetEnvironmentImage(PtTracer* t, const float* rgb, int w, int h) { return t != nullptr && t->tracer.setEnvironmentImage(rgb, w, h); }
void ptTracerSetEnvironmentIntensity(PtTracer* t, float v) { if (t) t->tracer.setEnvironmentIntensity(v); }
float ptTracerEnvironmentIntensity(const PtTracer* t) { return t ? t->tracer.environmentIntensity() : 0.0f; }
void ptTracerSetRussianRouletteMinDepth(PtTracer* t, int d) { if (t) t->tracer.setRussianRouletteMinDepth(d); }
int ptTracerRussianRouletteMinDepth(const PtTracer* t) { return t ? t->tracer.russianRouletteMinDepth() : 0; }
void ptTracerSetRegionRenderEnabled(PtTracer* t, bool e) { if (t) t->tracer.setRegionRenderEnabled(e); }
bool ptTracerRegionRenderEnabled(const PtTracer* t) { return t != nullptr && t->tracer.regionRenderEnabled(); }
void ptTracerSetRenderRegion(PtTracer* t, int a, int b, int c, int d) { if (t) t->tracer.setRenderRegion(a, b, c, d); }
void ptTracerRenderRegion(const PtTracer* t, int* minX, int* minY, int* maxX, int* maxY)
{
    if (t == nullptr) {
        return;
    }
    const PtRegion r = t->tracer.renderRegion();
    if (minX) *minX = r.minX;
    if (minY) *minY = r.minY;
    if (maxX) *maxX = r.maxX;
    if (maxY) *maxY = r.maxY;
}
bool ptTracerWaitUntilIdle(PtTracer* t, int timeoutMs) { return t != nullptr && t->tracer.waitUntilIdle(timeoutMs); }
bool ptTracerReadAccumulator(PtTracer* t, float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* conv)
{
    return t != nullptr && t->tracer.readAccumulator(acc4, counts, lumMean, m2, conv);
}
bool ptTracerReadPreviewLevel(PtTracer* t, int level, float* acc4, int* w, int* h, int* ds)
{
    return t != nullptr && t->tracer.readPreviewLevel(level, acc4, w, h, ds);
}
float ptTracerExposureMultiplier(const PtTracer* t) { return t ? t->tracer.exposureMultiplier() : 0.0f; }
void ptTracerSetMaxPathDepth(PtTracer* t, int depth) { if (t) t->tracer.setMaxPathDepth(depth); }
int ptTracerMaxPathDepth(const PtTracer* t) { return t ? t->tracer.maxPathDepth() : 0; }
void ptTracerSetPassesPerLaunch(PtTracer* t, int passes) { if (t) t->tracer.setPassesPerLaunch(passes); }
bool ptTracerLastWorkerError(const PtTracer* t, char* buffer, int capacity)
{
    if (t == nullptr || buffer == nullptr || capacity <= 0) {
        return false;
    }
    const std::string e = t->tracer.lastWorkerError();
    std::snprintf(buffer, static_cast<size_t>(capacity), "%s", e.c_str());
    return !e.empty();
}
int ptTracerPassesPerLaunch(const PtTracer* t) { return t ? t->tracer.passesPerLaunch() : 0; }
void ptTracerSetBvhBuildMode(PtTracer* t, int mode) { if (t) t->tracer.setBvhBuildMode(mode); }
int ptTracerBvhBuildMode(const PtTracer* t) { return t ? t->tracer.bvhBuildMode() : 0; }

} // extern "C"
