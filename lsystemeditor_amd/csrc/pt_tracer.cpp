// pt_tracer.cpp -- the headless renderer facade (class PathTracer of include/pt_tracer.h).
//
// Public behaviour follows the reference's PathTracer (PathTracerLib/src/PathTracer.h:30-152):
// the same methods, clamps, reset rules, sample-budget policy (PathTracerSampleBudget.h), preview
// pyramid (PathTracerPreviewLevels.h), exposure (PhysicalCamera.cpp:397-402) and per-pass
// schedule of its render loop (PathTracer.cpp:2220-2543: preview levels first, then adaptive
// passes with globalSeed = iteration + 1 and an active-list compaction every fourth
// full-resolution pass). The internals are this library's own:
//   - Controls: every user setting in one plain struct under one lock; the worker copies it once
//     per step, and device copies of the camera and render parameters are refreshed by comparing
//     them with what was last uploaded (no dirty flags);
//   - Surfaces: the device state of one resolution (accumulator, two active lists, preview
//     pyramid, two display buffers), owned through RAII arrays;
//   - two device parameter blocks, one per stream (sampling and display), so a display publish
//     never rewrites what a pass in flight reads;
//   - the worker is a loop of plan -> launch -> await -> commit over an explicit Step record,
//     with resets delivered as request bits.
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

#include "../../include/pt_procedural.h"
#include "../../include/pt_tracer.h"
#include "pt_internal.h"

namespace {

// ---- reference policies ---------------------------------------------------------------------

template <typename T>
T clampTo(T v, T lo, T hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

// setter ranges (PathTracer.cpp:126-131, 1710-1990) and physical-camera ranges (PhysicalCamera.h:36-51)
struct Range {
    int lo, hi;
};
constexpr Range kMaxSppRange{0, 1000000};
constexpr Range kMinSppRange{1, 10000};
constexpr Range kOverlayRange{0, 3};
constexpr Range kPreviewStepRange{0, 128};
constexpr Range kRouletteRange{0, 256};
constexpr Range kPassesRange{1, 256};
constexpr Range kBuildModeRange{0, 2};
constexpr float kThresholdLo = 0.001f, kThresholdHi = 1.0f;
constexpr float kIntensityLo = 0.0f, kIntensityHi = 100.0f;
constexpr int kDepthCap = 256;
constexpr float kFStopLo = 1.2f, kFStopHi = 22.0f;
constexpr float kShutterLo = 1.0f / 8000.0f, kShutterHi = 120.0f;
constexpr float kIsoLo = 100.0f, kIsoHi = 25600.0f;

constexpr int kCompactEvery = 4;      // PathTracerAdaptiveSampling.h:7
constexpr int kPreviewDepth = 4;      // PathTracerPreviewLevels.h:13
constexpr int kUvOverlay = 3;         // RenderViewOverlayMode::Uv

// PathTracerSampleBudget.h: a positive sample limit bounds preview steps + full passes.
int budgetTotal(int previewSteps, int maxSpp) { return maxSpp <= 0 ? -1 : previewSteps + maxSpp; }
bool budgetAllows(int count, int previewSteps, int maxSpp) { return maxSpp <= 0 || count < previewSteps + maxSpp; }

// PathTracerPreviewLevels.h: level i of n is 1/2^(n-i) of full resolution (at least 1 pixel).
struct LevelSize {
    int width, height, downscale;
};
LevelSize levelSize(int w, int h, int levels, int index)
{
    const int ds = levels > 0 && index >= 0 && index < levels ? 1 << (levels - index) : 1;
    auto axis = [ds](int full) { return full <= 0 ? 0 : std::max(1, full / ds); };
    return {axis(w), axis(h), ds};
}

// Camera model of PhysicalCamera.cpp: exposure (397-402), aperture from focal length and f-stop
// (291-298), focal length from the vertical field of view on a 23.9 mm sensor (348-363).
struct Optics {
    static constexpr float kSensorMm = 23.9f;
    static constexpr float kFocalLo = 14.0f, kFocalHi = 1000.0f;

    CameraGpu geometry{};
    float fStop = 2.8f;
    float shutter = 1.0f / 125.0f;
    float iso = 100.0f;
    float focalMm = 24.0f;
    float focusDistance = 1000.0f;
    bool focusOn = true;

    Optics()
    {
        geometry.orientation = PtFloat4{1.0f, 0.0f, 0.0f, 0.0f};
        geometry.fovY = 2.0f * std::atan(kSensorMm / (2.0f * focalMm));
        geometry.aspect = 1.0f;
        geometry.nearPlane = focalMm;
        geometry.farPlane = 1000000.0f;
    }

    float exposure() const
    {
        const float f = std::max(fStop, 0.1f);
        return shutter * (std::max(iso, 1.0f) / 100.0f) / (f * f);
    }

    // -Z of camera space in world space: (0, 0, -1) turned by the orientation quaternion
    // (w, x, y, z), v + w t + u x t with t = 2 u x v (CameraDevice.cuh:8-25 uses the same rotation).
    Vec3 viewAxis() const
    {
        const PtFloat4& q = geometry.orientation;
        const float w = q.x, x = q.y, y = q.z, z = q.w;
        const Vec3 v = {0.0f, 0.0f, -1.0f};
        const Vec3 t = {2.0f * (y * v.z - z * v.y), 2.0f * (z * v.x - x * v.z), 2.0f * (x * v.y - y * v.x)};
        return Vec3{v.x + w * t.x + (y * t.z - z * t.y), v.y + w * t.y + (z * t.x - x * t.z),
                    v.z + w * t.z + (x * t.y - y * t.x)};
    }

    void takeGeometry(const CameraGpu& c)
    {
        geometry = c;
        const float t = std::tan(c.fovY * 0.5f);
        focalMm = t <= 0.0f ? kFocalHi : clampTo(kSensorMm / (2.0f * t), kFocalLo, kFocalHi);
        focusDistance = c.focusDistance;
        focusOn = c.focusValid != 0;
    }

    CameraGpu forDevice(int width, int height) const
    {
        CameraGpu g = geometry;
        g.apertureRadius = focusOn && fStop > 0.0f ? focalMm / (2.0f * fStop) : 0.0f;
        g.focusDistance = focusOn ? focusDistance : 0.0f;
        const Vec3 a = viewAxis();
        g.focusPoint = Vec3{geometry.position.x + a.x * focusDistance, geometry.position.y + a.y * focusDistance,
                            geometry.position.z + a.z * focusDistance};
        g.focusValid = focusOn ? 1 : 0;
        if (width > 0 && height > 0) {
            g.aspect = static_cast<float>(width) / static_cast<float>(height);
        }
        return g;
    }
};

// Everything a caller can set (defaults: PathTracer.cpp:60-84, RenderTypes.h:15-29).
struct Controls {
    int maxSpp = 8;
    int minSpp = 16;
    float threshold = 0.02f;
    int overlay = 0;
    int brdfFlags = 0;
    int previewSteps = 0;
    int passesPerLaunch = 1;
    int buildMode = 0;
    int maxDepth = 32;
    int rouletteDepth = 3;
    float envIntensity = 1.0f;
    float clear[3] = {10.0f / 255.0f, 10.0f / 255.0f, 10.0f / 255.0f};
    bool regionOn = false;
    PtRegion region{0, 0, 0, 0};
    Optics optics;

    RenderParamsGpu params(bool preview) const
    {
        RenderParamsGpu p{};
        p.backgroundR = clear[0];
        p.backgroundG = clear[1];
        p.backgroundB = clear[2];
        p.environmentIntensity = envIntensity;
        p.maxPathDepth = preview ? kPreviewDepth : maxDepth;
        p.russianRouletteMinDepth = rouletteDepth;
        p.minSamples = minSpp;
        p.maxSamplesPerPixel = maxSpp;
        p.debugOverlayMode = overlay;
        p.brdfDebugFlags = brdfFlags;
        p.relativeErrorThreshold = threshold;
        return p;
    }
};

// ---- device memory --------------------------------------------------------------------------

template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept { swap(o); }
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        reset();
        swap(o);
        return *this;
    }
    ~DeviceArray() { reset(); }

    bool allocate(size_t n)
    {
        reset();
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
            return false;
        }
        ptr_ = static_cast<T*>(p);
        size_ = n;
        return true;
    }
    void reset()
    {
        if (ptr_ != nullptr) {
            (void)hipFree(ptr_);
        }
        ptr_ = nullptr;
        size_ = 0;
    }
    T* get() const { return ptr_; }
    size_t size() const { return size_; }
    explicit operator bool() const { return ptr_ != nullptr; }
    void swap(DeviceArray& o)
    {
        std::swap(ptr_, o.ptr_);
        std::swap(size_, o.size_);
    }

private:
    T* ptr_ = nullptr;
    size_t size_ = 0;
};

// A device copy of a small POD with the host value it was last given: upload() copies only when
// the value changed (or was never uploaded).
template <typename T>
class DeviceValue {
public:
    bool allocate()
    {
        known_ = false;
        return mem_.allocate(1);
    }
    void release()
    {
        mem_.reset();
        known_ = false;
    }
    void forget() { known_ = false; }
    const T* get() const { return mem_.get(); }
    bool upload(const T& v, hipStream_t s)
    {
        if (!mem_) {
            return false;
        }
        if (known_ && std::memcmp(&v, &last_, sizeof(T)) == 0) {
            return true;
        }
        last_ = v;
        if (hipMemcpyAsync(mem_.get(), &last_, sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            known_ = false;
            return false;
        }
        known_ = true;
        return true;
    }

private:
    DeviceArray<T> mem_;
    T last_{};
    bool known_ = false;
};

struct PreviewBuffer {
    DeviceArray<PtFloat4> texels;
    LevelSize size{0, 0, 1};
};

// Device state of one configured resolution.
struct Surfaces {
    int width = 0, height = 0;
    DeviceArray<PtFloat4> radiance;
    DeviceArray<uint32_t> counts;
    DeviceArray<float> mean, m2;
    DeviceArray<uint8_t> converged;
    DeviceArray<int> active, activeNext, activeCount;
    std::vector<PreviewBuffer> pyramid;
    DeviceArray<PtUchar4> display[2];
    DeviceValue<CameraGpu> sampleCamera, displayCamera;
    DeviceValue<RenderParamsGpu> sampleParams, displayParams;

    size_t pixels() const { return static_cast<size_t>(width) * static_cast<size_t>(height); }
    bool ready() const { return static_cast<bool>(radiance); }

    bool allocate(int w, int h, int previewLevels)
    {
        release();
        width = w;
        height = h;
        const size_t n = pixels();
        const bool ok = radiance.allocate(n) && counts.allocate(n) && mean.allocate(n) && m2.allocate(n) &&
                        converged.allocate(n) && active.allocate(n) && activeNext.allocate(n) &&
                        activeCount.allocate(1) && display[0].allocate(n) && display[1].allocate(n) &&
                        sampleCamera.allocate() && displayCamera.allocate() && sampleParams.allocate() &&
                        displayParams.allocate() && buildPyramid(previewLevels);
        if (!ok) {
            release();
        }
        return ok;
    }
    bool buildPyramid(int levels)
    {
        pyramid.clear();
        pyramid.resize(static_cast<size_t>(levels));
        for (int i = 0; i < levels; ++i) {
            PreviewBuffer& b = pyramid[static_cast<size_t>(i)];
            b.size = levelSize(width, height, levels, i);
            if (!b.texels.allocate(static_cast<size_t>(b.size.width) * static_cast<size_t>(b.size.height))) {
                pyramid.clear();
                return false;
            }
        }
        return true;
    }
    bool zeroPyramid(hipStream_t s) const
    {
        for (const PreviewBuffer& b : pyramid) {
            if (hipMemsetAsync(b.texels.get(), 0, b.texels.size() * sizeof(PtFloat4), s) != hipSuccess) {
                return false;
            }
        }
        return true;
    }
    void release()
    {
        radiance.reset();
        counts.reset();
        mean.reset();
        m2.reset();
        converged.reset();
        active.reset();
        activeNext.reset();
        activeCount.reset();
        pyramid.clear();
        display[0].reset();
        display[1].reset();
        sampleCamera.release();
        displayCamera.release();
        sampleParams.release();
        displayParams.release();
        width = height = 0;
    }
};

// One unit of worker work.
struct Step {
    bool preview = false;
    int iteration = 0;     // sample count when planned (preview level index for previews)
    int passes = 1;
    int active = 0;
    bool compact = false;
};

enum : unsigned { kFullReset = 1u, kRegionReset = 2u };
enum class Wait { Done, Interrupted, Stopped, Failed };

} // namespace

namespace ptdetail {

struct TracerImpl {
    // settings
    mutable std::mutex controlsLock;
    Controls controls;
    CameraGpu lastCamera{};

    // device state; streamLock orders everything enqueued on the two streams
    std::mutex streamLock;
    Surfaces surf;
    mutable std::mutex sceneLock;
    PtScene* scene = nullptr;
    PtEnvironment* env = nullptr;
    std::vector<float> envPixels;
    int envSize[2] = {0, 0};
    unsigned envRevision = 1, envUploaded = 0;   // guarded by streamLock
    hipStream_t sampling = nullptr, displaying = nullptr;
    hipEvent_t stepDone = nullptr, publishDone = nullptr;

    // worker
    std::thread worker;
    std::mutex wakeLock;
    std::condition_variable wake;
    std::atomic<bool> running{false}, configured{false}, idle{false};
    std::atomic<unsigned> requests{0};
    std::atomic<int> sampleCount{0}, activePixels{0}, finestPreview{-1};
    std::atomic<bool> displayable{false};
    int countReadback = 0;                 // written by the compaction's device-to-host copy
    mutable std::mutex errorLock;
    std::string workerError;
    std::mutex callbackLock;
    PathTracer::FrameReadyCallback frameReady;

    // display publish
    std::atomic<bool> publishing{false};
    std::atomic<int> publishSlot{-1};

    ~TracerImpl()
    {
        releaseDevice();
        if (stepDone != nullptr) {
            (void)hipEventDestroy(stepDone);
        }
        if (publishDone != nullptr) {
            (void)hipEventDestroy(publishDone);
        }
        if (sampling != nullptr) {
            (void)hipStreamDestroy(sampling);
        }
        if (displaying != nullptr) {
            (void)hipStreamDestroy(displaying);
        }
    }

    Controls snapshot() const
    {
        std::lock_guard<std::mutex> g(controlsLock);
        return controls;
    }

    template <typename F>
    void edit(F&& change)
    {
        {
            std::lock_guard<std::mutex> g(controlsLock);
            change(controls);
        }
        poke();
    }

    void poke()
    {
        { std::lock_guard<std::mutex> g(wakeLock); }
        wake.notify_all();
    }

    void request(unsigned bits)
    {
        requests.fetch_or(bits);
        poke();
    }

    // resets that setters ask for apply only once a scene is configured
    void resetIfConfigured()
    {
        if (configured.load()) {
            request(kFullReset);
        }
    }

    bool ensureStreams()
    {
        return (sampling != nullptr || hipStreamCreate(&sampling) == hipSuccess) &&
               (displaying != nullptr || hipStreamCreate(&displaying) == hipSuccess) &&
               (stepDone != nullptr || hipEventCreateWithFlags(&stepDone, hipEventDisableTiming) == hipSuccess) &&
               (publishDone != nullptr || hipEventCreateWithFlags(&publishDone, hipEventDisableTiming) == hipSuccess);
    }

    void releaseDevice()
    {
        surf.release();
        if (env != nullptr) {
            ptEnvironmentDestroy(env);
            env = nullptr;
        }
        envUploaded = 0;
        std::lock_guard<std::mutex> g(sceneLock);
        if (scene != nullptr) {
            ptSceneDestroy(scene);
            scene = nullptr;
        }
    }

    bool buildScene(const PtHostMesh& mesh, int buildMode)
    {
        std::lock_guard<std::mutex> g(sceneLock);
        if (scene == nullptr && (scene = ptSceneCreate()) == nullptr) {
            return false;
        }
        return ptSceneSetBuildMode(scene, buildMode) &&
               ptSceneBuild(scene, mesh.triangles, mesh.triangleCount, mesh.materials, mesh.materialCount,
                            mesh.textures, mesh.textureCount, sampling);
    }

    const MeshAccelSceneGpu* deviceScene() const
    {
        std::lock_guard<std::mutex> g(sceneLock);
        return ptSceneDevice(scene);
    }

    // ---- accumulator resets (PathTracer.cpp:511-632 behaviour), streamLock held -------------

    bool clearAll()
    {
        Surfaces& s = surf;
        if (!pathTracerClearAccumulator(s.radiance.get(), s.counts.get(), s.mean.get(), s.m2.get(), s.converged.get(),
                                        s.active.get(), s.activeNext.get(), s.activeCount.get(), s.width, s.height,
                                        sampling) ||
            !s.zeroPyramid(sampling)) {
            return false;
        }
        activePixels.store(s.width * s.height);
        return true;
    }

    bool clearRegion(const PtRegion& r)
    {
        Surfaces& s = surf;
        if (!pathTracerClearRegionAccumulator(s.radiance.get(), s.counts.get(), s.mean.get(), s.m2.get(),
                                              s.converged.get(), s.active.get(), s.activeCount.get(), s.width,
                                              s.height, r.minX, r.minY, r.maxX, r.maxY, sampling)) {
            return false;
        }
        activePixels.store((r.maxX - r.minX + 1) * (r.maxY - r.minY + 1));
        return true;
    }

    bool relistActive(int maxSpp)
    {
        Surfaces& s = surf;
        int n = 0;
        if (!pathTracerRebuildActiveList(s.converged.get(), s.counts.get(), s.active.get(), s.activeCount.get(),
                                         s.width, s.height, maxSpp, sampling) ||
            hipMemcpyAsync(&n, s.activeCount.get(), sizeof(int), hipMemcpyDeviceToHost, sampling) != hipSuccess ||
            hipStreamSynchronize(sampling) != hipSuccess) {
            return false;
        }
        activePixels.store(n);
        return true;
    }

    bool refreshEnvironment()
    {
        if (envUploaded == envRevision) {
            return true;
        }
        if (env != nullptr) {
            ptEnvironmentDestroy(env);
            env = nullptr;
        }
        if (!envPixels.empty() && (env = ptEnvironmentCreate(envPixels.data(), envSize[0], envSize[1], sampling)) == nullptr) {
            return false;
        }
        envUploaded = envRevision;
        return true;
    }

    // ---- worker ---------------------------------------------------------------------------

    bool wantsSample(const Controls& c) const
    {
        const int n = sampleCount.load();
        return n < c.previewSteps ? budgetAllows(n, c.previewSteps, c.maxSpp) : activePixels.load() > 0;
    }

    bool hasPictureToShow(const Controls& c) const
    {
        const int n = sampleCount.load();
        if (n <= 0) {
            return false;
        }
        return c.previewSteps == 0 || n > c.previewSteps || finestPreview.load() >= 0;
    }

    // Applies pending reset requests; returns whether the caller should tell the frame callback.
    void serveRequests()
    {
        const unsigned r = requests.exchange(0u);
        if (r == 0u) {
            return;
        }
        idle.store(false);
        const Controls c = snapshot();
        const bool full = (r & kFullReset) != 0;
        bool done = false, wiped = false;
        {
            std::lock_guard<std::mutex> g(streamLock);
            if (surf.ready()) {
                if (c.regionOn) {
                    done = clearRegion(c.region);
                } else if (full) {
                    done = wiped = clearAll();
                } else {
                    done = relistActive(c.maxSpp);
                }
            }
        }
        const bool synced = done && hipStreamSynchronize(sampling) == hipSuccess;
        if (wiped) {
            displayable.store(false);
        } else if (synced && sampleCount.load() > 0) {
            notifyFrame();
        }
        if (full) {
            sampleCount.store(0);
            finestPreview.store(-1);
        }
    }

    // Blocks until there is a sample to take, a request, or a stop; false on stop.
    bool awaitWork()
    {
        std::unique_lock<std::mutex> g(wakeLock);
        auto ready = [this] { return !running.load() || requests.load() != 0u || wantsSample(snapshot()); };
        if (!ready()) {
            idle.store(true);
            wake.wait(g, ready);
        }
        return running.load();
    }

    Step plan(const Controls& c) const
    {
        Step s;
        s.iteration = sampleCount.load();
        s.preview = !c.regionOn && c.previewSteps > 0 && s.iteration < c.previewSteps &&
                    s.iteration < static_cast<int>(surf.pyramid.size());
        if (!s.preview) {
            s.active = activePixels.load();
            const int wave = s.iteration - c.previewSteps + 1;   // full-resolution passes so far, this one included
            s.passes = s.active > 0 ? c.passesPerLaunch : 1;
            s.compact = s.active > 0 && (s.passes > 1 || wave % kCompactEvery == 0);
        }
        return s;
    }

    // Enqueues the step and its completion event; streamLock held.
    bool launch(const Step& st, const Controls& c)
    {
        Surfaces& s = surf;
        if (publishing.load() && hipStreamWaitEvent(sampling, publishDone, 0) != hipSuccess) {
            return false;
        }
        if (!s.sampleCamera.upload(c.optics.forDevice(s.width, s.height), sampling) ||
            !s.sampleParams.upload(c.params(st.preview), sampling) || !refreshEnvironment()) {
            return false;
        }
        const unsigned seed = static_cast<unsigned>(st.iteration) + 1u;
        const MeshAccelSceneGpu* scn = deviceScene();
        const EnvironmentMapGpu* e = env != nullptr ? ptEnvironmentDevice(env) : nullptr;
        bool ok = true;
        if (st.preview) {
            const PreviewBuffer& b = s.pyramid[static_cast<size_t>(st.iteration)];
            ok = pathTracerPreviewSample(b.texels.get(), b.size.width, b.size.height, s.sampleCamera.get(), scn, e,
                                         s.sampleParams.get(), seed, sampling);
        } else if (st.active > 0 && st.passes > 1) {
            const int depth = c.maxDepth > 0 ? c.maxDepth : (c.maxDepth == 0 ? 256 : 32);
            ok = pathTracerAdaptiveRenderWavefront(s.radiance.get(), s.counts.get(), s.mean.get(), s.m2.get(),
                                                   s.converged.get(), s.active.get(), st.active, s.width, s.height,
                                                   s.sampleCamera.get(), scn, e, s.sampleParams.get(), seed,
                                                   st.passes, depth, 0, sampling);
        } else if (st.active > 0) {
            ok = pathTracerAdaptiveSample(s.radiance.get(), s.counts.get(), s.mean.get(), s.m2.get(),
                                          s.converged.get(), s.active.get(), st.active, s.width, s.height,
                                          s.sampleCamera.get(), scn, e, s.sampleParams.get(), seed, sampling);
        }
        if (ok && st.compact) {
            countReadback = 0;
            ok = pathTracerCompactActiveList(s.active.get(), s.activeNext.get(), s.activeCount.get(),
                                             s.converged.get(), s.counts.get(), c.maxSpp, st.active, s.width,
                                             s.height, c.regionOn, c.region.minX, c.region.minY, c.region.maxX,
                                             c.region.maxY, sampling) &&
                 hipMemcpyAsync(&countReadback, s.activeCount.get(), sizeof(int), hipMemcpyDeviceToHost,
                                sampling) == hipSuccess;
        }
        return ok && hipEventRecord(stepDone, sampling) == hipSuccess;
    }

    // Polls the step's event; a reset that arrives meanwhile voids the step.
    Wait await()
    {
        bool interrupted = false;
        for (;;) {
            if (!running.load()) {
                return Wait::Stopped;
            }
            interrupted = interrupted || requests.load() != 0u;
            const hipError_t q = hipEventQuery(stepDone);
            if (q == hipSuccess) {
                return interrupted ? Wait::Interrupted : Wait::Done;
            }
            if (q != hipErrorNotReady) {
                return Wait::Failed;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    }

    void commit(const Step& st)
    {
        if (st.compact) {
            activePixels.store(countReadback);
            surf.active.swap(surf.activeNext);
        }
        if (st.preview) {
            finestPreview.store(st.iteration);
        }
        sampleCount.fetch_add(st.passes);
        displayable.store(true);
        notifyFrame();
    }

    void notifyFrame()
    {
        PathTracer::FrameReadyCallback cb;
        {
            std::lock_guard<std::mutex> g(callbackLock);
            cb = frameReady;
        }
        if (cb) {
            cb();
        }
    }

    void workerMain()
    {
        while (running.load()) {
            serveRequests();
            if (!surf.ready()) {
                break;
            }
            if (!awaitWork()) {
                break;
            }
            const Controls c = snapshot();
            if (!wantsSample(c)) {
                continue;
            }
            idle.store(false);
            const Step st = plan(c);
            bool queued;
            {
                std::lock_guard<std::mutex> g(streamLock);
                queued = launch(st, c);
            }
            if (!queued) {
                {
                    std::lock_guard<std::mutex> g(errorLock);
                    workerError = pathTracerLastError();
                }
                if (!st.preview && st.active > 0) {
                    std::this_thread::sleep_for(std::chrono::milliseconds(1));
                }
                continue;
            }
            const Wait w = await();
            if (w == Wait::Stopped) {
                break;
            }
            if (w == Wait::Done) {
                commit(st);
            }
        }
        idle.store(false);
    }

    // ---- display (PathTracer.cpp:1477-1708 behaviour) ----------------------------------------

    bool enqueuePublish(int slot)
    {
        Surfaces& s = surf;
        if (slot < 0 || slot > 1 || !s.display[slot]) {
            return false;
        }
        const Controls c = snapshot();
        const bool uv = c.overlay == kUvOverlay;
        if (!uv && (!displayable.load() || !hasPictureToShow(c))) {
            return false;
        }
        if (!uv && hipStreamWaitEvent(displaying, stepDone, 0) != hipSuccess) {
            return false;
        }
        if (!s.displayParams.upload(c.params(false), displaying) ||
            !s.displayCamera.upload(c.optics.forDevice(s.width, s.height), displaying)) {
            return false;
        }
        const int n = sampleCount.load();
        const int finest = finestPreview.load();
        const bool fromPyramid = !uv && !c.regionOn && c.previewSteps > 0 && finest >= 0 &&
                                 finest < static_cast<int>(s.pyramid.size()) && n <= c.previewSteps;
        const float exposure = c.optics.exposure();
        bool ok;
        if (uv) {
            ok = pathTracerWriteUvDebugToPbo(s.displayCamera.get(), deviceScene(), s.display[slot].get(), s.width,
                                             s.height, s.displayParams.get(), displaying);
        } else if (fromPyramid) {
            const PreviewBuffer& b = s.pyramid[static_cast<size_t>(finest)];
            ok = pathTracerUpsamplePreviewToPbo(b.texels.get(), b.size.width, b.size.height, b.size.downscale,
                                                s.display[slot].get(), s.width, s.height, s.displayParams.get(),
                                                exposure, displaying);
        } else {
            ok = pathTracerCopyToPbo(s.radiance.get(), s.counts.get(), s.converged.get(), s.display[slot].get(),
                                     s.width, s.height, s.displayParams.get(), exposure, displaying);
        }
        return ok && hipEventRecord(publishDone, displaying) == hipSuccess;
    }
};

} // namespace ptdetail

using ptdetail::TracerImpl;

// ---- PathTracer -----------------------------------------------------------------------------

PathTracer::PathTracer() : m_impl(new TracerImpl()) {}

PathTracer::~PathTracer() { stop(); }

void PathTracer::setFrameReadyCallback(FrameReadyCallback callback)
{
    std::lock_guard<std::mutex> g(m_impl->callbackLock);
    m_impl->frameReady = std::move(callback);
}

bool PathTracer::configure(int width, int height, const PtHostMesh& mesh)
{
    TracerImpl& t = *m_impl;
    if (width <= 0 || height <= 0) {
        ptSetError("PathTracer::configure: invalid dimensions");
        return false;
    }
    stop();
    t.configured.store(false);
    t.releaseDevice();
    if (!t.ensureStreams()) {
        ptSetError("PathTracer::configure: stream or event creation failed");
        return false;
    }
    t.publishing.store(false);
    t.publishSlot.store(-1);
    const Controls c = t.snapshot();
    if (!t.surf.allocate(width, height, c.previewSteps)) {
        ptSetError("PathTracer::configure: device allocation failed");
        return false;
    }
    bool ok;
    {
        std::lock_guard<std::mutex> g(t.streamLock);
        ok = t.buildScene(mesh, c.buildMode) && t.refreshEnvironment() &&
             t.surf.sampleCamera.upload(c.optics.forDevice(width, height), t.sampling) &&
             t.surf.sampleParams.upload(c.params(false), t.sampling) && t.clearAll() &&
             hipStreamSynchronize(t.sampling) == hipSuccess && hipStreamSynchronize(t.displaying) == hipSuccess;
    }
    if (!ok) {
        if (pathTracerLastError()[0] == '\0') {
            ptSetError("PathTracer::configure: scene build or upload failed");
        }
        t.releaseDevice();
        return false;
    }
    t.requests.store(0u);
    t.sampleCount.store(0);
    t.finestPreview.store(-1);
    t.displayable.store(false);
    t.configured.store(true);
    return true;
}

namespace {

// meshSceneBuild (MeshSceneContent.cpp:5-52): the instances' HostMesh, or NULL with the error set.
PtHostMeshBuf* proceduralMesh(const std::vector<PtProceduralInstance>& instances, const PtMeshSceneBuildParams& params)
{
    PtProceduralBuildParams p;
    ptProceduralDefaultParams(&p);
    p.circularSegments = params.circularSegments;
    p.creaseAngleDeg = params.creaseAngleDeg;
    return ptProceduralBuildScene(instances.empty() ? nullptr : instances.data(),
                                  static_cast<uint32_t>(instances.size()), &p);
}

PtHostMesh viewOf(const PtHostMeshBuf* m)
{
    PtHostMesh v;
    ptHostMeshCounts(m, &v.triangleCount, &v.materialCount, &v.textureCount);
    v.triangles = ptHostMeshTriangles(m);
    v.materials = ptHostMeshMaterials(m);
    v.textures = ptHostMeshTextures(m);
    return v;
}

} // namespace

bool PathTracer::configure(int width, int height, const std::vector<PtProceduralInstance>& instances,
                           const PtMeshSceneBuildParams& params)
{
    PtHostMeshBuf* m = proceduralMesh(instances, params);
    if (m == nullptr) {
        return false;
    }
    const bool ok = configure(width, height, viewOf(m));
    ptHostMeshDestroy(m);
    return ok;
}

bool PathTracer::rebuildMeshScene(const PtHostMesh& mesh)
{
    TracerImpl& t = *m_impl;
    if (!t.configured.load()) {
        return false;
    }
    const int mode = t.snapshot().buildMode;
    bool ok;
    {
        std::lock_guard<std::mutex> g(t.streamLock);
        ok = hipStreamSynchronize(t.sampling) == hipSuccess && t.buildScene(mesh, mode);
    }
    if (ok) {
        resetAccumulation();
    }
    return ok;
}

bool PathTracer::rebuildMeshScene(const std::vector<PtProceduralInstance>& instances,
                                  const PtMeshSceneBuildParams& params)
{
    PtHostMeshBuf* m = proceduralMesh(instances, params);
    if (m == nullptr) {
        return false;
    }
    const bool ok = rebuildMeshScene(viewOf(m));
    ptHostMeshDestroy(m);
    return ok;
}

void PathTracer::start()
{
    TracerImpl& t = *m_impl;
    if (t.running.load() || !t.configured.load()) {
        return;
    }
    if (t.worker.joinable()) {
        t.worker.join();
    }
    t.running.store(true);
    t.worker = std::thread([this] { renderLoop(); });
}

void PathTracer::stop()
{
    TracerImpl& t = *m_impl;
    t.running.store(false);
    t.poke();
    if (t.worker.joinable()) {
        t.worker.join();
    }
}

void PathTracer::resetAccumulation() { m_impl->request(kFullReset); }
void PathTracer::resetRegionAccumulation() { m_impl->request(kRegionReset); }
bool PathTracer::isRunning() const { return m_impl->running.load(); }

void PathTracer::releaseOutputSurfaces()
{
    stop();
    m_impl->surf.display[0].reset();
    m_impl->surf.display[1].reset();
    m_impl->configured.store(false);
}

bool PathTracer::beginDisplayPublish(int slot)
{
    TracerImpl& t = *m_impl;
    if (slot < 0 || slot > 1 || !t.configured.load() || t.publishing.load()) {
        return false;
    }
    std::lock_guard<std::mutex> g(t.streamLock);
    if (!t.enqueuePublish(slot)) {
        return false;
    }
    t.publishSlot.store(slot);
    t.publishing.store(true);
    return true;
}

bool PathTracer::isDisplayPublishReady() const
{
    return m_impl->publishing.load() && m_impl->publishDone != nullptr && hipEventQuery(m_impl->publishDone) == hipSuccess;
}

bool PathTracer::hasDisplayPublishInFlight() const { return m_impl->publishing.load(); }

int PathTracer::finishDisplayPublish()
{
    if (!isDisplayPublishReady()) {
        return -1;
    }
    const int slot = m_impl->publishSlot.exchange(-1);
    m_impl->publishing.store(false);
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
    const TracerImpl& t = *m_impl;
    if (slot < 0 || slot > 1 || hostRgba == nullptr || !t.surf.display[slot]) {
        return false;
    }
    return hipStreamSynchronize(t.displaying) == hipSuccess &&
           hipMemcpy(hostRgba, t.surf.display[slot].get(), t.surf.pixels() * sizeof(PtUchar4),
                     hipMemcpyDeviceToHost) == hipSuccess;
}

void PathTracer::setMaxSamplesPerPixel(int max)
{
    m_impl->edit([max](Controls& c) { c.maxSpp = clampTo(max, kMaxSppRange.lo, kMaxSppRange.hi); });
}
int PathTracer::maxSamplesPerPixel() const { return m_impl->snapshot().maxSpp; }

void PathTracer::setMinSamples(int min)
{
    m_impl->edit([min](Controls& c) { c.minSpp = clampTo(min, kMinSppRange.lo, kMinSppRange.hi); });
    m_impl->resetIfConfigured();
}
int PathTracer::minSamples() const { return m_impl->snapshot().minSpp; }

void PathTracer::setRelativeErrorThreshold(float threshold)
{
    m_impl->edit([threshold](Controls& c) { c.threshold = clampTo(threshold, kThresholdLo, kThresholdHi); });
    m_impl->resetIfConfigured();
}
float PathTracer::relativeErrorThreshold() const { return m_impl->snapshot().threshold; }

void PathTracer::setDebugOverlayMode(int mode)
{
    m_impl->edit([mode](Controls& c) { c.overlay = clampTo(mode, kOverlayRange.lo, kOverlayRange.hi); });
}
int PathTracer::debugOverlayMode() const { return m_impl->snapshot().overlay; }

void PathTracer::setBrdfDebugFlags(int flags)
{
    m_impl->edit([flags](Controls& c) { c.brdfFlags = flags; });
}
int PathTracer::brdfDebugFlags() const { return m_impl->snapshot().brdfFlags; }

void PathTracer::setPreviewStepsPerLevel(int steps)
{
    TracerImpl& t = *m_impl;
    const int n = clampTo(steps, kPreviewStepRange.lo, kPreviewStepRange.hi);
    t.edit([n](Controls& c) { c.previewSteps = n; });
    if (t.configured.load() && t.surf.ready()) {
        std::lock_guard<std::mutex> g(t.streamLock);
        if (hipStreamSynchronize(t.sampling) == hipSuccess && t.surf.buildPyramid(n) && t.surf.zeroPyramid(t.sampling)) {
            (void)hipStreamSynchronize(t.sampling);
        }
    }
    t.poke();
}
int PathTracer::previewStepsPerLevel() const { return m_impl->snapshot().previewSteps; }

int PathTracer::currentSampleCount() const { return m_impl->sampleCount.load(); }
int PathTracer::currentActivePixelCount() const { return m_impl->activePixels.load(); }

int PathTracer::sampleBudgetTotalIterations() const
{
    const Controls c = m_impl->snapshot();
    return budgetTotal(c.previewSteps, c.maxSpp);
}

bool PathTracer::isSampleBudgetExhausted() const
{
    const Controls c = m_impl->snapshot();
    const int n = m_impl->sampleCount.load();
    return n < c.previewSteps ? !budgetAllows(n, c.previewSteps, c.maxSpp) : m_impl->activePixels.load() <= 0;
}

void PathTracer::setCamera(const CameraGpu& camera)
{
    TracerImpl& t = *m_impl;
    CameraGpu cam = camera;
    if (t.surf.width > 0 && t.surf.height > 0) {
        cam.aspect = static_cast<float>(t.surf.width) / static_cast<float>(t.surf.height);
    }
    t.edit([&cam](Controls& c) { c.optics.takeGeometry(cam); });
    std::lock_guard<std::mutex> g(t.controlsLock);
    t.lastCamera = cam;
}

CameraGpu PathTracer::lastSampleCamera() const
{
    std::lock_guard<std::mutex> g(m_impl->controlsLock);
    return m_impl->lastCamera;
}

void PathTracer::setClearColor(float r, float g, float b)
{
    m_impl->edit([r, g, b](Controls& c) {
        c.clear[0] = r;
        c.clear[1] = g;
        c.clear[2] = b;
    });
    m_impl->resetIfConfigured();
}

void PathTracer::setPhysicalCamera(float fStop, float shutterSpeedSeconds, float iso)
{
    m_impl->edit([=](Controls& c) {
        c.optics.fStop = clampTo(fStop, kFStopLo, kFStopHi);
        c.optics.shutter = clampTo(shutterSpeedSeconds, kShutterLo, kShutterHi);
        c.optics.iso = clampTo(iso, kIsoLo, kIsoHi);
    });
}

void PathTracer::setFocusPoint(const Vec3& p)
{
    m_impl->edit([&p](Controls& c) {
        Optics& o = c.optics;
        const Vec3 a = o.viewAxis();
        const Vec3& e = o.geometry.position;
        const float d = (p.x - e.x) * a.x + (p.y - e.y) * a.y + (p.z - e.z) * a.z;
        o.focusDistance = d > 0.0f ? d : 0.0f;
        o.focusOn = o.focusDistance > 0.0f;
    });
    m_impl->resetIfConfigured();
}

void PathTracer::clearFocusPoint()
{
    m_impl->edit([](Controls& c) { c.optics.focusOn = false; });
    m_impl->resetIfConfigured();
}

// Closest hit in [0.001, 1e30) along normalize(rd) (PathTracer.cpp:1991-2013), traced on the GPU.
bool PathTracer::pickSurface(const Vec3& ro, const Vec3& rd, Vec3* hitPoint) const
{
    const MeshAccelSceneGpu* scn = m_impl->deviceScene();
    const float len = std::sqrt(rd.x * rd.x + rd.y * rd.y + rd.z * rd.z);
    if (hitPoint == nullptr || scn == nullptr || !(len > 0.0f)) {
        return false;
    }
    const float ray[6] = {ro.x, ro.y, ro.z, rd.x / len, rd.y / len, rd.z / len};
    DeviceArray<float> dev;
    float out[8] = {};
    const bool ok = dev.allocate(14) && hipMemcpy(dev.get(), ray, sizeof(ray), hipMemcpyHostToDevice) == hipSuccess &&
                    pathTracerTraceRays(scn, dev.get(), dev.get() + 3, 1, 0.001f, 1.0e30f, dev.get() + 6, nullptr) &&
                    hipMemcpy(out, dev.get() + 6, sizeof(out), hipMemcpyDeviceToHost) == hipSuccess;
    int hit = 0;
    std::memcpy(&hit, &out[0], sizeof(int));
    if (!ok || hit == 0) {
        return false;
    }
    *hitPoint = Vec3{ro.x + ray[3] * out[1], ro.y + ray[4] * out[1], ro.z + ray[5] * out[1]};
    return true;
}

bool PathTracer::setEnvironmentImage(const float* rgb, int width, int height)
{
    TracerImpl& t = *m_impl;
    if (rgb != nullptr && (width <= 0 || height <= 0)) {
        return false;
    }
    {
        std::lock_guard<std::mutex> g(t.streamLock);
        if (rgb == nullptr) {
            t.envPixels.clear();
            t.envSize[0] = t.envSize[1] = 0;
        } else {
            t.envPixels.assign(rgb, rgb + 3 * static_cast<size_t>(width) * static_cast<size_t>(height));
            t.envSize[0] = width;
            t.envSize[1] = height;
        }
        ++t.envRevision;
    }
    t.resetIfConfigured();
    return true;
}

void PathTracer::setEnvironmentIntensity(float intensity)
{
    const float v = clampTo(intensity, kIntensityLo, kIntensityHi);
    bool changed = false;
    m_impl->edit([&](Controls& c) {
        changed = c.envIntensity != v;
        c.envIntensity = v;
    });
    if (changed) {
        m_impl->resetIfConfigured();
    }
}
float PathTracer::environmentIntensity() const { return m_impl->snapshot().envIntensity; }

void PathTracer::setRussianRouletteMinDepth(int depth)
{
    const int v = clampTo(depth, kRouletteRange.lo, kRouletteRange.hi);
    bool changed = false;
    m_impl->edit([&](Controls& c) {
        changed = c.rouletteDepth != v;
        c.rouletteDepth = v;
    });
    if (changed) {
        m_impl->resetIfConfigured();
    }
}
int PathTracer::russianRouletteMinDepth() const { return m_impl->snapshot().rouletteDepth; }

void PathTracer::setRegionRenderEnabled(bool enabled)
{
    m_impl->edit([enabled](Controls& c) { c.regionOn = enabled; });
}
bool PathTracer::regionRenderEnabled() const { return m_impl->snapshot().regionOn; }

void PathTracer::setRenderRegion(int minX, int minY, int maxX, int maxY)
{
    const int xMax = std::max(0, m_impl->surf.width - 1);
    const int yMax = std::max(0, m_impl->surf.height - 1);
    const PtRegion r{std::max(0, std::min(minX, maxX)), std::max(0, std::min(minY, maxY)),
                     std::min(xMax, std::max(minX, maxX)), std::min(yMax, std::max(minY, maxY))};
    std::lock_guard<std::mutex> g(m_impl->controlsLock);
    m_impl->controls.region = r;
}

PtRegion PathTracer::renderRegion() const { return m_impl->snapshot().region; }

bool PathTracer::waitUntilIdle(int timeoutMs)
{
    TracerImpl& t = *m_impl;
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeoutMs);
    for (;;) {
        if (t.idle.load() && t.requests.load() == 0u) {
            return true;
        }
        if (!t.running.load()) {
            return !t.wantsSample(t.snapshot());
        }
        if (std::chrono::steady_clock::now() >= deadline) {
            return false;
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
}

bool PathTracer::readAccumulator(float* acc4, uint32_t* counts, float* lumMean, float* m2, uint8_t* converged) const
{
    TracerImpl& t = *m_impl;
    if (!t.surf.ready()) {
        return false;
    }
    const size_t n = t.surf.pixels();
    std::lock_guard<std::mutex> g(t.streamLock);
    auto fetch = [](void* dst, const void* src, size_t bytes) {
        return dst == nullptr || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    };
    return hipStreamSynchronize(t.sampling) == hipSuccess && fetch(acc4, t.surf.radiance.get(), n * sizeof(PtFloat4)) &&
           fetch(counts, t.surf.counts.get(), n * sizeof(uint32_t)) && fetch(lumMean, t.surf.mean.get(), n * sizeof(float)) &&
           fetch(m2, t.surf.m2.get(), n * sizeof(float)) && fetch(converged, t.surf.converged.get(), n);
}

bool PathTracer::readPreviewLevel(int level, float* acc4, int* width, int* height, int* downscale) const
{
    TracerImpl& t = *m_impl;
    std::lock_guard<std::mutex> g(t.streamLock);
    if (level < 0 || level >= static_cast<int>(t.surf.pyramid.size())) {
        return false;
    }
    const PreviewBuffer& b = t.surf.pyramid[static_cast<size_t>(level)];
    if (width != nullptr) {
        *width = b.size.width;
    }
    if (height != nullptr) {
        *height = b.size.height;
    }
    if (downscale != nullptr) {
        *downscale = b.size.downscale;
    }
    return hipStreamSynchronize(t.sampling) == hipSuccess &&
           (acc4 == nullptr ||
            hipMemcpy(acc4, b.texels.get(), b.texels.size() * sizeof(PtFloat4), hipMemcpyDeviceToHost) == hipSuccess);
}

float PathTracer::exposureMultiplier() const { return m_impl->snapshot().optics.exposure(); }

void PathTracer::setMaxPathDepth(int depth)
{
    const int v = std::min(depth, kDepthCap);
    bool changed = false;
    m_impl->edit([&](Controls& c) {
        changed = c.maxDepth != v;
        c.maxDepth = v;
    });
    if (changed) {
        m_impl->resetIfConfigured();
    }
}
int PathTracer::maxPathDepth() const { return m_impl->snapshot().maxDepth; }

std::string PathTracer::lastWorkerError() const
{
    std::lock_guard<std::mutex> g(m_impl->errorLock);
    return m_impl->workerError;
}

void PathTracer::setPassesPerLaunch(int passes)
{
    m_impl->edit([passes](Controls& c) { c.passesPerLaunch = clampTo(passes, kPassesRange.lo, kPassesRange.hi); });
}
int PathTracer::passesPerLaunch() const { return m_impl->snapshot().passesPerLaunch; }

void PathTracer::setBvhBuildMode(int mode)
{
    m_impl->edit([mode](Controls& c) { c.buildMode = clampTo(mode, kBuildModeRange.lo, kBuildModeRange.hi); });
}
int PathTracer::bvhBuildMode() const { return m_impl->snapshot().buildMode; }

void PathTracer::notifyWorker() { m_impl->poke(); }
void PathTracer::invokeFrameReadyCallback() { m_impl->notifyFrame(); }
void PathTracer::renderLoop() { m_impl->workerMain(); }

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

std::vector<PtProceduralInstance> instancesOf(const PtProceduralInstance* p, uint32_t n)
{
    return p == nullptr ? std::vector<PtProceduralInstance>{} : std::vector<PtProceduralInstance>(p, p + n);
}

PtMeshSceneBuildParams meshParamsOf(const PtMeshSceneBuildParams* p)
{
    return p != nullptr ? *p : PtMeshSceneBuildParams{64, 50.0f};
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
bool ptTracerConfigureProcedural(PtTracer* t, int width, int height, const PtProceduralInstance* instances,
                                 uint32_t count, const PtMeshSceneBuildParams* params)
{
    if (t == nullptr || (instances == nullptr && count != 0)) {
        ptSetError("ptTracerConfigureProcedural: invalid arguments");
        return false;
    }
    return t->tracer.configure(width, height, instancesOf(instances, count), meshParamsOf(params));
}
bool ptTracerRebuildMeshScene(PtTracer* t, const HostTriangle* triangles, uint32_t triangleCount,
                              const MaterialGpu* materials, uint32_t materialCount, const TextureDescGpu* textures,
                              uint32_t textureCount)
{
    return t != nullptr &&
           t->tracer.rebuildMeshScene(meshOf(triangles, triangleCount, materials, materialCount, textures, textureCount));
}
bool ptTracerRebuildMeshSceneProcedural(PtTracer* t, const PtProceduralInstance* instances, uint32_t count,
                                        const PtMeshSceneBuildParams* params)
{
    if (t == nullptr || (instances == nullptr && count != 0)) {
        ptSetError("ptTracerRebuildMeshSceneProcedural: invalid arguments");
        return false;
    }
    return t->tracer.rebuildMeshScene(instancesOf(instances, count), meshParamsOf(params));
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
