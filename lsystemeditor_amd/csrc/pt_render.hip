// pt_render.hip -- gfx950 path-tracing kernels and the launcher C ABI
// (PathTracerLib/src/PathTracerCuda.h:12-126, implemented in the reference at
// PathTracer.cu:614-936).
#include "pt_device.h"
#include "pt_integrator.h"
#include "pt_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

using namespace pt;

namespace {

constexpr int kBlock = 256;

// PathTracer.cu:69-105. SPEC: hero-wavelength sample (one more draw after the lens draws picks
// the four wavelengths; the spectral radiance is converted to linear sRGB before the clamp).
template <bool SPEC>
PT_DEV V3 pixel_radiance(int idx, int x, int y, int width, int height, bool jitter, int sampleIndex,
                         const CameraGpu& cam, const MeshAccelSceneGpu& s, const EnvironmentMapGpu* env,
                         const RenderParamsGpu& p, uint32_t seed, TraceStats* st, const SpecModel& sm, int lm)
{
    Rng rng;
    rng_init(rng, idx, sampleIndex, seed);
    const float ju = jitter ? rand01(rng) : 0.5f;
    const float jv = jitter ? rand01(rng) : 0.5f;
    const float u = (static_cast<float>(x) + ju) / static_cast<float>(width);
    const float v = (static_cast<float>(y) + jv) / static_cast<float>(height);
    const float l1 = rand01(rng);
    const float l2 = rand01(rng);
    V3 ro, rd;
    camera_ray(cam, u, v, l1, l2, ro, rd);
    if constexpr (SPEC) {
        const HeroColors pol = hero_colors(sm, rand01(rng));
        return clamp_firefly_spectral(pol.to_rgb(trace_path_t(s, p, env, ro, rd, pol, rng, st, lm)));
    } else {
        (void)sm;
        return clamp_firefly(trace_path(s, p, env, ro, rd, rng, st, lm));
    }
}

PT_DEV const EnvironmentMapGpu* env_or_null(const EnvironmentMapGpu* env) { return env; }

// PathTracer.cu:233-272: Welford on luminance, running mean of RGB, convergence.
// Returns true when the pixel converged with this sample.
PT_DEV bool accumulate(uint32_t n, V3 r, float4& acc, float& mean, float& m2, const RenderParamsGpu& p)
{
    const float sl = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    const float delta = sl - mean;
    const float nm = mean + delta / static_cast<float>(n + 1u);
    const float nM2 = m2 + delta * (sl - nm);
    mean = nm;
    m2 = nM2;
    const float inv = 1.0f / static_cast<float>(n + 1u);
    acc = make_float4((acc.x * static_cast<float>(n) + r.x) * inv, (acc.y * static_cast<float>(n) + r.y) * inv,
                      (acc.z * static_cast<float>(n) + r.z) * inv, 1.0f);
    const uint32_t nc = n + 1u;
    const int maxS = p.maxSamplesPerPixel;
    if (maxS > 0 && nc >= static_cast<uint32_t>(maxS)) {
        return true;
    }
    if ((nc % 4u) != 0u) {
        return false;
    }
    // PathTracer.cu:107-130
    if (nc < static_cast<uint32_t>(p.minSamples) || nc == 0u || nm < 1.0e-5f) {
        return false;
    }
    const float var = nM2 / static_cast<float>(nc);
    const float sd = sqrtf(var);
    const float denom = fmaxf(nm, 1.0e-4f);
    return (sd / denom) < p.relativeErrorThreshold;
}

// PathTracer.cu:186-273
template <bool SPEC>
__global__ __launch_bounds__(kBlock) void sample_adaptive_kernel(
    float4* __restrict__ acc, uint32_t* __restrict__ counts, float* __restrict__ lumMean, float* __restrict__ m2,
    uint8_t* __restrict__ converged, const int* __restrict__ active, int activeCount, const CameraGpu* camera,
    const MeshAccelSceneGpu* scene, const EnvironmentMapGpu* env, const RenderParamsGpu* params, int width,
    int height, uint32_t seed, SpecModel sm, int lm)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= activeCount) {
        return;
    }
    const int idx = active[t];
    if (idx < 0 || converged[idx] != 0) {
        return;
    }
    const MeshAccelSceneGpu s = *scene;
    const RenderParamsGpu p = *params;
    const CameraGpu cam = *camera;
    const uint32_t n = counts[idx];
    const V3 r = pixel_radiance<SPEC>(idx, idx % width, idx / width, width, height, true, static_cast<int>(n), cam, s,
                                      env, p, seed, nullptr, sm, lm);
    float4 a = acc[idx];
    float mean = lumMean[idx];
    float mm = m2[idx];
    const bool conv = accumulate(n, r, a, mean, mm, p);
    lumMean[idx] = mean;
    m2[idx] = mm;
    acc[idx] = a;
    counts[idx] = n + 1u;
    if (conv) {
        converged[idx] = 1;
    }
}

// Several passes in one launch; each lane takes list entries from a work counter and runs
// the pixel's consecutive samples (pass p: seed firstSeed + p) until it converges or the
// passes run out. Per-pixel results equal passCount calls of the single-pass kernel.
// STATS adds the traversal counters of all samples into totals[0..3] (nodes, triangles,
// rays, closest hits) and totals[4] (samples).
template <bool STATS, bool SPEC>
__global__ __launch_bounds__(kBlock) void render_passes_kernel(
    float4* __restrict__ acc, uint32_t* __restrict__ counts, float* __restrict__ lumMean, float* __restrict__ m2,
    uint8_t* __restrict__ converged, const int* __restrict__ active, int activeCount, const CameraGpu* camera,
    const MeshAccelSceneGpu* scene, const EnvironmentMapGpu* env, const RenderParamsGpu* params, int width,
    int height, uint32_t firstSeed, int passCount, int* workCounter, unsigned long long* totals, SpecModel sm, int lm)
{
    const MeshAccelSceneGpu s = *scene;
    const RenderParamsGpu p = *params;
    const CameraGpu cam = *camera;
    unsigned long long tot[5] = {0ull, 0ull, 0ull, 0ull, 0ull};
    while (true) {
        const int t = atomicAdd(workCounter, 1);
        if (t >= activeCount) {
            break;
        }
        const int idx = active[t];
        if (idx < 0 || converged[idx] != 0) {
            continue;
        }
        uint32_t n = counts[idx];
        float4 a = acc[idx];
        float mean = lumMean[idx];
        float mm = m2[idx];
        bool conv = false;
        const int x = idx % width;
        const int y = idx / width;
        for (int pass = 0; pass < passCount && !conv; ++pass) {
            TraceStats st{0u, 0u, 0u, 0u};
            const V3 r = pixel_radiance<SPEC>(idx, x, y, width, height, true, static_cast<int>(n), cam, s, env, p,
                                              firstSeed + static_cast<uint32_t>(pass), STATS ? &st : nullptr, sm, lm);
            if (STATS) {
                tot[0] += st.nodes;
                tot[1] += st.tris;
                tot[2] += st.rays;
                tot[3] += st.hits;
                tot[4] += 1ull;
            }
            conv = accumulate(n, r, a, mean, mm, p);
            n += 1u;
        }
        lumMean[idx] = mean;
        m2[idx] = mm;
        acc[idx] = a;
        counts[idx] = n;
        if (conv) {
            converged[idx] = 1;
        }
    }
    if (STATS) {
        for (int k = 0; k < 5; ++k) {
            atomicAdd(&totals[k], tot[k]);
        }
    }
}

// Traversal statistics for one pass (same sample as sample_adaptive_kernel, no writes to the
// accumulator): internal nodes fetched, triangles tested, rays, closest hits.
template <bool SPEC>
__global__ __launch_bounds__(kBlock) void count_traversal_kernel(const uint32_t* __restrict__ counts,
                                                                 const int* __restrict__ active, int activeCount,
                                                                 const CameraGpu* camera,
                                                                 const MeshAccelSceneGpu* scene,
                                                                 const EnvironmentMapGpu* env,
                                                                 const RenderParamsGpu* params, int width,
                                                                 int height, uint32_t seed, uint32_t* out,
                                                                 SpecModel sm, int lm)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= activeCount) {
        return;
    }
    const int idx = active[t];
    uint4 o = make_uint4(0u, 0u, 0u, 0u);
    if (idx >= 0) {
        TraceStats st{0u, 0u, 0u, 0u};
        const MeshAccelSceneGpu s = *scene;
        (void)pixel_radiance<SPEC>(idx, idx % width, idx / width, width, height, true,
                                   static_cast<int>(counts[idx]), *camera, s, env, *params, seed, &st, sm, lm);
        o = make_uint4(st.nodes, st.tris, st.rays, st.hits);
    }
    reinterpret_cast<uint4*>(out)[t] = o;
}

// PathTracer.cu:132-184, preview branch: overwrite, no jitter, sample index 0.
template <bool SPEC>
__global__ __launch_bounds__(kBlock) void preview_kernel(float4* __restrict__ buf, const CameraGpu* camera,
                                                         const MeshAccelSceneGpu* scene,
                                                         const EnvironmentMapGpu* env,
                                                         const RenderParamsGpu* params, int width, int height,
                                                         uint32_t seed, SpecModel sm, int lm)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= width * height) {
        return;
    }
    const RenderParamsGpu p = params != nullptr ? *params : RenderParamsGpu{10.0f / 255.0f, 10.0f / 255.0f,
                                                                            10.0f / 255.0f, 1.0f, 32, 3, 16,
                                                                            1024, 0, 0, 0.02f};
    const MeshAccelSceneGpu s = *scene;
    const V3 r =
        pixel_radiance<SPEC>(i, i % width, i / width, width, height, false, 0, *camera, s, env, p, seed, nullptr, sm, lm);
    buf[i] = make_float4(r.x, r.y, r.z, 1.0f);
}

// PathTracer.cu:275-303
__global__ void clear_kernel(float4* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, int* active,
                             int* activeCount, int n)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        acc[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        counts[i] = 0u;
        mean[i] = 0.0f;
        m2[i] = 0.0f;
        conv[i] = 0u;
        active[i] = i;
    }
    if (i == 0) {
        *activeCount = n;
    }
}

// PathTracer.cu:305-344
__global__ void clear_region_kernel(float4* acc, uint32_t* counts, float* mean, float* m2, uint8_t* conv, int* active,
                                    int* activeCount, int width, int minX, int minY, int regionW, int regionCount)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= regionCount) {
        return;
    }
    const int x = minX + t % regionW;
    const int y = minY + t / regionW;
    const int idx = y * width + x;
    acc[idx] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    counts[idx] = 0u;
    mean[idx] = 0.0f;
    m2[idx] = 0.0f;
    conv[idx] = 0u;
    active[t] = idx;
    if (t == 0) {
        *activeCount = regionCount;
    }
}

// Order-preserving compaction in three steps: per-tile survivor counts, a single-block scan
// of the tile counts (which also writes the total), and the in-order scatter.
constexpr int kTile = 1024;

PT_DEV bool keep_entry(int p, const uint8_t* conv, const uint32_t* counts, int maxSpp, int width, bool region, int rx0,
                       int ry0, int rx1, int ry1)
{
    if (p < 0) {
        return false;
    }
    if (conv[p] != 0) {
        return false;
    }
    if (maxSpp > 0 && counts[p] >= static_cast<uint32_t>(maxSpp)) {
        return false;
    }
    if (region) {
        const int x = p % width;
        const int y = p / width;
        if (x < rx0 || x > rx1 || y < ry0 || y > ry1) {
            return false;
        }
    }
    return true;
}

// `in == nullptr` means the identity list 0..count-1 (rebuild from the full frame).
__global__ __launch_bounds__(kBlock) void compact_count_kernel(const int* in, int count, const uint8_t* conv,
                                                               const uint32_t* counts, int maxSpp, int width,
                                                               int region, int rx0, int ry0, int rx1, int ry1,
                                                               int* tileCounts)
{
    __shared__ int warpSums[kBlock / 64];
    const int base = static_cast<int>(blockIdx.x) * kTile;
    int c = 0;
    for (int k = threadIdx.x; k < kTile; k += kBlock) {
        const int t = base + k;
        if (t < count) {
            const int p = in != nullptr ? in[t] : t;
            c += keep_entry(p, conv, counts, maxSpp, width, region != 0, rx0, ry0, rx1, ry1) ? 1 : 0;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_down(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        warpSums[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            sum += warpSums[w];
        }
        tileCounts[blockIdx.x] = sum;
    }
}

__global__ void compact_scan_kernel(int* tileCounts, int tiles, int* total)
{
    __shared__ int chunkSums[kBlock];
    // each thread scans a contiguous chunk serially, then the chunk sums are scanned
    const int per = (tiles + kBlock - 1) / kBlock;
    const int b = threadIdx.x * per;
    const int e = min(tiles, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) {
        s += tileCounts[i];
    }
    chunkSums[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < kBlock; ++i) {
            const int v = chunkSums[i];
            chunkSums[i] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    int run = chunkSums[threadIdx.x];
    for (int i = b; i < e; ++i) {
        const int v = tileCounts[i];
        tileCounts[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kBlock) void compact_scatter_kernel(const int* in, int count, const uint8_t* conv,
                                                                 const uint32_t* counts, int maxSpp, int width,
                                                                 int region, int rx0, int ry0, int rx1, int ry1,
                                                                 const int* tileOffsets, int* out)
{
    __shared__ int waveSums[kBlock / 64];
    __shared__ int runBase;
    const int base = static_cast<int>(blockIdx.x) * kTile;
    if (threadIdx.x == 0) {
        runBase = tileOffsets[blockIdx.x];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int k0 = 0; k0 < kTile; k0 += kBlock) {
        const int t = base + k0 + static_cast<int>(threadIdx.x);
        int p = -1;
        bool keep = false;
        if (t < count) {
            p = in != nullptr ? in[t] : t;
            keep = keep_entry(p, conv, counts, maxSpp, width, region != 0, rx0, ry0, rx1, ry1);
        }
        const unsigned long long mask = __ballot(keep);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) {
            waveSums[wave] = __popcll(mask);
        }
        __syncthreads();
        int waveBase = runBase;
        for (int w = 0; w < wave; ++w) {
            waveBase += waveSums[w];
        }
        if (keep) {
            out[waveBase + before] = p;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int add = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                add += waveSums[w];
            }
            runBase += add;
        }
        __syncthreads();
    }
}

// PathTracer.cu:421-512
PT_DEV uchar4 tone_u8(float r, float g, float b)
{
    return make_uchar4(static_cast<unsigned char>(fminf(fmaxf(r, 0.0f), 1.0f) * 255.0f),
                       static_cast<unsigned char>(fminf(fmaxf(g, 0.0f), 1.0f) * 255.0f),
                       static_cast<unsigned char>(fminf(fmaxf(b, 0.0f), 1.0f) * 255.0f), 255);
}

__global__ void display_kernel(const float4* src, const uint32_t* counts, const uint8_t* conv, uchar4* out, int outW,
                               int outH, int srcW, int srcH, int downscale, int useAcc, const RenderParamsGpu* params,
                               float exposure)
{
    const int x = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= outW || y >= outH) {
        return;
    }
    int si = y * outW + x;
    if (downscale > 1 && srcW > 0 && srcH > 0) {
        const int px = x / downscale;
        const int py = y / downscale;
        si = (py < srcH ? py : srcH - 1) * srcW + (px < srcW ? px : srcW - 1);
    }
    if (params != nullptr && params->debugOverlayMode == 2) {
        const bool green = useAcc && conv != nullptr && conv[si] != 0;
        out[y * outW + x] = green ? tone_u8(0.0f, 0.35f, 0.0f) : tone_u8(1.0f, 0.0f, 0.0f);
        return;
    }
    float r = params != nullptr ? params->backgroundR : 10.0f / 255.0f;
    float g = params != nullptr ? params->backgroundG : 10.0f / 255.0f;
    float b = params != nullptr ? params->backgroundB : 10.0f / 255.0f;
    if (src != nullptr) {
        if (!useAcc || (counts != nullptr && counts[si] > 0u)) {
            const float4 v = src[si];
            r = v.x;
            g = v.y;
            b = v.z;
        }
    }
    r = r * exposure;
    g = g * exposure;
    b = b * exposure;
    out[y * outW + x] = tone_u8(r / (1.0f + r), g / (1.0f + g), b / (1.0f + b));
}

// PathTracer.cu:514-554
__global__ void uv_debug_kernel(const CameraGpu* camera, const MeshAccelSceneGpu* scene, uchar4* out, int width,
                                int height, const RenderParamsGpu* params)
{
    const int x = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= width || y >= height) {
        return;
    }
    const float u = (static_cast<float>(x) + 0.5f) / static_cast<float>(width);
    const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(height);
    const CameraGpu cam = *camera;
    const V3 ro = mk(cam.position.x, cam.position.y, cam.position.z);
    const V3 rd = normalize(rot_quat(cam, film_local(cam, u, v)));
    float r = params->backgroundR, g = params->backgroundG, b = params->backgroundB;
    const MeshAccelSceneGpu s = *scene;
    const Hit h = trace_closest(s, ro, rd, 0.001f, 1.0e30f, nullptr);
    if (h.hit) {
        r = h.uvx - floorf(h.uvx);
        g = h.uvy - floorf(h.uvy);
        b = 0.0f;
    }
    out[y * width + x] = tone_u8(r, g, b);
}

__global__ __launch_bounds__(kBlock) void trace_rays_kernel(const MeshAccelSceneGpu* scene, const float* o,
                                                            const float* d, int n, float tMin, float tMax, float* out)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const MeshAccelSceneGpu s = *scene;
    const Hit h = trace_closest(s, mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                                tMin, tMax, nullptr);
    float* r = out + 8 * static_cast<size_t>(i);
    r[0] = __int_as_float(h.hit ? 1 : 0);
    r[1] = h.t;
    r[2] = h.normal.x;
    r[3] = h.normal.y;
    r[4] = h.normal.z;
    r[5] = h.uvx;
    r[6] = h.uvy;
    r[7] = __uint_as_float(h.tri);
}

unsigned grid1(long long n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

bool launched(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ptSetError("%s: launch failed: %s", what, hipGetErrorString(e));
        return false;
    }
    return true;
}

// Grow-only device scratch shared by the compaction and the work counter.
struct Scratch {
    int* ptr = nullptr;
    size_t ints = 0;
};
Scratch gScratch;

int* scratch_ints(size_t n)
{
    if (gScratch.ints >= n) {
        return gScratch.ptr;
    }
    if (gScratch.ptr != nullptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(gScratch.ptr);
        gScratch.ptr = nullptr;
        gScratch.ints = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(int)) != hipSuccess) {
        ptSetError("scratch allocation of %zu bytes failed", n * sizeof(int));
        return nullptr;
    }
    gScratch.ptr = static_cast<int*>(p);
    gScratch.ints = n;
    return gScratch.ptr;
}

// The three compaction launches over a caller-chosen tile-count buffer of at least
// ceil(count / kTile) ints; every entry the scan reads is written by the count kernel first.
bool compact_launch(const int* in, int count, int* out, int* activeCount, const uint8_t* conv,
                    const uint32_t* counts, int maxSpp, int width, bool region, int rx0, int ry0, int rx1, int ry1,
                    int* tileCounts, hipStream_t stream, const char* what)
{
    const int tiles = (count + kTile - 1) / kTile;
    hipLaunchKernelGGL(compact_count_kernel, dim3(tiles), dim3(kBlock), 0, stream, in, count, conv, counts, maxSpp,
                       width, region ? 1 : 0, rx0, ry0, rx1, ry1, tileCounts);
    hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(kBlock), 0, stream, tileCounts, tiles, activeCount);
    hipLaunchKernelGGL(compact_scatter_kernel, dim3(tiles), dim3(kBlock), 0, stream, in, count, conv, counts, maxSpp,
                       width, region ? 1 : 0, rx0, ry0, rx1, ry1, tileCounts, out);
    return launched(what);
}

bool compact_impl(const int* in, int count, int* out, int* activeCount, const uint8_t* conv, const uint32_t* counts,
                  int maxSpp, int width, bool region, int rx0, int ry0, int rx1, int ry1, hipStream_t stream,
                  const char* what)
{
    const int tiles = (count + kTile - 1) / kTile;
    int* tileCounts = scratch_ints(static_cast<size_t>(tiles) + 64);
    if (tileCounts == nullptr) {
        return false;
    }
    return compact_launch(in, count, out, activeCount, conv, counts, maxSpp, width, region, rx0, ry0, rx1, ry1,
                          tileCounts, stream, what);
}

// Tile counts of the allocation-free compaction: one int per kTile list entries, so one call
// compacts at most kNoAllocMaxTiles * kTile = 64M entries (an 8192 x 8192 frame). 256 KiB per
// GPU, held from the library's load until the process ends. Every call on a GPU shares its
// copy, so calls that overlap on one GPU (two streams, two threads) race on it.
constexpr int kNoAllocMaxTiles = 65536;
static __device__ __align__(256) int scratch[kNoAllocMaxTiles];

// Device address of `scratch` per device, looked up once per device with that device current.
// In host code the array's name is only a stand-in for the symbol, never a device address.
constexpr int kMaxDevices = 64;
std::atomic<int*> gNoAllocScratch[kMaxDevices];

int* no_alloc_scratch(const char* what)
{
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) {
        ptSetError("%s: no current device", what);
        return nullptr;
    }
    int* p = gNoAllocScratch[device].load(std::memory_order_acquire);
    if (p != nullptr) {
        return p;
    }
    void* addr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || addr == nullptr) {
        ptSetError("%s: scratch symbol lookup failed: %s", what, hipGetErrorString(e));
        return nullptr;
    }
    p = static_cast<int*>(addr);
    gNoAllocScratch[device].store(p, std::memory_order_release);
    return p;
}

} // namespace

// Light sampling of the forward renderers (pt_device.h kLightReference / kLightMis).
static std::atomic<int> gLightMode{kLightReference};

int pt_light_mode() { return gLightMode.load(); }

extern "C" {

bool pathTracerSetLightSampling(int mode)
{
    if (mode != kLightReference && mode != kLightMis) {
        ptSetError("pathTracerSetLightSampling: mode must be 0 (reference) or 1 (MIS on area lights), got %d", mode);
        return false;
    }
    gLightMode.store(mode);
    return true;
}

int pathTracerLightSampling(void) { return gLightMode.load(); }

bool pathTracerClearAccumulator(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                uint8_t* d_converged, int* d_activeIndices, int* d_activeScratch, int* d_activeCount,
                                int width, int height, hipStream_t stream)
{
    (void)d_activeScratch;
    if (d_buffer == nullptr || d_samples == nullptr || d_lumMean == nullptr || d_m2 == nullptr ||
        d_converged == nullptr || d_activeIndices == nullptr || d_activeCount == nullptr || width <= 0 ||
        height <= 0) {
        ptSetError("pathTracerClearAccumulator: invalid arguments");
        return false;
    }
    const int n = width * height;
    hipLaunchKernelGGL(clear_kernel, dim3(grid1(n, kBlock)), dim3(kBlock), 0, stream,
                       reinterpret_cast<float4*>(d_buffer), d_samples, d_lumMean, d_m2, d_converged, d_activeIndices,
                       d_activeCount, n);
    return launched("pathTracerClearAccumulator");
}

bool pathTracerClearRegionAccumulator(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                      uint8_t* d_converged, int* d_activeIndices, int* d_activeCount, int width,
                                      int height, int minX, int minY, int maxX, int maxY, hipStream_t stream)
{
    if (d_buffer == nullptr || d_samples == nullptr || d_lumMean == nullptr || d_m2 == nullptr ||
        d_converged == nullptr || d_activeIndices == nullptr || d_activeCount == nullptr || width <= 0 ||
        height <= 0 || minX > maxX || minY > maxY) {
        ptSetError("pathTracerClearRegionAccumulator: invalid arguments");
        return false;
    }
    const int rw = maxX - minX + 1;
    const int rh = maxY - minY + 1;
    const int rc = rw * rh;
    if (rc <= 0) {
        ptSetError("pathTracerClearRegionAccumulator: empty region");
        return false;
    }
    hipLaunchKernelGGL(clear_region_kernel, dim3(grid1(rc, kBlock)), dim3(kBlock), 0, stream,
                       reinterpret_cast<float4*>(d_buffer), d_samples, d_lumMean, d_m2, d_converged, d_activeIndices,
                       d_activeCount, width, minX, minY, rw, rc);
    return launched("pathTracerClearRegionAccumulator");
}

bool pathTracerRebuildActiveList(const uint8_t* d_converged, const uint32_t* d_counts, int* d_activeIndices,
                                 int* d_activeCount, int width, int height, int maxSamplesPerPixel,
                                 hipStream_t stream)
{
    if (d_converged == nullptr || d_counts == nullptr || d_activeIndices == nullptr || d_activeCount == nullptr ||
        width <= 0 || height <= 0) {
        ptSetError("pathTracerRebuildActiveList: invalid arguments");
        return false;
    }
    return compact_impl(nullptr, width * height, d_activeIndices, d_activeCount, d_converged, d_counts,
                        maxSamplesPerPixel, width, false, 0, 0, 0, 0, stream, "pathTracerRebuildActiveList");
}

bool pathTracerAdaptiveSample(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                              uint8_t* d_converged, const int* d_activeIndices, int activeCount, int width,
                              int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                              const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                              unsigned int globalSeed, hipStream_t stream)
{
    if (d_buffer == nullptr || d_samples == nullptr || d_lumMean == nullptr || d_m2 == nullptr ||
        d_converged == nullptr || d_activeIndices == nullptr || d_camera == nullptr || d_params == nullptr ||
        width <= 0 || height <= 0 || activeCount <= 0) {
        ptSetError("pathTracerAdaptiveSample: invalid arguments");
        return false;
    }
    if (d_scene == nullptr) {
        ptSetError("pathTracerAdaptiveSample: no scene");
        return false;
    }
    const bool spec = pt_spectral_enabled();
    const SpecModel sm = pt_spectral_model();
    hipLaunchKernelGGL(spec ? sample_adaptive_kernel<true> : sample_adaptive_kernel<false>,
                       dim3(grid1(activeCount, kBlock)), dim3(kBlock), 0, stream, reinterpret_cast<float4*>(d_buffer),
                       d_samples, d_lumMean, d_m2, d_converged, d_activeIndices, activeCount, d_camera, d_scene, d_env,
                       d_params, width, height, globalSeed, sm, pt_light_mode());
    return launched("pathTracerAdaptiveSample");
}

static bool render_passes_impl(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                               uint8_t* d_converged, const int* d_activeIndices, int activeCount, int width,
                               int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                               const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                               unsigned int firstGlobalSeed, int passCount, unsigned long long* d_totals,
                               hipStream_t stream)
{
    if (d_buffer == nullptr || d_samples == nullptr || d_lumMean == nullptr || d_m2 == nullptr ||
        d_converged == nullptr || d_activeIndices == nullptr || d_camera == nullptr || d_params == nullptr ||
        width <= 0 || height <= 0 || activeCount <= 0 || passCount <= 0 || d_scene == nullptr) {
        ptSetError("pathTracerAdaptiveRenderPasses: invalid arguments");
        return false;
    }
    int* counter = scratch_ints(64);
    if (counter == nullptr) {
        return false;
    }
    if (hipMemsetAsync(counter, 0, sizeof(int), stream) != hipSuccess) {
        ptSetError("pathTracerAdaptiveRenderPasses: counter reset failed");
        return false;
    }
    int device = 0;
    int cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    const long long wanted = static_cast<long long>(cus) * 16;
    const unsigned blocks = static_cast<unsigned>(
        std::min<long long>(wanted, (static_cast<long long>(activeCount) + kBlock - 1) / kBlock));
    const bool spec = pt_spectral_enabled();
    const SpecModel sm = pt_spectral_model();
    auto kernel = d_totals != nullptr ? (spec ? render_passes_kernel<true, true> : render_passes_kernel<true, false>)
                                      : (spec ? render_passes_kernel<false, true> : render_passes_kernel<false, false>);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, stream, reinterpret_cast<float4*>(d_buffer), d_samples,
                       d_lumMean, d_m2, d_converged, d_activeIndices, activeCount, d_camera, d_scene, d_env, d_params,
                       width, height, firstGlobalSeed, passCount, counter, d_totals, sm, pt_light_mode());
    return launched("pathTracerAdaptiveRenderPasses");
}

bool pathTracerAdaptiveRenderPasses(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                    uint8_t* d_converged, const int* d_activeIndices, int activeCount, int width,
                                    int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                    const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                    unsigned int firstGlobalSeed, int passCount, hipStream_t stream)
{
    return render_passes_impl(d_buffer, d_samples, d_lumMean, d_m2, d_converged, d_activeIndices, activeCount, width,
                              height, d_camera, d_scene, d_env, d_params, firstGlobalSeed, passCount, nullptr,
                              stream);
}

bool pathTracerAdaptiveRenderPassesCounted(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                           uint8_t* d_converged, const int* d_activeIndices, int activeCount,
                                           int width, int height, const CameraGpu* d_camera,
                                           const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                                           const RenderParamsGpu* d_params, unsigned int firstGlobalSeed,
                                           int passCount, uint64_t* d_totals, hipStream_t stream)
{
    if (d_totals == nullptr) {
        ptSetError("pathTracerAdaptiveRenderPassesCounted: invalid arguments");
        return false;
    }
    return render_passes_impl(d_buffer, d_samples, d_lumMean, d_m2, d_converged, d_activeIndices, activeCount, width,
                              height, d_camera, d_scene, d_env, d_params, firstGlobalSeed, passCount,
                              reinterpret_cast<unsigned long long*>(d_totals), stream);
}

bool pathTracerCompactActiveList(const int* d_activeIndicesIn, int* d_activeIndicesOut, int* d_activeCount,
                                 const uint8_t* d_converged, const uint32_t* d_counts, int maxSamplesPerPixel,
                                 int inCount, int width, int height, bool useRegionFilter, int regionMinX,
                                 int regionMinY, int regionMaxX, int regionMaxY, hipStream_t stream)
{
    if (d_activeIndicesIn == nullptr || d_activeIndicesOut == nullptr || d_activeCount == nullptr ||
        d_converged == nullptr || d_counts == nullptr || inCount <= 0 || width <= 0 || height <= 0) {
        ptSetError("pathTracerCompactActiveList: invalid arguments");
        return false;
    }
    return compact_impl(d_activeIndicesIn, inCount, d_activeIndicesOut, d_activeCount, d_converged, d_counts,
                        maxSamplesPerPixel, width, useRegionFilter, regionMinX, regionMinY, regionMaxX, regionMaxY,
                        stream, "pathTracerCompactActiveList");
}

bool pathTracerCompactActiveListNoAlloc(const int* d_activeIndicesIn, int* d_activeIndicesOut, int* d_activeCount,
                                        const uint8_t* d_converged, const uint32_t* d_counts,
                                        int maxSamplesPerPixel, int inCount, int width, int height,
                                        bool useRegionFilter, int regionMinX, int regionMinY, int regionMaxX,
                                        int regionMaxY, hipStream_t stream)
{
    if (d_activeIndicesIn == nullptr || d_activeIndicesOut == nullptr || d_activeCount == nullptr ||
        d_converged == nullptr || d_counts == nullptr || inCount <= 0 || width <= 0 || height <= 0) {
        ptSetError("pathTracerCompactActiveListNoAlloc: invalid arguments");
        return false;
    }
    if (static_cast<long long>(inCount) > static_cast<long long>(kNoAllocMaxTiles) * kTile) {
        ptSetError("pathTracerCompactActiveListNoAlloc: %d entries exceed the maximum of %lld", inCount,
                   static_cast<long long>(kNoAllocMaxTiles) * kTile);
        return false;
    }
    int* tileCounts = no_alloc_scratch("pathTracerCompactActiveListNoAlloc");
    if (tileCounts == nullptr) {
        return false;
    }
    return compact_launch(d_activeIndicesIn, inCount, d_activeIndicesOut, d_activeCount, d_converged, d_counts,
                          maxSamplesPerPixel, width, useRegionFilter, regionMinX, regionMinY, regionMaxX, regionMaxY,
                          tileCounts, stream, "pathTracerCompactActiveListNoAlloc");
}

bool pathTracerPreviewSample(PtFloat4* d_buffer, int width, int height, const CameraGpu* d_camera,
                             const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                             const RenderParamsGpu* d_params, unsigned int globalSeed, hipStream_t stream)
{
    if (d_buffer == nullptr || d_camera == nullptr || width <= 0 || height <= 0) {
        ptSetError("pathTracerPreviewSample: invalid arguments");
        return false;
    }
    if (d_scene == nullptr) {
        ptSetError("pathTracerPreviewSample: no scene");
        return false;
    }
    const bool spec = pt_spectral_enabled();
    hipLaunchKernelGGL(spec ? preview_kernel<true> : preview_kernel<false>,
                       dim3(grid1(static_cast<long long>(width) * height, kBlock)), dim3(kBlock), 0, stream,
                       reinterpret_cast<float4*>(d_buffer), d_camera, d_scene, d_env, d_params, width, height,
                       globalSeed, pt_spectral_model(), pt_light_mode());
    return launched("pathTracerPreviewSample");
}

bool pathTracerCopyToPbo(const PtFloat4* acc, const uint32_t* counts, const uint8_t* converged, PtUchar4* pbo,
                         int width, int height, const RenderParamsGpu* d_params, float exposure, hipStream_t stream)
{
    if (acc == nullptr || counts == nullptr || pbo == nullptr || d_params == nullptr || width <= 0 || height <= 0) {
        ptSetError("pathTracerCopyToPbo: invalid arguments");
        return false;
    }
    const dim3 block(16, 16);
    const dim3 grid((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(display_kernel, grid, block, 0, stream, reinterpret_cast<const float4*>(acc), counts, converged,
                       reinterpret_cast<uchar4*>(pbo), width, height, width, height, 1, 1, d_params, exposure);
    return launched("pathTracerCopyToPbo");
}

bool pathTracerUpsamplePreviewToPbo(const PtFloat4* preview, int previewWidth, int previewHeight, int downscale,
                                    PtUchar4* pbo, int fullWidth, int fullHeight, const RenderParamsGpu* d_params,
                                    float exposure, hipStream_t stream)
{
    if (preview == nullptr || pbo == nullptr || d_params == nullptr || previewWidth <= 0 || previewHeight <= 0 ||
        fullWidth <= 0 || fullHeight <= 0) {
        ptSetError("pathTracerUpsamplePreviewToPbo: invalid arguments");
        return false;
    }
    const dim3 block(16, 16);
    const dim3 grid((fullWidth + 15) / 16, (fullHeight + 15) / 16);
    hipLaunchKernelGGL(display_kernel, grid, block, 0, stream, reinterpret_cast<const float4*>(preview), nullptr,
                       nullptr, reinterpret_cast<uchar4*>(pbo), fullWidth, fullHeight, previewWidth, previewHeight,
                       downscale, 0, d_params, exposure);
    return launched("pathTracerUpsamplePreviewToPbo");
}

bool pathTracerWriteUvDebugToPbo(const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene, PtUchar4* pbo, int width,
                                 int height, const RenderParamsGpu* d_params, hipStream_t stream)
{
    if (d_camera == nullptr || d_scene == nullptr || pbo == nullptr || d_params == nullptr || width <= 0 ||
        height <= 0) {
        ptSetError("pathTracerWriteUvDebugToPbo: invalid arguments");
        return false;
    }
    const dim3 block(16, 16);
    const dim3 grid((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(uv_debug_kernel, grid, block, 0, stream, d_camera, d_scene, reinterpret_cast<uchar4*>(pbo),
                       width, height, d_params);
    return launched("pathTracerWriteUvDebugToPbo");
}

bool pathTracerTraceRays(const MeshAccelSceneGpu* d_scene, const float* d_origins, const float* d_dirs, int rayCount,
                         float tMin, float tMax, float* d_out, hipStream_t stream)
{
    if (d_scene == nullptr || d_origins == nullptr || d_dirs == nullptr || d_out == nullptr || rayCount <= 0) {
        ptSetError("pathTracerTraceRays: invalid arguments");
        return false;
    }
    hipLaunchKernelGGL(trace_rays_kernel, dim3(grid1(rayCount, kBlock)), dim3(kBlock), 0, stream, d_scene, d_origins,
                       d_dirs, rayCount, tMin, tMax, d_out);
    return launched("pathTracerTraceRays");
}

bool pathTracerCountTraversal(const uint32_t* d_samples, const int* d_activeIndices, int activeCount, int width,
                              int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                              const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                              unsigned int globalSeed, uint32_t* d_stats, hipStream_t stream)
{
    if (d_samples == nullptr || d_activeIndices == nullptr || activeCount <= 0 || width <= 0 || height <= 0 ||
        d_camera == nullptr || d_scene == nullptr || d_params == nullptr || d_stats == nullptr) {
        ptSetError("pathTracerCountTraversal: invalid arguments");
        return false;
    }
    const bool spec = pt_spectral_enabled();
    hipLaunchKernelGGL(spec ? count_traversal_kernel<true> : count_traversal_kernel<false>,
                       dim3(grid1(activeCount, kBlock)), dim3(kBlock), 0, stream, d_samples, d_activeIndices,
                       activeCount, d_camera, d_scene, d_env, d_params, width, height, globalSeed, d_stats,
                       pt_spectral_model(), pt_light_mode());
    return launched("pathTracerCountTraversal");
}

} // extern "C"
