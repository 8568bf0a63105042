// pt_wavefront.hip -- wavefront-scheduled adaptive renderer for gfx950.
//
// The per-pixel result is identical to the reference schedule (one adaptive pass per
// iteration, PathTracer.cpp:2379-2465 and PathTracer.cu:186-273): pass p uses globalSeed
// firstSeed + p, a pixel is sampled in every pass until it converges, and convergence is
// tested only when its count reaches a multiple of 4. The samples a pixel takes between two
// tests are independent paths, so a group of G consecutive samples of every active pixel is
// traced at once and the pixel state is then updated sample by sample in order; samples
// after a convergence inside the group are discarded. Each group runs the path loop
// (PathIntegratorRandCore.h:206-363) one bounce per launch pair:
//   trace    -- closest-hit for extension rays, any-hit for shadow rays (lean kernels: the
//               traversal stack lives in LDS, spilling to global memory past 8 entries);
//   shade    -- resolves the previous bounce's shadow results in the reference's order,
//               handles hits and misses, samples the lights and the BSDF, and appends the
//               next rays to queues (wave-aggregated atomics).
#include "pt_device.h"
#include "pt_adjoint_math.h"
#include "pt_integrator.h"
#include "pt_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cstring>
#include <vector>

using namespace pt;

namespace {

constexpr int kWfBlock = 256;
// LDS stack entries per thread (powers of two). Measured on the C1 frame: the closest-hit trace
// on the 4-wide tree is fastest with 16 (38.5 ms against 39.3 with 8, although the 32 KiB ring
// per block then holds it to 5 waves per SIMD); the any-hit trace (15.6 ms against 15.9 with 16)
// and the closest-hit trace on binary trees (median-split tree: 679 against 643 Msamples/s with
// 16) are fastest with 8; 4 entries spill to the global area too often (46.1 ms).
constexpr int kLdsEntriesWideClosest = 16;
constexpr int kLdsEntriesDefault = 8;
// stack entries per thread in LDS ring + global spill: binary trees push one entry per level
// (depth <= 48), 4-wide trees up to three (depth <= 24, so at most 72)
constexpr int kSpillEntries = 96;
constexpr int kFetchParts = 16;   // ray-queue partitions with a fetch counter each (trace kernels)
constexpr int kFetchStride = 16;  // ints between fetch counters: one 64-byte line each
// Per bounce: [0] ext, [1] shade, [2] shadow rays produced, [3] generated paths, then the
// extension and the shadow queue's fetch counters (kFetchParts each, kFetchStride apart).
constexpr int kFetchOffset = 16;
constexpr int kCountersPerBounce = kFetchOffset + 2 * kFetchParts * kFetchStride;
constexpr int kRefillLanes = 32;   // idle lanes that trigger a wave's ray refill
// Queue entries a trace wave claims per atomic. The queue is split into kFetchParts contiguous
// parts with a counter each; a wave starts on part blockIdx % kFetchParts and moves to the next
// part when its own runs dry. Chunks and parts keep every counter's atomic traffic low while
// leaving enough entries unclaimed for the waves to balance uneven ray costs.
constexpr int kFetchChunk = 64;
constexpr int kMaxBatchScenes = 4096;   // scenes per batch
// Paths in flight per group when the caller gives no bound: 64M paths of state (about 27 GB) are
// a tenth of the device's HBM, and groups this wide keep the number of groups (each paying the
// latency tail of its deep bounces) small.
constexpr int kDefaultPathsInFlight = 1 << 26;
constexpr int kMaxBatchGradRows = 512;  // material rows of a batch adjoint (LDS sums)
constexpr int kMaxBatchSpectralGradRows = 224;  // the same in hero-wavelength mode (36 doubles a row)

enum : uint32_t {
    kFlagExt = 1u,   // an extension ray was traced for this path at this bounce
    kFlagEnv = 2u,   // environment NEE shadow ray pending
    kFlagEmis = 4u,  // emissive NEE shadow ray pending
    kFlagReach = 8u, // the extension direction lies where the vertex's emissive NEE samples (kLightMis)
};

// Path state over the slots of one bounce's shade queue, in blocks of 64 slots (one wave's items):
// field k of slot i is the 4-byte word at base + (i / 64) * kPathBlockBytes + 256 k + 4 (i % 64).
// A wave's access to one field is one contiguous 256-byte segment, as in a structure of arrays,
// while a kernel needs one base pointer per state instead of one per field (the shade kernel reads
// one state and writes another: with a pointer per field the pointers alone outgrew its scalar
// registers). Two states alternate by bounce parity: the shade kernel reads item i of its queue at
// slot i of one and writes the paths that go on at their slot of the next queue in the other (the
// generation kernel fills state 0 in path order).
enum PathField : uint32_t {
    kPfTx, kPfTy, kPfTz, kPfTw, kPfLx, kPfLy, kPfLz, kPfLw, kPfEta, kPfPdf, kPfOx, kPfOy, kPfOz, kPfDx, kPfDy, kPfDz, kPfTmin, kPfCex, kPfCey, kPfCez, kPfCew, kPfCmx, kPfCmy, kPfCmz, kPfCmw, kPfHero, kPfPixel, kPfSample, kPfSeed, kPfDraws, kPfFlags, kPfSlot0, kPfScene,
    kPfOcc,  // the two shadow results of the slot: byte 0 environment, byte 1 emissive
    kPathFields
};
constexpr size_t kPathBlockBytes = 256u * kPathFields;

struct PathState {
    char* base;

    PT_DEV char* word(uint32_t k, uint32_t i) const
    {
        return base + static_cast<size_t>(i >> 6) * kPathBlockBytes + 256u * k + 4u * (i & 63u);
    }
    PT_DEV float& fl(uint32_t k, uint32_t i) const { return *reinterpret_cast<float*>(word(k, i)); }
    PT_DEV uint32_t& u32(uint32_t k, uint32_t i) const { return *reinterpret_cast<uint32_t*>(word(k, i)); }
    PT_DEV uint8_t& occ(uint32_t i, uint32_t type) const { return *reinterpret_cast<uint8_t*>(word(kPfOcc, i) + type); }
    PT_DEV float& Tx(uint32_t i) const { return fl(kPfTx, i); }
    PT_DEV float& Ty(uint32_t i) const { return fl(kPfTy, i); }
    PT_DEV float& Tz(uint32_t i) const { return fl(kPfTz, i); }
    PT_DEV float& Tw(uint32_t i) const { return fl(kPfTw, i); }
    PT_DEV float& Lx(uint32_t i) const { return fl(kPfLx, i); }
    PT_DEV float& Ly(uint32_t i) const { return fl(kPfLy, i); }
    PT_DEV float& Lz(uint32_t i) const { return fl(kPfLz, i); }
    PT_DEV float& Lw(uint32_t i) const { return fl(kPfLw, i); }
    PT_DEV float& eta(uint32_t i) const { return fl(kPfEta, i); }
    PT_DEV float& pdf(uint32_t i) const { return fl(kPfPdf, i); }
    PT_DEV float& ox(uint32_t i) const { return fl(kPfOx, i); }
    PT_DEV float& oy(uint32_t i) const { return fl(kPfOy, i); }
    PT_DEV float& oz(uint32_t i) const { return fl(kPfOz, i); }
    PT_DEV float& dx(uint32_t i) const { return fl(kPfDx, i); }
    PT_DEV float& dy(uint32_t i) const { return fl(kPfDy, i); }
    PT_DEV float& dz(uint32_t i) const { return fl(kPfDz, i); }
    PT_DEV float& tmin(uint32_t i) const { return fl(kPfTmin, i); }
    PT_DEV float& cex(uint32_t i) const { return fl(kPfCex, i); }
    PT_DEV float& cey(uint32_t i) const { return fl(kPfCey, i); }
    PT_DEV float& cez(uint32_t i) const { return fl(kPfCez, i); }
    PT_DEV float& cew(uint32_t i) const { return fl(kPfCew, i); }
    PT_DEV float& cmx(uint32_t i) const { return fl(kPfCmx, i); }
    PT_DEV float& cmy(uint32_t i) const { return fl(kPfCmy, i); }
    PT_DEV float& cmz(uint32_t i) const { return fl(kPfCmz, i); }
    PT_DEV float& cmw(uint32_t i) const { return fl(kPfCmw, i); }
    PT_DEV float& hero(uint32_t i) const { return fl(kPfHero, i); }
    PT_DEV uint32_t& pixel(uint32_t i) const { return u32(kPfPixel, i); }
    PT_DEV uint32_t& sample(uint32_t i) const { return u32(kPfSample, i); }
    PT_DEV uint32_t& seed(uint32_t i) const { return u32(kPfSeed, i); }
    PT_DEV uint32_t& draws(uint32_t i) const { return u32(kPfDraws, i); }
    PT_DEV uint32_t& flags(uint32_t i) const { return u32(kPfFlags, i); }
    PT_DEV uint32_t& slot0(uint32_t i) const { return u32(kPfSlot0, i); }
    PT_DEV uint32_t& scene(uint32_t i) const { return u32(kPfScene, i); }
};

struct Paths {
    PathState S[2];
    float *rx, *ry, *rz; // final clamped radiance per generation slot
    // Hit records in shade-queue order, double-buffered by bounce parity: (slot, triangle or -1,
    // material or -1, t bits) and (u, v), written by the extension trace at the ray's slot.
    int4* hitA[2];
    float2* hitB[2];
    int* extQ[2];   // extension ray queues (slots of the next shade queue)
    float4* shadowQ; // 2 per ray: (origin, slot*4 + type), (direction, -)
    int* counters;  // kCountersPerBounce per bounce (layout at kCountersPerBounce)
    int* pxBase;
    int* pxK;
    int2* spill;
    int spillThreads;
};

// One scene of a batch render, expanded on the device at every call from the caller's pointers
// (wf_batch_expand_kernel): the kernels of a batch read a path's scene, parameters, camera and
// environment here, indexed by the path's scene id.
struct BatchScene {
    MeshAccelSceneGpu scene;
    RenderParamsGpu params;
    CameraGpu camera;
    EnvironmentMapGpu env;  // valid = 0 where the scene has none (solid background)
    int width, height;
    int pixelOffset;        // first pixel of the scene in the batch's per-pixel buffers
    int gradRow;            // first gradient row of the scene's materials (adjoint)
    int maxDepth;           // the scene's resolved path depth
};

// The batch scene of global pixel gp: the last scene whose pixelOffset <= gp.
PT_DEV int batch_scene_of(const BatchScene* tab, int count, int gp)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].pixelOffset <= gp) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

PT_DEV int lane_id() { return static_cast<int>(threadIdx.x & 63u); }

// Appends one entry per lane with `want` set; returns the lane's slot (or -1).
PT_DEV int wave_append(int* counter, bool want)
{
    const unsigned long long mask = __ballot(want);
    if (mask == 0ull) {
        return -1;
    }
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    int base = 0;
    if (lane_id() == leader) {
        base = atomicAdd(counter, __popcll(mask));
    }
    base = __shfl(base, leader, 64);
    return want ? base + __popcll(mask & ((1ull << lane_id()) - 1ull)) : -1;
}

// Queue appends of one shade trip, aggregated over the block: one 64-bit atomic for the
// extension and shade queues (adjacent counters) and one for the shadow-ray queue, per 256
// items, instead of one atomic per wave and queue. Lanes get their slots in lane order within a
// wave and wave order within the block. Every thread of the block must call it.
struct BlockAppendLds {
    unsigned long long waveBase[kWfBlock / 64];
    int shadowBase[kWfBlock / 64];
    int counts[kWfBlock / 64][3];
};

PT_DEV void block_append(BlockAppendLds& lds, int* counters, bool wantExt, bool wantNext, bool wantEnv,
                         bool wantEmis, int& extSlot, int& nextSlot, int& envSlot, int& emisSlot)
{
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    const unsigned long long mExt = __ballot(wantExt);
    const unsigned long long mNext = __ballot(wantNext);
    const unsigned long long mEnv = __ballot(wantEnv);
    const unsigned long long mEmis = __ballot(wantEmis);
    if (lane_id() == 0) {
        lds.counts[wave][0] = __popcll(mExt);
        lds.counts[wave][1] = __popcll(mNext);
        lds.counts[wave][2] = __popcll(mEnv) + __popcll(mEmis);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int e = 0, n = 0, sh = 0;
        for (int w = 0; w < kWfBlock / 64; ++w) {
            e += lds.counts[w][0];
            n += lds.counts[w][1];
            sh += lds.counts[w][2];
        }
        unsigned long long base = 0ull;
        if (e + n > 0) {
            base = atomicAdd(reinterpret_cast<unsigned long long*>(counters),
                             (static_cast<unsigned long long>(n) << 32) | static_cast<unsigned long long>(e));
        }
        int sbase = sh > 0 ? atomicAdd(&counters[2], sh) : 0;
        unsigned long long eb = base & 0xFFFFFFFFull, nb = base >> 32;
        for (int w = 0; w < kWfBlock / 64; ++w) {
            lds.waveBase[w] = (nb << 32) | eb;
            lds.shadowBase[w] = sbase;
            eb += static_cast<unsigned long long>(lds.counts[w][0]);
            nb += static_cast<unsigned long long>(lds.counts[w][1]);
            sbase += lds.counts[w][2];
        }
    }
    __syncthreads();
    const unsigned long long wb = lds.waveBase[wave];
    const int sb = lds.shadowBase[wave];
    extSlot = static_cast<int>(wb & 0xFFFFFFFFull) + __popcll(mExt & below);
    nextSlot = static_cast<int>(wb >> 32) + __popcll(mNext & below);
    envSlot = sb + __popcll(mEnv & below);
    emisSlot = sb + __popcll(mEnv) + __popcll(mEmis & below);
    __syncthreads();  // the LDS record is reused by the next trip
}

// ----------------------------------------------------------------------------------------
// Traversal with an LDS stack. Same visit order and culling as pt::traverse.
// ----------------------------------------------------------------------------------------
// Typed address spaces keep the ring in ds_read/ds_write and the spill area in global
// loads/stores (a pointer that may point to either would become a flat access).
typedef int NativeInt2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) NativeInt2 LdsInt2;
typedef __attribute__((address_space(1))) NativeInt2 GlobalInt2;

template <int kLdsEntries>
struct LdsStack {
    LdsInt2* ring;     // this thread's column of the block's LDS ring
    GlobalInt2* spill; // this thread's column of the global spill area
    int spillStride;
    int sp, lo;

    PT_DEV void push(int ref, float t)
    {
        if (sp - lo == kLdsEntries) {
            spill[static_cast<size_t>(lo) * spillStride] = ring[(lo & (kLdsEntries - 1)) * kWfBlock];
            ++lo;
        }
        NativeInt2 e;
        e.x = ref;
        e.y = __float_as_int(t);
        ring[(sp & (kLdsEntries - 1)) * kWfBlock] = e;
        ++sp;
    }

    // Moves the oldest ring entries to the spill area until n more entries fit in the ring.
    PT_DEV void make_room(int n)
    {
        while (sp + n - lo > kLdsEntries) {
            spill[static_cast<size_t>(lo) * spillStride] = ring[(lo & (kLdsEntries - 1)) * kWfBlock];
            ++lo;
        }
    }

    // Writes stack position k (lo <= k < lo + kLdsEntries) without moving sp.
    PT_DEV void put(int k, int ref, float t)
    {
        NativeInt2 e;
        e.x = ref;
        e.y = __float_as_int(t);
        ring[(k & (kLdsEntries - 1)) * kWfBlock] = e;
    }

    // Requires sp > 0. When the ring is empty the entries below it are reloaded from the spill
    // area first, so the pop itself is always an LDS read.
    PT_DEV int2 pop()
    {
        if (sp == lo) {
            const int n = lo < kLdsEntries ? lo : kLdsEntries;
            for (int k = lo - n; k < lo; ++k) {
                ring[(k & (kLdsEntries - 1)) * kWfBlock] = spill[static_cast<size_t>(k) * spillStride];
            }
            lo -= n;
        }
        --sp;
        const NativeInt2 e = ring[(sp & (kLdsEntries - 1)) * kWfBlock];
        return make_int2(e.x, e.y);
    }
};

constexpr int kDone = static_cast<int>(0x80000000u); // traversal finished (leaf refs are ~slot >= -2^28)
constexpr float kInfDist = __builtin_huge_valf();     // a 4-wide child that was missed

// What a trace kernel reads of a scene: the tree, the leaf triangles and the root.
struct TraceScene {
    const PtFloat4* nodes;
    const PtFloat4* tris;
    const PtFloat4* wide;
    uint32_t triCount;
    int rootRef;
    int wideRoot;
    float rmin[3], rmax[3];
};

PT_DEV TraceScene trace_scene(const MeshAccelSceneGpu& m)
{
    return TraceScene{m.nodes,
                      m.triPos,
                      m.wideNodes,
                      m.triangleCount,
                      m.rootRef,
                      m.wideRootRef,
                      {m.rootMin[0], m.rootMin[1], m.rootMin[2]},
                      {m.rootMax[0], m.rootMax[1], m.rootMax[2]}};
}

// Orders two (entry distance, reference) pairs by distance (selects, no branch).
PT_DEV void order2(float& da, int& ra, float& db, int& rb)
{
    const bool swap = db < da;
    const float dLo = swap ? db : da;
    const float dHi = swap ? da : db;
    const int rLo = swap ? rb : ra;
    const int rHi = swap ? ra : rb;
    da = dLo;
    db = dHi;
    ra = rLo;
    rb = rHi;
}

// slab() for one child box of a 4-wide node: the entry distance, or kInfDist where the ray misses
// the box within [tMin, tMax]. The slab distances are the same floats as slab()'s: where slab()
// swaps the two distances of an axis for a negative reciprocal, the smaller of them is the near
// one (rounding is monotonic, so (max - o) * inv <= (min - o) * inv then), and min/max select
// the same values without a branch or a per-ray plane order. The running max and min are taken in
// another order, which gives the same values (no distance is NaN: the planes and the ray are
// finite, and a zero reciprocal gives zero distances; the sign of a zero changes no comparison),
// and te <= tx alone decides the hit, since te >= tMin.
PT_DEV float wide_child(const Ray& r, float x0, float y0, float z0, float x1, float y1, float z1, float tMin,
                        float tMax)
{
    const float ax = (x0 - r.o.x) * r.inv.x, bx = (x1 - r.o.x) * r.inv.x;
    const float ay = (y0 - r.o.y) * r.inv.y, by = (y1 - r.o.y) * r.inv.y;
    const float az = (z0 - r.o.z) * r.inv.z, bz = (z1 - r.o.z) * r.inv.z;
    const float te = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tMin));
    const float tx = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tMax));
    return te <= tx ? te : kInfDist;
}

// Persistent traversal: every lane takes the next ray from its wave's claimed range of the queue as
// soon as its current ray finishes (the wave claims kFetchChunk entries per atomic), so the lanes
// of a wave stay busy while per-ray node counts vary widely. One loop iteration advances a lane by one node: an
// internal node (both child boxes tested, near child first, far child deferred on the LDS stack
// with its entry distance), or a leaf triangle, followed by pops that skip deferred children whose
// entry distance exceeds the current closest t. Visit order and culling equal pt::traverse.
//   ANY_HIT = false: extension rays of the extension queue (path ids; identity when queue is
//                    null), results to the hit record at the path's shade slot.
//   ANY_HIT = true:  shadow rays of shadowQ, result to occ.
//   BATCH:           rays of a batch render; each ray traverses its own scene's tree (the scene
//                    id of its path, or the one its shadow-queue entry carries).
//   WIDE:            traverse the scene's 4-wide tree: the four child boxes of a node are tested
//                    at once (the binary tree's box floats, so the same entry distances), the
//                    nearest hit child is visited next and the others are pushed farthest first.
// The kernels are bound by the latency of their dependent node fetches, so they are held to 72
// VGPRs: 7 waves per SIMD instead of 6 (C1: extension trace 43.8 -> 42.2 ms per frame); at 64
// VGPRs (8 waves) the compiler spills and the extension trace takes 57 ms. The extension-ray
// kernel on the 4-wide tree is held to 5 waves per SIMD by its 32 KiB LDS ring instead (16 stack
// entries per lane), which leaves it 77 VGPRs and was faster still.
template <bool ANY_HIT, bool STATS, bool BATCH, bool WIDE>
__global__ __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(7))) void wf_trace_kernel(const MeshAccelSceneGpu* scene, Paths P,
                                                            PathState PS, const int* __restrict__ queue,
                                                            const int* __restrict__ countPtr, int* fetchCounters,
                                                            unsigned long long* stats, int4* __restrict__ hitA,
                                                            float2* __restrict__ hitB,
                                                            const BatchScene* __restrict__ tab, int lightMode)
{
    constexpr int kLdsEntries = (WIDE && !ANY_HIT) ? kLdsEntriesWideClosest : kLdsEntriesDefault;
    __shared__ int2 ring[kLdsEntries * kWfBlock];
    TraceScene s = BATCH ? TraceScene{} : trace_scene(*scene);
    const int n = *countPtr;
    if (n == 0) {
        return; // deep bounces of a group often have no rays left
    }
    const int gtid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    // Waves beyond what the queue can feed one chunk each leave at once: otherwise every wave of
    // the persistent grid probes every part's counter on its way out, a fixed cost per launch
    // (about a millisecond per 1M-path render) that small queues cannot amortise.
    if (static_cast<long long>(gtid >> 6) * kFetchChunk >= n) {
        return;
    }
    LdsStack<kLdsEntries> st{(LdsInt2*)(ring + threadIdx.x), (GlobalInt2*)(P.spill + gtid), P.spillThreads, 0, 0};
    TraceStats ts{0u, 0u, 0u, 0u};
    const unsigned long long lowerLanes = (1ull << lane_id()) - 1ull;
    int item = -1;        // queue index of the lane's current ray
    int wNext = 0, wEnd = 0; // the wave's unclaimed range of queue entries (same in every lane)
    int part = static_cast<int>(blockIdx.x % kFetchParts);
    int tried = 0;           // parts found drained
    bool exhausted = false;
    uint32_t tag = 0u;    // slot (extension) or slot * 4 + type (shadow)
    int bestMat = -1;
    Ray r;
    r.o = mk(0.0f, 0.0f, 0.0f);
    r.d = r.o;
    r.inv = r.o;
    float tMin = 0.0f, bestT = kRayTMax, bestU = 0.0f, bestV = 0.0f;
    float rayMax = kRayTMax;  // the ray's tMax: kRayTMax, or short of the light point (kLightMis)
    float cullT = kRayTMax * kCullSlack;  // box culling distance: the closest hit widened (kCullSlack)
    uint32_t bestTri = 0xFFFFFFFFu;
    int cur = kDone;
    while (true) {
        // Refill once half the wave is idle (or all of it): one atomic then serves >= 32 lanes.
        const bool need = item < 0 && !exhausted;
        const unsigned long long m = __ballot(need);
        const unsigned long long busy = __ballot(item >= 0);
        if (m != 0ull && (__popcll(m) >= kRefillLanes || busy == 0ull)) {
            if (wNext >= wEnd && tried < kFetchParts) {
                // a new range of at most kFetchChunk entries from the current part
                while (true) {
                    const int lo = static_cast<int>((static_cast<long long>(n) * part) / kFetchParts);
                    const int hi = static_cast<int>((static_cast<long long>(n) * (part + 1)) / kFetchParts);
                    int base = 0;
                    if (lane_id() == 0) {
                        base = atomicAdd(&fetchCounters[kFetchStride * part], kFetchChunk);
                    }
                    base = __shfl(base, 0, 64);
                    const int start = lo + base;
                    if (start < hi) {
                        wNext = start;
                        wEnd = min(start + kFetchChunk, hi);
                        break;
                    }
                    part = part + 1 == kFetchParts ? 0 : part + 1;
                    if (++tried == kFetchParts) {
                        wNext = wEnd = n;  // every part is drained
                        break;
                    }
                }
            }
            const int take = min(__popcll(m), wEnd - wNext);
            const int rank = __popcll(m & lowerLanes);
            const int idx = wNext + rank;
            wNext += take;
            if (need && rank < take) {
                if (idx < wEnd) {
                    item = idx;
                    if (ANY_HIT) {
                        const float4 a = P.shadowQ[2 * idx];
                        const float4 b = P.shadowQ[2 * idx + 1];
                        tag = __float_as_uint(a.w);
                        if (lightMode == kLightMis && (tag & 3u) == 2u) {
                            // emissive NEE under kLightMis: (origin, light point), bounded
                            const V3 o = mk(a.x, a.y, a.z);
                            const V3 toP = sub(mk(b.x, b.y, b.z), o);
                            r = make_ray(o, toP);
                            rayMax = length(toP) * kBoundedShadowScale;
                        } else {
                            r = make_ray(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z));
                            rayMax = kRayTMax;
                        }
                        tMin = kShadowBias;
                        if constexpr (BATCH) {
                            s = trace_scene(tab[__float_as_uint(b.w)].scene);
                        }
                    } else {
                        const int q = queue != nullptr ? queue[idx] : idx;
                        tag = static_cast<uint32_t>(q);
                        r = make_ray(mk(PS.ox(q), PS.oy(q), PS.oz(q)), mk(PS.dx(q), PS.dy(q), PS.dz(q)));
                        tMin = PS.tmin(q);
                        if constexpr (BATCH) {
                            s = trace_scene(tab[PS.scene(q)].scene);
                        }
                    }
                    bestT = rayMax;
                    cullT = rayMax * kCullSlack;
                    bestTri = 0xFFFFFFFFu;
                    bestMat = -1;
                    bestU = 0.0f;
                    bestV = 0.0f;
                    st.sp = 0;
                    st.lo = 0;
                    if (STATS) {
                        ts.rays += 1u;
                    }
                    float te;
                    cur = (s.triCount != 0u && slab(r, s.rmin[0], s.rmin[1], s.rmin[2], s.rmax[0], s.rmax[1],
                                                    s.rmax[2], tMin, cullT, te))
                        ? (WIDE ? s.wideRoot : s.rootRef)
                        : kDone;
                } else {
                    exhausted = true;
                }
            } else if (need && tried >= kFetchParts) {
                exhausted = true;
            }
        }
        if (__ballot(item >= 0) == 0ull) {
            if (__ballot(!exhausted) == 0ull) {
                break; // every lane is idle and the queue is exhausted
            }
            continue;
        }
        if (item < 0) {
            continue;
        }
        bool finished = cur == kDone;
        bool pop = false;
        if (!finished) {
            if constexpr (WIDE) {
                // One memory round trip per step for the whole wave: lanes at a leaf load their
                // triangle (3 float4) and lanes at a node its 4-wide record (7 float4) through the
                // same load instructions, all issued before either test, so a wave whose lanes are
                // split between triangles and nodes waits for memory once instead of twice.
                const bool leaf = cur < 0;
                const PtFloat4* src = leaf ? s.tris + 3u * static_cast<uint32_t>(~cur)
                                           : s.wide + 8u * static_cast<uint32_t>(cur);
                const float4 q0 = ld4(src, 0u);
                const float4 q1 = ld4(src, 1u);
                const float4 q2 = ld4(src, 2u);
                float4 q3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), q4 = q3, q5 = q3, q6 = q3;
                if (!leaf) {
                    q3 = ld4(src, 3u);
                    q4 = ld4(src, 4u);
                    q5 = ld4(src, 5u);
                    q6 = ld4(src, 6u);
                }
                if (leaf) {
                    const float4 a = q0;
                    const float4 b = q1;
                    const float4 c = q2;
                    asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w));
                    if (STATS) {
                        ts.tris += 1u;
                    }
                    float t, u, v;
                    const uint32_t id = __float_as_uint(a.w);
                    if (tri_test(r, a, b, c, tMin, bestT, t, u, v) && closer(t, id, bestT, bestTri)) {
                        bestT = t;
                        cullT = t * kCullSlack;
                        bestU = u;
                        bestV = v;
                        bestTri = id;
                        bestMat = __float_as_int(b.w);
                        if (ANY_HIT && t < rayMax) {
                            finished = true;
                        }
                    }
                    pop = !finished;
                } else {
                    const float4 mnx = q0;
                    const float4 mny = q1;
                    const float4 mnz = q2;
                    const float4 mxx = q3;
                    const float4 mxy = q4;
                    const float4 mxz = q5;
                    const float4 refs = q6;
                    if (STATS) {
                        ts.nodes += 1u;
                    }
                    int r0 = __float_as_int(refs.x), r1 = __float_as_int(refs.y), r2 = __float_as_int(refs.z),
                        r3 = __float_as_int(refs.w);
                    const float c0 = wide_child(r, mnx.x, mny.x, mnz.x, mxx.x, mxy.x, mxz.x, tMin, cullT);
                    const float c1 = wide_child(r, mnx.y, mny.y, mnz.y, mxx.y, mxy.y, mxz.y, tMin, cullT);
                    const float c2 = wide_child(r, mnx.z, mny.z, mnz.z, mxx.z, mxy.z, mxz.z, tMin, cullT);
                    const float c3 = wide_child(r, mnx.w, mny.w, mnz.w, mxx.w, mxy.w, mxz.w, tMin, cullT);
                    float d0 = r0 != kDone ? c0 : kInfDist;
                    float d1 = r1 != kDone ? c1 : kInfDist;
                    float d2 = r2 != kDone ? c2 : kInfDist;
                    float d3 = r3 != kDone ? c3 : kInfDist;
                    // sorting network for four; the children hit come first
                    order2(d0, r0, d1, r1);
                    order2(d2, r2, d3, r3);
                    order2(d0, r0, d2, r2);
                    order2(d1, r1, d3, r3);
                    order2(d1, r1, d2, r2);
                    if (d0 == kInfDist) {
                        pop = true;
                    } else {
                        // the others farthest first, so the nearest of them is popped first
                        // (the stack never holds more than 3 entries per level, 72 in all)
                        const int np = min(static_cast<int>(d1 != kInfDist) + static_cast<int>(d2 != kInfDist) +
                                               static_cast<int>(d3 != kInfDist),
                                           kSpillEntries - st.sp);
                        st.make_room(np);
                        if (np == 3) {
                            st.put(st.sp, r3, d3);
                        }
                        if (np >= 2) {
                            st.put(st.sp + np - 2, r2, d2);
                        }
                        if (np >= 1) {
                            st.put(st.sp + np - 1, r1, d1);
                        }
                        st.sp += np;
                        cur = r0;
                    }
                }
            } else {
                // binary trees: the same single round trip (3 float4 per triangle, 4 per node)
                const bool leaf = cur < 0;
                const PtFloat4* src = leaf ? s.tris + 3u * static_cast<uint32_t>(~cur)
                                           : s.nodes + 4u * static_cast<uint32_t>(cur);
                const float4 q0 = ld4(src, 0u);
                const float4 q1 = ld4(src, 1u);
                const float4 q2 = ld4(src, 2u);
                float4 q3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (!leaf) {
                    q3 = ld4(src, 3u);
                }
                if (leaf) {
                const float4 a = q0;
                const float4 b = q1;
                const float4 c = q2;
                asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w));
                if (STATS) {
                    ts.tris += 1u;
                }
                float t, u, v;
                const uint32_t id = __float_as_uint(a.w);
                if (tri_test(r, a, b, c, tMin, bestT, t, u, v) && closer(t, id, bestT, bestTri)) {
                    bestT = t;
                    cullT = t * kCullSlack;
                    bestU = u;
                    bestV = v;
                    bestTri = id;
                    bestMat = __float_as_int(b.w);
                    if (ANY_HIT && t < rayMax) {
                        finished = true;
                    }
                }
                pop = !finished;
                } else {
                const float4 n0 = q0;
                const float4 n1 = q1;
                const float4 n2 = q2;
                const float4 n3 = q3;
                if (STATS) {
                    ts.nodes += 1u;
                }
                float le, re;
                const bool lh = slab(r, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, tMin, cullT, le);
                const bool rh = slab(r, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, tMin, cullT, re);
                const int left = __float_as_int(n3.x);
                const int right = __float_as_int(n3.y);
                if (lh && rh) {
                    const bool leftFirst = le < re;
                    if (st.sp < kSpillEntries) {
                        st.push(leftFirst ? right : left, leftFirst ? re : le);
                    }
                    cur = leftFirst ? left : right;
                } else if (lh) {
                    cur = left;
                } else if (rh) {
                    cur = right;
                } else {
                    pop = true;
                }
                }
            }
        }
        if (pop) {
            finished = true;
            while (st.sp > 0) {
                const int2 e = st.pop();
                if (!(__int_as_float(e.y) > cullT)) {
                    cur = e.x;
                    finished = false;
                    break;
                }
            }
        }
        if (finished) {
            const bool hit = bestT < rayMax;
            if (ANY_HIT) {
                PS.occ(tag >> 2, (tag & 3u) - 1u) = hit ? 1u : 0u;
            } else {
                hitA[tag] = make_int4(static_cast<int>(tag), hit ? static_cast<int>(bestTri) : -1,
                                      hit ? bestMat : -1, __float_as_int(hit ? bestT : 0.0f));
                hitB[tag] = make_float2(bestU, bestV);
            }
            item = -1;
        }
    }
    if (STATS) {
        const int o = ANY_HIT ? 3 : 0;
        atomicAdd(&stats[o + 0], static_cast<unsigned long long>(ts.nodes));
        atomicAdd(&stats[o + 1], static_cast<unsigned long long>(ts.tris));
        atomicAdd(&stats[o + 2], static_cast<unsigned long long>(ts.rays));
    }
}

// ----------------------------------------------------------------------------------------
// Generation: camera rays for G consecutive samples of every active pixel (PathTracer.cu:69-99).
// In a batch (tab != null) active entries are batch pixels; each path keeps its scene-local pixel,
// which keys its random numbers as in a render of that scene alone, and its scene id.
// ----------------------------------------------------------------------------------------
__global__ __launch_bounds__(kWfBlock) void wf_gen_kernel(const int* __restrict__ active,
                                                          const int* __restrict__ activeCountPtr,
                                                          const uint32_t* __restrict__ counts,
                                                          const uint8_t* __restrict__ converged,
                                                          const CameraGpu* camera, const RenderParamsGpu* params,
                                                          int width, int height, uint32_t firstSeed, int passBase,
                                                          int passCount, int groupSize, Paths P, bool spectral,
                                                          int sampleLimit, const BatchScene* __restrict__ tab,
                                                          int tabCount)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int activeCount = *activeCountPtr;
    int idx = -1;
    int k = 0;
    uint32_t n = 0u;
    int sid = 0;
    int local = 0;
    if (t < activeCount) {
        idx = active[t];
        local = idx;
        if (tab != nullptr && idx >= 0) {
            sid = batch_scene_of(tab, tabCount, idx);
            local = idx - tab[sid].pixelOffset;
        }
        if (idx >= 0 && converged[idx] == 0) {
            n = counts[idx];
            k = min(groupSize, passCount - passBase);
            const int maxS = sampleLimit >= 0 ? sampleLimit
                                              : (tab != nullptr ? tab[sid].params.maxSamplesPerPixel
                                                                : params->maxSamplesPerPixel);
            if (maxS > 0) {
                k = min(k, maxS - static_cast<int>(n));
            }
            k = max(k, 0);
        }
    }
    // wave-level exclusive scan of k, one atomic per wave
    int incl = k;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane_id() >= off) {
            incl += v;
        }
    }
    const int total = __shfl(incl, 63, 64);
    int base = 0;
    if (lane_id() == 63 && total > 0) {
        base = atomicAdd(&P.counters[3], total);
    }
    const int waveBase = __shfl(base, 63, 64);
    if (t < activeCount) {
        P.pxBase[t] = waveBase + incl - k;
        P.pxK[t] = k;
    }
    // The wave's samples occupy the contiguous slots [waveBase, waveBase + total), pixel by pixel.
    // Lanes take consecutive slots so every store of the path state is coalesced; a slot's pixel
    // is the lane whose inclusive prefix first exceeds the slot offset.
    for (int q0 = 0; q0 < total; q0 += 64) {
        const int q = q0 + lane_id();
        int owner = 0;
        for (int step = 32; step > 0; step >>= 1) {
            const int probe = __shfl(incl, owner + step - 1, 64);
            if (probe <= q) {
                owner += step;
            }
        }
        const int oIdx = __shfl(local, owner, 64);
        const int oSid = __shfl(sid, owner, 64);
        const uint32_t oN = __shfl(n, owner, 64);
        const int oExcl = __shfl(incl - k, owner, 64);
        if (q >= total) {
            continue;
        }
        const int j = q - oExcl;
        const int p = waveBase + q;
        const int w = tab != nullptr ? tab[oSid].width : width;
        const int h = tab != nullptr ? tab[oSid].height : height;
        const CameraGpu cam = tab != nullptr ? tab[oSid].camera : *camera;
        const int x = oIdx % w;
        const int y = oIdx / w;
        const uint32_t seed = firstSeed + static_cast<uint32_t>(passBase + j);
        Rng rng;
        rng_init(rng, oIdx, static_cast<int>(oN) + j, seed);
        const float ju = rand01(rng);
        const float jv = rand01(rng);
        const float u = (static_cast<float>(x) + ju) / static_cast<float>(w);
        const float v = (static_cast<float>(y) + jv) / static_cast<float>(h);
        const float l1 = rand01(rng);
        const float l2 = rand01(rng);
        V3 ro, rd;
        camera_ray(cam, u, v, l1, l2, ro, rd);
        PathState& O = P.S[0];
        if (tab != nullptr) {
            O.scene(p) = static_cast<uint32_t>(oSid);
        }
        if (spectral) {
            O.hero(p) = rand01(rng);  // the hero-wavelength draw follows the lens draws
            O.Tw(p) = 1.0f;
            O.Lw(p) = 0.0f;
        }
        O.pixel(p) = static_cast<uint32_t>(oIdx);
        O.sample(p) = oN + static_cast<uint32_t>(j);
        O.seed(p) = seed;
        O.draws(p) = rng_draws(rng);
        O.slot0(p) = static_cast<uint32_t>(p);
        O.ox(p) = ro.x;
        O.oy(p) = ro.y;
        O.oz(p) = ro.z;
        O.dx(p) = rd.x;
        O.dy(p) = rd.y;
        O.dz(p) = rd.z;
        O.tmin(p) = 0.0f;
        O.Tx(p) = 1.0f;
        O.Ty(p) = 1.0f;
        O.Tz(p) = 1.0f;
        O.Lx(p) = 0.0f;
        O.Ly(p) = 0.0f;
        O.Lz(p) = 0.0f;
        O.eta(p) = 1.0f;
        O.pdf(p) = 0.0f;
        O.flags(p) = kFlagExt;
    }
}

// A shadow-queue entry: origin and slot * 4 + type, direction and the batch scene id (0 outside
// batches).
PT_DEV void store_shadow(Paths& P, int slot, V3 o, V3 d, int item, uint32_t type, uint32_t sceneId)
{
    P.shadowQ[2 * slot] = make_float4(o.x, o.y, o.z, __uint_as_float((static_cast<uint32_t>(item) << 2) | type));
    P.shadowQ[2 * slot + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(sceneId));
}

// One bounce of PathIntegratorRandCore.h:241-347 for every path in the shade queue.
// Colours stored as four consecutive fields from k (RGB uses the first three).
PT_DEV V3 ld_col(const PathState& S, uint32_t k, uint32_t p, const V3*)
{
    return mk(S.fl(k, p), S.fl(k + 1u, p), S.fl(k + 2u, p));
}
PT_DEV S4 ld_col(const PathState& S, uint32_t k, uint32_t p, const S4*)
{
    return mk4(S.fl(k, p), S.fl(k + 1u, p), S.fl(k + 2u, p), S.fl(k + 3u, p));
}
PT_DEV void st_col(const PathState& S, uint32_t k, uint32_t p, V3 v)
{
    S.fl(k, p) = v.x;
    S.fl(k + 1u, p) = v.y;
    S.fl(k + 2u, p) = v.z;
}
PT_DEV void st_col(const PathState& S, uint32_t k, uint32_t p, S4 v)
{
    S.fl(k, p) = v.x;
    S.fl(k + 1u, p) = v.y;
    S.fl(k + 2u, p) = v.z;
    S.fl(k + 3u, p) = v.w;
}
PT_DEV RgbColors make_colors(float, const SpecModel&, const RgbColors*) { return RgbColors{}; }
PT_DEV HeroColors make_colors(float hero, const SpecModel& sm, const HeroColors*) { return hero_colors(sm, hero); }
PT_DEV V3 finish_sample(const RgbColors&, V3 L) { return clamp_firefly(L); }  // PathTracer.cu:101-104
PT_DEV V3 finish_sample(const HeroColors& pol, S4 L) { return clamp_firefly_spectral(pol.to_rgb(L)); }

// BATCH: paths of a batch render. Each path reads its scene's tree, materials, parameters and
// environment through its own pointer into the batch table (lanes of a wave nearly always share
// a scene, so these reads share cache lines), and its depth limit is its scene's.
template <class Pol, bool BATCH>
PT_DEV void shade_items(const MeshAccelSceneGpu* scene, const RenderParamsGpu* params, const EnvironmentMapGpu* env,
                        PathState I, PathState O, Paths P, int bounce, int maxDepth, const int4* __restrict__ inA,
                        const float2* __restrict__ inB, const int* __restrict__ inCountPtr, int* extOut,
                        int4* __restrict__ outA, int* counters, float* capture, int captureSpp, SpecModel sm,
                        const BatchScene* __restrict__ tab, int lightMode)
{
    using C = typename Pol::Col;
    constexpr bool kSpectral = ColN<C>::n == 4;
    const C* ct = nullptr;
    __shared__ BlockAppendLds appendLds;
    const int n = *inCountPtr;
    // one scene: its description and parameters are read once, before the loop
    MeshAccelSceneGpu scene1{};
    RenderParamsGpu params1{};
    if constexpr (!BATCH) {
        scene1 = *scene;
        params1 = *params;
    }
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    // uniform trip count so the wave-level appends see every lane of the wave
    const int trips = (n + stride - 1) / stride;
    for (int it = 0; it < trips; ++it) {
        const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + it * stride;
        const bool live = i < n;
        // what goes to the item's slot of the next shade queue
        bool wantExt = false, wantEnv = false, wantEmis = false, reach = false;
        V3 envO = mk(0.0f, 0.0f, 0.0f), envD = envO, emO = envO, emD = envO, nextO = envO, nextD = envO;
        C L = ColN<C>::splat(0.0f), T = L, ce = L, cm = L;
        float nextTmin = 0.0f, nextEta = 1.0f, nextPdf = 0.0f, hero = 0.0f;
        uint32_t pixel = 0u, sample = 0u, seed = 0u, slot0 = 0u, draws = 0u, newFlags = 0u, sid = 0u;
        auto body = [&](const MeshAccelSceneGpu& s, const RenderParamsGpu& prm, const EnvironmentMapGpu* env,
                        int depthLimit, uint32_t pixelOffset) {
            const int rrMin = prm.russianRouletteMinDepth >= 0 ? prm.russianRouletteMinDepth : 3;
            const int4 ha = inA[i];
            const float2 hb = inB[i];
            const uint32_t flags = I.flags(i);
            pixel = I.pixel(i);
            sample = I.sample(i);
            seed = I.seed(i);
            slot0 = I.slot0(i);
            if constexpr (kSpectral) {
                hero = I.hero(i);
            }
            const Pol pol = make_colors(hero, sm, static_cast<const Pol*>(nullptr));
            L = ld_col(I, kPfLx, i, ct);
            // shadow results of the previous bounce, in the reference's order (env, then emissive)
            if ((flags & kFlagEnv) != 0u && I.occ(i, 0) == 0u) {
                L = add(L, ld_col(I, kPfCex, i, ct));
            }
            if ((flags & kFlagEmis) != 0u && I.occ(i, 1) == 0u) {
                L = add(L, ld_col(I, kPfCmx, i, ct));
            }
            if ((flags & kFlagExt) != 0u) {
                const V3 origin = mk(I.ox(i), I.oy(i), I.oz(i));
                const V3 dir = mk(I.dx(i), I.dy(i), I.dz(i));
                const uint32_t tri = static_cast<uint32_t>(ha.y);
                if (tri == 0xFFFFFFFFu) {
                    if (bounce == 0) {
                        L = pol.radiance(env_eval(env, prm, dir));  // PathIntegratorRandCore.h:215-217
                    } else {                                         // :338-345
                        const C envL = pol.radiance(env_eval(env, prm, dir));
                        const float lpdf = env_pdf(env, dir);
                        const float mis = mis_balance(I.pdf(i), lpdf);
                        T = ld_col(I, kPfTx, i, ct);
                        L = add(L, mul(T, scale(envL, mis)));
                    }
                } else if (bounce < depthLimit) {
                    const float t = __int_as_float(ha.w);
                    const Ray ray = make_ray(origin, dir);
                    Hit h;
                    h.hit = true;
                    h.t = t;
                    finish_hit(s, ray, tri, hb.x, hb.y, h);
                    Rng rng;
                    rng_resume(rng, static_cast<int>(pixel), static_cast<int>(sample), seed, I.draws(i));
                    T = ld_col(I, kPfTx, i, ct);
                    const V3 pos = add(origin, scale(dir, t));
                    Ctx c;
                    c.n = h.normal;
                    c.wo = scale(dir, -1.0f);
                    const uint32_t mi = static_cast<uint32_t>(ha.z);
                    c.m = resolve_material_index(s, mi, s.materials != nullptr && tri < s.triangleCount &&
                                                            mi < s.materialCount, h.uvx, h.uvy);
                    c.w = lobe_weights(c.m);
                    c.eta = I.eta(i);
                    c.flags = 0;
                    const C base = pol.base(c.m);
                    const C rad = c.w.subsurface != 0.0f ? pol.radii(c.m) : ColN<C>::splat(0.0f);
                    if (lightMode == kLightMis && bounce > 0 && c.m.emission > 0.0f) {
                        const float w = emission_mis_weight(s, tri, dir, t, I.pdf(i), (flags & kFlagReach) != 0u);
                        L = add(L, mul(T, scale(pol.emission(c.m), w)));
                    } else {
                        L = add(L, mul(T, pol.emission(c.m)));
                    }
                    // environment NEE (:65-108): value computed now, added if the shadow ray is clear
                    if (!skips_env_nee(c.m)) {
                        const float u1 = rand01(rng);
                        const float u2 = rand01(rng);
                        float lightPdf = 0.0f;
                        const V3 wi = env_sample(env, u1, u2, lightPdf);
                        if (lightPdf > kMinPdf) {
                            const float cosT = max2(0.0f, dot(c.n, wi));
                            if (cosT > 0.0f) {
                                const C Le = pol.radiance(env_eval(env, prm, wi));
                                const C f = brdf_eval_t(c, wi, base, rad);
                                const float bpdf = brdf_pdf(c, wi);
                                const float mis = mis_balance(lightPdf, bpdf);
                                const float sc = mis * cosT / lightPdf;
                                ce = mul(T, mul(f, scale(Le, sc)));
                                wantEnv = true;
                                envO = add(pos, scale(c.n, kShadowBias));
                                envD = wi;
                            }
                        }
                    }
                    // emissive-triangle NEE (:110-167); under kLightMis the path's last vertex
                    // takes the draws and adds nothing (pt_integrator.h emissive_nee_t)
                    if (s.emissiveTriangleCount != 0u) {
                        const float uTri = rand01(rng);
                        const float u1 = rand01(rng);
                        const float u2 = rand01(rng);
                        V3 lp, ln, lr;
                        float areaPdf = 0.0f;
                        int entry = 0;
                        if (!(lightMode == kLightMis && bounce + 1 == depthLimit) &&
                            sample_emissive(s, uTri, u1, u2, lp, ln, lr, areaPdf, &entry)) {
                            const V3 toL = sub(lp, pos);
                            const float d2 = dot(toL, toL);
                            if (d2 > kMinPdf) {
                                const V3 wi = normalize(toL);
                                const float cosT = max2(0.0f, dot(c.n, wi));
                                const float cosL = max2(0.0f, dot(ln, scale(wi, -1.0f)));
                                if (cosT > 0.0f && cosL > 0.0f) {
                                    const float lightPdf = areaPdf * d2 / cosL;
                                    if (lightPdf > kMinPdf) {
                                        emO = add(pos, scale(wi, kShadowBias));
                                        // kLightMis: the shadow ray runs to the light point (the
                                        // trace kernel bounds it); otherwise along wi to kRayTMax,
                                        // where the sampled triangle itself nearly always stops
                                        // it: then no ray is queued, which the next bounce reads
                                        // exactly as a traced ray found occluded, and the
                                        // contribution it would gate is not evaluated
                                        emD = lightMode == kLightMis ? lp : wi;
                                        wantEmis = lightMode == kLightMis ||
                                                   !emissive_blocks(s, entry, make_ray(emO, wi), kShadowBias, kRayTMax);
                                        if (wantEmis) {
                                            const C f = brdf_eval_t(c, wi, base, rad);
                                            const float bpdf = brdf_pdf(c, wi);
                                            const float mis = mis_balance(lightPdf, bpdf);
                                            const float sc = mis * cosT / lightPdf;
                                            cm = mul(T, mul(f, scale(pol.light(s, entry, lr), sc)));
                                        }
                                    }
                                }
                            }
                        }
                    }
                    // BSDF sampling, throughput, Russian roulette, continuation (:280-337)
                    c.flags = prm.brdfDebugFlags;
                    const float u1 = rand01(rng);
                    const float u2 = rand01(rng);
                    const Sample smp = brdf_sample(c, u1, u2);
                    bool cont = smp.valid && smp.pdf > kMinPdf;
                    if (cont) {
                        const C f = brdf_eval_t(c, smp.dir, base, rad);
                        T = apply_throughput_t(T, c, smp, f, base);
                        T = subsurface_tier_b_t(T, c.m, base, smp, rng);
                        if ((c.flags & 16) != 0) {
                            T = smp.transmitted ? mul(T, pol.reflectance(mk(0.2f, 1.0f, 0.2f)))
                                                : mul(T, pol.reflectance(mk(1.0f, 0.2f, 0.2f)));
                        }
                        const float tl = pol.lum(T);
                        if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
                            cont = false;
                        }
                        if (cont && bounce >= rrMin) {
                            const float surv = min2(0.95f, max2(tl, 0.05f));
                            if (rand01(rng) > surv) {
                                cont = false;
                            } else {
                                T = scale(T, 1.0f / surv);
                            }
                        }
                    }
                    if (cont) {
                        nextTmin = max2(kShadowBias, kRelShadowBias * t);
                        const float sgn = dot(c.n, smp.dir) >= 0.0f ? 1.0f : -1.0f;
                        nextO = add(pos, scale(c.n, sgn * nextTmin));
                        nextD = smp.dir;
                        nextEta = smp.nextEta;
                        nextPdf = smp.pdf;
                        wantExt = true;
                        reach = lightMode == kLightMis && dot(c.n, smp.dir) > 0.0f;
                    }
                    draws = rng_draws(rng);
                }
            }
            newFlags = (wantExt ? kFlagExt : 0u) | (wantEnv ? kFlagEnv : 0u) | (wantEmis ? kFlagEmis : 0u) |
                       (reach ? kFlagReach : 0u);
            if (newFlags == 0u) {
                if (capture != nullptr && sample < static_cast<uint32_t>(captureSpp)) {
                    // unclamped radiance of (pixel, sample) for the adjoint (RGB, or the four
                    // hero wavelengths)
                    const size_t slot =
                        ColN<C>::n * ((static_cast<size_t>(pixelOffset) + pixel) * captureSpp + sample);
#pragma unroll
                    for (int k = 0; k < ColN<C>::n; ++k) {
                        capture[slot + k] = cget(L, k);
                    }
                }
                const V3 r = finish_sample(pol, L);
                P.rx[slot0] = r.x;
                P.ry[slot0] = r.y;
                P.rz[slot0] = r.z;
            }
        };
        if constexpr (BATCH) {
            sid = live ? I.scene(i) : 0u;
            if (live) {
                const BatchScene* B = tab + sid;
                body(B->scene, B->params, &B->env, B->maxDepth, static_cast<uint32_t>(B->pixelOffset));
            }
        } else if (live) {
            body(scene1, params1, env, maxDepth, 0u);
        }
        const bool next = wantExt || wantEnv || wantEmis;
        int ext, sh, envSlot, emisSlot;
        block_append(appendLds, counters, wantExt, next, wantEnv, wantEmis, ext, sh, envSlot, emisSlot);
        if (wantExt) {
            extOut[ext] = sh;
        }
        if (next) {
            // the next bounce's record (the extension trace fills in its hit) and state
            outA[sh] = make_int4(sh, -1, -1, 0);
            O.flags(sh) = newFlags;
            O.pixel(sh) = pixel;
            O.sample(sh) = sample;
            O.seed(sh) = seed;
            O.slot0(sh) = slot0;
            if constexpr (BATCH) {
                O.scene(sh) = sid;
            }
            if constexpr (kSpectral) {
                O.hero(sh) = hero;
            }
            st_col(O, kPfLx, sh, L);
            if (wantEnv) {
                st_col(O, kPfCex, sh, ce);
            }
            if (wantEmis) {
                st_col(O, kPfCmx, sh, cm);
            }
            if (wantExt) {
                O.ox(sh) = nextO.x;
                O.oy(sh) = nextO.y;
                O.oz(sh) = nextO.z;
                O.dx(sh) = nextD.x;
                O.dy(sh) = nextD.y;
                O.dz(sh) = nextD.z;
                O.tmin(sh) = nextTmin;
                O.eta(sh) = nextEta;
                O.pdf(sh) = nextPdf;
                O.draws(sh) = draws;
                st_col(O, kPfTx, sh, T);
            }
        }
        if (wantEnv) {
            store_shadow(P, envSlot, envO, envD, sh, 1u, sid);
        }
        if (wantEmis) {
            store_shadow(P, emisSlot, emO, emD, sh, 2u, sid);
        }
    }
}

// The shade kernels. The RGB instantiation fits 128 VGPRs with one register spilled, so it is held
// to 4 waves per SIMD (C1: shade 26.2 -> 23.7 ms per frame); the hero-wavelength one would spill
// about 40 registers there and runs unbounded at 3 waves.
#define PT_SHADE_PARAMS                                                                                          \
    const MeshAccelSceneGpu *scene, const RenderParamsGpu *params, const EnvironmentMapGpu *env, PathState I,      \
        PathState O, Paths P, int bounce, int maxDepth, const int4 *__restrict__ inA,                            \
        const float2 *__restrict__ inB, const int *__restrict__ inCountPtr, int *extOut,                         \
        int4 *__restrict__ outA, int *counters, float *capture, int captureSpp, SpecModel sm,                    \
        const BatchScene *__restrict__ tab, int lightMode
#define PT_SHADE_ARGS                                                                                            \
    scene, params, env, I, O, P, bounce, maxDepth, inA, inB, inCountPtr, extOut, outA, counters, capture,        \
        captureSpp, sm, tab, lightMode

template <bool BATCH>
__global__ __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(4))) void wf_shade_rgb_kernel(
    PT_SHADE_PARAMS)
{
    shade_items<RgbColors, BATCH>(PT_SHADE_ARGS);
}

template <bool BATCH>
__global__ __launch_bounds__(kWfBlock) void wf_shade_hero_kernel(PT_SHADE_PARAMS)
{
    shade_items<HeroColors, BATCH>(PT_SHADE_ARGS);
}
#undef PT_SHADE_PARAMS
#undef PT_SHADE_ARGS

// ----------------------------------------------------------------------------------------
// Adjoint (path replay) of the fixed-spp render on the same schedule. Every path is the primal
// path (same RNG use and decisions as shade_items); per vertex the adjoint accumulates the
// derivatives of that vertex's contributions with respect to its material: emission, the
// next-event BSDF values and the BSDF value of the sampled continuation, which scales everything
// the path gathers afterwards, R = C - (what the path has gathered so far). A next-event
// contribution only counts when its shadow ray is clear, which the next bounce learns; since the
// BSDF derivative is linear in its weight vector, the vertex's continuation term is taken with R
// excluding its own next-event values and the two corrections (next-event derivative minus the
// continuation term of the next-event value) wait with the shadow rays as pending terms.
// Derivatives are FP32 per vertex; the sums are FP64 (replicated LDS per block, then the block's
// row of a partial-sum buffer, reduced once per adjoint).
// ----------------------------------------------------------------------------------------
struct AdjointState {
    float *gx, *gy, *gz, *gw;        // per colour component k: dLoss/dRadiance_k of the path's sample
                                     // (RGB: after the clamp VJP; hero wavelengths: through the
                                     // conversion to linear sRGB as well; gw is the fourth wavelength's)
    float* pe[kGradParams];          // pending terms of the previous vertex's environment NEE
    float* pm[kGradParams];          // pending terms of its emissive NEE
    int* pmi;                        // that vertex's material (-1: nothing pending)
};

// Loss weights and R = C for the paths of a group (slot p = generation order). RGB: C is the
// sample's radiance and the weights the clamp VJP of dLoss/dPixel / spp. Hero wavelengths: C is
// the radiance at the four wavelengths, and wavelength k's weight is the clamp VJP contracted with
// its linear-sRGB response, 0.25 cie(lambda_k) / (white * pdf) (zero where the conversion zeroes
// the wavelength: L_k <= 0 or not finite).
template <class Pol>
__global__ __launch_bounds__(kWfBlock) void wf_adjoint_init_kernel(PathState S, AdjointState A,
                                                                   const int* __restrict__ countPtr,
                                                                   const float* __restrict__ gradImg,
                                                                   const float* __restrict__ radiance, int spp,
                                                                   float invSpp, const BatchScene* __restrict__ tab,
                                                                   SpecModel sm)
{
    using C = typename Pol::Col;
    constexpr int n = ColN<C>::n;
    const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= *countPtr) {
        return;
    }
    const uint32_t px = S.pixel(p) + (tab != nullptr ? static_cast<uint32_t>(tab[S.scene(p)].pixelOffset) : 0u);
    const size_t q = n * (static_cast<size_t>(px) * spp + S.sample(p));
    const V3 gs = mk(gradImg[3 * px] * invSpp, gradImg[3 * px + 1] * invSpp, gradImg[3 * px + 2] * invSpp);
    bool any = false;
    if constexpr (n == 3) {
        const V3 Cs = mk(radiance[q], radiance[q + 1], radiance[q + 2]);
        const V3 gw = clamp_vjp(Cs, gs);
        A.gx[p] = gw.x;
        A.gy[p] = gw.y;
        A.gz[p] = gw.z;
        st_col(S, kPfLx, p, Cs);
        any = gw.x != 0.0f || gw.y != 0.0f || gw.z != 0.0f;
    } else {
        const S4 L = mk4(radiance[q], radiance[q + 1], radiance[q + 2], radiance[q + 3]);
        const HeroColors pol = hero_colors(sm, S.hero(p));
        const V3 gw = clamp_vjp_spectral(pol.to_rgb(L), gs);
        const float k0 = 0.25f / sp_max(1.0f / sm.lambdaRange, 1.0e-8f);
        const float wr = gw.x / sp_max(sm.whiteR, 1.0e-8f), wg = gw.y / sp_max(sm.whiteG, 1.0e-8f),
                    wb = gw.z / sp_max(sm.whiteB, 1.0e-8f);
        float g[4];
        for (int k = 0; k < 4; ++k) {
            const float Lk = cget(L, k);
            g[k] = 0.0f;
            if (isfinite(Lk) && Lk > 0.0f) {
                float rr, rg, rb;
                cie_rgb_response(cget(pol.lambda, k), rr, rg, rb);
                g[k] = k0 * (wr * rr + wg * rg + wb * rb);
            }
            any = any || g[k] != 0.0f;
        }
        A.gx[p] = g[0];
        A.gy[p] = g[1];
        A.gz[p] = g[2];
        A.gw[p] = g[3];
        st_col(S, kPfLx, p, L);
    }
    A.pmi[p] = -1;
    if (!any) {
        S.flags(p) = 0u;  // nothing to differentiate: the path stops at its first shade
    }
}

template <class C>
PT_DEV C safe_ratio(const C& g, const C& x, const C& f)
{
    C r;
#pragma unroll
    for (int k = 0; k < ColN<C>::n; ++k) {
        cset(r, k, cget(f, k) != 0.0f ? cget(g, k) * cget(x, k) / cget(f, k) : 0.0f);
    }
    return r;
}

// A vertex's BSDF record (RGB: r, g, b, roughness, metallic, specular, emission; hero
// wavelengths: the base colour at the four wavelengths, roughness, metallic, specular) into
// gradient row mi.
PT_DEV void flush_record(double* lg, int mi, const RgbColors&, const float acc[kGradParams]) { flush(lg, mi, acc); }
PT_DEV void flush_record(double* lg, int mi, const HeroColors& pol, const float acc[kGradParams])
{
    flush_spectral(lg, mi, pol.bin, acc);
}

// Pol: RgbColors (7 parameters per material) or HeroColors (kSpecGradParams: the binned
// reflectance and emission, roughness, metallic, specular, emission scale; a material's
// reflectance bins are parameters where it has binned reflectance, its emission bins and scale
// where it has binned emission). BATCH as in shade_items; a scene's material m sums into gradient
// row gradRow + m, and materialCount is the batch's row count.
template <class Pol, bool BATCH>
__global__ __launch_bounds__(kWfBlock) void wf_adjoint_shade_kernel(
    const MeshAccelSceneGpu* scene, const RenderParamsGpu* params, const EnvironmentMapGpu* env, PathState I,
    PathState O, AdjointState AI, AdjointState AO, Paths P, int bounce, int maxDepth, const int4* __restrict__ inA,
    const float2* __restrict__ inB, const int* __restrict__ inCountPtr, int* extOut, int4* __restrict__ outA,
    int* counters, int materialCount, int copies, double* __restrict__ partial, const BatchScene* __restrict__ tab,
    SpecModel sm)
{
    using C = typename Pol::Col;
    constexpr int n = ColN<C>::n;
    constexpr bool kSpectral = n == 4;
    constexpr int kParams = kSpectral ? kSpecGradParams : kGradParams;
    const C* ct = nullptr;
    __shared__ BlockAppendLds appendLds;
    extern __shared__ double lgAll[];
    const int perCopy = materialCount * kParams;
    for (int k = static_cast<int>(threadIdx.x); k < copies * perCopy; k += kWfBlock) {
        lgAll[k] = 0.0;
    }
    double* lg = lgAll + (static_cast<int>(threadIdx.x) % copies) * perCopy;
    __syncthreads();
    MeshAccelSceneGpu scene1{};
    RenderParamsGpu params1{};
    if constexpr (!BATCH) {
        scene1 = *scene;
        params1 = *params;
    }
    const int nItems = *inCountPtr;
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    const int trips = (nItems + stride - 1) / stride;
    for (int it = 0; it < trips; ++it) {
        const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + it * stride;
        const bool live = i < nItems;
        bool wantExt = false, wantEnv = false, wantEmis = false;
        V3 envO = mk(0.0f, 0.0f, 0.0f), envD = envO, emO = envO, emD = envO, nextO = envO, nextD = envO;
        C R = ColN<C>::splat(0.0f), T = R, ce = R, cm = R, g = R;
        float e1[kGradParams], e2[kGradParams];
        for (int k = 0; k < kGradParams; ++k) {
            e1[k] = 0.0f;
            e2[k] = 0.0f;
        }
        float nextTmin = 0.0f, nextEta = 1.0f, nextPdf = 0.0f, hero = 0.0f;
        uint32_t pixel = 0u, sample = 0u, seed = 0u, slot0 = 0u, draws = 0u, newFlags = 0u, sid = 0u;
        int pendMat = -1;
        auto body = [&](const MeshAccelSceneGpu& s, const RenderParamsGpu& prm, const EnvironmentMapGpu* env,
                        int depthLimit, int gradRow) {
            const int rrMin = prm.russianRouletteMinDepth >= 0 ? prm.russianRouletteMinDepth : 3;
            const int4 ha = inA[i];
            const float2 hb = inB[i];
            const uint32_t flags = I.flags(i);
            pixel = I.pixel(i);
            sample = I.sample(i);
            seed = I.seed(i);
            slot0 = I.slot0(i);
            if constexpr (kSpectral) {
                hero = I.hero(i);
                g = mk4(AI.gx[i], AI.gy[i], AI.gz[i], AI.gw[i]);
            } else {
                g = mk(AI.gx[i], AI.gy[i], AI.gz[i]);
            }
            const Pol pol = make_colors(hero, sm, static_cast<const Pol*>(nullptr));
            R = ld_col(I, kPfLx, i, ct);
            // the previous vertex's next-event terms, in the forward's order (env, then emissive)
            const int pmi = AI.pmi[i];
            float prev[kGradParams];
            for (int k = 0; k < kGradParams; ++k) {
                prev[k] = 0.0f;
            }
            if ((flags & kFlagEnv) != 0u && I.occ(i, 0) == 0u) {
                R = sub(R, ld_col(I, kPfCex, i, ct));
                if (pmi >= 0) {
                    for (int k = 0; k < kGradParams; ++k) {
                        prev[k] += AI.pe[k][i];
                    }
                }
            }
            if ((flags & kFlagEmis) != 0u && I.occ(i, 1) == 0u) {
                R = sub(R, ld_col(I, kPfCmx, i, ct));
                if (pmi >= 0) {
                    for (int k = 0; k < kGradParams; ++k) {
                        prev[k] += AI.pm[k][i];
                    }
                }
            }
            if (pmi >= 0) {
                flush_record(lg, pmi, pol, prev);
            }
            const uint32_t tri = static_cast<uint32_t>(ha.y);
            if ((flags & kFlagExt) != 0u && tri != 0xFFFFFFFFu && bounce < depthLimit) {
                const V3 origin = mk(I.ox(i), I.oy(i), I.oz(i));
                const V3 dir = mk(I.dx(i), I.dy(i), I.dz(i));
                const float t = __int_as_float(ha.w);
                const Ray ray = make_ray(origin, dir);
                Hit h;
                h.hit = true;
                h.t = t;
                finish_hit(s, ray, tri, hb.x, hb.y, h);
                Rng rng;
                rng_resume(rng, static_cast<int>(pixel), static_cast<int>(sample), seed, I.draws(i));
                T = ld_col(I, kPfTx, i, ct);
                const V3 pos = add(origin, scale(dir, t));
                Ctx c;
                c.n = h.normal;
                c.wo = scale(dir, -1.0f);
                const uint32_t mi = static_cast<uint32_t>(ha.z);
                const bool valid = s.materials != nullptr && tri < s.triangleCount && mi < s.materialCount;
                c.m = resolve_material_index(s, mi, valid, h.uvx, h.uvy);
                c.w = lobe_weights(c.m);
                c.eta = I.eta(i);
                c.flags = 0;
                const C base = pol.base(c.m);
                const C rad = c.w.subsurface != 0.0f ? pol.radii(c.m) : ColN<C>::splat(0.0f);
                int gmi = -1;
                Live lm{false, false, false, false, false, false, false};
                if (valid && gradRow + static_cast<int>(mi) < materialCount) {
                    gmi = gradRow + static_cast<int>(mi);
                    lm = live_of(s.materials[mi], c.m, s);
                    if constexpr (kSpectral) {
                        lm.rgb = c.m.spec != nullptr && c.m.spec[0] >= 0.0f;  // the reflectance bins
                    }
                }
                float acc[kGradParams];
                for (int k = 0; k < kGradParams; ++k) {
                    acc[k] = 0.0f;
                }
                // emission
                R = sub(R, mul(T, pol.emission(c.m)));
                if constexpr (kSpectral) {
                    if (gmi >= 0 && c.m.spec != nullptr && c.m.spec[kPtSpectrumBins] >= 0.0f) {
                        double* row = lg + gmi * kSpecGradParams;
                        float dScale = 0.0f;
                        for (int k = 0; k < 4; ++k) {
                            const float gt = cget(g, k) * cget(T, k);
                            if (gt != 0.0f) {
                                atomicAdd(&row[kPtSpectrumBins + pol.bin[k]], static_cast<double>(gt * c.m.emission));
                                dScale += gt * c.m.spec[kPtSpectrumBins + pol.bin[k]];
                            }
                        }
                        if (lm.emis && dScale != 0.0f) {
                            atomicAdd(&row[kSpecEmission], static_cast<double>(dScale));
                        }
                    }
                } else {
                    if (lm.emis) {
                        acc[6] += g.x * T.x * c.m.r + g.y * T.y * c.m.g + g.z * T.z * c.m.b;
                    }
                    if (lm.rgb) {
                        acc[0] += g.x * T.x * c.m.emission;
                        acc[1] += g.y * T.y * c.m.emission;
                        acc[2] += g.z * T.z * c.m.emission;
                    }
                }
                // environment NEE
                if (!skips_env_nee(c.m)) {
                    const float u1 = rand01(rng);
                    const float u2 = rand01(rng);
                    float lightPdf = 0.0f;
                    const V3 wi = env_sample(env, u1, u2, lightPdf);
                    if (lightPdf > kMinPdf) {
                        const float cosT = max2(0.0f, dot(c.n, wi));
                        if (cosT > 0.0f) {
                            const C f = brdf_eval_t(c, wi, base, rad);
                            const float mis = mis_balance(lightPdf, brdf_pdf(c, wi));
                            const C w = scale(pol.radiance(env_eval(env, prm, wi)), mis * cosT / lightPdf);
                            ce = mul(T, mul(f, w));
                            if (gmi >= 0) {
                                brdf_grad_t(c, wi, base, mul(g, mul(T, w)), lm, e1, e1 + n);
                            }
                            wantEnv = true;
                            envO = add(pos, scale(c.n, kShadowBias));
                            envD = wi;
                        }
                    }
                }
                // emissive-triangle NEE (under the reference light sampling the light itself nearly
                // always blocks it, so the light's own emission derivative is not taken here)
                if (s.emissiveTriangleCount != 0u) {
                    const float uTri = rand01(rng);
                    const float u1 = rand01(rng);
                    const float u2 = rand01(rng);
                    V3 lp, ln, lr;
                    float areaPdf = 0.0f;
                    int entry = 0;
                    if (sample_emissive(s, uTri, u1, u2, lp, ln, lr, areaPdf, &entry)) {
                        const V3 toL = sub(lp, pos);
                        const float d2 = dot(toL, toL);
                        if (d2 > kMinPdf) {
                            const V3 wi = normalize(toL);
                            const float cosT = max2(0.0f, dot(c.n, wi));
                            const float cosL = max2(0.0f, dot(ln, scale(wi, -1.0f)));
                            if (cosT > 0.0f && cosL > 0.0f) {
                                const float lightPdf = areaPdf * d2 / cosL;
                                if (lightPdf > kMinPdf) {
                                    const C f = brdf_eval_t(c, wi, base, rad);
                                    const float mis = mis_balance(lightPdf, brdf_pdf(c, wi));
                                    const C w = scale(pol.light(s, entry, lr), mis * cosT / lightPdf);
                                    cm = mul(T, mul(f, w));
                                    if (gmi >= 0) {
                                        brdf_grad_t(c, wi, base, mul(g, mul(T, w)), lm, e2, e2 + n);
                                    }
                                    wantEmis = true;
                                    emO = add(pos, scale(wi, kShadowBias));
                                    emD = wi;
                                }
                            }
                        }
                    }
                }
                // BSDF sampling: the continuation term, then throughput and Russian roulette
                c.flags = prm.brdfDebugFlags;
                const float u1 = rand01(rng);
                const float u2 = rand01(rng);
                const Sample smp = brdf_sample(c, u1, u2);
                bool cont = smp.valid && smp.pdf > kMinPdf;
                if (cont) {
                    const C f = brdf_eval_t(c, smp.dir, base, rad);
                    if (gmi >= 0) {
                        brdf_grad_t(c, smp.dir, base, safe_ratio(g, R, f), lm, acc, acc + n);
                        if (wantEnv) {
                            float gg[kGradParams] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                            brdf_grad_t(c, smp.dir, base, safe_ratio(g, ce, f), lm, gg, gg + n);
                            for (int k = 0; k < kGradParams; ++k) {
                                e1[k] -= gg[k];
                            }
                        }
                        if (wantEmis) {
                            float gg[kGradParams] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                            brdf_grad_t(c, smp.dir, base, safe_ratio(g, cm, f), lm, gg, gg + n);
                            for (int k = 0; k < kGradParams; ++k) {
                                e2[k] -= gg[k];
                            }
                        }
                    }
                    T = apply_throughput_t(T, c, smp, f, base);
                    T = subsurface_tier_b_t(T, c.m, base, smp, rng);
                    if ((c.flags & 16) != 0) {
                        T = smp.transmitted ? mul(T, pol.reflectance(mk(0.2f, 1.0f, 0.2f)))
                                            : mul(T, pol.reflectance(mk(1.0f, 0.2f, 0.2f)));
                    }
                    const float tl = pol.lum(T);
                    if (!isfinite(tl) || tl <= 0.0f || tl > 1.0e8f) {
                        cont = false;
                    }
                    if (cont && bounce >= rrMin) {
                        const float surv = min2(0.95f, max2(tl, 0.05f));
                        if (rand01(rng) > surv) {
                            cont = false;
                        } else {
                            T = scale(T, 1.0f / surv);
                        }
                    }
                }
                if (gmi >= 0) {
                    flush_record(lg, gmi, pol, acc);
                    if (wantEnv || wantEmis) {
                        pendMat = gmi;
                    }
                }
                if (cont) {
                    nextTmin = max2(kShadowBias, kRelShadowBias * t);
                    const float sgn = dot(c.n, smp.dir) >= 0.0f ? 1.0f : -1.0f;
                    nextO = add(pos, scale(c.n, sgn * nextTmin));
                    nextD = smp.dir;
                    nextEta = smp.nextEta;
                    nextPdf = smp.pdf;
                    wantExt = true;
                }
                draws = rng_draws(rng);
            }
            newFlags = (wantExt ? kFlagExt : 0u) | (wantEnv ? kFlagEnv : 0u) | (wantEmis ? kFlagEmis : 0u);
        };
        if constexpr (BATCH) {
            sid = live ? I.scene(i) : 0u;
            if (live) {
                const BatchScene* B = tab + sid;
                body(B->scene, B->params, &B->env, B->maxDepth, B->gradRow);
            }
        } else if (live) {
            body(scene1, params1, env, maxDepth, 0);
        }
        const bool next = wantExt || wantEnv || wantEmis;
        int ext, sh, envSlot, emisSlot;
        block_append(appendLds, counters, wantExt, next, wantEnv, wantEmis, ext, sh, envSlot, emisSlot);
        if (wantExt) {
            extOut[ext] = sh;
        }
        if (next) {
            outA[sh] = make_int4(sh, -1, -1, 0);
            O.flags(sh) = newFlags;
            O.pixel(sh) = pixel;
            O.sample(sh) = sample;
            O.seed(sh) = seed;
            O.slot0(sh) = slot0;
            if constexpr (BATCH) {
                O.scene(sh) = sid;
            }
            st_col(O, kPfLx, sh, R);
            AO.gx[sh] = cget(g, 0);
            AO.gy[sh] = cget(g, 1);
            AO.gz[sh] = cget(g, 2);
            if constexpr (kSpectral) {
                O.hero(sh) = hero;
                AO.gw[sh] = cget(g, 3);
            }
            AO.pmi[sh] = pendMat;
            if (wantEnv) {
                st_col(O, kPfCex, sh, ce);
                for (int k = 0; k < kGradParams; ++k) {
                    AO.pe[k][sh] = e1[k];
                }
            }
            if (wantEmis) {
                st_col(O, kPfCmx, sh, cm);
                for (int k = 0; k < kGradParams; ++k) {
                    AO.pm[k][sh] = e2[k];
                }
            }
            if (wantExt) {
                O.ox(sh) = nextO.x;
                O.oy(sh) = nextO.y;
                O.oz(sh) = nextO.z;
                O.dx(sh) = nextD.x;
                O.dy(sh) = nextD.y;
                O.dz(sh) = nextD.z;
                O.tmin(sh) = nextTmin;
                O.eta(sh) = nextEta;
                O.pdf(sh) = nextPdf;
                O.draws(sh) = draws;
                st_col(O, kPfTx, sh, T);
            }
        }
        if (wantEnv) {
            store_shadow(P, envSlot, envO, envD, sh, 1u, sid);
        }
        if (wantEmis) {
            store_shadow(P, emisSlot, emO, emD, sh, 2u, sid);
        }
    }
    // this block's sums into its row of the partial-sum buffer (rows are block-private)
    __syncthreads();
    for (int k = static_cast<int>(threadIdx.x); k < perCopy; k += kWfBlock) {
        double sum = 0.0;
        for (int c = 0; c < copies; ++c) {
            sum += lgAll[c * perCopy + k];
        }
        partial[static_cast<size_t>(blockIdx.x) * perCopy + k] += sum;
    }
}

// grad[k] += sum over block rows of partial[row][k]
__global__ void wf_adjoint_reduce_kernel(const double* __restrict__ partial, int rows, int perRow,
                                         double* __restrict__ grad)
{
    const int k = static_cast<int>(blockIdx.x);
    double sum = 0.0;
    for (int r = static_cast<int>(threadIdx.x); r < rows; r += static_cast<int>(blockDim.x)) {
        sum += partial[static_cast<size_t>(r) * perRow + k];
    }
    __shared__ double red[kWfBlock];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int off = kWfBlock / 2; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            red[threadIdx.x] += red[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        grad[k] += red[0];
    }
}

__global__ void wf_fill_kernel(uint32_t* p, int n, uint32_t v)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        p[i] = v;
    }
}

__global__ void wf_zero_f64_kernel(double* p, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0.0;
    }
}

// Applies each active pixel's group samples in order (PathTracer.cu:233-272). The samples of the
// 64 pixels of a wave occupy one contiguous slot range (the generation kernel's wave-level
// scan), which the wave streams through LDS in chunks with coalesced reads; each lane then runs
// its pixel's Welford updates in sample order from LDS. (Reading each pixel's samples straight
// from global memory puts 64 far-apart streams per wave in flight, and the L2 evicts them
// between consecutive samples.)
constexpr int kAccChunk = 512;                          // samples per wave and chunk
constexpr int kAccSkew = kAccChunk + kAccChunk / 32;    // one skip word per 32 against bank conflicts

PT_DEV void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kWfBlock) void wf_accumulate_kernel(const int* __restrict__ active,
                                                                 const int* __restrict__ activeCountPtr, Paths P,
                                                                 float4* __restrict__ acc, uint32_t* __restrict__ counts,
                                                                 float* __restrict__ lumMean, float* __restrict__ m2,
                                                                 uint8_t* __restrict__ converged,
                                                                 const RenderParamsGpu* params,
                                                                 const BatchScene* __restrict__ tab, int tabCount)
{
    __shared__ float stage[kWfBlock / 64][3][kAccSkew];
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = *activeCountPtr;
    const int waveFirst = t - lane_id();
    if (waveFirst >= n) {
        return;  // the whole wave is past the list
    }
    float* sx = stage[threadIdx.x >> 6][0];
    float* sy = stage[threadIdx.x >> 6][1];
    float* sz = stage[threadIdx.x >> 6][2];
    const bool valid = t < n;
    const int k = valid ? P.pxK[t] : 0;
    const int base = valid ? P.pxBase[t] : 0;
    const int lastLane = min(63, n - 1 - waveFirst);
    const int w0 = __shfl(base, 0, 64);
    const int w1 = __shfl(base + k, lastLane, 64);
    const bool mine = valid && k > 0;
    int idx = 0;
    RenderParamsGpu prm{};
    uint32_t cnt = 0u;
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float mean = 0.0f, mm = 0.0f;
    bool conv = false;
    if (mine) {
        idx = active[t];
        prm = tab != nullptr ? tab[batch_scene_of(tab, tabCount, idx)].params : *params;
        cnt = counts[idx];
        a = acc[idx];
        mean = lumMean[idx];
        mm = m2[idx];
    }
    for (int c0 = w0; c0 < w1; c0 += kAccChunk) {
        const int len = min(kAccChunk, w1 - c0);
        wave_sync();  // the previous chunk has been read
        for (int e = lane_id(); e < len; e += 64) {
            const int q = e + (e >> 5);
            sx[q] = P.rx[c0 + e];
            sy[q] = P.ry[c0 + e];
            sz[q] = P.rz[c0 + e];
        }
        wave_sync();
        if (mine && !conv) {
            const int lo = max(base, c0);
            const int hi = min(base + k, c0 + len);
            for (int p = lo; p < hi && !conv; ++p) {
                const int e = p - c0;
                const int q = e + (e >> 5);
                const V3 r = mk(sx[q], sy[q], sz[q]);
                const float sl = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
                const float delta = sl - mean;
                const float nm = mean + delta / static_cast<float>(cnt + 1u);
                const float nM2 = mm + delta * (sl - nm);
                mean = nm;
                mm = nM2;
                const float inv = 1.0f / static_cast<float>(cnt + 1u);
                a = make_float4((a.x * static_cast<float>(cnt) + r.x) * inv, (a.y * static_cast<float>(cnt) + r.y) * inv,
                                (a.z * static_cast<float>(cnt) + r.z) * inv, 1.0f);
                const uint32_t nc = cnt + 1u;
                cnt = nc;
                if (prm.maxSamplesPerPixel > 0 && nc >= static_cast<uint32_t>(prm.maxSamplesPerPixel)) {
                    conv = true;
                } else if ((nc % 4u) == 0u && nc >= static_cast<uint32_t>(prm.minSamples) && nm >= 1.0e-5f) {
                    const float sd = sqrtf(nM2 / static_cast<float>(nc));
                    conv = (sd / fmaxf(nm, 1.0e-4f)) < prm.relativeErrorThreshold;
                }
            }
        }
    }
    if (mine) {
        acc[idx] = a;
        counts[idx] = cnt;
        lumMean[idx] = mean;
        m2[idx] = mm;
        if (conv) {
            converged[idx] = 1u;
        }
    }
}

// Keeps the active entries that are still unconverged, in order (device-side count).
__global__ void wf_compact_kernel(const int* __restrict__ in, const int* __restrict__ inCountPtr,
                                  const uint8_t* __restrict__ converged, int* __restrict__ out, int* outCount)
{
    const int n = *inCountPtr;
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    const int trips = (n + stride - 1) / stride;
    for (int it = 0; it < trips; ++it) {
        const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + it * stride;
        const int idx = t < n ? in[t] : -1;
        const bool keep = idx >= 0 && converged[idx] == 0u;
        const int slot = wave_append(outCount, keep);
        if (keep) {
            out[slot] = idx;
        }
    }
}

// Active list 0..n-1 and its device count (fixed-spp renders sample every pixel every pass).
__global__ void wf_iota_kernel(int* list, int* count, int n)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        list[i] = i;
    }
    if (i == 0) {
        *count = n;
    }
}

__global__ void wf_zero_kernel(int* p, int n)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        p[i] = 0;
    }
}

// ----------------------------------------------------------------------------------------
// Host side: workspace and the group loop.
// ----------------------------------------------------------------------------------------
struct Workspace {
    void* mem = nullptr;
    size_t bytes = 0;
    int capacity = 0;  // paths
    int pixels = 0;    // active-list entries
    int depth = 0;
    int spillThreads = 0;
    Paths P{};
    int* lists[2] = {nullptr, nullptr};
    int* listCount = nullptr; // [2]
};
Workspace gWs;

// The workspace's host-side record is shared by every entry point below: one recursive lock
// serialises them while they size the workspace and enqueue (kernels of calls on different
// streams still must not overlap on the device). A graph captured over a fixed-spp render keeps
// the workspace's addresses, so each captured call adds a pin, and while pins are held the
// workspace does not move: a call that would need a larger one fails instead.
std::recursive_mutex gWsLock;
int gWsPins = 0;

// Opt-in measurement: hipEvents around every launch of the render stream, accumulated per
// kernel class, and (counted mode) traversal counters of the trace kernels.
struct Profile {
    bool timing = false;
    bool counted = false;
    double ms[4] = {0, 0, 0, 0};      // trace ext, trace shadow, shade, other
    long long launches[4] = {0, 0, 0, 0};
    unsigned long long* dStats = nullptr;  // 6 counters
    unsigned long long stats[6] = {0, 0, 0, 0, 0, 0};
    long long paths = 0;               // samples started (path slots generated)
    double bounceMs[3][258] = {};      // per bounce: trace ext, trace shadow, shade
    long long bounceRays[2][258] = {}; // per bounce: extension rays, shadow rays
    hipEvent_t ev[2 * 2048] = {};
    int cls[2048] = {};
    int bounce[2048] = {};
    int used = 0;
};
Profile gProf;

// Paths per group once pixels may converge mid-group (0 = the workspace capacity). Pixels of an
// adaptive render mostly converge at minSamples (constant radiance) or run to maxSamplesPerPixel,
// so wide groups discard few speculative samples and save many nearly empty deep-bounce launches.
int gSpeculativeBudget = 0;

// Optional per-sample radiance capture (pathTracerWavefrontCaptureRadiance).
float* gCapture = nullptr;
int gCaptureSpp = 0;

bool gCapturing = false; // set while a fixed-spp render is being captured into a graph

void prof_mark(hipStream_t stream, int cls, bool end, int bounce = -1)
{
    if (!gProf.timing || gCapturing || gProf.used >= 2048) {
        return;
    }
    const int i = gProf.used;
    if (gProf.ev[2 * i + (end ? 1 : 0)] == nullptr) {
        (void)hipEventCreate(&gProf.ev[2 * i + (end ? 1 : 0)]);
    }
    (void)hipEventRecord(gProf.ev[2 * i + (end ? 1 : 0)], stream);
    if (end) {
        gProf.cls[i] = cls;
        gProf.bounce[i] = bounce;
        gProf.used = i + 1;
    }
}

void prof_collect()
{
    for (int i = 0; i < gProf.used; ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, gProf.ev[2 * i], gProf.ev[2 * i + 1]) == hipSuccess) {
            gProf.ms[gProf.cls[i]] += ms;
            gProf.launches[gProf.cls[i]] += 1;
            if (gProf.cls[i] < 3 && gProf.bounce[i] >= 0 && gProf.bounce[i] < 258) {
                gProf.bounceMs[gProf.cls[i]][gProf.bounce[i]] += ms;
            }
        }
    }
    gProf.used = 0;
}

template <typename T>
T* carve(char*& cur, size_t count)
{
    cur = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(cur) + 255u) & ~static_cast<uintptr_t>(255u));
    T* p = reinterpret_cast<T*>(cur);
    cur += count * sizeof(T);
    return p;
}

size_t layout(Workspace& w, char* base, int cap, int pixels, int depth, int spillThreads)
{
    char* cur = base;
    Paths& P = w.P;
    const size_t c = static_cast<size_t>(cap);
    for (PathState& S : P.S) {
        S.base = carve<char>(cur, (c + 63u) / 64u * kPathBlockBytes);
    }
    P.rx = carve<float>(cur, c);
    P.ry = carve<float>(cur, c);
    P.rz = carve<float>(cur, c);
    P.extQ[0] = carve<int>(cur, c);
    P.extQ[1] = carve<int>(cur, c);
    P.hitA[0] = carve<int4>(cur, c);
    P.hitA[1] = carve<int4>(cur, c);
    P.hitB[0] = carve<float2>(cur, c);
    P.hitB[1] = carve<float2>(cur, c);
    P.shadowQ = carve<float4>(cur, 4 * c);
    P.counters = carve<int>(cur, kCountersPerBounce * static_cast<size_t>(depth + 2));
    P.pxBase = carve<int>(cur, static_cast<size_t>(pixels));
    P.pxK = carve<int>(cur, static_cast<size_t>(pixels));
    P.spill = carve<int2>(cur, static_cast<size_t>(spillThreads) * kSpillEntries);
    P.spillThreads = spillThreads;
    w.lists[0] = carve<int>(cur, static_cast<size_t>(pixels));
    w.lists[1] = carve<int>(cur, static_cast<size_t>(pixels));
    w.listCount = carve<int>(cur, 4);
    return static_cast<size_t>(cur - base) + 256;
}

bool ensure_workspace(int cap, int pixels, int depth, int spillThreads)
{
    if (gWs.mem != nullptr && gWs.capacity >= cap && gWs.pixels >= pixels && gWs.depth >= depth &&
        gWs.spillThreads >= spillThreads) {
        return true;
    }
    if (gWsPins > 0) {
        ptSetError("wavefront workspace: %d graph capture(s) hold it at %d paths, %d pixels, depth %d; it cannot grow "
                   "to %d paths, %d pixels, depth %d until they are released (pathTracerWavefrontUnpin). Reserve "
                   "enough with pathTracerWavefrontReserve before capturing",
                   gWsPins, gWs.capacity, gWs.pixels, gWs.depth, std::max(cap, gWs.capacity),
                   std::max(pixels, gWs.pixels), std::max(depth, gWs.depth));
        return false;
    }
    cap = std::max(cap, gWs.capacity);
    pixels = std::max(pixels, gWs.pixels);
    depth = std::max(depth, gWs.depth);
    spillThreads = std::max(spillThreads, gWs.spillThreads);
    Workspace probe;
    const size_t bytes = layout(probe, nullptr, cap, pixels, depth, spillThreads);
    if (gWs.mem != nullptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(gWs.mem);
        gWs.mem = nullptr;
    }
    void* mem = nullptr;
    if (hipMalloc(&mem, bytes) != hipSuccess) {
        ptSetError("wavefront workspace allocation of %zu bytes failed", bytes);
        gWs = Workspace{};
        return false;
    }
    gWs.mem = mem;
    gWs.bytes = bytes;
    gWs.capacity = cap;
    gWs.pixels = pixels;
    gWs.depth = depth;
    gWs.spillThreads = spillThreads;
    layout(gWs, static_cast<char*>(mem), cap, pixels, depth, spillThreads);
    return true;
}

unsigned blocks_for(long long n) { return static_cast<unsigned>(std::max<long long>(1, (n + kWfBlock - 1) / kWfBlock)); }

} // namespace

// One group's bounce loop: per bounce, the closest-hit trace of the extension queue, the any-hit
// trace of the shadow queue (from bounce 1 on) and the shade launch that fills the next queues.
// Launch counts depend only on the host-known depth and path bound (no readbacks).
// shade(b, shadeCount, cOut, traceBlocks) launches bounce b's shade kernel.
// The trace kernel of a launch: counted (profiling) or not, batch or single scene, binary or
// 4-wide tree.
template <bool ANY_HIT>
auto trace_kernel_for(bool counted, bool batch, bool wide)
{
    if (batch) {
        return wide ? wf_trace_kernel<ANY_HIT, false, true, true> : wf_trace_kernel<ANY_HIT, false, true, false>;
    }
    if (counted) {
        return wide ? wf_trace_kernel<ANY_HIT, true, false, true> : wf_trace_kernel<ANY_HIT, true, false, false>;
    }
    return wide ? wf_trace_kernel<ANY_HIT, false, false, true> : wf_trace_kernel<ANY_HIT, false, false, false>;
}

template <class ShadeLaunch>
void launch_bounces_with(hipStream_t stream, Paths& P, const MeshAccelSceneGpu* d_scene, int maxPathDepth,
                         long long maxPaths, int persistentBlocks, ShadeLaunch shade, const BatchScene* tab,
                         bool wide, int lightMode)
{
    const unsigned traceBlocks = static_cast<unsigned>(std::min<long long>(persistentBlocks, blocks_for(maxPaths)));
    const unsigned shadowBlocks =
        static_cast<unsigned>(std::min<long long>(persistentBlocks, blocks_for(2 * maxPaths)));
    const bool counted = gProf.counted && tab == nullptr;
    const auto extKernel = trace_kernel_for<false>(counted, tab != nullptr, wide);
    const auto shadowKernel = trace_kernel_for<true>(counted, tab != nullptr, wide);
    unsigned long long* stats = counted ? gProf.dStats : nullptr;
    for (int b = 0; b <= maxPathDepth; ++b) {
        int* cIn = P.counters + kCountersPerBounce * b;        // produced by bounce b-1 (or gen)
        int* extFetch = cIn + kFetchOffset;
        int* shadowFetch = extFetch + kFetchParts * kFetchStride;
        int* cOut = P.counters + kCountersPerBounce * (b + 1); // produced by this bounce
        const int* extQueue = b == 0 ? nullptr : P.extQ[b & 1];
        const int* extCount = b == 0 ? &P.counters[3] : &cIn[0];
        prof_mark(stream, 0, false, b);
        hipLaunchKernelGGL(extKernel, dim3(traceBlocks), dim3(kWfBlock), 0, stream, d_scene, P, P.S[b & 1], extQueue,
                           extCount, extFetch, stats, P.hitA[b & 1], P.hitB[b & 1], tab, lightMode);
        prof_mark(stream, 0, true, b);
        if (b > 0) {
            prof_mark(stream, 1, false, b);
            hipLaunchKernelGGL(shadowKernel, dim3(shadowBlocks), dim3(kWfBlock), 0, stream, d_scene, P, P.S[b & 1],
                               nullptr, &cIn[2], shadowFetch, stats, nullptr, nullptr, tab, lightMode);
            prof_mark(stream, 1, true, b);
        }
        const int* shadeCount = b == 0 ? &P.counters[3] : &cIn[1];
        prof_mark(stream, 2, false, b);
        shade(b, shadeCount, cOut, traceBlocks);
        prof_mark(stream, 2, true, b);
    }
}

void launch_bounces(hipStream_t stream, Paths& P, const MeshAccelSceneGpu* d_scene, const RenderParamsGpu* d_params,
                    const EnvironmentMapGpu* d_env, int maxPathDepth, long long maxPaths, int persistentBlocks,
                    bool spec, const SpecModel& sm, const BatchScene* tab, bool wide)
{
    const int lightMode = pt_light_mode();
    const auto kernel = tab != nullptr ? (spec ? wf_shade_hero_kernel<true> : wf_shade_rgb_kernel<true>)
                                       : (spec ? wf_shade_hero_kernel<false> : wf_shade_rgb_kernel<false>);
    launch_bounces_with(
        stream, P, d_scene, maxPathDepth, maxPaths, persistentBlocks,
        [&](int b, const int* shadeCount, int* cOut, unsigned traceBlocks) {
            hipLaunchKernelGGL(kernel, dim3(traceBlocks), dim3(kWfBlock), 0, stream, d_scene, d_params, d_env,
                               P.S[b & 1], P.S[(b + 1) & 1], P, b, maxPathDepth, P.hitA[b & 1], P.hitB[b & 1],
                               shadeCount, P.extQ[(b + 1) & 1], P.hitA[(b + 1) & 1], cOut, gCapture, gCaptureSpp, sm,
                               tab, lightMode);
        },
        tab, wide, lightMode);
}

extern "C" bool pathTracerAdaptiveRenderWavefront(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean,
                                                  float* d_m2, uint8_t* d_converged, const int* d_activeIndices,
                                                  int activeCount, int width, int height, const CameraGpu* d_camera,
                                                  const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                                                  const RenderParamsGpu* d_params, unsigned int firstGlobalSeed,
                                                  int passCount, int maxPathDepth, int maxPathsInFlight,
                                                  hipStream_t stream)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    if (d_buffer == nullptr || d_samples == nullptr || d_lumMean == nullptr || d_m2 == nullptr ||
        d_converged == nullptr || d_activeIndices == nullptr || d_camera == nullptr || d_params == nullptr ||
        d_scene == nullptr || width <= 0 || height <= 0 || activeCount <= 0 || passCount <= 0 ||
        maxPathDepth < 0 || maxPathDepth > 256) {
        ptSetError("pathTracerAdaptiveRenderWavefront: invalid arguments");
        return false;
    }
    // maxPathDepth must equal the depth the params resolve to (PathIntegratorRandCore.h:219-227);
    // it is passed by value because the params live in device memory.
    const int target = maxPathsInFlight > 0 ? maxPathsInFlight : kDefaultPathsInFlight;
    // Samples every active pixel is sure to take: convergence is only tested at counts that are
    // multiples of 4 and at least minSamples, or reached at maxSamplesPerPixel. A first group of
    // that size starts no speculative sample.
    RenderParamsGpu hp{};
    if (hipMemcpyAsync(&hp, d_params, sizeof(hp), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        ptSetError("pathTracerAdaptiveRenderWavefront: params read failed");
        return false;
    }
    int safeEnd = std::max(4, (hp.minSamples + 3) & ~3);
    if (hp.maxSamplesPerPixel > 0) {
        safeEnd = std::min(safeEnd, hp.maxSamplesPerPixel);
    }
    int device = 0, cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    const int persistentBlocks = cus * 8;
    const bool spec = pt_spectral_enabled();
    const SpecModel sm = pt_spectral_model();
    const bool wide = pt_scene_is_wide(d_scene);
    const int spillThreads = persistentBlocks * kWfBlock;
    const int cap = std::max(4 * activeCount, std::min(target, 64 * activeCount));
    if (!ensure_workspace(cap, activeCount, maxPathDepth, spillThreads)) {
        return false;
    }
    Paths P = gWs.P;
    int* list = gWs.lists[0];
    int* listNext = gWs.lists[1];
    int* countCur = gWs.listCount;
    int* countNext = gWs.listCount + 1;
    if (hipMemcpyAsync(list, d_activeIndices, sizeof(int) * static_cast<size_t>(activeCount), hipMemcpyDeviceToDevice,
                       stream) != hipSuccess ||
        hipMemcpyAsync(countCur, &activeCount, sizeof(int), hipMemcpyHostToDevice, stream) != hipSuccess) {
        ptSetError("pathTracerAdaptiveRenderWavefront: list copy failed");
        return false;
    }
    int active = activeCount;
    int passBase = 0;
    const int counterInts = kCountersPerBounce * (maxPathDepth + 2);
    while (active > 0 && passBase < passCount) {
        // group size: 4 samples per pixel (one convergence interval), widened speculatively up
        // to 256 while the group fits the paths-in-flight bound: pixels of an adaptive render
        // mostly converge at minSamples or run to maxSamplesPerPixel, so few speculative samples
        // are discarded, and every group saved is one set of deep-bounce launch tails saved
        int group = 0;
        if (passBase < safeEnd) {
            group = safeEnd - passBase;
            const int fit = std::max(1, gWs.capacity / std::max(1, active));
            group = std::min(group, fit >= 4 ? fit & ~3 : fit);
        } else {
            const int budget = gSpeculativeBudget > 0 ? std::min(gWs.capacity, gSpeculativeBudget) : gWs.capacity;
            group = 4 * std::max(1, std::min(64, budget / std::max(1, 4 * active)));
        }
        group = std::min(group, passCount - passBase);
        const long long maxPaths = static_cast<long long>(group) * active;
        hipLaunchKernelGGL(wf_zero_kernel, dim3(blocks_for(counterInts)), dim3(kWfBlock), 0, stream, P.counters,
                           counterInts);
        prof_mark(stream, 3, false);
        hipLaunchKernelGGL(wf_gen_kernel, dim3(blocks_for(active)), dim3(kWfBlock), 0, stream, list, countCur,
                           d_samples, d_converged, d_camera, d_params, width, height, firstGlobalSeed, passBase,
                           passCount, group, P, spec, -1, nullptr, 0);
        prof_mark(stream, 3, true);
        launch_bounces(stream, P, d_scene, d_params, d_env, maxPathDepth, maxPaths, persistentBlocks, spec, sm, nullptr,
                       wide);
        prof_mark(stream, 3, false);
        hipLaunchKernelGGL(wf_accumulate_kernel, dim3(blocks_for(active)), dim3(kWfBlock), 0, stream, list, countCur, P,
                           reinterpret_cast<float4*>(d_buffer), d_samples, d_lumMean, d_m2, d_converged, d_params, nullptr, 0);
        prof_mark(stream, 3, true);
        hipLaunchKernelGGL(wf_zero_kernel, dim3(1), dim3(kWfBlock), 0, stream, countNext, 1);
        hipLaunchKernelGGL(wf_compact_kernel, dim3(std::min<unsigned>(blocks_for(active), 4096u)), dim3(kWfBlock), 0,
                           stream, list, countCur, d_converged, listNext, countNext);
        if (hipGetLastError() != hipSuccess) {
            ptSetError("pathTracerAdaptiveRenderWavefront: launch failed");
            return false;
        }
        if (hipMemcpyAsync(&active, countNext, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            ptSetError("pathTracerAdaptiveRenderWavefront: count readback failed");
            return false;
        }
        if (gProf.timing || gProf.counted) {
            std::vector<int> cnt(static_cast<size_t>(counterInts));
            if (hipMemcpy(cnt.data(), P.counters, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost) == hipSuccess) {
                gProf.paths += cnt[3];
                for (int b = 0; b <= maxPathDepth && b < 258; ++b) {
                    gProf.bounceRays[0][b] += b == 0 ? cnt[3] : cnt[kCountersPerBounce * b + 0];
                    gProf.bounceRays[1][b] += b == 0 ? 0 : cnt[kCountersPerBounce * b + 2];
                }
            }
        }
        prof_collect();
        std::swap(list, listNext);
        std::swap(countCur, countNext);
        passBase += group;
    }
    return true;
}

extern "C" bool pathTracerWavefrontProfile(int timing, int counted)
{
    gProf.timing = timing != 0;
    gProf.counted = counted != 0;
    for (int k = 0; k < 4; ++k) {
        gProf.ms[k] = 0.0;
        gProf.launches[k] = 0;
    }
    gProf.used = 0;
    gProf.paths = 0;
    std::memset(gProf.bounceMs, 0, sizeof(gProf.bounceMs));
    std::memset(gProf.bounceRays, 0, sizeof(gProf.bounceRays));
    if (gProf.counted) {
        if (gProf.dStats == nullptr && hipMalloc(reinterpret_cast<void**>(&gProf.dStats), 6 * sizeof(unsigned long long)) != hipSuccess) {
            ptSetError("pathTracerWavefrontProfile: allocation failed");
            gProf.counted = false;
            return false;
        }
        if (hipMemset(gProf.dStats, 0, 6 * sizeof(unsigned long long)) != hipSuccess) {
            ptSetError("pathTracerWavefrontProfile: reset failed");
            return false;
        }
    }
    return true;
}

// out[0..3]: milliseconds of trace-ext, trace-shadow, shade, other launches; out[4..7]: launch
// counts; out[8..13]: ext nodes, ext triangles, ext rays, shadow nodes, shadow triangles,
// shadow rays (counted mode). Collected since the last pathTracerWavefrontProfile call.
extern "C" bool pathTracerWavefrontProfileRead(double* out, int n)
{
    if (out == nullptr || n < 14) {
        ptSetError("pathTracerWavefrontProfileRead: need 14 doubles");
        return false;
    }
    unsigned long long st[6] = {0, 0, 0, 0, 0, 0};
    if (gProf.dStats != nullptr &&
        hipMemcpy(st, gProf.dStats, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) {
        ptSetError("pathTracerWavefrontProfileRead: copy failed");
        return false;
    }
    for (int k = 0; k < 4; ++k) {
        out[k] = gProf.ms[k];
        out[4 + k] = static_cast<double>(gProf.launches[k]);
    }
    for (int k = 0; k < 6; ++k) {
        out[8 + k] = static_cast<double>(st[k]);
    }
    if (n >= 15) {
        out[14] = static_cast<double>(gProf.paths);
    }
    return true;
}

// Per-bounce breakdown of the profiled renders: for bounce b < n, out[5 * b + 0..4] = ms of the
// extension trace, shadow trace and shade launches, extension rays, shadow rays.
extern "C" bool pathTracerWavefrontProfileBounces(double* out, int n)
{
    if (out == nullptr || n <= 0 || n > 258) {
        ptSetError("pathTracerWavefrontProfileBounces: invalid arguments");
        return false;
    }
    for (int b = 0; b < n; ++b) {
        out[5 * b + 0] = gProf.bounceMs[0][b];
        out[5 * b + 1] = gProf.bounceMs[1][b];
        out[5 * b + 2] = gProf.bounceMs[2][b];
        out[5 * b + 3] = static_cast<double>(gProf.bounceRays[0][b]);
        out[5 * b + 4] = static_cast<double>(gProf.bounceRays[1][b]);
    }
    return true;
}

extern "C" bool pathTracerWavefrontSchedule(int speculativePaths)
{
    if (speculativePaths < 0) {
        ptSetError("pathTracerWavefrontSchedule: invalid arguments");
        return false;
    }
    gSpeculativeBudget = speculativePaths;
    return true;
}

extern "C" bool pathTracerWavefrontCaptureRadiance(float* d_out, int samplesPerPixel)
{
    if (d_out != nullptr && samplesPerPixel <= 0) {
        ptSetError("pathTracerWavefrontCaptureRadiance: invalid arguments");
        return false;
    }
    gCapture = d_out;
    gCaptureSpp = d_out != nullptr ? samplesPerPixel : 0;
    return true;
}

extern "C" int pathTracerWavefrontPins(void)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    return gWsPins;
}

extern "C" bool pathTracerWavefrontPin(void)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    ++gWsPins;
    return true;
}

extern "C" bool pathTracerWavefrontUnpin(int count)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    if (count < 0 || count > gWsPins) {
        ptSetError("pathTracerWavefrontUnpin: %d pins requested, %d held", count, gWsPins);
        return false;
    }
    gWsPins -= count;
    return true;
}

extern "C" bool pathTracerWavefrontReserve(int maxPathsInFlight, int pixels, int maxPathDepth)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    if (maxPathsInFlight <= 0 || pixels <= 0 || maxPathDepth < 0 || maxPathDepth > 256) {
        ptSetError("pathTracerWavefrontReserve: invalid arguments");
        return false;
    }
    int device = 0, cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    return ensure_workspace(std::max(maxPathsInFlight, 4), pixels, maxPathDepth, cus * 8 * kWfBlock);
}

namespace {

// One scene (tab == null), or the batch whose expanded table is tab (tabCount scenes,
// batchPixels pixels in all; the single-scene arguments are then unused).
bool render_fixed_impl(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2, uint8_t* d_converged,
                       int width, int height, const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                       const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params, int samplesPerPixel,
                       int maxPathDepth, int maxPathsInFlight, unsigned int firstSeed, hipStream_t stream,
                       const BatchScene* tab = nullptr, int tabCount = 0, int batchPixels = 0, bool batchWide = false)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    const bool single = tab == nullptr;
    const bool wide = single ? pt_scene_is_wide(d_scene) : batchWide;
    if (d_buffer == nullptr || d_samples == nullptr || d_lumMean == nullptr || d_m2 == nullptr ||
        d_converged == nullptr || samplesPerPixel <= 0 || maxPathDepth < 0 || maxPathDepth > 256 ||
        (single && (d_camera == nullptr || d_scene == nullptr || d_params == nullptr || width <= 0 || height <= 0)) ||
        (!single && (tabCount <= 0 || batchPixels <= 0))) {
        ptSetError(single ? "pathTracerRenderFixedWavefront: invalid arguments"
                          : "pathTracerRenderFixedBatch: invalid arguments");
        return false;
    }
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(stream, &capture);
    const bool capturing = capture == hipStreamCaptureStatusActive;
    const int pixels = single ? width * height : batchPixels;
    const int target = maxPathsInFlight > 0 ? maxPathsInFlight : kDefaultPathsInFlight;
    int device = 0, cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    const int persistentBlocks = cus * 8;
    const bool spec = pt_spectral_enabled();
    const SpecModel sm = pt_spectral_model();
    const int spillThreads = persistentBlocks * kWfBlock;
    const int cap = std::max(pixels, std::min(target, samplesPerPixel * pixels));
    const bool fits = gWs.mem != nullptr && gWs.capacity >= cap && gWs.pixels >= pixels &&
                      gWs.depth >= maxPathDepth && gWs.spillThreads >= spillThreads;
    if (!fits) {
        if (capturing) {
            ptSetError("pathTracerRenderFixedWavefront: workspace too small for graph capture "
                       "(call pathTracerWavefrontReserve first)");
            return false;
        }
        if (!ensure_workspace(cap, pixels, maxPathDepth, spillThreads)) {
            return false;
        }
    }
    gCapturing = capturing;
    if (capturing) {
        ++gWsPins;   // the graph being captured keeps the workspace's addresses
    }
    Paths P = gWs.P;
    int* list = gWs.lists[0];
    int* count = gWs.listCount;
    const int counterInts = kCountersPerBounce * (maxPathDepth + 2);
    const int group = std::max(1, std::min(samplesPerPixel, std::min(gWs.capacity, target) / pixels));
    hipLaunchKernelGGL(wf_iota_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream, list, count, pixels);
    for (int passBase = 0; passBase < samplesPerPixel; passBase += group) {
        const int g = std::min(group, samplesPerPixel - passBase);
        hipLaunchKernelGGL(wf_zero_kernel, dim3(blocks_for(counterInts)), dim3(kWfBlock), 0, stream, P.counters,
                           counterInts);
        hipLaunchKernelGGL(wf_gen_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream, list, count, d_samples,
                           d_converged, d_camera, d_params, width, height, firstSeed, passBase, samplesPerPixel, g, P,
                           spec, -1, tab, tabCount);
        launch_bounces(stream, P, d_scene, d_params, d_env, maxPathDepth, static_cast<long long>(g) * pixels,
                       persistentBlocks, spec, sm, tab, wide);
        hipLaunchKernelGGL(wf_accumulate_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream, list, count, P,
                           reinterpret_cast<float4*>(d_buffer), d_samples, d_lumMean, d_m2, d_converged, d_params, tab,
                           tabCount);
    }
    gCapturing = false;
    if (hipGetLastError() != hipSuccess) {
        ptSetError(single ? "pathTracerRenderFixedWavefront: launch failed" : "pathTracerRenderFixedBatch: launch failed");
        return false;
    }
    return true;
}

} // namespace

extern "C" bool pathTracerRenderFixedWavefront(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                               uint8_t* d_converged, int width, int height,
                                               const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene,
                                               const EnvironmentMapGpu* d_env, const RenderParamsGpu* d_params,
                                               int samplesPerPixel, int maxPathDepth, int maxPathsInFlight,
                                               hipStream_t stream)
{
    return render_fixed_impl(d_buffer, d_samples, d_lumMean, d_m2, d_converged, width, height, d_camera, d_scene,
                             d_env, d_params, samplesPerPixel, maxPathDepth, maxPathsInFlight, 1u, stream);
}

// ---- adjoint on the wavefront schedule ------------------------------------------------------

namespace {

struct AdjointWs {
    void* mem = nullptr;
    int capacity = 0;
    int pixels = 0;
    size_t partialDoubles = 0;
    AdjointState A[2]{};
    uint32_t* counts = nullptr;  // per pixel: samples taken so far (the generation kernel's n)
    uint8_t* conv = nullptr;     // per pixel: 0
    double* partial = nullptr;   // one row of gradient sums per block of the shade grid
    // the primal per-sample radiance and the accumulators of the render that records it, when
    // the caller has none (grow-only: adjoint calls on a device do not overlap)
    void* tmp = nullptr;
    size_t tmpBytes = 0;
};
AdjointWs gAd;

bool ensure_adjoint_tmp(size_t bytes)
{
    if (gAd.tmp != nullptr && gAd.tmpBytes >= bytes) {
        return true;
    }
    if (gAd.tmp != nullptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(gAd.tmp);
        gAd.tmp = nullptr;
        gAd.tmpBytes = 0;
    }
    if (hipMalloc(&gAd.tmp, bytes) != hipSuccess) {
        gAd.tmp = nullptr;
        ptSetError("pathTracerAdjointRender: radiance buffer allocation of %zu bytes failed", bytes);
        return false;
    }
    gAd.tmpBytes = bytes;
    return true;
}

bool ensure_adjoint(int cap, int pixels, size_t partialDoubles)
{
    if (gAd.mem != nullptr && gAd.capacity >= cap && gAd.pixels >= pixels && gAd.partialDoubles >= partialDoubles) {
        return true;
    }
    cap = std::max(cap, gAd.capacity);
    pixels = std::max(pixels, gAd.pixels);
    partialDoubles = std::max(partialDoubles, gAd.partialDoubles);
    auto lay = [&](char* base) {
        char* cur = base;
        for (AdjointState& A : gAd.A) {
            A.gx = carve<float>(cur, cap);
            A.gy = carve<float>(cur, cap);
            A.gz = carve<float>(cur, cap);
            A.gw = carve<float>(cur, cap);
            for (int k = 0; k < kGradParams; ++k) {
                A.pe[k] = carve<float>(cur, cap);
                A.pm[k] = carve<float>(cur, cap);
            }
            A.pmi = carve<int>(cur, cap);
        }
        gAd.counts = carve<uint32_t>(cur, pixels);
        gAd.conv = carve<uint8_t>(cur, pixels);
        gAd.partial = carve<double>(cur, partialDoubles);
        return static_cast<size_t>(cur - base) + 256;
    };
    const size_t bytes = lay(nullptr);
    if (gAd.mem != nullptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(gAd.mem);
        gAd.mem = nullptr;
    }
    void* mem = nullptr;
    if (hipMalloc(&mem, bytes) != hipSuccess) {
        ptSetError("adjoint workspace allocation of %zu bytes failed", bytes);
        gAd.capacity = 0;
        gAd.pixels = 0;
        gAd.partialDoubles = 0;
        return false;
    }
    gAd.mem = mem;
    gAd.capacity = cap;
    gAd.pixels = pixels;
    gAd.partialDoubles = partialDoubles;
    lay(static_cast<char*>(mem));
    return true;
}

} // namespace

namespace {

// maxPathDepth as the integrator resolves it (PathIntegratorRandCore.h:219-227)
__host__ __device__ inline int resolve_depth(int maxPathDepth)
{
    return maxPathDepth > 0 ? maxPathDepth : (maxPathDepth == 0 ? 256 : 32);
}

// Renders the samples of seeds firstSeed.. and records their per-sample radiance (RGB, or the
// four hero wavelengths in spectral mode) into the adjoint's grow-only buffer (one scene, or the
// batch of tab); returns the radiance or null.
const float* adjoint_primal(int width, int height, int pixels, int spp, int maxDepth, const CameraGpu* d_camera,
                            const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                            const RenderParamsGpu* d_params, unsigned int firstSeed, hipStream_t stream,
                            const BatchScene* tab, int tabCount, bool batchWide)
{
    const size_t nRad = static_cast<size_t>(pixels) * spp * (pt_spectral_enabled() ? 4 : 3);
    const size_t bytes = nRad * sizeof(float) + static_cast<size_t>(pixels) * (16 + 4 + 4 + 4 + 1) + 6 * 256;
    if (!ensure_adjoint_tmp(bytes)) {
        return nullptr;
    }
    float* rad = static_cast<float*>(gAd.tmp);
    char* cur = reinterpret_cast<char*>(rad + nRad);
    PtFloat4* acc = carve<PtFloat4>(cur, pixels);
    uint32_t* cnt = carve<uint32_t>(cur, pixels);
    float* mean = carve<float>(cur, pixels);
    float* m2 = carve<float>(cur, pixels);
    uint8_t* cv = carve<uint8_t>(cur, pixels);
    float* savedCapture = gCapture;
    const int savedSpp = gCaptureSpp;
    gCapture = rad;
    gCaptureSpp = spp;
    const bool ok = hipMemsetAsync(cnt, 0, pixels * sizeof(uint32_t), stream) == hipSuccess &&
                    hipMemsetAsync(cv, 0, pixels, stream) == hipSuccess &&
                    hipMemsetAsync(mean, 0, pixels * sizeof(float), stream) == hipSuccess &&
                    hipMemsetAsync(m2, 0, pixels * sizeof(float), stream) == hipSuccess &&
                    hipMemsetAsync(acc, 0, pixels * sizeof(PtFloat4), stream) == hipSuccess &&
                    render_fixed_impl(acc, cnt, mean, m2, cv, width, height, d_camera, d_scene, d_env, d_params, spp,
                                      maxDepth, std::max(1 << 24, gWs.capacity), firstSeed, stream, tab, tabCount,
                                      pixels, batchWide);
    gCapture = savedCapture;
    gCaptureSpp = savedSpp;
    return ok ? rad : nullptr;
}

// The replay of every sample (seeds firstSeed..) against the primal radiance: gradient rows
// (7 doubles each; kSpecGradParams in spectral mode) summed into d_grad.
bool adjoint_replay(const float* d_gradImage, int width, int height, int pixels, int spp, int maxDepth,
                    const CameraGpu* d_camera, const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                    const RenderParamsGpu* d_params, int gradRows, const float* radiance, double* d_grad,
                    unsigned int firstSeed, hipStream_t stream, const BatchScene* tab, int tabCount, bool batchWide)
{
    const bool wide = tab != nullptr ? batchWide : pt_scene_is_wide(d_scene);
    int device = 0, cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    const int persistentBlocks = cus * 8;
    const int spillThreads = persistentBlocks * kWfBlock;
    // paths per group: 16M, or as many as a larger reserved workspace holds
    const int target = std::max(1 << 24, gWs.capacity);
    const int cap = std::max(pixels, std::min(target, spp * pixels));
    const bool spec = pt_spectral_enabled();
    const SpecModel sm = pt_spectral_model();
    const int perCopy = gradRows * (spec ? kSpecGradParams : kGradParams);
    if (!ensure_workspace(cap, pixels, maxDepth, spillThreads) ||
        !ensure_adjoint(cap, pixels, static_cast<size_t>(persistentBlocks) * perCopy)) {
        return false;
    }
    Paths P = gWs.P;
    int* list = gWs.lists[0];
    int* count = gWs.listCount;
    const int counterInts = kCountersPerBounce * (maxDepth + 2);
    const int group = std::max(1, std::min(spp, std::min(gWs.capacity, target) / pixels));
    // LDS copies of the per-row sums: as many as fit 24 KB (at least one)
    int copies = 16;
    while (copies > 1 && static_cast<size_t>(copies) * perCopy * sizeof(double) > 24576) {
        copies >>= 1;
    }
    const size_t lds = static_cast<size_t>(copies) * perCopy * sizeof(double);
    const float invSpp = 1.0f / static_cast<float>(spp);
    hipLaunchKernelGGL(wf_iota_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream, list, count, pixels);
    hipLaunchKernelGGL(wf_fill_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream,
                       reinterpret_cast<uint32_t*>(gAd.conv), (pixels + 3) / 4, 0u);
    hipLaunchKernelGGL(wf_zero_f64_kernel, dim3(blocks_for(static_cast<long long>(persistentBlocks) * perCopy)),
                       dim3(kWfBlock), 0, stream, gAd.partial, static_cast<long long>(persistentBlocks) * perCopy);
    const auto initKernel = spec ? wf_adjoint_init_kernel<HeroColors> : wf_adjoint_init_kernel<RgbColors>;
    const auto shadeKernel =
        tab != nullptr ? (spec ? wf_adjoint_shade_kernel<HeroColors, true> : wf_adjoint_shade_kernel<RgbColors, true>)
                       : (spec ? wf_adjoint_shade_kernel<HeroColors, false> : wf_adjoint_shade_kernel<RgbColors, false>);
    for (int passBase = 0; passBase < spp; passBase += group) {
        const int g = std::min(group, spp - passBase);
        hipLaunchKernelGGL(wf_fill_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream, gAd.counts, pixels,
                           static_cast<uint32_t>(passBase));
        hipLaunchKernelGGL(wf_zero_kernel, dim3(blocks_for(counterInts)), dim3(kWfBlock), 0, stream, P.counters,
                           counterInts);
        hipLaunchKernelGGL(wf_gen_kernel, dim3(blocks_for(pixels)), dim3(kWfBlock), 0, stream, list, count,
                           gAd.counts, gAd.conv, d_camera, d_params, width, height, firstSeed, passBase, spp, g, P,
                           spec, 0, tab, tabCount);
        hipLaunchKernelGGL(initKernel, dim3(blocks_for(static_cast<long long>(g) * pixels)), dim3(kWfBlock), 0,
                           stream, P.S[0], gAd.A[0], &P.counters[3], d_gradImage, radiance, spp, invSpp, tab, sm);
        launch_bounces_with(
            stream, P, d_scene, maxDepth, static_cast<long long>(g) * pixels, persistentBlocks,
            [&](int b, const int* shadeCount, int* cOut, unsigned traceBlocks) {
                hipLaunchKernelGGL(shadeKernel, dim3(traceBlocks), dim3(kWfBlock), lds, stream, d_scene, d_params,
                                   d_env, P.S[b & 1], P.S[(b + 1) & 1], gAd.A[b & 1], gAd.A[(b + 1) & 1], P, b,
                                   maxDepth, P.hitA[b & 1], P.hitB[b & 1], shadeCount, P.extQ[(b + 1) & 1],
                                   P.hitA[(b + 1) & 1], cOut, gradRows, copies, gAd.partial, tab, sm);
            },
            tab, wide, kLightReference);
    }
    hipLaunchKernelGGL(wf_adjoint_reduce_kernel, dim3(static_cast<unsigned>(perCopy)), dim3(kWfBlock), 0, stream,
                       gAd.partial, persistentBlocks, perCopy, d_grad);
    if (hipGetLastError() != hipSuccess) {
        ptSetError("pathTracerAdjointRender: launch failed");
        return false;
    }
    return true;
}

} // namespace

bool pt_adjoint_wavefront(const float* d_gradImage, int width, int height, int spp, const CameraGpu* d_camera,
                          const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                          const RenderParamsGpu* d_params, int materialCount, const float* d_sampleRadiance,
                          double* d_grad, unsigned int firstSeed, hipStream_t stream)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    RenderParamsGpu hp;
    if (hipMemcpyAsync(&hp, d_params, sizeof(hp), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        ptSetError("pathTracerAdjointRender: reading the render parameters failed");
        return false;
    }
    const int maxDepth = resolve_depth(hp.maxPathDepth);
    const int pixels = width * height;
    // the primal per-sample radiance: the caller's (stored by the forward) or a fixed-spp render
    const float* radiance = d_sampleRadiance;
    if (radiance == nullptr) {
        radiance = adjoint_primal(width, height, pixels, spp, maxDepth, d_camera, d_scene, d_env, d_params, firstSeed,
                                  stream, nullptr, 0, false);
        if (radiance == nullptr) {
            return false;
        }
    }
    return adjoint_replay(d_gradImage, width, height, pixels, spp, maxDepth, d_camera, d_scene, d_env, d_params,
                          materialCount, radiance, d_grad, firstSeed, stream, nullptr, 0, false);
}

// ---- batches of scenes --------------------------------------------------------------------

namespace {

// The caller's pointers for one scene plus its place in the batch (device copy of the table).
struct BatchRef {
    const MeshAccelSceneGpu* scene;
    const CameraGpu* camera;
    const RenderParamsGpu* params;
    const EnvironmentMapGpu* env;
    int width, height, pixelOffset, gradRow, maxDepth;
};

// Copies every scene's description, parameters, camera and environment into the batch table, so
// the render reads what the caller's device records hold at this call.
__global__ void wf_batch_expand_kernel(const BatchRef* __restrict__ refs, int count, BatchScene* __restrict__ tab)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= count) {
        return;
    }
    const BatchRef r = refs[i];
    BatchScene b;
    b.scene = *r.scene;
    b.params = *r.params;
    b.camera = *r.camera;
    if (r.env != nullptr) {
        b.env = *r.env;
    } else {
        b.env.pixels = nullptr;
        b.env.marginalCdf = nullptr;
        b.env.rowCdf = nullptr;
        b.env.width = 0;
        b.env.height = 0;
        b.env.valid = 0;
    }
    b.width = r.width;
    b.height = r.height;
    b.pixelOffset = r.pixelOffset;
    b.gradRow = r.gradRow;
    b.maxDepth = min(resolve_depth(b.params.maxPathDepth), r.maxDepth);
    tab[i] = b;
}

} // namespace

struct PtBatch {
    bool wide = true;  // every scene has a 4-wide tree
    int count = 0;
    int totalPixels = 0;
    int gradRows = 0;
    int maxDepth = 0;
    std::vector<int> pixelOffsets;
    std::vector<int> gradRowOffsets;
    BatchRef* refs = nullptr;  // device
    BatchScene* tab = nullptr; // device, filled at every call
};

namespace {

void batch_free(PtBatch* b)
{
    if (b->refs != nullptr) {
        (void)hipFree(b->refs);
    }
    if (b->tab != nullptr) {
        (void)hipFree(b->tab);
    }
    delete b;
}

bool batch_expand(const PtBatch* b, hipStream_t stream)
{
    hipLaunchKernelGGL(wf_batch_expand_kernel, dim3(static_cast<unsigned>((b->count + kWfBlock - 1) / kWfBlock)),
                       dim3(kWfBlock), 0, stream, b->refs, b->count, b->tab);
    return hipGetLastError() == hipSuccess;
}

} // namespace

extern "C" PtBatch* ptBatchCreate(const PtBatchSceneGpu* scenes, int sceneCount)
{
    if (scenes == nullptr || sceneCount <= 0 || sceneCount > kMaxBatchScenes) {
        ptSetError("ptBatchCreate: invalid arguments (1..%d scenes)", kMaxBatchScenes);
        return nullptr;
    }
    std::vector<BatchRef> refs(static_cast<size_t>(sceneCount));
    auto* b = new PtBatch();
    b->count = sceneCount;
    long long pixels = 0;
    long long rows = 0;
    for (int i = 0; i < sceneCount; ++i) {
        const PtBatchSceneGpu& e = scenes[i];
        if (e.scene == nullptr || e.camera == nullptr || e.params == nullptr || e.width <= 0 || e.height <= 0) {
            ptSetError("ptBatchCreate: scene %d: invalid arguments", i);
            batch_free(b);
            return nullptr;
        }
        MeshAccelSceneGpu desc;
        RenderParamsGpu prm;
        if (hipMemcpy(&desc, e.scene, sizeof(desc), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&prm, e.params, sizeof(prm), hipMemcpyDeviceToHost) != hipSuccess) {
            ptSetError("ptBatchCreate: scene %d: reading its description or parameters failed", i);
            batch_free(b);
            return nullptr;
        }
        const int depth = resolve_depth(prm.maxPathDepth);
        b->maxDepth = std::max(b->maxDepth, depth);
        b->wide = b->wide && desc.wideNodes != nullptr && pt_scene_is_wide(e.scene);
        refs[static_cast<size_t>(i)] = BatchRef{e.scene, e.camera, e.params, e.env, e.width, e.height,
                                               static_cast<int>(pixels), static_cast<int>(rows), 0};
        b->pixelOffsets.push_back(static_cast<int>(pixels));
        b->gradRowOffsets.push_back(static_cast<int>(rows));
        pixels += static_cast<long long>(e.width) * e.height;
        rows += desc.materialCount;
        if (pixels > (1ll << 30)) {
            ptSetError("ptBatchCreate: more than 2^30 pixels in one batch");
            batch_free(b);
            return nullptr;
        }
    }
    if (b->maxDepth > 256) {
        ptSetError("ptBatchCreate: path depth above 256");
        batch_free(b);
        return nullptr;
    }
    for (BatchRef& r : refs) {
        r.maxDepth = b->maxDepth;
    }
    b->totalPixels = static_cast<int>(pixels);
    b->gradRows = static_cast<int>(rows);
    if (hipMalloc(&b->refs, refs.size() * sizeof(BatchRef)) != hipSuccess ||
        hipMalloc(&b->tab, refs.size() * sizeof(BatchScene)) != hipSuccess ||
        hipMemcpy(b->refs, refs.data(), refs.size() * sizeof(BatchRef), hipMemcpyHostToDevice) != hipSuccess) {
        ptSetError("ptBatchCreate: device table allocation failed");
        batch_free(b);
        return nullptr;
    }
    return b;
}

extern "C" void ptBatchDestroy(PtBatch* batch)
{
    if (batch != nullptr) {
        (void)hipDeviceSynchronize();
        batch_free(batch);
    }
}

extern "C" bool ptBatchInfo(const PtBatch* batch, int* totalPixels, int* gradRows, int* maxPathDepth,
                            int* pixelOffsets, int* gradRowOffsets)
{
    if (batch == nullptr) {
        ptSetError("ptBatchInfo: no batch");
        return false;
    }
    if (totalPixels != nullptr) {
        *totalPixels = batch->totalPixels;
    }
    if (gradRows != nullptr) {
        *gradRows = batch->gradRows;
    }
    if (maxPathDepth != nullptr) {
        *maxPathDepth = batch->maxDepth;
    }
    for (int i = 0; i < batch->count; ++i) {
        if (pixelOffsets != nullptr) {
            pixelOffsets[i] = batch->pixelOffsets[static_cast<size_t>(i)];
        }
        if (gradRowOffsets != nullptr) {
            gradRowOffsets[i] = batch->gradRowOffsets[static_cast<size_t>(i)];
        }
    }
    return true;
}

extern "C" bool pathTracerRenderFixedBatch(PtFloat4* d_buffer, uint32_t* d_samples, float* d_lumMean, float* d_m2,
                                           uint8_t* d_converged, const PtBatch* batch, int samplesPerPixel,
                                           int maxPathsInFlight, unsigned int firstSeed, hipStream_t stream)
{
    if (batch == nullptr) {
        ptSetError("pathTracerRenderFixedBatch: no batch");
        return false;
    }
    if (!batch_expand(batch, stream)) {
        ptSetError("pathTracerRenderFixedBatch: launch failed");
        return false;
    }
    return render_fixed_impl(d_buffer, d_samples, d_lumMean, d_m2, d_converged, 0, 0, nullptr, nullptr, nullptr,
                             nullptr, samplesPerPixel, batch->maxDepth, maxPathsInFlight, firstSeed, stream,
                             batch->tab, batch->count, batch->totalPixels, batch->wide);
}

extern "C" bool pathTracerAdjointRenderBatch(const float* d_gradImage, const PtBatch* batch, int samplesPerPixel,
                                             unsigned int firstSeed, double* d_grad, hipStream_t stream)
{
    std::lock_guard<std::recursive_mutex> wsLock(gWsLock);
    if (d_gradImage == nullptr || batch == nullptr || d_grad == nullptr || samplesPerPixel <= 0) {
        ptSetError("pathTracerAdjointRenderBatch: invalid arguments");
        return false;
    }
    // hero-wavelength mode: 36 parameters per row, so fewer rows fit the LDS sums (64 KB)
    const int maxRows = pt_spectral_enabled() ? kMaxBatchSpectralGradRows : kMaxBatchGradRows;
    if (batch->gradRows <= 0 || batch->gradRows > maxRows) {
        ptSetError("pathTracerAdjointRenderBatch: %d material rows (1..%d)", batch->gradRows, maxRows);
        return false;
    }
    if (pt_light_mode() != 0) {
        ptSetError("pathTracerAdjointRenderBatch: the reference light sampling only (pathTracerSetLightSampling(0))");
        return false;
    }
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(stream, &capture);
    if (capture == hipStreamCaptureStatusActive) {
        ptSetError("pathTracerAdjointRenderBatch: not capturable");
        return false;
    }
    if (!batch_expand(batch, stream)) {
        ptSetError("pathTracerAdjointRenderBatch: launch failed");
        return false;
    }
    const float* radiance = adjoint_primal(0, 0, batch->totalPixels, samplesPerPixel, batch->maxDepth, nullptr,
                                           nullptr, nullptr, nullptr, firstSeed, stream, batch->tab, batch->count,
                                           batch->wide);
    if (radiance == nullptr) {
        return false;
    }
    return adjoint_replay(d_gradImage, 0, 0, batch->totalPixels, samplesPerPixel, batch->maxDepth, nullptr, nullptr,
                          nullptr, nullptr, batch->gradRows, radiance, d_grad, firstSeed, stream, batch->tab,
                          batch->count, batch->wide);
}
