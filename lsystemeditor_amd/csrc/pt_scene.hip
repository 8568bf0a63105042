// pt_scene.hip -- scene build on the GPU: median-split BVH (MeshAccel/MeshBvhBuilder.cpp:52-142),
// device layout, emissive list (MeshAccel/MeshAccelScene.cpp:62-95) and upload
// (MeshAccelScene.cpp:141-341).
//
// BVH build. The reference splits each node at count/2 along the longest axis of the node's
// triangle bounds, using nth_element on triangle centroids (leaf = one triangle, nodes in
// pre-order, left subtree first). Because every split is at count/2, the tree's topology is
// a function of N alone: the segment of positions covered by a node, its pre-order index and
// its depth follow from N by a walk from the root. Only which triangles fall on which side is
// data dependent. The build therefore runs level by level over all nodes of a depth at once:
//   1. triangles are sorted once per axis by (centroid, index) -- three lists;
//   2. per level: segment bounds by wave-aggregated min/max atomics, one thread per segment
//      picks the axis and writes the node records, the left half of the chosen axis' list
//      marks its triangles, and all three lists are stable-partitioned by that mark within
//      their segments (a scan over the three lists together).
// Stable partitioning keeps every list sorted within every segment, so the left set of a
// split is exactly the count/2 smallest (centroid, index) pairs -- the nth_element partition
// with ties broken by triangle index.
#include "pt_device.h"
#include "pt_internal.h"

#include "../../include/pt_hip.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_set>
#include <vector>

namespace {

constexpr int kB = 256;
constexpr int kWave = 64;
constexpr int kBoundsChunk = 8;  // positions per block of bounds_kernel: kBoundsChunk x kB

__device__ __forceinline__ uint32_t to_sortable(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float from_sortable(uint32_t u)
{
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(b);
}

__device__ __forceinline__ float bmin2(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float bmax2(float a, float b) { return a > b ? a : b; }

struct Seg {
    int begin, count, node, j, parentInternal;
    bool isLeft, exists;
};

// Segment containing position `pos` at depth d (exists == false if it became a leaf earlier).
__device__ __forceinline__ Seg walk_pos(int pos, int d, int n)
{
    Seg s{0, n, 0, 0, -1, false, true};
    for (int step = 0; step < d; ++step) {
        if (s.count == 1) {
            s.exists = false;
            return s;
        }
        const int mid = s.count / 2;
        s.parentInternal = s.node - s.begin;
        if (pos < s.begin + mid) {
            s.node += 1;
            s.count = mid;
            s.j = 2 * s.j;
            s.isLeft = true;
        } else {
            s.node += 2 * mid;
            s.begin += mid;
            s.count -= mid;
            s.j = 2 * s.j + 1;
            s.isLeft = false;
        }
    }
    return s;
}

// Segment number j (left to right) at depth d.
__device__ __forceinline__ Seg walk_seg(int j, int d, int n)
{
    Seg s{0, n, 0, j, -1, false, true};
    for (int step = 0; step < d; ++step) {
        if (s.count == 1) {
            s.exists = false;
            return s;
        }
        const int mid = s.count / 2;
        s.parentInternal = s.node - s.begin;
        if (((j >> (d - 1 - step)) & 1) == 0) {
            s.node += 1;
            s.count = mid;
            s.isLeft = true;
        } else {
            s.node += 2 * mid;
            s.begin += mid;
            s.count -= mid;
            s.isLeft = false;
        }
    }
    return s;
}

// HostTriangle is 25 packed 32-bit words: v0 v1 v2 n0 n1 n2 (3 each), uv0 uv1 uv2 (2 each), material.
__global__ void prep_kernel(const float* __restrict__ tris, int n, uint32_t* __restrict__ keys,
                            int* __restrict__ values, float4* __restrict__ bounds, float4* __restrict__ shade)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const float* t = tris + 25 * static_cast<size_t>(i);
    const float v0x = t[0], v0y = t[1], v0z = t[2];
    const float v1x = t[3], v1y = t[4], v1z = t[5];
    const float v2x = t[6], v2y = t[7], v2z = t[8];
    // MeshBvhBuilder.cpp:10-13: (v0 + v1 + v2) * (1/3)
    const float third = 1.0f / 3.0f;
    const float cx = ((v0x + v1x) + v2x) * third;
    const float cy = ((v0y + v1y) + v2y) * third;
    const float cz = ((v0z + v1z) + v2z) * third;
    // sort keys: +0 and -0 compare equal, so -0 is folded into +0 before the bit transform
    keys[i] = to_sortable(cx + 0.0f);
    keys[n + i] = to_sortable(cy + 0.0f);
    keys[2 * n + i] = to_sortable(cz + 0.0f);
    values[i] = i;
    values[n + i] = i;
    values[2 * n + i] = i;
    // MeshBvhBuilder.cpp:15-29
    bounds[2 * i] =
        make_float4(bmin2(bmin2(v0x, v1x), v2x), bmin2(bmin2(v0y, v1y), v2y), bmin2(bmin2(v0z, v1z), v2z), 0.0f);
    bounds[2 * i + 1] =
        make_float4(bmax2(bmax2(v0x, v1x), v2x), bmax2(bmax2(v0y, v1y), v2y), bmax2(bmax2(v0z, v1z), v2z), 0.0f);
    // shading record: (n0.xyz, n1.x) (n1.yz, n2.xy) (n2.z, uv0.xy, uv1.x) (uv1.y, uv2.xy, material)
    shade[4 * i + 0] = make_float4(t[9], t[10], t[11], t[12]);
    shade[4 * i + 1] = make_float4(t[13], t[14], t[15], t[16]);
    shade[4 * i + 2] = make_float4(t[17], t[18], t[19], t[20]);
    shade[4 * i + 3] = make_float4(t[21], t[22], t[23], t[24]);
}

__global__ void fill_bounds_kernel(uint32_t* segBounds, int segs)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= segs) {
        return;
    }
    for (int k = 0; k < 3; ++k) {
        segBounds[6 * i + k] = 0xFFFFFFFFu;
        segBounds[6 * i + 3 + k] = 0u;
    }
}

// Segment bounds at depth d. Min/max are taken in the sortable-integer order, in which -0 < +0:
// the union is then independent of the visiting order (the reference's vecMin2/vecMax2 loop
// keeps whichever signed zero it met first; both compare equal in every traversal test).
// Segments at depth d are contiguous position ranges, so one segment's lanes within a wave form
// one run: a segmented inclusive scan (shuffles) leaves each run's union in the run's last lane,
// which alone updates the block's LDS copy (blocks inside one segment, flushed to global once per
// block) or the segment's record. A block covers kBoundsChunk x kB consecutive positions.
__global__ __launch_bounds__(kB) void bounds_kernel(const int* __restrict__ list0, const float4* __restrict__ bounds,
                                                    int n, int d, uint32_t* __restrict__ segBounds)
{
    __shared__ uint32_t lds[6];
    const int base = static_cast<int>(blockIdx.x) * kB * kBoundsChunk;
    const int last = min(base + kB * kBoundsChunk, n) - 1;
    const Seg first = walk_pos(base, d, n);
    const int j0 = first.exists && last < first.begin + first.count ? first.j : -1;
    const bool uniform = j0 >= 0;
    if (uniform && threadIdx.x < 6) {
        lds[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
    }
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    for (int it = 0; it < kBoundsChunk; ++it) {
        const int i = base + it * kB + static_cast<int>(threadIdx.x);
        int j = -1;
        uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
        if (i < n) {
            const Seg s = uniform ? first : walk_pos(i, d, n);
            if (s.exists) {
                j = s.j;
                const int tri = list0[i];
                const float4 lo = bounds[2 * tri];
                const float4 hi = bounds[2 * tri + 1];
                v[0] = to_sortable(lo.x);
                v[1] = to_sortable(lo.y);
                v[2] = to_sortable(lo.z);
                v[3] = to_sortable(hi.x);
                v[4] = to_sortable(hi.y);
                v[5] = to_sortable(hi.z);
            }
        }
        for (int dd = 1; dd < kWave; dd <<= 1) {
            const int jo = __shfl_up(j, dd, kWave);
            const bool take = lane >= dd && jo == j;
            for (int k = 0; k < 6; ++k) {
                const uint32_t o = __shfl_up(v[k], dd, kWave);
                if (take) {
                    v[k] = k < 3 ? min(v[k], o) : max(v[k], o);
                }
            }
        }
        const int jn = __shfl_down(j, 1, kWave);
        if (j >= 0 && (lane == kWave - 1 || jn != j)) {
            uint32_t* dst = uniform ? lds : segBounds + 6 * j;
            for (int k = 0; k < 3; ++k) {
                atomicMin(&dst[k], v[k]);
                atomicMax(&dst[3 + k], v[3 + k]);
            }
        }
    }
    __syncthreads();
    if (uniform && threadIdx.x < 6) {
        const int k = static_cast<int>(threadIdx.x);
        if (k < 3) {
            atomicMin(&segBounds[6 * j0 + k], lds[k]);
        } else {
            atomicMax(&segBounds[6 * j0 + k], lds[k]);
        }
    }
}

// One thread per segment at depth d: node records in both layouts, split axis, leaf triangles.
__global__ void seg_kernel(const uint32_t* __restrict__ segBounds, int d, int n, int segs,
                           const int* __restrict__ list0, const float* __restrict__ tris, MeshBvhNode* refNodes,
                           float* nodes, float4* triPos, unsigned char* segAxis, float* rootBox)
{
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= segs) {
        return;
    }
    const Seg s = walk_seg(j, d, n);
    if (!s.exists) {
        return;
    }
    const float bx0 = from_sortable(segBounds[6 * j + 0]);
    const float by0 = from_sortable(segBounds[6 * j + 1]);
    const float bz0 = from_sortable(segBounds[6 * j + 2]);
    const float bx1 = from_sortable(segBounds[6 * j + 3]);
    const float by1 = from_sortable(segBounds[6 * j + 4]);
    const float bz1 = from_sortable(segBounds[6 * j + 5]);
    MeshBvhNode rec;
    rec.boundsMin.x = bx0;
    rec.boundsMin.y = by0;
    rec.boundsMin.z = bz0;
    rec.pad0 = 0.0f;
    rec.boundsMax.x = bx1;
    rec.boundsMax.y = by1;
    rec.boundsMax.z = bz1;
    rec.pad1 = 0.0f;
    int ref;
    if (s.count == 1) {
        const int tri = list0[s.begin];
        rec.leftIndex = 0u;
        rec.rightIndex = 0u;
        rec.triStart = static_cast<uint32_t>(tri);
        rec.triCount = 1u;
        const float* t = tris + 25 * static_cast<size_t>(tri);
        // (v0, original index) (v1 - v0, material index) (v2 - v0): the edges the reference
        // computes per test; the material index rides along so a hit record can carry it
        triPos[3 * s.begin + 0] = make_float4(t[0], t[1], t[2], __int_as_float(tri));
        triPos[3 * s.begin + 1] = make_float4(t[3] - t[0], t[4] - t[1], t[5] - t[2], t[24]); // .w: material
        triPos[3 * s.begin + 2] = make_float4(t[6] - t[0], t[7] - t[1], t[8] - t[2], 0.0f);
        ref = ~s.begin;
    } else {
        const int mid = s.count / 2;
        rec.leftIndex = static_cast<uint32_t>(s.node + 1);
        rec.rightIndex = static_cast<uint32_t>(s.node + 2 * mid);
        rec.triStart = 0u;
        rec.triCount = 0u;
        // MeshBvhBuilder.cpp:31-40
        const float ex = bx1 - bx0, ey = by1 - by0, ez = bz1 - bz0;
        int axis = 2;
        if (ex >= ey && ex >= ez) {
            axis = 0;
        } else if (ey >= ez) {
            axis = 1;
        }
        segAxis[j] = static_cast<unsigned char>(axis);
        ref = s.node - s.begin;
    }
    refNodes[s.node] = rec;
    if (d == 0) {
        rootBox[0] = bx0;
        rootBox[1] = by0;
        rootBox[2] = bz0;
        rootBox[3] = bx1;
        rootBox[4] = by1;
        rootBox[5] = bz1;
        rootBox[6] = __int_as_float(ref);
    } else {
        float* p = nodes + 16 * static_cast<size_t>(s.parentInternal);
        const int o = s.isLeft ? 0 : 6;
        p[o + 0] = bx0;
        p[o + 1] = by0;
        p[o + 2] = bz0;
        p[o + 3] = bx1;
        p[o + 4] = by1;
        p[o + 5] = bz1;
        p[s.isLeft ? 12 : 13] = __int_as_float(ref);
    }
}

// Marks the left half (count/2 smallest on the split axis) of every splitting segment.
__global__ void flags_kernel(const int* __restrict__ lists, const unsigned char* __restrict__ segAxis, int n, int d,
                             unsigned char* __restrict__ flag)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const Seg s = walk_pos(i, d, n);
    if (!s.exists || s.count == 1) {
        return;
    }
    const int axis = segAxis[s.j];
    const int tri = lists[static_cast<size_t>(axis) * n + i];
    flag[tri] = (i - s.begin < s.count / 2) ? 1u : 0u;
}

__global__ void gather_kernel(const int* __restrict__ lists, const unsigned char* __restrict__ flag, int n, int d,
                              int* __restrict__ f)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const Seg s = walk_pos(i, d, n);
    const bool split = s.exists && s.count > 1;
    for (int b = 0; b < 3; ++b) {
        f[static_cast<size_t>(b) * n + i] = split ? flag[lists[static_cast<size_t>(b) * n + i]] : 0;
    }
}

__global__ void scatter_kernel(const int* __restrict__ lists, const unsigned char* __restrict__ flag,
                               const int* __restrict__ scan, int n, int d, int* __restrict__ out)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const Seg s = walk_pos(i, d, n);
    const bool split = s.exists && s.count > 1;
    const int mid = s.count / 2;
    for (int b = 0; b < 3; ++b) {
        const size_t off = static_cast<size_t>(b) * n;
        const int tri = lists[off + i];
        int pos = i;
        if (split) {
            const int rank = scan[off + i] - scan[off + s.begin];
            pos = flag[tri] ? s.begin + rank : s.begin + mid + (i - s.begin - rank);
        }
        out[off + pos] = tri;
    }
}

// Device-side deep copy helper.
template <typename T>
bool upload(T** dst, const T* src, size_t count, hipStream_t stream)
{
    *dst = nullptr;
    if (count == 0) {
        return true;
    }
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
        ptSetError("device allocation of %zu bytes failed", count * sizeof(T));
        return false;
    }
    *dst = static_cast<T*>(p);
    if (src != nullptr && hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, stream) != hipSuccess) {
        ptSetError("host to device copy failed");
        return false;
    }
    return true;
}

template <typename T>
bool alloc(T** dst, size_t count)
{
    *dst = nullptr;
    if (count == 0) {
        return true;
    }
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
        ptSetError("device allocation of %zu bytes failed", count * sizeof(T));
        return false;
    }
    *dst = static_cast<T*>(p);
    return true;
}

void release(void* p)
{
    if (p != nullptr) {
        (void)hipFree(p);
    }
}

unsigned g1(long long n) { return static_cast<unsigned>((n + kB - 1) / kB); }

// MeshAccelScene.cpp:27-46
MaterialGpu default_material()
{
    MaterialGpu m;
    std::memset(&m, 0, sizeof(m));
    m.r = 0.8f;
    m.g = 0.8f;
    m.b = 0.8f;
    m.roughness = 0.5f;
    m.ior = 1.5f;
    m.diffuseRoughness = -1.0f;
    m.specular = 1.0f;
    return m;
}

// Scene descriptions (device addresses) that carry a 4-wide tree: the launchers run the wide
// trace kernels for these.
std::mutex gWideMutex;
std::unordered_set<const MeshAccelSceneGpu*> gWideScenes;

void wide_registry(const MeshAccelSceneGpu* d, bool add)
{
    if (d == nullptr) {
        return;
    }
    std::lock_guard<std::mutex> lock(gWideMutex);
    if (add) {
        gWideScenes.insert(d);
    } else {
        gWideScenes.erase(d);
    }
}

} // namespace

bool pt_scene_is_wide(const MeshAccelSceneGpu* d)
{
    std::lock_guard<std::mutex> lock(gWideMutex);
    return gWideScenes.count(d) != 0;
}

struct PtScene {
    MeshAccelSceneGpu desc{};
    MeshAccelSceneGpu* d_desc = nullptr;
    float4* nodes = nullptr;
    float4* triPos = nullptr;
    float4* triShade = nullptr;
    MaterialGpu* materials = nullptr;
    uint32_t* emIdx = nullptr;
    float* emCdf = nullptr;
    float4* emTris = nullptr;
    TextureDescGpu* textures = nullptr;
    MeshBvhNode* refNodes = nullptr;
    uint32_t refCount = 0;
    float buildMs = 0.0f;
    int buildMode = 0; // 0: median split (reference topology), 1: binned SAH + 4-wide tree, 2: binned SAH
    float4* wide = nullptr; // 4-wide tree (mode 1)
    int wideCount = 0;
    int levels = 0;
    float* srcTris = nullptr; // the triangles (25-word records) the SAH modes keep for ptSceneRebuild
    int triangleCount = 0;
    int builtMode = -1;        // the build mode of the last ptSceneBuild (what the device layout holds)
    int* rebuildStatus = nullptr; // this scene's copy of its last rebuild's status (pt_sah_snapshot)
    // host copies for ptSceneUpdateMaterials: materials, and per emissive entry its area and
    // material index
    std::vector<MaterialGpu> matsHost;
    std::vector<float> emAreaHost;
    std::vector<uint32_t> emMatHost;
    std::vector<float4> emTrisHost;
    // binned spectra (ptSceneSetSpectra); kept across builds that keep the material count
    float* spectra = nullptr;
    uint32_t spectraCount = 0;

    void clear()
    {
        wide_registry(d_desc, false);
        release(d_desc);
        release(wide);
        wide = nullptr;
        wideCount = 0;
        release(nodes);
        release(triPos);
        release(triShade);
        release(materials);
        release(emIdx);
        release(emCdf);
        release(emTris);
        release(textures);
        release(refNodes);
        release(srcTris);
        srcTris = nullptr;
        release(rebuildStatus);
        rebuildStatus = nullptr;
        builtMode = -1;
        triangleCount = 0;
        d_desc = nullptr;
        nodes = nullptr;
        triPos = nullptr;
        triShade = nullptr;
        materials = nullptr;
        emIdx = nullptr;
        emCdf = nullptr;
        emTris = nullptr;
        textures = nullptr;
        refNodes = nullptr;
        refCount = 0;
        std::memset(&desc, 0, sizeof(desc));
    }
};

namespace {

// Runs the level-by-level build over device triangles (25-word records).
bool build_bvh(PtScene* sc, const float* dTris, int n, hipStream_t stream)
{
    uint32_t* keys = nullptr;
    uint32_t* keysOut = nullptr;
    int* lists = nullptr;
    int* listsAlt = nullptr;
    int* values = nullptr;
    float4* bounds = nullptr;
    uint32_t* segBounds = nullptr;
    unsigned char* segAxis = nullptr;
    unsigned char* flag = nullptr;
    int* f = nullptr;
    int* scan = nullptr;
    float* rootBox = nullptr;
    void* tmp = nullptr;
    bool ok = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    int depth = 0;
    while ((1ll << depth) < n) {
        ++depth;
    }
    const int maxSegs = 1 << depth;
    const size_t n3 = 3 * static_cast<size_t>(n);

    do {
        if (!alloc(&keys, n3) || !alloc(&keysOut, n3) || !alloc(&lists, n3) || !alloc(&listsAlt, n3) ||
            !alloc(&values, n3) || !alloc(&bounds, 2 * static_cast<size_t>(n)) ||
            !alloc(&segBounds, 6 * static_cast<size_t>(maxSegs)) || !alloc(&segAxis, static_cast<size_t>(maxSegs)) ||
            !alloc(&flag, static_cast<size_t>(n)) || !alloc(&f, n3) || !alloc(&scan, n3) || !alloc(&rootBox, 8)) {
            break;
        }
        size_t sortBytes = 0, scanBytes = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, keys, keysOut, values, lists, n, 0, 32, stream);
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, f, scan, static_cast<int>(n3), stream);
        const size_t tmpBytes = std::max(sortBytes, scanBytes);
        if (hipMalloc(&tmp, tmpBytes) != hipSuccess) {
            ptSetError("BVH build: temporary allocation failed");
            break;
        }
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, stream);

        hipLaunchKernelGGL(prep_kernel, dim3(g1(n)), dim3(kB), 0, stream, dTris, n, keys, values, bounds,
                           sc->triShade);
        bool sortOk = true;
        for (int a = 0; a < 3 && sortOk; ++a) {
            size_t bytes = tmpBytes;
            sortOk = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys + static_cast<size_t>(a) * n,
                                                        keysOut + static_cast<size_t>(a) * n,
                                                        values + static_cast<size_t>(a) * n,
                                                        lists + static_cast<size_t>(a) * n, n, 0, 32,
                                                        stream) == hipSuccess;
        }
        if (!sortOk) {
            ptSetError("BVH build: radix sort failed");
            break;
        }
        bool levelOk = true;
        for (int d = 0; d <= depth && levelOk; ++d) {
            const int segs = 1 << d;
            hipLaunchKernelGGL(fill_bounds_kernel, dim3(g1(segs)), dim3(kB), 0, stream, segBounds, segs);
            hipLaunchKernelGGL(bounds_kernel, dim3(g1((n + kBoundsChunk - 1) / kBoundsChunk)), dim3(kB), 0, stream, lists,
                               bounds, n, d, segBounds);
            hipLaunchKernelGGL(seg_kernel, dim3(g1(segs)), dim3(kB), 0, stream, segBounds, d, n, segs, lists, dTris,
                               sc->refNodes, reinterpret_cast<float*>(sc->nodes), sc->triPos, segAxis, rootBox);
            if (d == depth) {
                break;
            }
            hipLaunchKernelGGL(flags_kernel, dim3(g1(n)), dim3(kB), 0, stream, lists, segAxis, n, d, flag);
            hipLaunchKernelGGL(gather_kernel, dim3(g1(n)), dim3(kB), 0, stream, lists, flag, n, d, f);
            size_t bytes = tmpBytes;
            levelOk = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, f, scan, static_cast<int>(n3), stream) ==
                      hipSuccess;
            hipLaunchKernelGGL(scatter_kernel, dim3(g1(n)), dim3(kB), 0, stream, lists, flag, scan, n, d, listsAlt);
            std::swap(lists, listsAlt);
        }
        if (!levelOk || hipGetLastError() != hipSuccess) {
            ptSetError("BVH build: kernel launch failed");
            break;
        }
        (void)hipEventRecord(e1, stream);
        float box[8] = {0};
        if (hipMemcpyAsync(box, rootBox, 7 * sizeof(float), hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            ptSetError("BVH build: synchronisation failed");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        sc->buildMs = ms;
        for (int k = 0; k < 3; ++k) {
            sc->desc.rootMin[k] = box[k];
            sc->desc.rootMax[k] = box[3 + k];
        }
        int rootRef = 0;
        std::memcpy(&rootRef, &box[6], sizeof(int));
        sc->desc.rootRef = rootRef;
        ok = true;
    } while (false);

    if (e0 != nullptr) {
        (void)hipEventDestroy(e0);
    }
    if (e1 != nullptr) {
        (void)hipEventDestroy(e1);
    }
    release(keys);
    release(keysOut);
    release(lists);
    release(listsAlt);
    release(values);
    release(bounds);
    release(segBounds);
    release(segAxis);
    release(flag);
    release(f);
    release(scan);
    release(rootBox);
    release(tmp);
    return ok;
}

} // namespace

extern "C" {

PtScene* ptSceneCreate(void) { return new PtScene(); }

void ptSceneDestroy(PtScene* scene)
{
    if (scene != nullptr) {
        (void)hipDeviceSynchronize();
        scene->clear();
        release(scene->spectra);
        delete scene;
    }
}

bool ptSceneBuild(PtScene* sc, const HostTriangle* triangles, uint32_t triangleCount, const MaterialGpu* materials,
                  uint32_t materialCount, const TextureDescGpu* textures, uint32_t textureCount, hipStream_t stream)
{
    static_assert(sizeof(HostTriangle) == 100, "HostTriangle must be 100 bytes");
    static_assert(sizeof(MaterialGpu) == 96, "MaterialGpu must be 96 bytes");
    static_assert(sizeof(MeshAccelSceneGpu) == 144, "MeshAccelSceneGpu must be 144 bytes");
    static_assert(sizeof(TextureDescGpu) == 64, "TextureDescGpu must be 64 bytes");
    static_assert(sizeof(MeshBvhNode) == 48, "MeshBvhNode must be 48 bytes");
    static_assert(sizeof(CameraGpu) == 80, "CameraGpu must be 80 bytes");
    static_assert(sizeof(RenderParamsGpu) == 44, "RenderParamsGpu must be 44 bytes");
    if (sc == nullptr || (triangleCount > 0 && triangles == nullptr) || (materialCount > 0 && materials == nullptr) ||
        (textureCount > 0 && textures == nullptr)) {
        ptSetError("ptSceneBuild: invalid arguments");
        return false;
    }
    if (triangleCount > (1u << 28)) {
        ptSetError("ptSceneBuild: %u triangles exceed the supported maximum of 2^28", triangleCount);
        return false;
    }
    (void)hipStreamSynchronize(stream);
    sc->clear();
    sc->emAreaHost.clear();
    sc->emMatHost.clear();
    const int n = static_cast<int>(triangleCount);

    // Materials: default material when none is given (MeshAccelScene.cpp:156-159).
    std::vector<MaterialGpu> mats(materials, materials + materialCount);
    if (n > 0 && mats.empty()) {
        mats.push_back(default_material());
    }

    // Emissive list and prefix CDF, sequential float sum as MeshAccelScene.cpp:62-95.
    std::vector<uint32_t> emIdx;
    std::vector<float> emCdf(1, 0.0f);
    std::vector<float4> emTris;
    float total = 0.0f;
    for (uint32_t i = 0; i < triangleCount; ++i) {
        const HostTriangle& t = triangles[i];
        if (t.materialIndex >= mats.size()) {
            continue;
        }
        const MaterialGpu& m = mats[t.materialIndex];
        const float lum = m.emission * (0.2126f * m.r + 0.7152f * m.g + 0.0722f * m.b);
        if (lum <= 0.0f) {
            continue;
        }
        const float e1x = t.v1.x - t.v0.x, e1y = t.v1.y - t.v0.y, e1z = t.v1.z - t.v0.z;
        const float e2x = t.v2.x - t.v0.x, e2y = t.v2.y - t.v0.y, e2z = t.v2.z - t.v0.z;
        const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        const float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
        const float w = area * lum;
        if (w <= 0.0f) {
            continue;
        }
        emIdx.push_back(i);
        sc->emAreaHost.push_back(area);
        sc->emMatHost.push_back(t.materialIndex);
        total += w;
        emCdf.push_back(total);
        emTris.push_back(make_float4(t.v0.x, t.v0.y, t.v0.z, m.r * m.emission));
        emTris.push_back(make_float4(t.v1.x, t.v1.y, t.v1.z, m.g * m.emission));
        emTris.push_back(make_float4(t.v2.x, t.v2.y, t.v2.z, m.b * m.emission));
    }
    if (total <= 0.0f) {
        emCdf.assign(1, 0.0f);
    }
    sc->matsHost = mats;
    sc->emTrisHost = emTris;

    float* dTris = nullptr;
    bool ok = true;
    if (n > 0) {
        ok = upload(&dTris, reinterpret_cast<const float*>(triangles), 25 * static_cast<size_t>(n), stream) &&
             alloc(&sc->triShade, 4 * static_cast<size_t>(n)) && alloc(&sc->triPos, 3 * static_cast<size_t>(n)) &&
             alloc(&sc->nodes, 4 * static_cast<size_t>(n > 1 ? n - 1 : 1)) &&
             alloc(&sc->refNodes, 2 * static_cast<size_t>(n) - 1);
        if (ok && (sc->buildMode == 1 || sc->buildMode == 2)) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, stream);
            float root[7] = {0};
            std::vector<int> levelStarts;
            ok = pt_build_bvh_sah(dTris, n, sc->nodes, sc->triPos, sc->triShade, stream, root, &sc->levels,
                                  &levelStarts);
            if (ok && sc->buildMode == 1) {
                ok = pt_build_wide(sc->nodes, n - 1, levelStarts, &sc->wide, &sc->wideCount, stream);
            }
            (void)hipEventRecord(e1, stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&sc->buildMs, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            for (int k = 0; k < 3; ++k) {
                sc->desc.rootMin[k] = root[k];
                sc->desc.rootMax[k] = root[3 + k];
            }
            std::memcpy(&sc->desc.rootRef, &root[6], sizeof(int));
            sc->refCount = 0u; // the reference's 48-byte pre-order layout describes median-split trees only
        } else if (ok) {
            ok = build_bvh(sc, dTris, n, stream);
            sc->refCount = ok ? static_cast<uint32_t>(2 * n - 1) : 0u;
        }
    }
    if (ok && n > 0 && (sc->buildMode == 1 || sc->buildMode == 2)) {
        sc->srcTris = dTris;  // kept for ptSceneRebuild
        sc->triangleCount = static_cast<int>(n);
        ok = alloc(&sc->rebuildStatus, pt_sah_status_words()) && pt_sah_status_consume(sc->rebuildStatus);
    } else {
        release(dTris);
    }
    ok = ok && upload(&sc->materials, mats.data(), mats.size(), stream) &&
         upload(&sc->textures, textures, textureCount, stream) && upload(&sc->emIdx, emIdx.data(), emIdx.size(), stream) &&
         upload(&sc->emCdf, emCdf.data(), emCdf.size(), stream) && upload(&sc->emTris, emTris.data(), emTris.size(), stream);
    if (!ok) {
        sc->clear();
        return false;
    }
    MeshAccelSceneGpu& d = sc->desc;
    d.nodes = reinterpret_cast<const PtFloat4*>(sc->nodes);
    d.triPos = reinterpret_cast<const PtFloat4*>(sc->triPos);
    d.triShade = reinterpret_cast<const PtFloat4*>(sc->triShade);
    d.materials = sc->materials;
    d.emissiveTriangleIndices = sc->emIdx;
    d.emissiveTriangleCdf = sc->emCdf;
    d.emissiveTris = reinterpret_cast<const PtFloat4*>(sc->emTris);
    d.textures = textureCount > 0 ? sc->textures : nullptr;
    d.nodeCount = n > 1 ? static_cast<uint32_t>(n - 1) : 0u;
    d.triangleCount = triangleCount;
    d.materialCount = n > 0 ? static_cast<uint32_t>(mats.size()) : 0u;
    d.emissiveTriangleCount = static_cast<uint32_t>(emIdx.size());
    d.textureCount = textureCount;
    if (n == 0) {
        d.rootRef = 0;
    }
    // the wide tree: node 0 where the binary root is internal, the same leaf otherwise
    const bool wide = sc->buildMode == 1 && n > 0;
    d.wideNodes = wide && sc->wideCount > 0 ? reinterpret_cast<const PtFloat4*>(sc->wide) : nullptr;
    d.wideNodeCount = wide ? static_cast<uint32_t>(sc->wideCount) : 0u;
    d.wideRootRef = wide ? (d.rootRef >= 0 ? 0 : d.rootRef) : 0;
    d.materialSpectra = sc->spectra != nullptr && sc->spectraCount == d.materialCount && n > 0 ? sc->spectra : nullptr;
    if (!upload(&sc->d_desc, &d, 1, stream) || hipStreamSynchronize(stream) != hipSuccess) {
        sc->clear();
        ptSetError("ptSceneBuild: scene upload failed");
        return false;
    }
    if (wide) {
        wide_registry(sc->d_desc, true);
    }
    sc->builtMode = sc->buildMode;
    return true;
}

bool ptSceneReadWide(const PtScene* scene, float* nodes, int* count, int* rootRef)
{
    if (scene == nullptr || scene->d_desc == nullptr || count == nullptr) {
        ptSetError("ptSceneReadWide: invalid arguments");
        return false;
    }
    if (scene->builtMode != 1) {
        ptSetError("ptSceneReadWide: the scene has no 4-wide tree (build mode 1)");
        return false;
    }
    if (nodes != nullptr) {
        if (*count < scene->wideCount) {
            ptSetError("ptSceneReadWide: buffer of %d nodes for %d", *count, scene->wideCount);
            return false;
        }
        if (scene->wideCount > 0 &&
            hipMemcpy(nodes, scene->wide, static_cast<size_t>(scene->wideCount) * 8 * sizeof(float4),
                      hipMemcpyDeviceToHost) != hipSuccess) {
            ptSetError("ptSceneReadWide: copy failed");
            return false;
        }
    }
    *count = scene->wideCount;
    if (rootRef != nullptr) {
        *rootRef = scene->desc.wideRootRef;
    }
    return true;
}

const MeshAccelSceneGpu* ptSceneDevice(const PtScene* scene) { return scene != nullptr ? scene->d_desc : nullptr; }

bool ptSceneUpdateMaterials(PtScene* sc, const MaterialGpu* materials, uint32_t materialCount, hipStream_t stream)
{
    if (sc == nullptr || sc->d_desc == nullptr || materials == nullptr ||
        materialCount != static_cast<uint32_t>(sc->matsHost.size())) {
        ptSetError("ptSceneUpdateMaterials: invalid arguments (the material count must not change)");
        return false;
    }
    auto lum = [](const MaterialGpu& m) { return m.emission * (0.2126f * m.r + 0.7152f * m.g + 0.0722f * m.b); };
    for (uint32_t k = 0; k < materialCount; ++k) {
        if ((lum(sc->matsHost[k]) > 0.0f) != (lum(materials[k]) > 0.0f)) {
            ptSetError("ptSceneUpdateMaterials: material %u changes whether it emits; rebuild the scene", k);
            return false;
        }
    }
    // same sequential float CDF as the build (MeshAccelScene.cpp:62-95)
    const size_t ne = sc->emAreaHost.size();
    std::vector<float> cdf(1, 0.0f);
    std::vector<float> radiance(3 * ne);
    float total = 0.0f;
    for (size_t e = 0; e < ne; ++e) {
        const MaterialGpu& m = materials[sc->emMatHost[e]];
        total += sc->emAreaHost[e] * lum(m);
        cdf.push_back(total);
        radiance[3 * e + 0] = m.r * m.emission;
        radiance[3 * e + 1] = m.g * m.emission;
        radiance[3 * e + 2] = m.b * m.emission;
    }
    if (total <= 0.0f) {
        cdf.assign(1, 0.0f);
    }
    bool ok = hipMemcpyAsync(sc->materials, materials, materialCount * sizeof(MaterialGpu), hipMemcpyHostToDevice,
                             stream) == hipSuccess;
    if (ok && ne > 0) {
        ok = hipMemcpyAsync(sc->emCdf, cdf.data(), cdf.size() * sizeof(float), hipMemcpyHostToDevice, stream) ==
             hipSuccess;
        // radiance lives in the .w lane of the three float4 records of each emissive entry
        for (size_t e = 0; e < ne; ++e) {
            for (int q = 0; q < 3; ++q) {
                sc->emTrisHost[3 * e + q].w = radiance[3 * e + q];
            }
        }
        ok = ok && hipMemcpyAsync(sc->emTris, sc->emTrisHost.data(), sc->emTrisHost.size() * sizeof(float4),
                                  hipMemcpyHostToDevice, stream) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(stream) == hipSuccess;
    if (!ok) {
        ptSetError("ptSceneUpdateMaterials: upload failed");
        return false;
    }
    sc->matsHost.assign(materials, materials + materialCount);
    return true;
}

bool ptSceneSetSpectra(PtScene* sc, const float* spectra, uint32_t materialCount, hipStream_t stream)
{
    if (sc == nullptr || (spectra != nullptr && materialCount > 0 && sc->d_desc != nullptr &&
                          materialCount != sc->desc.materialCount)) {
        ptSetError("ptSceneSetSpectra: invalid arguments (one spectrum record per material of the scene)");
        return false;
    }
    (void)hipStreamSynchronize(stream);
    if (spectra == nullptr || materialCount == 0) {
        release(sc->spectra);
        sc->spectra = nullptr;
        sc->spectraCount = 0;
    } else {
        if (materialCount != sc->spectraCount) {
            release(sc->spectra);
            sc->spectra = nullptr;
            sc->spectraCount = 0;
            if (!alloc(&sc->spectra, static_cast<size_t>(materialCount) * kPtSpectrumWords)) {
                return false;
            }
            sc->spectraCount = materialCount;
        }
        if (hipMemcpyAsync(sc->spectra, spectra, static_cast<size_t>(materialCount) * kPtSpectrumWords * sizeof(float),
                           hipMemcpyHostToDevice, stream) != hipSuccess) {
            ptSetError("ptSceneSetSpectra: upload failed");
            return false;
        }
    }
    if (sc->d_desc != nullptr) {
        sc->desc.materialSpectra = sc->spectra;
        if (hipMemcpyAsync(&sc->d_desc->materialSpectra, &sc->desc.materialSpectra, sizeof(const float*),
                           hipMemcpyHostToDevice, stream) != hipSuccess) {
            ptSetError("ptSceneSetSpectra: upload failed");
            return false;
        }
    }
    if (hipStreamSynchronize(stream) != hipSuccess) {
        ptSetError("ptSceneSetSpectra: upload failed");
        return false;
    }
    return true;
}

bool ptSceneSetBuildMode(PtScene* scene, int mode)
{
    if (scene == nullptr || mode < 0 || mode > 2) {
        ptSetError("ptSceneSetBuildMode: invalid arguments");
        return false;
    }
    scene->buildMode = mode;
    return true;
}

bool ptSceneReadLayout(const PtScene* scene, float* nodes, float* triPos, int* levels)
{
    if (scene == nullptr || scene->d_desc == nullptr) {
        ptSetError("ptSceneReadLayout: no scene");
        return false;
    }
    const size_t nn = scene->desc.nodeCount, nt = scene->desc.triangleCount;
    if (levels != nullptr) {
        *levels = scene->levels;
    }
    if ((nodes != nullptr && nn > 0 &&
         hipMemcpy(nodes, scene->nodes, nn * 64, hipMemcpyDeviceToHost) != hipSuccess) ||
        (triPos != nullptr && nt > 0 &&
         hipMemcpy(triPos, scene->triPos, nt * 48, hipMemcpyDeviceToHost) != hipSuccess)) {
        ptSetError("ptSceneReadLayout: copy failed");
        return false;
    }
    return true;
}

bool ptSceneDescription(const PtScene* scene, MeshAccelSceneGpu* out)
{
    if (scene == nullptr || out == nullptr || scene->d_desc == nullptr) {
        ptSetError("ptSceneDescription: no scene");
        return false;
    }
    *out = scene->desc;
    return true;
}

bool ptSceneReadBvh(const PtScene* scene, MeshBvhNode* hostNodes, uint32_t capacity, uint32_t* outCount,
                    uint32_t* outRoot)
{
    if (scene == nullptr || outCount == nullptr || outRoot == nullptr) {
        ptSetError("ptSceneReadBvh: invalid arguments");
        return false;
    }
    *outCount = scene->refCount;
    *outRoot = 0u;
    if (hostNodes == nullptr || scene->refCount == 0) {
        return true;
    }
    if (capacity < scene->refCount) {
        ptSetError("ptSceneReadBvh: capacity %u < %u nodes", capacity, scene->refCount);
        return false;
    }
    if (hipMemcpy(hostNodes, scene->refNodes, scene->refCount * sizeof(MeshBvhNode), hipMemcpyDeviceToHost) !=
        hipSuccess) {
        ptSetError("ptSceneReadBvh: copy failed");
        return false;
    }
    return true;
}

bool ptSceneReadEmissive(const PtScene* scene, uint32_t* indices, float* cdf, float* records, uint32_t capacity,
                         uint32_t* outCount)
{
    if (scene == nullptr || outCount == nullptr) {
        ptSetError("ptSceneReadEmissive: invalid arguments");
        return false;
    }
    const uint32_t n = scene->desc.emissiveTriangleCount;
    *outCount = n;
    if (n == 0 || (indices == nullptr && cdf == nullptr && records == nullptr)) {
        if (cdf != nullptr && capacity >= 1 && scene->emCdf != nullptr &&
            hipMemcpy(cdf, scene->emCdf, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
            ptSetError("ptSceneReadEmissive: copy failed");
            return false;
        }
        return true;
    }
    if (capacity < n) {
        ptSetError("ptSceneReadEmissive: capacity %u < %u entries", capacity, n);
        return false;
    }
    bool ok = true;
    if (indices != nullptr) {
        ok = ok && hipMemcpy(indices, scene->emIdx, n * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (cdf != nullptr) {
        ok = ok && hipMemcpy(cdf, scene->emCdf, (n + 1) * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (records != nullptr) {
        ok = ok && hipMemcpy(records, scene->emTris, 3 * n * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (!ok) {
        ptSetError("ptSceneReadEmissive: copy failed");
    }
    return ok;
}

float ptSceneLastBuildMs(const PtScene* scene) { return scene != nullptr ? scene->buildMs : 0.0f; }

PtBuildWorkspace* ptBuildWorkspaceCreate(int maxTriangles)
{
    return reinterpret_cast<PtBuildWorkspace*>(pt_sah_work_create(maxTriangles));
}

void ptBuildWorkspaceDestroy(PtBuildWorkspace* workspace)
{
    pt_sah_work_destroy(reinterpret_cast<PtSahWork*>(workspace));
}

namespace {

bool rebuild_levels(PtScene* scene, PtBuildWorkspace* workspace, int levels, hipStream_t stream, const char* who)
{
    if (scene == nullptr || workspace == nullptr) {
        ptSetError("%s: null scene or workspace", who);
        return false;
    }
    if (scene->srcTris == nullptr || scene->triangleCount < 1 || (scene->builtMode != 1 && scene->builtMode != 2)) {
        ptSetError("%s: the scene was not built with a binned-SAH build mode", who);
        return false;
    }
    // the layout the scene was built with (a later ptSceneSetBuildMode applies to the next build)
    PtSahWork* w = reinterpret_cast<PtSahWork*>(workspace);
    return pt_sah_enqueue(w, scene->srcTris, scene->triangleCount, scene->nodes, scene->triPos, scene->triShade,
                          scene->builtMode == 1 ? scene->wide : nullptr, scene->wideCount, levels, stream) &&
           pt_sah_snapshot(w, scene->rebuildStatus, stream);
}

} // namespace

bool ptSceneRebuildCold(PtScene* scene, PtBuildWorkspace* workspace, hipStream_t stream)
{
    return rebuild_levels(scene, workspace, 0, stream, "ptSceneRebuildCold");
}

bool ptSceneRebuild(PtScene* scene, PtBuildWorkspace* workspace, hipStream_t stream)
{
    return rebuild_levels(scene, workspace, scene != nullptr ? scene->levels : 0, stream, "ptSceneRebuild");
}

bool ptSceneRebuildCheck(PtScene* scene, PtBuildWorkspace* workspace, hipStream_t stream)
{
    (void)workspace;   // each scene keeps its own copy of its last rebuild's status
    if (scene == nullptr || scene->srcTris == nullptr || scene->rebuildStatus == nullptr) {
        ptSetError("ptSceneRebuildCheck: null scene, or a scene without a binned-SAH build");
        return false;
    }
    const int n = scene->triangleCount;
    float root[7] = {0};
    int levels = 0, wideCount = 0;
    const bool read = pt_sah_status_result(scene->rebuildStatus, n, stream, root, &levels, &wideCount);
    if (!pt_sah_status_consume(scene->rebuildStatus) && read) {
        ptSetError("ptSceneRebuildCheck: resetting the scene's rebuild status failed");
        return false;
    }
    if (!read) {
        return false;
    }
    int rootRef = 0;
    std::memcpy(&rootRef, &root[6], sizeof(int));
    bool same = levels == scene->levels && rootRef == scene->desc.rootRef &&
                (scene->builtMode != 1 || n < 2 || wideCount == scene->wideCount);
    for (int k = 0; k < 3; ++k) {
        same = same && root[k] == scene->desc.rootMin[k] && root[3 + k] == scene->desc.rootMax[k];
    }
    if (!same) {
        ptSetError("ptSceneRebuildCheck: the rebuilt tree's root or shape differs from the scene's");
        return false;
    }
    return true;
}

} // extern "C"

// ---- environment map (EnvironmentMap.cpp:76-127 importance tables + upload) -----------------

struct PtEnvironment {
    EnvironmentMapGpu desc{};
    EnvironmentMapGpu* d_desc = nullptr;
    float* pixels = nullptr;
    float* marginal = nullptr;
    float* rows = nullptr;
    std::vector<float> hostMarginal;
    std::vector<float> hostRows;
};

extern "C" {

PtEnvironment* ptEnvironmentCreate(const float* rgb, int width, int height, hipStream_t stream)
{
    if (rgb == nullptr || width <= 0 || height <= 0 || static_cast<long long>(width) * height > (1ll << 28)) {
        ptSetError("ptEnvironmentCreate: invalid arguments");
        return nullptr;
    }
    PtEnvironment* e = new PtEnvironment();
    const size_t w1 = static_cast<size_t>(width) + 1;
    e->hostMarginal.assign(static_cast<size_t>(height) + 1, 0.0f);
    e->hostRows.assign(static_cast<size_t>(height) * w1, 0.0f);
    // rows: prefix sums of max(r,g,b) (floored at 1e-6) * sin(theta), normalised per row;
    // marginal: prefix sums of the row sums, normalised (sequential float sums as the reference)
    float total = 0.0f;
    for (int y = 0; y < height; ++y) {
        const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(height);
        const float st = pt::sinf_pt(v * 3.14159265f);
        float rowSum = 0.0f;
        float* row = e->hostRows.data() + static_cast<size_t>(y) * w1;
        row[0] = 0.0f;
        for (int x = 0; x < width; ++x) {
            const float* px = rgb + 3 * (static_cast<size_t>(y) * width + x);
            const float bc = px[1] < px[2] ? px[2] : px[1]; // std::max(g, b)
            const float lum = px[0] < bc ? bc : px[0];       // std::max(r, .)
            const float w = (lum < 1.0e-6f ? 1.0e-6f : lum) * st;
            rowSum += w;
            row[x + 1] = rowSum;
        }
        if (rowSum > 0.0f) {
            for (int x = 0; x < width; ++x) {
                row[x + 1] /= rowSum;
            }
        }
        total += rowSum;
        e->hostMarginal[static_cast<size_t>(y) + 1] = total;
    }
    if (total > 0.0f) {
        for (int y = 0; y < height; ++y) {
            e->hostMarginal[static_cast<size_t>(y) + 1] /= total;
        }
    }
    const size_t npx = 3 * static_cast<size_t>(width) * height;
    bool ok = upload(&e->pixels, rgb, npx, stream) && upload(&e->marginal, e->hostMarginal.data(), e->hostMarginal.size(), stream) &&
              upload(&e->rows, e->hostRows.data(), e->hostRows.size(), stream);
    if (ok) {
        e->desc.pixels = e->pixels;
        e->desc.marginalCdf = e->marginal;
        e->desc.rowCdf = e->rows;
        e->desc.width = width;
        e->desc.height = height;
        e->desc.valid = 1;
        ok = upload(&e->d_desc, &e->desc, 1, stream) && hipStreamSynchronize(stream) == hipSuccess;
    }
    if (!ok) {
        release(e->pixels);
        release(e->marginal);
        release(e->rows);
        release(e->d_desc);
        delete e;
        return nullptr;
    }
    return e;
}

const EnvironmentMapGpu* ptEnvironmentDevice(const PtEnvironment* env) { return env != nullptr ? env->d_desc : nullptr; }

bool ptEnvironmentReadTables(const PtEnvironment* env, float* marginal, float* rows)
{
    if (env == nullptr || marginal == nullptr || rows == nullptr) {
        ptSetError("ptEnvironmentReadTables: invalid arguments");
        return false;
    }
    std::memcpy(marginal, env->hostMarginal.data(), env->hostMarginal.size() * sizeof(float));
    std::memcpy(rows, env->hostRows.data(), env->hostRows.size() * sizeof(float));
    return true;
}

void ptEnvironmentDestroy(PtEnvironment* env)
{
    if (env != nullptr) {
        (void)hipDeviceSynchronize();
        release(env->pixels);
        release(env->marginal);
        release(env->rows);
        release(env->d_desc);
        delete env;
    }
}

} // extern "C"
