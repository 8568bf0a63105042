// pt_sah.hip -- binned-SAH BVH build on the GPU (performance build mode).
//
// Produces the same device layout as the median-split build (64-byte internal nodes holding both
// child boxes, one triangle per leaf slot), so every traversal kernel runs on it unchanged.
//   1. Triangles are put in Morton order of their centroids (30-bit codes, radix sort).
//   2. The tree is built top down, one level of segments at a time. Per level: each segment's
//      triangle and centroid bounds (segmented wave scans, then LDS or global atomics); 32 bins
//      along the segment's longest centroid axis, staged in LDS per block; one thread per
//      segment sweeps the bins for the lowest surface-area cost; a stable partition (scan)
//      moves each segment's left side to its front. The partition is stable, so the Morton
//      order survives inside every child: when the bins cannot separate a segment (all
//      centroids in one bin) it is split at its middle position, which is a spatial split of a
//      Morton-ordered range.
//   3. Internal nodes are numbered level by level (a scan over the level's segments), so the
//      tree is deterministic. Depth is capped: a segment that could not otherwise finish within
//      kMaxSahDepth levels is split at its middle position.
// Unlike the median-split build, the topology depends on the geometry; closest hits are the same
// set of (t, triangle) minima, and only exact ties between triangles could resolve differently.
#include "pt_device.h"
#include "pt_internal.h"

#include "../../include/pt_hip.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

namespace {

constexpr int kB = 256;
constexpr int kBins = 32;
constexpr int kBinWords = 7;   // count, min xyz, max xyz (sortable integers)
constexpr int kSegWords = 12;  // triangle bounds (6) and centroid bounds (6), sortable integers
constexpr int kWave = 64;
constexpr unsigned kPrepBlocks = 1024;
constexpr int kChunk = 8;       // positions per block of seg_bounds_kernel: kChunk x kB
static_assert(kMaxSahDepth <= 56, "traversal stacks hold 64 entries");

__device__ __forceinline__ uint32_t sortable(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float unsortable(uint32_t u)
{
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(b);
}

__device__ __forceinline__ float fmin2(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax2(float a, float b) { return a > b ? a : b; }

struct SegRec {
    int begin, count, parent, side;
};

// Per-segment split decision of the current level.
struct SplitRec {
    int axis;      // 0..2, or -1: split at the middle position
    float k0;      // centroid minimum on the axis
    float scale;   // bins per unit length
    int splitBin;  // left side: bins 0..splitBin
    int leftCount;
    int internal;  // 1 when the segment is an internal node
};

// Device-resident state of a build. The host enqueues a fixed number of levels (enough for the
// depth cap); every per-level kernel reads the live segment count here, and a level with no
// segments does nothing, so a build needs no host readback.
struct SahCtl {
    int segCount;   // segments of the current level
    int levelBase;  // internal nodes numbered before the current level
    int level;      // levels that produced internal nodes so far
    int internals;  // internal segments of the current level
    int error;      // 1: segment table overflow, 2: the 4-wide tree does not fit its buffer
    int wideCount;  // 4-wide nodes of the collapse
    int levelStart[kMaxSahDepth + 2];
    int wideStart[kMaxSahDepth + 2];
};
constexpr int kSahLevels = kMaxSahDepth + 1;  // leaves sit at depth <= kMaxSahDepth

// Grid-stride over the triangles (at most kPrepBlocks blocks): per-triangle bounds, centroid and
// shading record; the centroid bounds reduce per lane, per wave (shuffles) and per block (LDS),
// so the scene bounds take one set of global atomics per block.
__global__ __launch_bounds__(kB) void sah_prep_kernel(const float* __restrict__ tris, int n,
                                                      float4* __restrict__ bounds, float4* __restrict__ cent,
                                                      float4* __restrict__ shade, uint32_t* __restrict__ gcb)
{
    __shared__ uint32_t lds[6];
    if (threadIdx.x < 6) {
        lds[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
    }
    __syncthreads();
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += stride) {
        const float* t = tris + 25 * static_cast<size_t>(i);
        const float third = 1.0f / 3.0f;
        const float cx = ((t[0] + t[3]) + t[6]) * third;
        const float cy = ((t[1] + t[4]) + t[7]) * third;
        const float cz = ((t[2] + t[5]) + t[8]) * third;
        cent[i] = make_float4(cx, cy, cz, 0.0f);
        bounds[2 * i] = make_float4(fmin2(fmin2(t[0], t[3]), t[6]), fmin2(fmin2(t[1], t[4]), t[7]),
                                    fmin2(fmin2(t[2], t[5]), t[8]), 0.0f);
        bounds[2 * i + 1] = make_float4(fmax2(fmax2(t[0], t[3]), t[6]), fmax2(fmax2(t[1], t[4]), t[7]),
                                        fmax2(fmax2(t[2], t[5]), t[8]), 0.0f);
        shade[4 * i + 0] = make_float4(t[9], t[10], t[11], t[12]);
        shade[4 * i + 1] = make_float4(t[13], t[14], t[15], t[16]);
        shade[4 * i + 2] = make_float4(t[17], t[18], t[19], t[20]);
        shade[4 * i + 3] = make_float4(t[21], t[22], t[23], t[24]);
        const uint32_t c[3] = {sortable(cx), sortable(cy), sortable(cz)};
        for (int k = 0; k < 3; ++k) {
            v[k] = min(v[k], c[k]);
            v[3 + k] = max(v[3 + k], c[k]);
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        for (int k = 0; k < 3; ++k) {
            v[k] = min(v[k], static_cast<uint32_t>(__shfl_xor(static_cast<int>(v[k]), off, kWave)));
            v[3 + k] = max(v[3 + k], static_cast<uint32_t>(__shfl_xor(static_cast<int>(v[3 + k]), off, kWave)));
        }
    }
    if ((threadIdx.x & (kWave - 1)) == 0u) {
        for (int k = 0; k < 3; ++k) {
            atomicMin(&lds[k], v[k]);
            atomicMax(&lds[3 + k], v[3 + k]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&gcb[threadIdx.x], lds[threadIdx.x]);
    } else if (threadIdx.x < 6) {
        atomicMax(&gcb[threadIdx.x], lds[threadIdx.x]);
    }
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void morton_kernel(const float4* __restrict__ cent, int n, const uint32_t* __restrict__ gcb,
                              uint32_t* __restrict__ keys, int* __restrict__ vals)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const float4 c = cent[i];
    const float cc[3] = {c.x, c.y, c.z};
    uint32_t code = 0u;
    for (int k = 0; k < 3; ++k) {
        const float lo = unsortable(gcb[k]);
        const float hi = unsortable(gcb[3 + k]);
        const float ext = hi - lo;
        float u = ext > 0.0f ? (cc[k] - lo) / ext : 0.0f;
        u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
        code |= expand_bits(static_cast<uint32_t>(u)) << (2 - k);
    }
    keys[i] = code;
    vals[i] = i;
}

// Grid-stride over the level's segments (the grid is sized for the largest level, at most
// kPrepBlocks blocks).
__global__ void seg_reset_kernel(uint32_t* __restrict__ segStats, const SahCtl* __restrict__ ctl)
{
    const long long words = static_cast<long long>(ctl->segCount) * kSegWords;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) {
        const int w = static_cast<int>(i % kSegWords);
        segStats[i] = (w % 6) < 3 ? 0xFFFFFFFFu : 0u;
    }
}

// A segment occupies one contiguous range of positions, so its lanes within a wave form one run.
// A block covers kChunk x kB consecutive positions. A segmented inclusive scan over each wave
// (shuffles) leaves each run's bounds in the run's last lane, and only those lanes update the
// block's LDS copy (blocks inside one segment, flushed to global once per block) or the segment's
// global record.
__global__ __launch_bounds__(kB) void seg_bounds_kernel(const int* __restrict__ perm, const int* __restrict__ posSeg,
                                                        const float4* __restrict__ bounds,
                                                        const float4* __restrict__ cent, int n,
                                                        uint32_t* __restrict__ segStats, const SahCtl* __restrict__ ctl,
                                                        const SegRec* __restrict__ segs)
{
    __shared__ uint32_t lds[kSegWords];
    if (ctl->segCount == 0) {
        return;
    }
    const int base = static_cast<int>(blockIdx.x) * kB * kChunk;
    const int last = min(base + kB * kChunk, n) - 1;
    const int s0 = posSeg[base];
    const bool uniform = s0 >= 0 && posSeg[last] == s0;
    if (uniform && threadIdx.x < kSegWords) {
        lds[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0xFFFFFFFFu : 0u;
    }
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    for (int it = 0; it < kChunk; ++it) {
        const int i = base + it * kB + static_cast<int>(threadIdx.x);
        const int s = i < n ? posSeg[i] : -1;
        uint32_t v[kSegWords];
        if (s >= 0) {
            const int tri = perm[i];
            const float4 lo = bounds[2 * tri];
            const float4 hi = bounds[2 * tri + 1];
            const float4 c = cent[tri];
            v[0] = sortable(lo.x);
            v[1] = sortable(lo.y);
            v[2] = sortable(lo.z);
            v[3] = sortable(hi.x);
            v[4] = sortable(hi.y);
            v[5] = sortable(hi.z);
            v[6] = sortable(c.x);
            v[7] = sortable(c.y);
            v[8] = sortable(c.z);
            v[9] = v[6];
            v[10] = v[7];
            v[11] = v[8];
        } else {
            for (int k = 0; k < kSegWords; ++k) {
                v[k] = (k % 6) < 3 ? 0xFFFFFFFFu : 0u;
            }
        }
        for (int d = 1; d < kWave; d <<= 1) {
            const int so = __shfl_up(s, d, kWave);
            const bool take = lane >= d && so == s;
            for (int k = 0; k < kSegWords; ++k) {
                const uint32_t o = __shfl_up(v[k], d, kWave);
                if (take) {
                    v[k] = (k % 6) < 3 ? min(v[k], o) : max(v[k], o);
                }
            }
        }
        const int sn = __shfl_down(s, 1, kWave);
        if (s >= 0 && (lane == kWave - 1 || sn != s)) {
            uint32_t* dst = uniform ? lds : segStats + static_cast<size_t>(kSegWords) * s;
            // a segment that lies inside this wave is this run alone: no other lane or block
            // touches its record, so plain stores replace the atomics
            const SegRec sg = uniform ? SegRec{0, 0, 0, 0} : segs[s];
            if (!uniform && sg.begin >= i - lane && sg.begin + sg.count - 1 == i) {
                uint4* q = reinterpret_cast<uint4*>(dst);
                q[0] = make_uint4(v[0], v[1], v[2], v[3]);
                q[1] = make_uint4(v[4], v[5], v[6], v[7]);
                q[2] = make_uint4(v[8], v[9], v[10], v[11]);
            } else {
                for (int k = 0; k < kSegWords; ++k) {
                    if ((k % 6) < 3) {
                        atomicMin(&dst[k], v[k]);
                    } else {
                        atomicMax(&dst[k], v[k]);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (uniform && threadIdx.x < kSegWords) {
        const int k = static_cast<int>(threadIdx.x);
        uint32_t* g = segStats + static_cast<size_t>(kSegWords) * s0 + k;
        if ((k % 6) < 3) {
            atomicMin(g, lds[k]);
        } else {
            atomicMax(g, lds[k]);
        }
    }
}

__device__ __forceinline__ int ceil_log2(int c)
{
    int d = 0;
    while ((1 << d) < c) {
        ++d;
    }
    return d;
}

// One thread per segment: box into the parent (or the root), leaves, bin setup.
__global__ void seg_setup_kernel(const SegRec* __restrict__ segs, const SahCtl* __restrict__ ctl,
                                 const uint32_t* __restrict__ segStats,
                                 const int* __restrict__ perm, const float* __restrict__ tris, int level,
                                 float* __restrict__ nodes, float4* __restrict__ triPos, float* __restrict__ rootBox,
                                 SplitRec* __restrict__ split, uint32_t* __restrict__ bins)
{
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= ctl->segCount) {
        return;
    }
    const SegRec sg = segs[j];
    const uint32_t* st = segStats + static_cast<size_t>(kSegWords) * j;
    float box[6];
    for (int k = 0; k < 6; ++k) {
        box[k] = unsortable(st[k]);
    }
    float* dstBox = sg.parent < 0 ? rootBox : nodes + 16 * static_cast<size_t>(sg.parent) + 6 * sg.side;
    for (int k = 0; k < 6; ++k) {
        dstBox[k] = box[k];
    }
    SplitRec r{-1, 0.0f, 0.0f, -1, 0, 0};
    if (sg.count == 1) {
        const int slot = sg.begin;
        const int tri = perm[slot];
        const float* t = tris + 25 * static_cast<size_t>(tri);
        triPos[3 * slot + 0] = make_float4(t[0], t[1], t[2], __int_as_float(tri));
        triPos[3 * slot + 1] = make_float4(t[3] - t[0], t[4] - t[1], t[5] - t[2], t[24]); // .w: material
        triPos[3 * slot + 2] = make_float4(t[6] - t[0], t[7] - t[1], t[8] - t[2], 0.0f);
        const int ref = ~slot;
        if (sg.parent < 0) {
            rootBox[6] = __int_as_float(ref);
        } else {
            nodes[16 * static_cast<size_t>(sg.parent) + 12 + sg.side] = __int_as_float(ref);
        }
        split[j] = r;
        return;
    }
    r.internal = 1;
    const float ext[3] = {unsortable(st[9]) - unsortable(st[6]), unsortable(st[10]) - unsortable(st[7]),
                          unsortable(st[11]) - unsortable(st[8])};
    int axis = 0;
    if (ext[1] > ext[axis]) {
        axis = 1;
    }
    if (ext[2] > ext[axis]) {
        axis = 2;
    }
    const bool depthCap = level + ceil_log2(sg.count) >= kMaxSahDepth;
    if (ext[axis] > 0.0f && !depthCap) {
        r.axis = axis;
        r.k0 = unsortable(st[6 + axis]);
        r.scale = static_cast<float>(kBins) / ext[axis];
        uint32_t* b = bins + static_cast<size_t>(kBins * kBinWords) * j;
        for (int k = 0; k < kBins; ++k) {
            b[k * kBinWords + 0] = 0u;
            for (int q = 0; q < 3; ++q) {
                b[k * kBinWords + 1 + q] = 0xFFFFFFFFu;
                b[k * kBinWords + 4 + q] = 0u;
            }
        }
    }
    split[j] = r;
}

__device__ __forceinline__ int bin_of(const SplitRec& r, const float4& c)
{
    const float v = r.axis == 0 ? c.x : (r.axis == 1 ? c.y : c.z);
    const int b = static_cast<int>((v - r.k0) * r.scale);
    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

// Blocks inside one segment histogram in LDS and flush once. Blocks that span at most
// kLocalSegs segments histogram per segment in LDS and flush the cells they filled; others
// update the segments' global bins directly. Consecutive lanes often share a (segment, bin) key
// (positions are in Morton order), so each wave first reduces its runs of equal keys: a max-scan
// of run heads gives every lane its run's first lane, a segmented scan over those run ids leaves
// each run's count and bounds in its last lane, and only those lanes update a histogram.
constexpr int kLocalSegs = 32;

__device__ __forceinline__ void flush_cell_word(uint32_t* g, int w, uint32_t v)
{
    if (w == 0) {
        atomicAdd(g, v);
    } else if (w < 4) {
        atomicMin(g, v);
    } else {
        atomicMax(g, v);
    }
}

__global__ __launch_bounds__(kB) void bin_kernel(const int* __restrict__ perm, const int* __restrict__ posSeg,
                                                 const float4* __restrict__ bounds, const float4* __restrict__ cent,
                                                 const SplitRec* __restrict__ split, int n, uint32_t* __restrict__ bins,
                                                 const SahCtl* __restrict__ ctl)
{
    __shared__ uint32_t lds[kLocalSegs * kBins * kBinWords];
    __shared__ int range[2];
    if (ctl->segCount == 0) {
        return;
    }
    constexpr int kCell = kBins * kBinWords;
    const int base = static_cast<int>(blockIdx.x * blockDim.x);
    const int i = base + static_cast<int>(threadIdx.x);
    const int last = min(base + kB, n) - 1;
    const int s0 = posSeg[base];
    const bool uniform = s0 >= 0 && posSeg[last] == s0 && split[s0].axis >= 0;
    if (threadIdx.x == 0) {
        range[0] = 0x7FFFFFFF;
        range[1] = -1;
    }
    int s = i < n ? posSeg[i] : -1;
    int b = -1;  // (s, b): the lane's histogram cell; b < 0 for lanes with nothing to add
    uint32_t v[kBinWords] = {0u, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    if (s >= 0) {
        const SplitRec r = split[s];
        if (r.axis >= 0) {
            const int tri = perm[i];
            b = bin_of(r, cent[tri]);
            const float4 lo = bounds[2 * tri];
            const float4 hi = bounds[2 * tri + 1];
            v[0] = 1u;
            v[1] = sortable(lo.x);
            v[2] = sortable(lo.y);
            v[3] = sortable(lo.z);
            v[4] = sortable(hi.x);
            v[5] = sortable(hi.y);
            v[6] = sortable(hi.z);
        }
    }
    if (b < 0) {
        s = -1;
    }
    // the block's segment range (wave minimum and maximum, then LDS)
    int sMin = s >= 0 ? s : 0x7FFFFFFF;
    int sMax = s;
    for (int off = kWave / 2; off > 0; off >>= 1) {
        sMin = min(sMin, __shfl_xor(sMin, off, kWave));
        sMax = max(sMax, __shfl_xor(sMax, off, kWave));
    }
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0 && sMax >= 0) {
        atomicMin(&range[0], sMin);
        atomicMax(&range[1], sMax);
    }
    __syncthreads();
    const int sLo = uniform ? s0 : range[0];
    const int span = uniform ? 1 : range[1] - range[0] + 1;
    const bool local = span >= 1 && span <= kLocalSegs;
    if (local) {
        for (int k = static_cast<int>(threadIdx.x); k < span * kCell; k += kB) {
            const int w = k % kBinWords;
            lds[k] = w == 0 ? 0u : (w < 4 ? 0xFFFFFFFFu : 0u);
        }
    }
    __syncthreads();
    // every shuffle runs with the whole wave active (a shuffle under a lane-dependent condition
    // would read lanes that did not take part)
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int sPrev = __shfl_up(s, 1, kWave);
    const int bPrev = __shfl_up(b, 1, kWave);
    int run = (lane == 0 || sPrev != s || bPrev != b) ? lane : 0;
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(run, d, kWave);
        if (lane >= d) {
            run = max(run, o);
        }
    }
    for (int d = 1; d < kWave; d <<= 1) {
        const int runPrev = __shfl_up(run, d, kWave);
        const bool take = lane >= d && runPrev == run;
        for (int k = 0; k < kBinWords; ++k) {
            const uint32_t o = __shfl_up(v[k], d, kWave);
            if (take) {
                v[k] = k == 0 ? v[k] + o : (k < 4 ? min(v[k], o) : max(v[k], o));
            }
        }
    }
    const int sNext = __shfl_down(s, 1, kWave);
    const int bNext = __shfl_down(b, 1, kWave);
    if (b >= 0 && (lane == kWave - 1 || sNext != s || bNext != b)) {
        uint32_t* dst = local ? lds + (s - sLo) * kCell + b * kBinWords
                              : bins + static_cast<size_t>(kCell) * s + b * kBinWords;
        atomicAdd(&dst[0], v[0]);
        atomicMin(&dst[1], v[1]);
        atomicMin(&dst[2], v[2]);
        atomicMin(&dst[3], v[3]);
        atomicMax(&dst[4], v[4]);
        atomicMax(&dst[5], v[5]);
        atomicMax(&dst[6], v[6]);
    }
    __syncthreads();
    if (local) {
        // flush the cells this block filled (an empty cell's words would not change the record)
        uint32_t* g = bins + static_cast<size_t>(kCell) * sLo;
        for (int k = static_cast<int>(threadIdx.x); k < span * kCell; k += kB) {
            const int w = k % kBinWords;
            if (lds[k - w] != 0u) {
                flush_cell_word(&g[k], w, lds[k]);
            }
        }
    }
}

__device__ __forceinline__ float half_area(float x0, float y0, float z0, float x1, float y1, float z1)
{
    const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    return dx * dy + dy * dz + dz * dx;
}

// One thread per internal segment: lowest-cost bin boundary (left-to-right sweep of prefix areas,
// right-to-left of suffix areas); the middle position when no boundary separates the segment.
// Also clears internalFlag past the level's segments, so the scan over all maxSegs entries counts
// this level's internal segments only.
__global__ void split_kernel(const SegRec* __restrict__ segs, const SahCtl* __restrict__ ctl, int maxSegs,
                             const uint32_t* __restrict__ bins, SplitRec* __restrict__ split,
                             int* __restrict__ internalFlag)
{
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= ctl->segCount) {
        if (j < maxSegs) {
            internalFlag[j] = 0;
        }
        return;
    }
    SplitRec r = split[j];
    internalFlag[j] = r.internal;
    if (r.internal == 0) {
        return;
    }
    const int count = segs[j].count;
    if (r.axis >= 0) {
        const uint32_t* b = bins + static_cast<size_t>(kBins * kBinWords) * j;
        float suffix[kBins];
        int suffixN[kBins];
        float x0 = INFINITY, y0 = INFINITY, z0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, z1 = -INFINITY;
        int acc = 0;
        for (int k = kBins - 1; k >= 0; --k) {
            const uint32_t* e = b + k * kBinWords;
            if (e[0] != 0u) {
                x0 = fminf(x0, unsortable(e[1]));
                y0 = fminf(y0, unsortable(e[2]));
                z0 = fminf(z0, unsortable(e[3]));
                x1 = fmaxf(x1, unsortable(e[4]));
                y1 = fmaxf(y1, unsortable(e[5]));
                z1 = fmaxf(z1, unsortable(e[6]));
                acc += static_cast<int>(e[0]);
            }
            suffixN[k] = acc;
            suffix[k] = acc > 0 ? half_area(x0, y0, z0, x1, y1, z1) : 0.0f;
        }
        x0 = y0 = z0 = INFINITY;
        x1 = y1 = z1 = -INFINITY;
        acc = 0;
        float best = INFINITY;
        int bestBin = -1, bestLeft = 0;
        for (int k = 0; k < kBins - 1; ++k) {
            const uint32_t* e = b + k * kBinWords;
            if (e[0] != 0u) {
                x0 = fminf(x0, unsortable(e[1]));
                y0 = fminf(y0, unsortable(e[2]));
                z0 = fminf(z0, unsortable(e[3]));
                x1 = fmaxf(x1, unsortable(e[4]));
                y1 = fmaxf(y1, unsortable(e[5]));
                z1 = fmaxf(z1, unsortable(e[6]));
                acc += static_cast<int>(e[0]);
            }
            const int right = suffixN[k + 1];
            if (acc == 0 || right == 0) {
                continue;
            }
            const float cost = half_area(x0, y0, z0, x1, y1, z1) * static_cast<float>(acc) +
                               suffix[k + 1] * static_cast<float>(right);
            if (cost < best) {
                best = cost;
                bestBin = k;
                bestLeft = acc;
            }
        }
        if (bestBin >= 0) {
            r.splitBin = bestBin;
            r.leftCount = bestLeft;
        } else {
            r.axis = -1;
        }
    }
    if (r.axis < 0) {
        r.splitBin = -1;
        r.leftCount = count / 2;
    }
    split[j] = r;
}

// Internal nodes of this level get consecutive indices; each emits its two child segments.
__global__ void emit_kernel(const SegRec* __restrict__ segs, const SahCtl* __restrict__ ctl,
                            const SplitRec* __restrict__ split, const int* __restrict__ rank, float* __restrict__ nodes,
                            float* __restrict__ rootBox, SegRec* __restrict__ next)
{
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (ctl->internals == 0 || j >= ctl->segCount || split[j].internal == 0) {
        return;
    }
    const SegRec sg = segs[j];
    const int idx = ctl->levelBase + rank[j];
    if (sg.parent < 0) {
        rootBox[6] = __int_as_float(idx);
    } else {
        nodes[16 * static_cast<size_t>(sg.parent) + 12 + sg.side] = __int_as_float(idx);
    }
    const int lc = split[j].leftCount;
    next[2 * rank[j]] = SegRec{sg.begin, lc, idx, 0};
    next[2 * rank[j] + 1] = SegRec{sg.begin + lc, sg.count - lc, idx, 1};
}

__global__ void flag_kernel(const int* __restrict__ perm, const int* __restrict__ posSeg, const SegRec* __restrict__ segs,
                            const SplitRec* __restrict__ split, const float4* __restrict__ cent, int n,
                            int* __restrict__ flags, const SahCtl* __restrict__ ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || ctl->segCount == 0) {
        return;
    }
    const int s = posSeg[i];
    int f = 0;
    if (s >= 0) {
        const SplitRec r = split[s];
        if (r.internal != 0) {
            f = r.axis >= 0 ? (bin_of(r, cent[perm[i]]) <= r.splitBin ? 1 : 0)
                            : (i - segs[s].begin < r.leftCount ? 1 : 0);
        }
    }
    flags[i] = f;
}

__global__ void sah_scatter_kernel(const int* __restrict__ perm, const int* __restrict__ posSeg,
                                   const SegRec* __restrict__ segs, const SplitRec* __restrict__ split,
                                   const int* __restrict__ rank, const int* __restrict__ flags,
                                   const int* __restrict__ scan, int n, int* __restrict__ permOut,
                                   int* __restrict__ posSegOut, const SahCtl* __restrict__ ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || ctl->segCount == 0) {
        return;
    }
    const int s = posSeg[i];
    if (s < 0 || ctl->internals == 0 || split[s].internal == 0) {
        permOut[i] = perm[i];
        posSegOut[i] = -1;
        return;
    }
    const int begin = segs[s].begin;
    const int lc = split[s].leftCount;
    const int before = scan[i] - scan[begin];
    const bool left = flags[i] != 0;
    const int pos = left ? begin + before : begin + lc + (i - begin - before);
    permOut[pos] = perm[i];
    posSegOut[pos] = 2 * rank[s] + (left ? 0 : 1);
}

// Start of a build: one root segment holding every triangle, empty scene bounds.
__global__ void sah_init_kernel(SahCtl* __restrict__ ctl, SegRec* __restrict__ segs, uint32_t* __restrict__ gcb, int n)
{
    if (threadIdx.x == 0) {
        ctl->segCount = 1;
        ctl->levelBase = 0;
        ctl->level = 0;
        ctl->internals = 0;
        ctl->error = 0;
        ctl->wideCount = 0;
        segs[0] = SegRec{0, n, -1, 0};
    }
    if (threadIdx.x < 6) {
        gcb[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
    }
}

// After the scan of internalFlag: the level's internal segments, and the first node number of a
// level that has any.
__global__ void level_tail_kernel(SahCtl* __restrict__ ctl, const int* __restrict__ rank,
                                  const int* __restrict__ internalFlag, int maxSegs)
{
    const int sc = ctl->segCount;
    int internals = sc > 0 ? rank[sc - 1] + internalFlag[sc - 1] : 0;
    if (2 * static_cast<long long>(internals) > maxSegs) {
        ctl->error = 1;
        internals = 0;
    }
    ctl->internals = internals;
    if (internals > 0) {
        ctl->levelStart[ctl->level] = ctl->levelBase;
    }
}

__global__ void level_advance_kernel(SahCtl* __restrict__ ctl)
{
    const int internals = ctl->internals;
    if (internals > 0) {
        ctl->levelBase += internals;
        ctl->level += 1;
    }
    ctl->segCount = 2 * internals;
    ctl->internals = 0;
}

// First 4-wide node of every level (the even levels hold the wide nodes) from the binary tree's
// level starts, with internalCount as the end sentinel.
__global__ void wide_starts_kernel(SahCtl* __restrict__ ctl, int internalCount, int wideCapacity)
{
    const int levels = ctl->level;
    ctl->levelStart[levels] = internalCount;
    int count = 0;
    for (int L = 0; L < levels; ++L) {
        ctl->wideStart[L] = count;
        if ((L & 1) == 0) {
            count += ctl->levelStart[L + 1] - ctl->levelStart[L];
        }
    }
    ctl->wideStart[levels] = count;
    ctl->wideCount = count;
    if (count > wideCapacity) {
        ctl->error = 2;
    }
}

__global__ void fill_int_kernel(int* p, int n, int v)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        p[i] = v;
    }
}

__global__ void fill_u32_kernel(uint32_t* p, int n, uint32_t v0, uint32_t v1, int split)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        p[i] = i < split ? v0 : v1;
    }
}

unsigned grid(long long n) { return static_cast<unsigned>((n + kB - 1) / kB); }
unsigned chunks(long long n, int chunk) { return static_cast<unsigned>((n + kB * chunk - 1) / (kB * chunk)); }

template <typename T>
bool dalloc(T** p, size_t count)
{
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)) == hipSuccess;
}

template <typename T>
void dfree(T*& p)
{
    if (p != nullptr) {
        (void)hipFree(p);
        p = nullptr;
    }
}

} // namespace


// ---- 4-wide tree -------------------------------------------------------------------------
//
// Collapses the level-numbered binary tree: every internal node on an even level becomes a wide
// node holding up to four children, its grandchildren where its children are internal and its
// children where they are leaves. The boxes are the binary tree's own floats, so a traversal's
// box tests give the same entry distances. Wide node (8 float4, 7 used): min x / min y / min z /
// max x / max y / max z of the four children, then their references (int bits: >= 0 wide node,
// < 0 leaf slot ~ref, kWideEmpty unused).

namespace {

constexpr int kWideEmpty = static_cast<int>(0x80000000u);

struct ChildBox {
    float mn[3], mx[3];
    int ref;
};

__device__ void binary_children(const float4* __restrict__ nodes, int node, ChildBox out[2])
{
    const float4 n0 = nodes[4 * node + 0];
    const float4 n1 = nodes[4 * node + 1];
    const float4 n2 = nodes[4 * node + 2];
    const float4 n3 = nodes[4 * node + 3];
    out[0] = ChildBox{{n0.x, n0.y, n0.z}, {n0.w, n1.x, n1.y}, __float_as_int(n3.x)};
    out[1] = ChildBox{{n1.z, n1.w, n2.x}, {n2.y, n2.z, n2.w}, __float_as_int(n3.y)};
}

// levelStart[L]: first internal node of level L (L < levels; levelStart[levels] = internal count);
// wideStart[L]: first wide node made from level L (even L).
__global__ void wide_collapse_kernel(const float4* __restrict__ nodes, int internalCount,
                                     const int* __restrict__ levelStart, const int* __restrict__ wideStart,
                                     const int* __restrict__ levelsPtr, const int* __restrict__ errorPtr,
                                     float4* __restrict__ wide)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= internalCount || (errorPtr != nullptr && *errorPtr != 0)) {
        return;
    }
    const int levels = *levelsPtr;
    int level = 0;
    while (level + 1 < levels && levelStart[level + 1] <= node) {
        ++level;
    }
    if ((level & 1) != 0) {
        return;
    }
    ChildBox kids[2];
    binary_children(nodes, node, kids);
    ChildBox c[4];
    int k = 0;
    for (int side = 0; side < 2; ++side) {
        if (kids[side].ref >= 0) {
            ChildBox g[2];
            binary_children(nodes, kids[side].ref, g);
            for (int q = 0; q < 2; ++q) {
                if (g[q].ref >= 0) {  // internal on level + 2: a wide node
                    g[q].ref = wideStart[level + 2] + (g[q].ref - levelStart[level + 2]);
                }
                c[k++] = g[q];
            }
        } else {
            c[k++] = kids[side];
        }
    }
    for (; k < 4; ++k) {
        c[k] = ChildBox{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, kWideEmpty};
    }
    const int w = wideStart[level] + (node - levelStart[level]);
    float4* o = wide + 8 * static_cast<size_t>(w);
    o[0] = make_float4(c[0].mn[0], c[1].mn[0], c[2].mn[0], c[3].mn[0]);
    o[1] = make_float4(c[0].mn[1], c[1].mn[1], c[2].mn[1], c[3].mn[1]);
    o[2] = make_float4(c[0].mn[2], c[1].mn[2], c[2].mn[2], c[3].mn[2]);
    o[3] = make_float4(c[0].mx[0], c[1].mx[0], c[2].mx[0], c[3].mx[0]);
    o[4] = make_float4(c[0].mx[1], c[1].mx[1], c[2].mx[1], c[3].mx[1]);
    o[5] = make_float4(c[0].mx[2], c[1].mx[2], c[2].mx[2], c[3].mx[2]);
    o[6] = make_float4(__int_as_float(c[0].ref), __int_as_float(c[1].ref), __int_as_float(c[2].ref),
                       __int_as_float(c[3].ref));
    o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

} // namespace

struct PtSahWork {
    int capacity = 0;  // triangles
    float4* bounds = nullptr;
    float4* cent = nullptr;
    uint32_t* gcb = nullptr;
    uint32_t* keys = nullptr;
    uint32_t* keysOut = nullptr;
    int* vals = nullptr;
    int* perm = nullptr;
    int* permAlt = nullptr;
    int* posSeg = nullptr;
    int* posSegAlt = nullptr;
    SegRec* segs = nullptr;
    SegRec* segsNext = nullptr;
    SplitRec* split = nullptr;
    uint32_t* segStats = nullptr;
    uint32_t* bins = nullptr;
    int* internalFlag = nullptr;
    int* rank = nullptr;
    int* flags = nullptr;
    int* scan = nullptr;
    float* rootBox = nullptr;
    SahCtl* ctl = nullptr;
    void* tmp = nullptr;
    size_t tmpBytes = 0;
};

void pt_sah_work_destroy(PtSahWork* w)
{
    if (w == nullptr) {
        return;
    }
    dfree(w->bounds);
    dfree(w->cent);
    dfree(w->gcb);
    dfree(w->keys);
    dfree(w->keysOut);
    dfree(w->vals);
    dfree(w->perm);
    dfree(w->permAlt);
    dfree(w->posSeg);
    dfree(w->posSegAlt);
    dfree(w->segs);
    dfree(w->segsNext);
    dfree(w->split);
    dfree(w->segStats);
    dfree(w->bins);
    dfree(w->internalFlag);
    dfree(w->rank);
    dfree(w->flags);
    dfree(w->scan);
    dfree(w->rootBox);
    dfree(w->ctl);
    if (w->tmp != nullptr) {
        (void)hipFree(w->tmp);
    }
    delete w;
}

PtSahWork* pt_sah_work_create(int capacity)
{
    if (capacity < 1) {
        ptSetError("SAH build workspace: capacity %d < 1", capacity);
        return nullptr;
    }
    PtSahWork* w = new PtSahWork;
    w->capacity = capacity;
    const size_t n = static_cast<size_t>(capacity);
    const size_t maxSegs = n;  // segments of a level <= triangles
    if (!dalloc(&w->bounds, 2 * n) || !dalloc(&w->cent, n) || !dalloc(&w->gcb, 6) || !dalloc(&w->keys, n) ||
        !dalloc(&w->keysOut, n) || !dalloc(&w->vals, n) || !dalloc(&w->perm, n) || !dalloc(&w->permAlt, n) ||
        !dalloc(&w->posSeg, n) || !dalloc(&w->posSegAlt, n) || !dalloc(&w->segs, maxSegs) ||
        !dalloc(&w->segsNext, maxSegs) || !dalloc(&w->split, maxSegs) ||
        !dalloc(&w->segStats, static_cast<size_t>(kSegWords) * maxSegs) ||
        !dalloc(&w->bins, static_cast<size_t>(kBins * kBinWords) * maxSegs) || !dalloc(&w->internalFlag, maxSegs) ||
        !dalloc(&w->rank, maxSegs) || !dalloc(&w->flags, n) || !dalloc(&w->scan, n) || !dalloc(&w->rootBox, 8) ||
        !dalloc(&w->ctl, 1)) {
        ptSetError("SAH build: device allocation failed");
        pt_sah_work_destroy(w);
        return nullptr;
    }
    size_t sortBytes = 0, scanBytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, w->keys, w->keysOut, w->vals, w->perm, capacity, 0,
                                             30);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, w->flags, w->scan, capacity);
    w->tmpBytes = std::max(sortBytes, scanBytes);
    if (hipMalloc(&w->tmp, std::max<size_t>(w->tmpBytes, 1)) != hipSuccess) {
        ptSetError("SAH build: temporary allocation failed");
        pt_sah_work_destroy(w);
        return nullptr;
    }
    return w;
}

bool pt_sah_enqueue(PtSahWork* w, const float* dTris, int n, float4* nodes, float4* triPos, float4* triShade,
                    float4* wide, int wideCapacity, int levels, hipStream_t stream)
{
    const int levelCount = levels > 0 ? std::min(levels, kSahLevels) : kSahLevels;
    if (w == nullptr || n < 1 || n > w->capacity) {
        ptSetError("SAH build: %d triangles for a workspace of %d", n, w != nullptr ? w->capacity : 0);
        return false;
    }
    size_t sortBytes = 0, scanBytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, w->keys, w->keysOut, w->vals, w->perm, n, 0, 30,
                                             stream);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, w->flags, w->scan, n, stream);
    if (std::max(sortBytes, scanBytes) > w->tmpBytes) {
        ptSetError("SAH build: temporary storage too small");
        return false;
    }
    const int maxSegs = n;
    int* perm = w->perm;
    int* permAlt = w->permAlt;
    int* posSeg = w->posSeg;
    int* posSegAlt = w->posSegAlt;
    SegRec* segs = w->segs;
    SegRec* segsNext = w->segsNext;
    SahCtl* ctl = w->ctl;
    hipLaunchKernelGGL(sah_init_kernel, dim3(1), dim3(64), 0, stream, ctl, segs, w->gcb, n);
    hipLaunchKernelGGL(sah_prep_kernel, dim3(std::min(grid(n), kPrepBlocks)), dim3(kB), 0, stream, dTris, n,
                       w->bounds, w->cent, triShade, w->gcb);
    hipLaunchKernelGGL(morton_kernel, dim3(grid(n)), dim3(kB), 0, stream, w->cent, n, w->gcb, w->keys, w->vals);
    size_t bytes = w->tmpBytes;
    if (hipcub::DeviceRadixSort::SortPairs(w->tmp, bytes, w->keys, w->keysOut, w->vals, perm, n, 0, 30, stream) !=
        hipSuccess) {
        ptSetError("SAH build: radix sort failed");
        return false;
    }
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid(n)), dim3(kB), 0, stream, posSeg, n, 0);
    bool ok = true;
    // every level the depth cap allows (or the given count); levels past the last one find no
    // segments and do nothing
    for (int level = 0; level < levelCount && ok; ++level) {
        hipLaunchKernelGGL(seg_reset_kernel, dim3(std::min(grid(static_cast<long long>(maxSegs) * kSegWords), kPrepBlocks)),
                           dim3(kB), 0, stream, w->segStats, ctl);
        hipLaunchKernelGGL(seg_bounds_kernel, dim3(chunks(n, kChunk)), dim3(kB), 0, stream, perm, posSeg, w->bounds,
                           w->cent, n, w->segStats, ctl, segs);
        hipLaunchKernelGGL(seg_setup_kernel, dim3(grid(maxSegs)), dim3(kB), 0, stream, segs, ctl, w->segStats, perm,
                           dTris, level, reinterpret_cast<float*>(nodes), triPos, w->rootBox, w->split, w->bins);
        hipLaunchKernelGGL(bin_kernel, dim3(grid(n)), dim3(kB), 0, stream, perm, posSeg, w->bounds, w->cent, w->split,
                           n, w->bins, ctl);
        hipLaunchKernelGGL(split_kernel, dim3(grid(maxSegs)), dim3(kB), 0, stream, segs, ctl, maxSegs, w->bins,
                           w->split, w->internalFlag);
        bytes = w->tmpBytes;
        ok = hipcub::DeviceScan::ExclusiveSum(w->tmp, bytes, w->internalFlag, w->rank, maxSegs, stream) == hipSuccess;
        hipLaunchKernelGGL(level_tail_kernel, dim3(1), dim3(1), 0, stream, ctl, w->rank, w->internalFlag, maxSegs);
        hipLaunchKernelGGL(emit_kernel, dim3(grid(maxSegs)), dim3(kB), 0, stream, segs, ctl, w->split, w->rank,
                           reinterpret_cast<float*>(nodes), w->rootBox, segsNext);
        hipLaunchKernelGGL(flag_kernel, dim3(grid(n)), dim3(kB), 0, stream, perm, posSeg, segs, w->split, w->cent, n,
                           w->flags, ctl);
        bytes = w->tmpBytes;
        ok = ok && hipcub::DeviceScan::ExclusiveSum(w->tmp, bytes, w->flags, w->scan, n, stream) == hipSuccess;
        hipLaunchKernelGGL(sah_scatter_kernel, dim3(grid(n)), dim3(kB), 0, stream, perm, posSeg, segs, w->split,
                           w->rank, w->flags, w->scan, n, permAlt, posSegAlt, ctl);
        hipLaunchKernelGGL(level_advance_kernel, dim3(1), dim3(1), 0, stream, ctl);
        std::swap(perm, permAlt);
        std::swap(posSeg, posSegAlt);
        std::swap(segs, segsNext);
    }
    if (ok && wide != nullptr && n > 1) {
        hipLaunchKernelGGL(wide_starts_kernel, dim3(1), dim3(1), 0, stream, ctl, n - 1, wideCapacity);
        // device addresses of the control block's fields (host pointer arithmetic only)
        char* c = reinterpret_cast<char*>(ctl);
        hipLaunchKernelGGL(wide_collapse_kernel, dim3(grid(n - 1)), dim3(kB), 0, stream,
                           reinterpret_cast<const float4*>(nodes), n - 1,
                           reinterpret_cast<const int*>(c + offsetof(SahCtl, levelStart)),
                           reinterpret_cast<const int*>(c + offsetof(SahCtl, wideStart)),
                           reinterpret_cast<const int*>(c + offsetof(SahCtl, level)),
                           reinterpret_cast<const int*>(c + offsetof(SahCtl, error)), wide);
    }
    if (!ok || hipGetLastError() != hipSuccess) {
        ptSetError("SAH build: kernel launch failed");
        return false;
    }
    return true;
}

namespace {

// Checks a host copy of a build's control block: false with the error set unless it describes
// a finished tree of n - 1 internal nodes.
bool sah_check(const SahCtl& ctl, int n)
{
    if (ctl.error == 3) {
        ptSetError("no rebuild of this scene has run since its build or its last check");
        return false;
    }
    if (ctl.error == 1) {
        ptSetError("SAH build: segment table overflow");
        return false;
    }
    if (ctl.error == 2) {
        ptSetError("SAH build: the 4-wide tree (%d nodes) does not fit its buffer", ctl.wideCount);
        return false;
    }
    if (ctl.levelBase != n - 1 || ctl.segCount != 0) {
        ptSetError("SAH build: %d internal nodes for %d triangles", ctl.levelBase, n);
        return false;
    }
    return true;
}

constexpr int kCtlWords = static_cast<int>(sizeof(SahCtl) / sizeof(int));

__global__ void sah_snapshot_kernel(const int* __restrict__ ctl, const int* __restrict__ box, int* __restrict__ out)
{
    for (int i = threadIdx.x; i < kCtlWords + 8; i += blockDim.x) {
        out[i] = i < kCtlWords ? ctl[i] : (i - kCtlWords < 7 ? box[i - kCtlWords] : 0);
    }
}

} // namespace

size_t pt_sah_status_words() { return static_cast<size_t>(kCtlWords) + 8; }

bool pt_sah_status_consume(int* dStatus)
{
    // the control block's error word 3: "no rebuild since" (a rebuild's snapshot overwrites it)
    const int none = 3;
    return hipMemcpy(dStatus + offsetof(SahCtl, error) / sizeof(int), &none, sizeof(int), hipMemcpyHostToDevice) ==
           hipSuccess;
}

bool pt_sah_snapshot(PtSahWork* w, int* dStatus, hipStream_t stream)
{
    if (w == nullptr || dStatus == nullptr) {
        ptSetError("SAH build: no status buffer");
        return false;
    }
    hipLaunchKernelGGL(sah_snapshot_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<const int*>(w->ctl),
                       reinterpret_cast<const int*>(w->rootBox), dStatus);
    return hipGetLastError() == hipSuccess;
}

bool pt_sah_status_result(const int* dStatus, int n, hipStream_t stream, float rootOut[7], int* levelsOut,
                          int* wideCount)
{
    std::vector<int> words(pt_sah_status_words());
    if (hipMemcpyAsync(words.data(), dStatus, words.size() * sizeof(int), hipMemcpyDeviceToHost, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        ptSetError("SAH build: synchronisation failed");
        return false;
    }
    SahCtl ctl{};
    std::memcpy(&ctl, words.data(), sizeof(ctl));
    if (!sah_check(ctl, n)) {
        return false;
    }
    if (rootOut != nullptr) {
        std::memcpy(rootOut, words.data() + kCtlWords, 7 * sizeof(float));
    }
    if (levelsOut != nullptr) {
        *levelsOut = ctl.level + 1;
    }
    if (wideCount != nullptr) {
        *wideCount = ctl.wideCount;
    }
    return true;
}

bool pt_sah_result(PtSahWork* w, int n, hipStream_t stream, float rootOut[7], int* levelsOut,
                   std::vector<int>* levelStarts, int* wideCount)
{
    SahCtl ctl{};
    if (hipMemcpyAsync(&ctl, w->ctl, sizeof(ctl), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        (rootOut != nullptr &&
         hipMemcpyAsync(rootOut, w->rootBox, 7 * sizeof(float), hipMemcpyDeviceToHost, stream) != hipSuccess) ||
        hipStreamSynchronize(stream) != hipSuccess) {
        ptSetError("SAH build: synchronisation failed");
        return false;
    }
    if (!sah_check(ctl, n)) {
        return false;
    }
    if (levelsOut != nullptr) {
        *levelsOut = ctl.level + 1;
    }
    if (levelStarts != nullptr) {
        levelStarts->assign(ctl.levelStart, ctl.levelStart + ctl.level);
    }
    if (wideCount != nullptr) {
        *wideCount = ctl.wideCount;
    }
    return true;
}

bool pt_build_bvh_sah(const float* dTris, int n, float4* nodes, float4* triPos, float4* triShade, hipStream_t stream,
                      float rootOut[7], int* levelsOut, std::vector<int>* levelStarts)
{
    PtSahWork* w = pt_sah_work_create(n);
    if (w == nullptr) {
        return false;
    }
    const bool ok = pt_sah_enqueue(w, dTris, n, nodes, triPos, triShade, nullptr, 0, 0, stream) &&
                    pt_sah_result(w, n, stream, rootOut, levelsOut, levelStarts, nullptr);
    pt_sah_work_destroy(w);
    return ok;
}

bool pt_build_wide(const float4* nodes, int internalCount, const std::vector<int>& levelStarts, float4** wideOut,
                   int* wideCount, hipStream_t stream)
{
    *wideOut = nullptr;
    *wideCount = 0;
    if (internalCount <= 0) {
        return true;  // a single leaf (or nothing): no internal node to collapse
    }
    const int levels = static_cast<int>(levelStarts.size());
    std::vector<int> starts(levelStarts);
    starts.push_back(internalCount);
    std::vector<int> wideStart(static_cast<size_t>(levels) + 1, 0);
    int count = 0;
    for (int L = 0; L < levels; ++L) {
        wideStart[static_cast<size_t>(L)] = count;
        if ((L & 1) == 0) {
            count += starts[static_cast<size_t>(L) + 1] - starts[static_cast<size_t>(L)];
        }
    }
    wideStart[static_cast<size_t>(levels)] = count;
    wideStart.push_back(levels);  // read by the kernel through a device pointer
    int* dStarts = nullptr;
    int* dWide = nullptr;
    float4* wide = nullptr;
    bool ok = hipMalloc(&dStarts, starts.size() * sizeof(int)) == hipSuccess &&
              hipMalloc(&dWide, wideStart.size() * sizeof(int)) == hipSuccess &&
              hipMalloc(&wide, 8 * static_cast<size_t>(count) * sizeof(float4)) == hipSuccess &&
              hipMemcpyAsync(dStarts, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice, stream) ==
                  hipSuccess &&
              hipMemcpyAsync(dWide, wideStart.data(), wideStart.size() * sizeof(int), hipMemcpyHostToDevice,
                             stream) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(wide_collapse_kernel, dim3(static_cast<unsigned>((internalCount + 255) / 256)), dim3(256), 0,
                           stream, nodes, internalCount, dStarts, dWide, dWide + levels + 1, nullptr, wide);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(stream) == hipSuccess;
    }
    if (dStarts != nullptr) {
        (void)hipFree(dStarts);
    }
    if (dWide != nullptr) {
        (void)hipFree(dWide);
    }
    if (!ok) {
        if (wide != nullptr) {
            (void)hipFree(wide);
        }
        ptSetError("wide BVH collapse failed");
        return false;
    }
    *wideOut = wide;
    *wideCount = count;
    return true;
}
