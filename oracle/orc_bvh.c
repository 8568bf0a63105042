/*
 * orc_bvh.c -- oracle BVH build, ray/box, ray/triangle and closest-hit traversal.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "orc_internal.h"

#include <stdlib.h>
#include <string.h>

/* MeshBvhBuilder.cpp:10-13 */
static OVec3 tri_centroid(const OTriangle* t)
{
    return vscale(vadd(vadd(t->v0, t->v1), t->v2), 1.0f / 3.0f);
}

/* MeshBvhBuilder.cpp:15-29 */
static OVec3 tri_min(const OTriangle* t)
{
    return v3(fmin3(t->v0.x, t->v1.x, t->v2.x), fmin3(t->v0.y, t->v1.y, t->v2.y), fmin3(t->v0.z, t->v1.z, t->v2.z));
}
static OVec3 tri_max(const OTriangle* t)
{
    return v3(fmax3(t->v0.x, t->v1.x, t->v2.x), fmax3(t->v0.y, t->v1.y, t->v2.y), fmax3(t->v0.z, t->v1.z, t->v2.z));
}

/* MeshBvhBuilder.cpp:31-40: ties prefer x, then y. */
static int longest_axis(OVec3 e)
{
    if (e.x >= e.y && e.x >= e.z) {
        return 0;
    }
    if (e.y >= e.z) {
        return 1;
    }
    return 2;
}

typedef struct {
    const OTriangle* tris;
    const OVec3* centroids;
    int* indices;
    ONode* nodes;
    uint32_t nodeCount;
} BuildCtx;

static float axis_of(OVec3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

/* Total order used for the median split: centroid on the axis, then triangle index.
 * The reference's std::nth_element compares centroids only (MeshBvhBuilder.cpp:91-105);
 * its placement of exact ties is library-specific, so the index tie-break is this
 * build's specification of the same partition. */
static int less_idx(const BuildCtx* c, int axis, int a, int b)
{
    const float ka = axis_of(c->centroids[a], axis);
    const float kb = axis_of(c->centroids[b], axis);
    if (ka < kb) {
        return 1;
    }
    if (kb < ka) {
        return 0;
    }
    return a < b;
}

/* Quickselect: after return, idx[begin..k) are the k-begin smallest in the total order. */
static void select_kth(const BuildCtx* c, int axis, int* idx, int begin, int end, int k)
{
    while (end - begin > 1) {
        const int mid = begin + (end - begin) / 2;
        /* median of three pivot */
        int a = idx[begin], b = idx[mid], d = idx[end - 1];
        int pivot;
        if (less_idx(c, axis, a, b)) {
            pivot = less_idx(c, axis, b, d) ? b : (less_idx(c, axis, a, d) ? d : a);
        } else {
            pivot = less_idx(c, axis, a, d) ? a : (less_idx(c, axis, b, d) ? d : b);
        }
        /* three-way partition around pivot (elements are distinct in the total order) */
        int lo = begin, hi = end - 1, i = begin;
        while (i <= hi) {
            const int v = idx[i];
            if (less_idx(c, axis, v, pivot)) {
                idx[i] = idx[lo];
                idx[lo] = v;
                ++lo;
                ++i;
            } else if (less_idx(c, axis, pivot, v)) {
                idx[i] = idx[hi];
                idx[hi] = v;
                --hi;
            } else {
                ++i;
            }
        }
        /* [begin, lo) < pivot, [lo, hi] == pivot, (hi, end) > pivot */
        if (k < lo) {
            end = lo;
        } else if (k > hi + 1) {
            begin = hi + 1;
        } else {
            return;
        }
    }
}

/* Union min/max with -0 ordered below +0. The reference's vecMin2/vecMax2 loop
 * (MeshBvhBuilder.cpp:62-74) keeps whichever signed zero it meets first, and its visiting
 * order is nth_element's unspecified permutation; ordering -0 < +0 makes the union
 * independent of that order. Both zeros compare equal in every traversal test. */
static float umin(float a, float b) { return (a < b || (a == b && signbit(a))) ? a : b; }
static float umax(float a, float b) { return (a > b || (a == b && !signbit(a))) ? a : b; }

/* MeshBvhBuilder.cpp:52-113 */
static int build_recursive(BuildCtx* c, int begin, int end)
{
    const OTriangle* first = &c->tris[c->indices[begin]];
    OVec3 bmin = tri_min(first);
    OVec3 bmax = tri_max(first);
    for (int i = begin + 1; i < end; ++i) {
        const OTriangle* t = &c->tris[c->indices[i]];
        const OVec3 tmin = tri_min(t);
        const OVec3 tmax = tri_max(t);
        bmin.x = umin(bmin.x, tmin.x);
        bmin.y = umin(bmin.y, tmin.y);
        bmin.z = umin(bmin.z, tmin.z);
        bmax.x = umax(bmax.x, tmax.x);
        bmax.y = umax(bmax.y, tmax.y);
        bmax.z = umax(bmax.z, tmax.z);
    }

    const int nodeIndex = (int)c->nodeCount++;
    ONode* node = &c->nodes[nodeIndex];
    memset(node, 0, sizeof(*node));
    node->boundsMin = bmin;
    node->boundsMax = bmax;

    const int count = end - begin;
    if (count == 1) {
        node->triStart = (uint32_t)c->indices[begin];
        node->triCount = 1u;
        return nodeIndex;
    }

    const int axis = longest_axis(vsub(bmax, bmin));
    const int mid = begin + count / 2;
    select_kth(c, axis, c->indices, begin, end, mid);

    const int left = build_recursive(c, begin, mid);
    const int right = build_recursive(c, mid, end);
    c->nodes[nodeIndex].leftIndex = (uint32_t)left;
    c->nodes[nodeIndex].rightIndex = (uint32_t)right;
    c->nodes[nodeIndex].triCount = 0u;
    return nodeIndex;
}

/* MeshBvhBuilder.cpp:118-142. outNodes must hold 2*count-1 nodes. Returns node count (0 on failure). */
int orc_bvh_build(const OTriangle* tris, uint32_t count, ONode* outNodes, uint32_t* outRoot)
{
    *outRoot = 0;
    if (count == 0) {
        return 0;
    }
    BuildCtx c;
    c.tris = tris;
    c.nodes = outNodes;
    c.nodeCount = 0;
    c.indices = (int*)malloc(sizeof(int) * count);
    OVec3* cent = (OVec3*)malloc(sizeof(OVec3) * count);
    for (uint32_t i = 0; i < count; ++i) {
        c.indices[i] = (int)i;
        cent[i] = tri_centroid(&tris[i]);
    }
    c.centroids = cent;
    const int root = build_recursive(&c, 0, (int)count);
    free(cent);
    free(c.indices);
    *outRoot = (uint32_t)root;
    return (int)c.nodeCount;
}

/* MeshAccelIntersectCore.h:12-50 */
int orc_ray_aabb(OVec3 ro, OVec3 rd, OVec3 invRd, OVec3 bmin, OVec3 bmax, float tMin, float tMax,
                 float* outEnter, float* outExit)
{
    (void)rd;
    float tEnter = tMin;
    float tExit = tMax;
    for (int axis = 0; axis < 3; ++axis) {
        const float origin = axis == 0 ? ro.x : (axis == 1 ? ro.y : ro.z);
        const float invD = axis == 0 ? invRd.x : (axis == 1 ? invRd.y : invRd.z);
        const float lo = axis == 0 ? bmin.x : (axis == 1 ? bmin.y : bmin.z);
        const float hi = axis == 0 ? bmax.x : (axis == 1 ? bmax.y : bmax.z);
        float t0 = (lo - origin) * invD;
        float t1 = (hi - origin) * invD;
        if (invD < 0.0f) {
            const float tmp = t0;
            t0 = t1;
            t1 = tmp;
        }
        tEnter = fmax2(tEnter, t0);
        tExit = fmin2(tExit, t1);
        if (tEnter > tExit) {
            return 0;
        }
    }
    *outEnter = tEnter;
    *outExit = tExit;
    return tExit >= tMin;
}

/* MeshAccelIntersectCore.h:52-108 (two-sided Moller-Trumbore) */
int orc_ray_triangle(OVec3 ro, OVec3 rd, const OTriangle* tri, float tMin, float tMax, float* outT,
                     OVec3* outNormal, OVec2* outUv)
{
    const OVec3 e1 = vsub(tri->v1, tri->v0);
    const OVec3 e2 = vsub(tri->v2, tri->v0);
    const OVec3 pvec = v3(rd.y * e2.z - rd.z * e2.y, rd.z * e2.x - rd.x * e2.z, rd.x * e2.y - rd.y * e2.x);
    const float det = vdot(e1, pvec);
    if (fabsf(det) < 1.0e-8f) {
        return 0;
    }
    const float invDet = 1.0f / det;
    const OVec3 tvec = vsub(ro, tri->v0);
    const float baryU = vdot(tvec, pvec) * invDet;
    if (baryU < 0.0f || baryU > 1.0f) {
        return 0;
    }
    const OVec3 qvec = v3(tvec.y * e1.z - tvec.z * e1.y, tvec.z * e1.x - tvec.x * e1.z, tvec.x * e1.y - tvec.y * e1.x);
    const float baryV = vdot(rd, qvec) * invDet;
    if (baryV < 0.0f || baryU + baryV > 1.0f) {
        return 0;
    }
    const float t = vdot(e2, qvec) * invDet;
    if (t < tMin || t > tMax) {
        return 0;
    }
    const float baryW = 1.0f - baryU - baryV;
    *outT = t;
    OVec3 n = vnorm(vadd(vadd(vscale(tri->n0, baryW), vscale(tri->n1, baryU)), vscale(tri->n2, baryV)));
    if (vdot(n, rd) > 0.0f) {
        n = vscale(n, -1.0f);
    }
    *outNormal = n;
    outUv->x = tri->uv0.x * baryW + tri->uv1.x * baryU + tri->uv2.x * baryV;
    outUv->y = tri->uv0.y * baryW + tri->uv1.y * baryU + tri->uv2.y * baryV;
    return 1;
}

#define K_CULL_SLACK 1.000244140625f /* 1 + 2^-12 */

/* MeshAccelIntersectCore.h:110-221 */
OHit orc_trace_ray_counted(OVec3 ro, OVec3 rd, const OScene* scene, float tMin, float tMax,
                           uint64_t* nodeFetches, uint64_t* triTests)
{
    OHit result;
    memset(&result, 0, sizeof(result));
    if (scene == NULL || scene->nodes == NULL || scene->triangles == NULL || scene->nodeCount == 0 ||
        scene->triangleCount == 0) {
        return result;
    }

    rd = vnorm(rd);
    const OVec3 invRd = v3(fabsf(rd.x) > 1.0e-8f ? 1.0f / rd.x : 0.0f, fabsf(rd.y) > 1.0e-8f ? 1.0f / rd.y : 0.0f,
                           fabsf(rd.z) > 1.0e-8f ? 1.0f / rd.z : 0.0f);

    int stack[64];
    int stackSize = 0;
    stack[stackSize++] = (int)scene->rootIndex;

    float closestT = tMax;
    /* Boxes are culled against the closest hit widened by 2^-12 of its distance. The reference
     * culls against the closest hit itself, but the ray-triangle t can fall short of the entry
     * distance into the triangle's own box by more than rounding (thin triangles seen from afar),
     * so a box whose entry lies just past the current hit can hold a closer one; which of them the
     * reference reports then depends on its visit order. The margin keeps every such triangle in
     * reach, so the closest hit does not depend on the tree. */
    float cullT = tMax * K_CULL_SLACK;
    OVec3 closestNormal = v3(0.0f, 0.0f, 0.0f);
    OVec2 closestUv = {0.0f, 0.0f};
    uint32_t closestTri = 0xFFFFFFFFu;
    uint64_t nodes = 0, tests = 0;

    while (stackSize > 0) {
        const int nodeIndex = stack[--stackSize];
        if (nodeIndex < 0 || nodeIndex >= (int)scene->nodeCount) {
            continue;
        }
        const ONode* node = &scene->nodes[nodeIndex];
        ++nodes;
        float tEnter = 0.0f, tExit = 0.0f;
        if (!orc_ray_aabb(ro, rd, invRd, node->boundsMin, node->boundsMax, tMin, cullT, &tEnter, &tExit)) {
            continue;
        }
        if (node->triCount > 0) {
            const uint32_t end = node->triStart + node->triCount;
            for (uint32_t ti = node->triStart; ti < end; ++ti) {
                if (ti >= scene->triangleCount) {
                    break;
                }
                float tHit = 0.0f;
                OVec3 n;
                OVec2 uv;
                ++tests;
                /* Exact ties of t go to the smaller original index (the reference keeps the last
                 * tied triangle it visits, an order its nth_element tie placement leaves open,
                 * MeshBvhBuilder.cpp:91-105); every tree then gives the same closest hit. */
                if (orc_ray_triangle(ro, rd, &scene->triangles[ti], tMin, closestT, &tHit, &n, &uv) &&
                    (tHit < closestT || ti < closestTri)) {
                    closestT = tHit;
                    cullT = tHit * K_CULL_SLACK;
                    closestNormal = n;
                    closestUv = uv;
                    closestTri = ti;
                }
            }
            continue;
        }
        if (stackSize + 2 > 64) {
            continue;
        }
        const int left = (int)node->leftIndex;
        const int right = (int)node->rightIndex;
        float le = 0.0f, lx = 0.0f, re = 0.0f, rx = 0.0f;
        int leftHit = 0, rightHit = 0;
        if (left >= 0 && left < (int)scene->nodeCount) {
            const ONode* ln = &scene->nodes[left];
            ++nodes;
            leftHit = orc_ray_aabb(ro, rd, invRd, ln->boundsMin, ln->boundsMax, tMin, cullT, &le, &lx);
        }
        if (right >= 0 && right < (int)scene->nodeCount) {
            const ONode* rn = &scene->nodes[right];
            ++nodes;
            rightHit = orc_ray_aabb(ro, rd, invRd, rn->boundsMin, rn->boundsMax, tMin, cullT, &re, &rx);
        }
        if (leftHit && rightHit) {
            if (le < re) {
                stack[stackSize++] = right;
                stack[stackSize++] = left;
            } else {
                stack[stackSize++] = left;
                stack[stackSize++] = right;
            }
        } else if (leftHit) {
            stack[stackSize++] = left;
        } else if (rightHit) {
            stack[stackSize++] = right;
        }
    }

    if (nodeFetches) {
        *nodeFetches += nodes;
    }
    if (triTests) {
        *triTests += tests;
    }
    if (closestT < tMax) {
        result.hit = 1;
        result.t = closestT;
        result.normal = vnorm(closestNormal);
        result.uv = closestUv;
        result.triangleIndex = closestTri;
    }
    return result;
}

OHit orc_trace_ray(OVec3 ro, OVec3 rd, const OScene* scene, float tMin, float tMax)
{
    return orc_trace_ray_counted(ro, rd, scene, tMin, tMax, NULL, NULL);
}

/* MeshAccelScene.cpp:48-60 */
static float tri_area(const OTriangle* t)
{
    const OVec3 e1 = vsub(t->v1, t->v0);
    const OVec3 e2 = vsub(t->v2, t->v0);
    return 0.5f * vlen(vcross(e1, e2));
}

static float emission_luminance(const OMaterial* m)
{
    return m->emission * (0.2126f * m->r + 0.7152f * m->g + 0.0722f * m->b);
}

/* MeshAccelScene.cpp:62-95. outCdf holds count+1 floats. Returns the emissive count. */
uint32_t orc_build_emissive(const OTriangle* tris, uint32_t count, const OMaterial* mats, uint32_t matCount,
                            uint32_t* outIndices, float* outCdf)
{
    uint32_t n = 0;
    float total = 0.0f;
    outCdf[0] = 0.0f;
    for (uint32_t i = 0; i < count; ++i) {
        const OTriangle* t = &tris[i];
        if (t->materialIndex >= matCount) {
            continue;
        }
        const float lum = emission_luminance(&mats[t->materialIndex]);
        if (lum <= 0.0f) {
            continue;
        }
        const float w = tri_area(t) * lum;
        if (w <= 0.0f) {
            continue;
        }
        outIndices[n] = i;
        total += w;
        outCdf[n + 1] = total;
        ++n;
    }
    return n;
}
