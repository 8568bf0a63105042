/*
 * orc_internal.h -- oracle vector helpers (Geometry/MathCore.h:13-115 restated).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef PT_ORC_INTERNAL_H
#define PT_ORC_INTERNAL_H

#include "oracle.h"

#include <math.h>

double orc_sin_d(double x);
double orc_cos_d(double x);
double orc_atan2_d(double y, double x);

/* Geometry/MathCore.h:13-16 */
static inline OVec3 v3(float x, float y, float z)
{
    OVec3 r = {x, y, z};
    return r;
}
/* MathCore.h:23-26 */
static inline float vdot(OVec3 a, OVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* MathCore.h:28-31 */
static inline float vlen(OVec3 v) { return sqrtf(vdot(v, v)); }
static inline OVec3 vadd(OVec3 a, OVec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline OVec3 vsub(OVec3 a, OVec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline OVec3 vscale(OVec3 v, float s) { return v3(v.x * s, v.y * s, v.z * s); }
static inline OVec3 vmul(OVec3 a, OVec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
/* MathCore.h:53-61: zero-length input returns (0, 0, -1). */
static inline OVec3 vnorm(OVec3 v)
{
    const float len = vlen(v);
    if (len <= 0.0f) {
        return v3(0.0f, 0.0f, -1.0f);
    }
    const float invLen = 1.0f / len;
    return vscale(v, invLen);
}
/* MathCore.h:63-66 */
static inline OVec3 vray(OVec3 ro, OVec3 rd, float t) { return vadd(ro, vscale(rd, t)); }
/* MathCore.h:68-76: comparisons, not fmaxf/fminf. */
static inline float fmax2(float a, float b) { return a > b ? a : b; }
static inline float fmin2(float a, float b) { return a < b ? a : b; }
static inline float fmax3(float a, float b, float c) { return fmax2(fmax2(a, b), c); }
static inline float fmin3(float a, float b, float c) { return fmin2(fmin2(a, b), c); }
static inline float fabs1(float v) { return v < 0.0f ? -v : v; }
/* MathCore.h:98-104 */
static inline OVec3 vcross(OVec3 a, OVec3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float clamp01f(float v) { return fmax2(0.0f, fmin2(1.0f, v)); }

#endif
