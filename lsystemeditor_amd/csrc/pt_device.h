// pt_device.h -- gfx950 device code shared by the path-tracing kernels.
//
// Arithmetic follows the reference's FP32 operation order exactly (no contraction:
// the library is built with -ffp-contract=off) so that per-pixel results are
// reproducible bit for bit. Memory layout and traversal structure are this library's
// own (64-byte internal nodes with both child boxes, leaf-order triangle positions,
// deferred shading-attribute fetch).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pt_hip.h"

#define PT_DEV __device__ __forceinline__
#define PT_HD __host__ __device__ __forceinline__

namespace pt {

// ----------------------------------------------------------------------------------------
// Vector helpers (Geometry/MathCore.h:13-115)
// ----------------------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};

PT_DEV V3 mk(float x, float y, float z) { return V3{x, y, z}; }
PT_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_DEV float length(V3 v) { return sqrtf(dot(v, v)); }
PT_DEV V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_DEV V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_DEV V3 scale(V3 v, float s) { return mk(v.x * s, v.y * s, v.z * s); }
PT_DEV V3 mul(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_DEV V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// MathCore.h:53-61: zero-length input maps to (0, 0, -1)
PT_DEV V3 normalize(V3 v)
{
    const float len = length(v);
    if (len <= 0.0f) {
        return mk(0.0f, 0.0f, -1.0f);
    }
    const float inv = 1.0f / len;
    return scale(v, inv);
}
// Per-wavelength values of the four hero wavelengths of a spectral sample. Colour code that is
// generic over RGB (V3) and spectral (S4) values works per component (cget/cset).
struct S4 {
    float x, y, z, w;
};

PT_DEV S4 mk4(float x, float y, float z, float w) { return S4{x, y, z, w}; }
PT_DEV S4 add(S4 a, S4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
PT_DEV S4 scale(S4 v, float s) { return mk4(v.x * s, v.y * s, v.z * s, v.w * s); }
PT_DEV S4 mul(S4 a, S4 b) { return mk4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
PT_DEV S4 sub(S4 a, S4 b) { return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }

template <class C> struct ColN;
template <> struct ColN<V3> {
    static constexpr int n = 3;
    PT_DEV static V3 splat(float v) { return mk(v, v, v); }
};
template <> struct ColN<S4> {
    static constexpr int n = 4;
    PT_DEV static S4 splat(float v) { return mk4(v, v, v, v); }
};
PT_DEV float cget(const V3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
PT_DEV float cget(const S4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }
PT_DEV void cset(V3& v, int k, float f)
{
    if (k == 0) {
        v.x = f;
    } else if (k == 1) {
        v.y = f;
    } else {
        v.z = f;
    }
}
PT_DEV void cset(S4& v, int k, float f)
{
    if (k == 0) {
        v.x = f;
    } else if (k == 1) {
        v.y = f;
    } else if (k == 2) {
        v.z = f;
    } else {
        v.w = f;
    }
}

// MathCore.h:68-92 (comparison forms, not fmaxf/fminf)
PT_DEV float max2(float a, float b) { return a > b ? a : b; }
PT_DEV float min2(float a, float b) { return a < b ? a : b; }
PT_DEV float absf(float v) { return v < 0.0f ? -v : v; }
PT_DEV float clamp01(float v) { return max2(0.0f, min2(1.0f, v)); }

// ----------------------------------------------------------------------------------------
// Transcendentals. Evaluated in IEEE double and rounded once to float: Cody-Waite
// reduction by pi/2 (or ln 2) with a split constant, then the Taylor series of the
// function, which converges far below float resolution on the reduced interval. Only
// correctly rounded operations are used, so results are identical on every IEEE target.
// ----------------------------------------------------------------------------------------
namespace dm {
constexpr double kPio2Hi = 1.57079632673412561417e+00;
constexpr double kPio2Lo = 6.07710050650619224932e-11;
constexpr double kTwoOverPi = 6.36619772367581382433e-01;
constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kLn2 = 6.93147180559945286227e-01;
constexpr double kPi = 3.14159265358979311600e+00;
constexpr double kPio2 = 1.57079632679489655800e+00;

PT_HD double sin_series(double r)
{
    const double r2 = r * r;
    double p = -1.0 / 1307674368000.0;
    p = p * r2 + 1.0 / 6227020800.0;
    p = p * r2 + -1.0 / 39916800.0;
    p = p * r2 + 1.0 / 362880.0;
    p = p * r2 + -1.0 / 5040.0;
    p = p * r2 + 1.0 / 120.0;
    p = p * r2 + -1.0 / 6.0;
    return r + r * (r2 * p);
}

PT_HD double cos_series(double r)
{
    const double r2 = r * r;
    double p = 1.0 / 20922789888000.0;
    p = p * r2 + -1.0 / 87178291200.0;
    p = p * r2 + 1.0 / 479001600.0;
    p = p * r2 + -1.0 / 3628800.0;
    p = p * r2 + 1.0 / 40320.0;
    p = p * r2 + -1.0 / 720.0;
    p = p * r2 + 1.0 / 24.0;
    p = p * r2 + -1.0 / 2.0;
    return 1.0 + r2 * p;
}

PT_HD double reduce(double x, int& q)
{
    const double k = rint(x * kTwoOverPi);
    const double r = (x - k * kPio2Hi) - k * kPio2Lo;
    q = static_cast<int>(static_cast<long long>(k) & 3);
    return r;
}

PT_HD double sin_d(double x)
{
    int q = 0;
    const double r = reduce(x, q);
    const double s = (q & 1) ? cos_series(r) : sin_series(r);
    return (q & 2) ? -s : s;
}

PT_HD double cos_d(double x)
{
    int q = 0;
    const double r = reduce(x, q);
    const double c = (q & 1) ? sin_series(r) : cos_series(r);
    return (q == 1 || q == 2) ? -c : c;
}

// sin and cos of one argument with a shared reduction.
PT_HD void sincos_d(double x, double& s, double& c)
{
    int q = 0;
    const double r = reduce(x, q);
    const double ss = sin_series(r);
    const double cc = cos_series(r);
    switch (q) {
    case 0: s = ss; c = cc; break;
    case 1: s = cc; c = -ss; break;
    case 2: s = -ss; c = -cc; break;
    default: s = -cc; c = ss; break;
    }
}

PT_HD double atan_unit(double z)
{
    double w = z / (1.0 + sqrt(1.0 + z * z));
    w = w / (1.0 + sqrt(1.0 + w * w));
    const double w2 = w * w;
    double p = -1.0 / 27.0;
    p = p * w2 + 1.0 / 25.0;
    p = p * w2 + -1.0 / 23.0;
    p = p * w2 + 1.0 / 21.0;
    p = p * w2 + -1.0 / 19.0;
    p = p * w2 + 1.0 / 17.0;
    p = p * w2 + -1.0 / 15.0;
    p = p * w2 + 1.0 / 13.0;
    p = p * w2 + -1.0 / 11.0;
    p = p * w2 + 1.0 / 9.0;
    p = p * w2 + -1.0 / 7.0;
    p = p * w2 + 1.0 / 5.0;
    p = p * w2 + -1.0 / 3.0;
    p = p * w2 + 1.0;
    return 4.0 * (w * p);
}

PT_HD double atan2_d(double y, double x)
{
    const double ay = fabs(y);
    const double ax = fabs(x);
    double a = 0.0;
    if (ay == 0.0 && ax == 0.0) {
        a = 0.0;
    } else if (ay <= ax) {
        a = atan_unit(ay / ax);
    } else {
        a = kPio2 - atan_unit(ax / ay);
    }
    if (x < 0.0 || (x == 0.0 && signbit(x))) {
        a = kPi - a;
    }
    return signbit(y) ? -a : a;
}
} // namespace dm

PT_HD float sinf_pt(float x)
{
    if (x == 0.0f) {
        return x; // keeps the sign of zero
    }
    return __builtin_isfinite(x) ? static_cast<float>(dm::sin_d(static_cast<double>(x))) : x - x;
}
PT_HD float cosf_pt(float x) { return __builtin_isfinite(x) ? static_cast<float>(dm::cos_d(static_cast<double>(x))) : x - x; }
PT_HD void sincosf_pt(float x, float& s, float& c)
{
    if (!__builtin_isfinite(x)) {
        s = x - x;
        c = x - x;
        return;
    }
    double sd, cd;
    dm::sincos_d(static_cast<double>(x), sd, cd);
    s = x == 0.0f ? x : static_cast<float>(sd);
    c = static_cast<float>(cd);
}
PT_HD float tanf_pt(float x)
{
    if (!__builtin_isfinite(x)) {
        return x - x;
    }
    if (x == 0.0f) {
        return x;
    }
    const double xd = static_cast<double>(x);
    return static_cast<float>(dm::sin_d(xd) / dm::cos_d(xd));
}

PT_HD float expf_pt(float x)
{
    if (x != x) {
        return x;
    }
    if (x > 100.0f) {
        return __builtin_inff();
    }
    if (x < -120.0f) {
        return 0.0f;
    }
    const double xd = static_cast<double>(x);
    const double k = rint(xd * dm::kInvLn2);
    const double r = (xd - k * dm::kLn2Hi) - k * dm::kLn2Lo;
    double p = 1.0 / 1307674368000.0;
    p = p * r + 1.0 / 87178291200.0;
    p = p * r + 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 1.0 / 2.0;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return static_cast<float>(ldexp(p, static_cast<int>(k)));
}

PT_HD float logf_pt(float x)
{
    if (x != x) {
        return x;
    }
    if (x < 0.0f) {
        return __builtin_nanf("");
    }
    if (x == 0.0f) {
        return -__builtin_inff();
    }
    if (__builtin_isinf(x)) {
        return x;
    }
    int e = 0;
    double m = frexp(static_cast<double>(x), &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    return static_cast<float>(static_cast<double>(e) * dm::kLn2 + 2.0 * s * p);
}

PT_HD float atan2f_pt(float y, float x)
{
    if (x != x || y != y) {
        return x + y;
    }
    return static_cast<float>(dm::atan2_d(static_cast<double>(y), static_cast<double>(x)));
}

PT_HD float acosf_pt(float x)
{
    if (x != x || x > 1.0f || x < -1.0f) {
        return __builtin_nanf("");
    }
    const double xd = static_cast<double>(x);
    return static_cast<float>(dm::atan2_d(sqrt((1.0 - xd) * (1.0 + xd)), xd));
}

// ----------------------------------------------------------------------------------------
// Counter-based RNG: Philox-4x32-10 (Salmon et al., SC'11). One stream per pixel sample:
// key = (seed hash of Sampling/RandCore.h:9-10, globalSeed), counter = (block, pixel, sample, 0).
// Uniforms are ((x >> 8) + 1) * 2^-24 in (0, 1], the range of the reference's rand01.
// ----------------------------------------------------------------------------------------
struct Rng {
    uint32_t k0, k1;
    uint32_t block;
    uint32_t pixel, sample;
    uint32_t buf[4];
    int avail;
};

PT_DEV void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

PT_DEV uint32_t seed_hash(int pixel, int sample, uint32_t globalSeed)
{
    return static_cast<uint32_t>(pixel) * 747796405u + static_cast<uint32_t>(sample) + globalSeed * 2891336453u;
}

PT_DEV void rng_init(Rng& r, int pixel, int sample, uint32_t globalSeed)
{
    r.k0 = seed_hash(pixel, sample, globalSeed);
    r.k1 = globalSeed;
    r.block = 0;
    r.pixel = static_cast<uint32_t>(pixel);
    r.sample = static_cast<uint32_t>(sample);
    r.avail = 0;
}

PT_DEV float u32_to_unit(uint32_t x) { return static_cast<float>((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// Number of uniforms drawn so far, and resumption of a stream after `draws` uniforms (the
// wavefront kernels keep only this count between launches).
PT_DEV uint32_t rng_draws(const Rng& r) { return r.block * 4u - static_cast<uint32_t>(r.avail); }

PT_DEV void rng_resume(Rng& r, int pixel, int sample, uint32_t globalSeed, uint32_t draws)
{
    rng_init(r, pixel, sample, globalSeed);
    r.block = draws >> 2;
    if ((draws & 3u) != 0u) {
        philox(r.block, r.pixel, r.sample, 0u, r.k0, r.k1, r.buf);
        r.block += 1u;
        r.avail = 4 - static_cast<int>(draws & 3u);
    }
}

PT_DEV float rand01(Rng& r)
{
    if (r.avail == 0) {
        philox(r.block, r.pixel, r.sample, 0u, r.k0, r.k1, r.buf);
        r.block += 1u;
        r.avail = 4;
    }
    const uint32_t x = r.buf[4 - r.avail];
    r.avail -= 1;
    return u32_to_unit(x);
}

// ----------------------------------------------------------------------------------------
// Scene access
// ----------------------------------------------------------------------------------------
constexpr float kRayTMax = 1.0e6f;
// Boxes are culled against the closest hit widened by 2^-12 of its distance: the ray-triangle t
// can fall short of the entry distance into the triangle's own box by more than rounding (thin
// triangles seen from afar), so a box entered just past the current hit can hold a closer one.
// The margin keeps such triangles in reach, so the closest hit does not depend on the tree or the
// visit order (the reference culls at the hit itself).
constexpr float kCullSlack = 1.000244140625f;
constexpr float kShadowBias = 1.0e-4f;
constexpr float kRelShadowBias = 1.0e-3f;

// Light sampling of the forward renderers. kLightReference is the reference's: the emissive
// next-event shadow ray runs to kRayTMax from p + wi * bias (LightSamplingCore.h:360-369), so it
// meets the sampled light itself and reports it occluded, and emission found by BSDF sampling is
// added unweighted (PathIntegratorRandCore.h:248): area lights are lit by BSDF sampling alone.
// kLightMis is multiple importance sampling on area lights: the shadow ray runs from p + wi * bias
// towards the light point and stops at kBoundedShadowScale of that distance, and emission that a
// BSDF-sampled ray finds takes the balance weight pdf_bsdf / (pdf_bsdf + pdf_nee), with pdf_nee
// the solid-angle density with which the previous vertex's next-event estimate picks that point
// (zero on the light's back face and for directions below the previous vertex's normal, which
// next-event estimation never samples). Both modes estimate the same image.
constexpr int kLightReference = 0;
constexpr int kLightMis = 1;
constexpr float kBoundedShadowScale = 0.9999f;
constexpr float kMinPdf = 1.0e-8f;
constexpr float kPi = 3.14159265f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kMinRadius = 1.0e-5f;
constexpr int kStackSize = 64; // traversal stack entries (tree depth <= 48 for every build mode)
constexpr int kRingEntries = 8; // entries of a thread's traversal stack kept in LDS
constexpr int kMaxTraceBlock = 256; // threads per block of every kernel that traverses

typedef int LdsNativeInt2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) LdsNativeInt2 LdsStackEntry;

// The block's LDS area for the first kRingEntries entries of every thread's traversal stack
// (entry e of thread t at e * kMaxTraceBlock + t, so a wave's accesses are conflict-free). One
// allocation per kernel, shared by every traversal in it: a thread runs one traversal at a time.
PT_DEV LdsStackEntry* traversal_ring()
{
    __shared__ LdsNativeInt2 ring[kRingEntries * kMaxTraceBlock];
    const int tid = static_cast<int>(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    return (LdsStackEntry*)(ring + tid);
}

// Loads through the global address space: global_load instructions (vector memory counter
// only) instead of flat ones, which also count against the LDS counter.
typedef float NativeFloat4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const NativeFloat4 GlobalFloat4;
PT_DEV float4 ld4(const PtFloat4* p, uint32_t i)
{
    const NativeFloat4 v = ((GlobalFloat4*)(p))[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

struct Ray {
    V3 o, d, inv;
};

// MeshAccelIntersectCore.h:124-128
PT_DEV Ray make_ray(V3 ro, V3 rdIn)
{
    Ray r;
    r.o = ro;
    r.d = normalize(rdIn);
    r.inv = mk(fabsf(r.d.x) > 1.0e-8f ? 1.0f / r.d.x : 0.0f, fabsf(r.d.y) > 1.0e-8f ? 1.0f / r.d.y : 0.0f,
               fabsf(r.d.z) > 1.0e-8f ? 1.0f / r.d.z : 0.0f);
    return r;
}

// MeshAccelIntersectCore.h:12-50. The per-axis early exit of the reference is equivalent to
// one comparison after all three axes (the running max only grows, the running min only
// shrinks), so the three slabs are evaluated unconditionally; the max/min chains keep the
// reference's order, which fixes signed zeros.
PT_DEV bool slab(const Ray& r, float bx0, float by0, float bz0, float bx1, float by1, float bz1, float tMin,
                 float tMax, float& tEnter)
{
    float t0 = (bx0 - r.o.x) * r.inv.x;
    float t1 = (bx1 - r.o.x) * r.inv.x;
    if (r.inv.x < 0.0f) {
        const float t = t0;
        t0 = t1;
        t1 = t;
    }
    float te = max2(tMin, t0);
    float tx = min2(tMax, t1);
    t0 = (by0 - r.o.y) * r.inv.y;
    t1 = (by1 - r.o.y) * r.inv.y;
    if (r.inv.y < 0.0f) {
        const float t = t0;
        t0 = t1;
        t1 = t;
    }
    te = max2(te, t0);
    tx = min2(tx, t1);
    t0 = (bz0 - r.o.z) * r.inv.z;
    t1 = (bz1 - r.o.z) * r.inv.z;
    if (r.inv.z < 0.0f) {
        const float t = t0;
        t0 = t1;
        t1 = t;
    }
    te = max2(te, t0);
    tx = min2(tx, t1);
    tEnter = te;
    return !(te > tx) && tx >= tMin;
}

// MeshAccelIntersectCore.h:52-95 on precomputed (v0, e1 = v1 - v0, e2 = v2 - v0); the edges
// are the same floats the reference computes per test. Returns t and the barycentrics.
PT_DEV bool tri_test(const Ray& r, const float4& a, const float4& b, const float4& c, float tMin, float tMax,
                     float& tOut, float& uOut, float& vOut)
{
    const V3 e1 = mk(b.x, b.y, b.z);
    const V3 e2 = mk(c.x, c.y, c.z);
    const V3 rd = r.d;
    const V3 pvec = mk(rd.y * e2.z - rd.z * e2.y, rd.z * e2.x - rd.x * e2.z, rd.x * e2.y - rd.y * e2.x);
    const float det = dot(e1, pvec);
    if (fabsf(det) < 1.0e-8f) {
        return false;
    }
    const float invDet = 1.0f / det;
    const V3 tvec = sub(r.o, mk(a.x, a.y, a.z));
    const float u = dot(tvec, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) {
        return false;
    }
    const V3 qvec = mk(tvec.y * e1.z - tvec.z * e1.y, tvec.z * e1.x - tvec.x * e1.z, tvec.x * e1.y - tvec.y * e1.x);
    const float v = dot(rd, qvec) * invDet;
    if (v < 0.0f || u + v > 1.0f) {
        return false;
    }
    const float t = dot(e2, qvec) * invDet;
    if (t < tMin || t > tMax) {
        return false;
    }
    tOut = t;
    uOut = u;
    vOut = v;
    return true;
}

// Closest-hit order: the smaller distance, and among triangles at the same distance the one
// with the smaller original index. tri_test accepts t <= the current best, so this decides
// exact ties. The reference keeps whichever tied triangle its traversal reaches last
// (MeshAccelIntersectCore.h:160-166 with the `t > tMax` rejection of :87), which depends on the
// tree's leaf order; its median split places equal centroids by std::nth_element
// (MeshBvhBuilder.cpp:91-105), so that order is not fixed by its source. With this rule the
// closest hit is a function of the scene alone, the same for every tree and traversal here.
PT_DEV bool closer(float t, uint32_t tri, float bestT, uint32_t bestTri)
{
    return t < bestT || tri < bestTri;
}

struct Hit {
    bool hit;
    float t;
    V3 normal;
    float uvx, uvy;
    uint32_t tri; // original triangle index
};

// Shading attributes of the closest hit (MeshAccelIntersectCore.h:95-106 and :212-218):
// interpolated normal, normalised, flipped against the ray, normalised again; UV.
PT_DEV void finish_hit(const MeshAccelSceneGpu& s, const Ray& r, uint32_t tri, float u, float v, Hit& h)
{
    const float4 s0 = ld4(s.triShade, 4u * tri + 0u);
    const float4 s1 = ld4(s.triShade, 4u * tri + 1u);
    const float4 s2 = ld4(s.triShade, 4u * tri + 2u);
    const float4 s3 = ld4(s.triShade, 4u * tri + 3u);
    const V3 n0 = mk(s0.x, s0.y, s0.z);
    const V3 n1 = mk(s0.w, s1.x, s1.y);
    const V3 n2 = mk(s1.z, s1.w, s2.x);
    const float w = 1.0f - u - v;
    V3 n = normalize(add(add(scale(n0, w), scale(n1, u)), scale(n2, v)));
    if (dot(n, r.d) > 0.0f) {
        n = scale(n, -1.0f);
    }
    h.normal = normalize(n);
    h.uvx = s2.y * w + s2.w * u + s3.y * v;
    h.uvy = s2.z * w + s3.x * u + s3.z * v;
    h.tri = tri;
}

PT_DEV uint32_t tri_material(const MeshAccelSceneGpu& s, uint32_t tri)
{
    return __float_as_uint(ld4(s.triShade, 4u * tri + 3u).w);
}

// Traversal statistics (only the counting kernel passes a non-null pointer).
struct TraceStats {
    uint32_t nodes, tris, rays, hits;
};

// Closest hit (MeshAccelIntersectCore.h:110-221). Visit order equals the reference's: the
// child with the smaller entry distance first, the right child first on ties; a deferred
// child is skipped when its entry distance exceeds the current closest t (the reference's
// re-test on pop). Leaves hold one triangle, tested with the current closest t.
// ANY_HIT: stop at the first triangle with t < tMax (the reference's occlusion result).
template <bool ANY_HIT>
PT_DEV bool traverse(const MeshAccelSceneGpu& s, const Ray& r, float tMin, float tMax, float& bestT, uint32_t& bestTri,
                     float& bestU, float& bestV, TraceStats* stats)
{
    bestT = tMax;
    bestTri = 0xFFFFFFFFu;
    bestU = 0.0f;
    bestV = 0.0f;
    if (s.triangleCount == 0) {
        return false;
    }
    float cullT = tMax * kCullSlack;
    float te;
    if (!slab(r, s.rootMin[0], s.rootMin[1], s.rootMin[2], s.rootMax[0], s.rootMax[1], s.rootMax[2], tMin, cullT, te)) {
        return false;
    }
    // stack entries below kRingEntries live in LDS, deeper ones in a private array
    LdsStackEntry* ring = traversal_ring();
    int stackRef[kStackSize - kRingEntries];
    float stackT[kStackSize - kRingEntries];
    int sp = 0;
    int cur = s.rootRef;
    while (true) {
        // One memory round trip per step for the whole wave: lanes at a leaf load its triangle
        // (3 float4) and lanes at a node the node (4 float4) through the same load instructions,
        // before either test, so a wave split between the two waits once.
        const bool leaf = cur < 0;
        const PtFloat4* src = leaf ? s.triPos + 3u * static_cast<uint32_t>(~cur)
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
            if (stats) {
                stats->tris += 1;
            }
            float t, u, v;
            const uint32_t id = __float_as_uint(a.w);
            if (tri_test(r, a, b, c, tMin, bestT, t, u, v) && closer(t, id, bestT, bestTri)) {
                bestT = t;
                cullT = t * kCullSlack;
                bestU = u;
                bestV = v;
                bestTri = id;
                if (ANY_HIT && t < tMax) {
                    return true;
                }
            }
        } else {
            const float4 n0 = q0;
            const float4 n1 = q1;
            const float4 n2 = q2;
            const float4 n3 = q3;
            if (stats) {
                stats->nodes += 1;
            }
            float le, re;
            const bool lh = slab(r, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, tMin, cullT, le);
            const bool rh = slab(r, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, tMin, cullT, re);
            const int left = __float_as_int(n3.x);
            const int right = __float_as_int(n3.y);
            if (lh && rh) {
                const bool leftFirst = le < re;
                // The median-split tree of at most 2^28 triangles is at most 28 levels deep, so
                // the deferred-child stack never holds more than 28 entries; the guard only
                // keeps the private array in bounds.
                if (sp < kRingEntries) {
                    LdsNativeInt2 e;
                    e.x = leftFirst ? right : left;
                    e.y = __float_as_int(leftFirst ? re : le);
                    ring[sp * kMaxTraceBlock] = e;
                    ++sp;
                } else if (sp < kStackSize) {
                    stackRef[sp - kRingEntries] = leftFirst ? right : left;
                    stackT[sp - kRingEntries] = leftFirst ? re : le;
                    ++sp;
                }
                cur = leftFirst ? left : right;
                continue;
            }
            if (lh) {
                cur = left;
                continue;
            }
            if (rh) {
                cur = right;
                continue;
            }
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            int ref;
            float t;
            if (sp < kRingEntries) {
                const LdsNativeInt2 e = ring[sp * kMaxTraceBlock];
                ref = e.x;
                t = __int_as_float(e.y);
            } else {
                ref = stackRef[sp - kRingEntries];
                t = stackT[sp - kRingEntries];
            }
            if (!(t > cullT)) {
                cur = ref;
                found = true;
                break;
            }
        }
        if (!found) {
            break;
        }
    }
    return bestT < tMax;
}

PT_DEV Hit trace_closest(const MeshAccelSceneGpu& s, V3 ro, V3 rd, float tMin, float tMax, TraceStats* stats)
{
    const Ray r = make_ray(ro, rd);
    Hit h;
    h.hit = false;
    h.t = 0.0f;
    h.normal = mk(0.0f, 0.0f, 0.0f);
    h.uvx = 0.0f;
    h.uvy = 0.0f;
    h.tri = 0;
    float t, u, v;
    uint32_t tri;
    if (stats) {
        stats->rays += 1;
    }
    if (traverse<false>(s, r, tMin, tMax, t, tri, u, v, stats)) {
        if (stats) {
            stats->hits += 1;
        }
        h.hit = true;
        h.t = t;
        finish_hit(s, r, tri, u, v, h);
    }
    return h;
}

PT_DEV bool trace_any(const MeshAccelSceneGpu& s, V3 ro, V3 rd, float tMin, float tMax, TraceStats* stats)
{
    const Ray r = make_ray(ro, rd);
    float t, u, v;
    uint32_t tri;
    if (stats) {
        stats->rays += 1;
    }
    return traverse<true>(s, r, tMin, tMax, t, tri, u, v, stats);
}

// ----------------------------------------------------------------------------------------
// Materials and textures (Texture/ProceduralTexture.h:138-207, LightSamplingCore.h:199-213)
// ----------------------------------------------------------------------------------------
struct Mat {
    float r, g, b, roughness, metallic, transmission, thin, ior, subsurface, emission, diffuseRoughness, srR, srG,
        srB, specular;
    const float* spec;  // the material's binned spectra record (MeshAccelSceneGpu::materialSpectra) or null
};

PT_DEV float tex_fract(float v) { return v - floorf(v); }
PT_DEV float tex_clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

PT_DEV float tex_scalar(const TextureDescGpu& d, float uvx, float uvy)
{
    switch (d.kind) {
    case 0u: return d.p0.x;
    case 3u: {
        const float c = tex_fract(uvx * d.p0.x);
        return c < d.p0.y ? d.p0.z : d.p0.w;
    }
    case 2u: {
        const float fx = tex_fract(uvx * d.p0.x);
        const float fy = tex_fract(uvy * d.p0.y);
        return (fx < d.p0.z || fy < d.p0.z) ? 1.0f : 0.0f;
    }
    default: return 0.0f;
    }
}

PT_DEV float chan(const MeshAccelSceneGpu& s, float inl, uint32_t ti, float uvx, float uvy)
{
    if (ti == 0u || s.textures == nullptr || ti >= s.textureCount) {
        return inl;
    }
    return tex_scalar(s.textures[ti], uvx, uvy);
}

// lightFetchMaterial + resolveMaterial + materialFromResolved
// Material `mi` resolved at uv (the fallback grey when !valid, LightSamplingCore.h:199-213).
PT_DEV Mat resolve_material_index(const MeshAccelSceneGpu& s, uint32_t mi, bool valid, float uvx, float uvy)
{
    MaterialGpu m;
    if (valid) {
        m = s.materials[mi];
    } else {
        m.r = 0.8f; m.g = 0.8f; m.b = 0.8f; m.roughness = 0.5f; m.metallic = 0.0f; m.transmission = 0.0f;
        m.thin = 0.0f; m.ior = 1.5f; m.subsurface = 0.0f; m.emission = 0.0f; m.diffuseRoughness = -1.0f;
        m.scatterRadiusR = 0.0f; m.scatterRadiusG = 0.0f; m.scatterRadiusB = 0.0f; m.specular = 1.0f;
        m.albedoTex = 0u; m.roughnessTex = 0u; m.metallicTex = 0u; m.transmissionTex = 0u; m.thinTex = 0u;
        m.iorTex = 0u; m.subsurfaceTex = 0u; m.emissionTex = 0u;
    }
    Mat r;
    if (m.albedoTex == 0u || s.textures == nullptr || m.albedoTex >= s.textureCount) {
        r.r = m.r;
        r.g = m.g;
        r.b = m.b;
    } else {
        const TextureDescGpu d = s.textures[m.albedoTex];
        if (d.kind == 1u) {
            r.r = d.p0.x; r.g = d.p0.y; r.b = d.p0.z;
        } else if (d.kind == 2u) {
            const float fx = tex_fract(uvx * d.p0.x);
            const float fy = tex_fract(uvy * d.p0.y);
            const float g = (fx < d.p0.z || fy < d.p0.z) ? 1.0f : 0.0f;
            r.r = d.p1.x + (d.p2.x - d.p1.x) * g;
            r.g = d.p1.y + (d.p2.y - d.p1.y) * g;
            r.b = d.p1.z + (d.p2.z - d.p1.z) * g;
        } else {
            r.r = 0.0f; r.g = 0.0f; r.b = 0.0f;
        }
    }
    r.roughness = tex_clamp01(chan(s, m.roughness, m.roughnessTex, uvx, uvy));
    r.metallic = tex_clamp01(chan(s, m.metallic, m.metallicTex, uvx, uvy));
    r.transmission = tex_clamp01(chan(s, m.transmission, m.transmissionTex, uvx, uvy));
    r.thin = tex_clamp01(chan(s, m.thin, m.thinTex, uvx, uvy));
    r.ior = fmaxf(1.0e-3f, chan(s, m.ior, m.iorTex, uvx, uvy));
    r.subsurface = tex_clamp01(chan(s, m.subsurface, m.subsurfaceTex, uvx, uvy));
    r.emission = fmaxf(0.0f, chan(s, m.emission, m.emissionTex, uvx, uvy));
    r.diffuseRoughness = m.diffuseRoughness < 0.0f ? r.roughness : tex_clamp01(m.diffuseRoughness);
    r.srR = fmaxf(0.0f, m.scatterRadiusR);
    r.srG = fmaxf(0.0f, m.scatterRadiusG);
    r.srB = fmaxf(0.0f, m.scatterRadiusB);
    r.specular = tex_clamp01(m.specular);
    r.spec = valid && s.materialSpectra != nullptr ? s.materialSpectra + static_cast<size_t>(mi) * kPtSpectrumWords
                                                    : nullptr;
    return r;
}

PT_DEV Mat resolve_material(const MeshAccelSceneGpu& s, uint32_t tri, float uvx, float uvy)
{
    bool valid = s.materials != nullptr && tri < s.triangleCount;
    uint32_t mi = 0;
    if (valid) {
        mi = tri_material(s, tri);
        valid = mi < s.materialCount;
    }
    return resolve_material_index(s, mi, valid, uvx, uvy);
}

// ----------------------------------------------------------------------------------------
// Principled BSDF (Brdf/BrdfBase.h, Brdf/PrincipledBrdf.h, Brdf/SubsurfaceCore.h,
// Brdf/BrdfDispatch.h). Lobe weights depend only on the material, so they are computed
// once per shading context; lobes with weight exactly +0 contribute exactly +0 to every sum
// and are skipped.
// ----------------------------------------------------------------------------------------
struct Lobes {
    float diffuse, specular, transmit, subsurface;
};

// PrincipledBrdf.h:75-100
PT_DEV Lobes lobe_weights(const Mat& m)
{
    Lobes w;
    const float metallic = clamp01(m.metallic);
    const float transmission = clamp01(m.transmission);
    const float subsurface = clamp01(m.subsurface);
    const float dielectric = 1.0f - metallic;
    w.transmit = dielectric * transmission;
    w.subsurface = dielectric * (1.0f - transmission) * subsurface;
    w.diffuse = dielectric * (1.0f - transmission) * (1.0f - subsurface);
    w.specular = metallic + dielectric * (1.0f - transmission);
    const float sum = w.diffuse + w.specular + w.transmit + w.subsurface;
    if (sum <= kMinPdf) {
        w.diffuse = 1.0f;
        return w;
    }
    const float inv = 1.0f / sum;
    w.diffuse *= inv;
    w.specular *= inv;
    w.transmit *= inv;
    w.subsurface *= inv;
    return w;
}

struct Ctx {
    V3 n, wo;
    Mat m;
    Lobes w;
    float eta;
    int flags;
};

struct Sample {
    V3 dir;
    float pdf;
    bool valid, transmitted, sss;
    float nextEta;
    V3 exitOffset;
    int steps;
};

PT_DEV void basis(V3 n, V3& t, V3& b)
{
    const V3 up = fabsf(n.y) < 0.999f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    t = normalize(cross(up, n));
    b = normalize(cross(n, t));
}

PT_DEV V3 to_world(V3 l, V3 n, V3 t, V3 b) { return normalize(add(add(scale(t, l.x), scale(b, l.y)), scale(n, l.z))); }

PT_DEV V3 cosine_local(float u1, float u2)
{
    const float phi = 2.0f * kPi * u1;
    const float r = sqrtf(u2);
    float sp, cp;
    sincosf_pt(phi, sp, cp);
    return mk(r * cp, r * sp, sqrtf(max2(0.0f, 1.0f - u2)));
}

PT_DEV float alpha_of(float rough)
{
    const float c = max2(rough, 0.04f);
    return c * c;
}

PT_DEV float ggx_d(float cosH, float alpha)
{
    const float a2 = alpha * alpha;
    const float c2 = cosH * cosH;
    const float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}

PT_DEV float smith_g1(float c, float alpha)
{
    const float a2 = alpha * alpha;
    const float c2 = max2(c * c, 1.0e-8f);
    const float tan2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * tan2));
}

PT_DEV V3 ggx_half_local(float alpha, float u1, float u2, float& pdf)
{
    const float phi = 2.0f * kPi * u1;
    const float a2 = alpha * alpha;
    const float ct = sqrtf((1.0f - u2) / (1.0f + (a2 - 1.0f) * u2));
    const float st = sqrtf(max2(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf_pt(phi, sp, cp);
    pdf = (a2 * ct) / (kPi * (ct * ct * (a2 - 1.0f) + 1.0f) * (ct * ct * (a2 - 1.0f) + 1.0f));
    return mk(st * cp, st * sp, ct);
}

PT_DEV float dielectric_fresnel(float cosThetaI, float etaI, float etaT)
{
    float cosI = max2(-1.0f, min2(1.0f, cosThetaI));
    float eta = etaI / max2(etaT, 1.0e-8f);
    if (cosI < 0.0f) {
        cosI = -cosI;
        eta = 1.0f / eta;
    }
    const float sin2t = eta * eta * (1.0f - cosI * cosI);
    if (sin2t >= 1.0f) {
        return 1.0f;
    }
    const float cosT = sqrtf(max2(0.0f, 1.0f - sin2t));
    const float rPar = (eta * cosI - cosT) / max2(eta * cosI + cosT, 1.0e-8f);
    const float rPerp = (cosI - eta * cosT) / max2(cosI + eta * cosT, 1.0e-8f);
    return 0.5f * (rPar * rPar + rPerp * rPerp);
}

PT_DEV bool refract_rel(V3 wo, V3 normal, float etaI, float etaT, V3& out)
{
    float cosO = dot(normal, wo);
    V3 n = normal;
    float etaRel = etaI / max2(etaT, 1.0e-8f);
    if (cosO > 0.0f) {
        n = scale(normal, -1.0f);
        etaRel = 1.0f / etaRel;
        cosO = -cosO;
    }
    const float sin2t = etaRel * etaRel * (1.0f - cosO * cosO);
    if (sin2t >= 1.0f) {
        return false;
    }
    const float cosT = sqrtf(max2(0.0f, 1.0f - sin2t));
    out = normalize(sub(scale(wo, -etaRel), scale(n, etaRel * cosO + cosT)));
    return true;
}

PT_DEV float oren_nayar(float sigma, float cosI, float cosO, float dotWoWi)
{
    if (sigma <= 1.0e-4f) {
        return 1.0f;
    }
    const float s2 = sigma * sigma;
    const float A = 1.0f - 0.5f * s2 / (s2 + 0.33f);
    const float B = 0.45f * s2 / (s2 + 0.09f);
    const float sinI = sqrtf(max2(0.0f, 1.0f - cosI * cosI));
    const float sinO = sqrtf(max2(0.0f, 1.0f - cosO * cosO));
    float sa, sb;
    if (cosI > cosO) {
        sa = sinI;
        sb = sinO;
    } else {
        sa = sinO;
        sb = sinI;
    }
    float cosPhi = 1.0f;
    const float d = dotWoWi - cosI * cosO;
    if (d <= 0.0f) {
        const float denom = max2(sa * sb, 1.0e-8f);
        cosPhi = max2(-1.0f, min2(1.0f, d / denom));
    } else {
        sa = 0.0f;
        sb = 0.0f;
    }
    const float tanB = sb / max2(cosI, cosO);
    return A + B * max2(0.0f, cosPhi) * sa * tanB;
}

PT_DEV float energy_comp(float rough, float cosTheta)
{
    const float alpha = alpha_of(rough);
    const float c = max2(cosTheta, 1.0e-4f);
    const float Ess = 1.0f - alpha * (0.3316f + 0.3237f * alpha);
    const float Ems = (1.0f - Ess) * (1.0f - c);
    return 1.0f + Ems / max2(Ess, 1.0e-4f);
}

PT_DEV float max_radius(const Mat& m) { return max2(max2(max2(0.0f, m.srR), max2(0.0f, m.srG)), max2(0.0f, m.srB)); }

PT_DEV float burley_sample_radius(float u, float d)
{
    const float sd = max2(d, kMinRadius);
    return sd * sqrtf(-logf_pt(max2(1.0e-8f, 1.0f - u * (1.0f - expf_pt(-1.0f / sd)))));
}
PT_DEV float burley_profile(float r, float d)
{
    const float sd = max2(d, kMinRadius);
    return (expf_pt(-r / sd) + expf_pt(-r / (3.0f * sd))) / (4.0f * kPi * sd);
}
PT_DEV float burley_pdf_radius(float r, float d)
{
    const float sd = max2(d, kMinRadius);
    const float t0 = expf_pt(-r / sd);
    const float t1 = expf_pt(-r / (3.0f * sd));
    const float norm = 1.0f - expf_pt(-1.0f / sd);
    return max2(0.0f, (t0 + t1) / max2(2.0f * sd * norm, 1.0e-8f));
}

PT_DEV float material_ior(const Ctx& c) { return max2(c.m.ior, 1.0e-3f); }

PT_DEV V3 mirror(V3 wo, V3 n) { return normalize(sub(scale(n, 2.0f * dot(n, wo)), wo)); }

PT_DEV float on_factor(const Ctx& c, V3 wi)
{
    const float cosWi = max2(0.0f, dot(c.n, wi));
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return 0.0f;
    }
    const float dr = max2(0.0f, min2(1.0f, c.m.diffuseRoughness < 0.0f ? c.m.roughness : c.m.diffuseRoughness));
    return oren_nayar(dr * 0.5f * kPi, cosWi, cosWo, dot(wi, c.wo));
}

PT_DEV bool refract_dir(const Ctx& c, V3 n, V3& out, float& nextEta)
{
    const float cosWo = dot(n, c.wo);
    const float ior = material_ior(c);
    float etaT;
    if (cosWo > 0.0f) {
        etaT = ior;
        nextEta = ior;
    } else {
        etaT = 1.0f;
        nextEta = 1.0f;
    }
    return refract_rel(c.wo, n, c.eta, etaT, out);
}

PT_DEV bool transmit_valid(const Ctx& c, float cosWoS)
{
    if (absf(cosWoS) <= kMinPdf) {
        return false;
    }
    return cosWoS > 0.0f || c.eta > 1.0f + 1.0e-4f;
}

// The lobe evaluations are generic over the colour type: `base` is the base colour (RGB, or
// its spectrum at the hero wavelengths) and `rad` the subsurface radii; every colour operation
// is per component and in the reference's order, so the RGB instantiation is the reference's
// arithmetic.
template <class C>
PT_DEV C eval_diffuse_t(const Ctx& c, V3 wi, const C& base)
{
    const float cosWi = max2(0.0f, dot(c.n, wi));
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return ColN<C>::splat(0.0f);
    }
    const float on = on_factor(c, wi);
    return scale(base, c.w.diffuse * on * kInvPi);
}

template <class C>
PT_DEV C eval_specular_t(const Ctx& c, V3 wi, const C& base)
{
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    const float cosWi = max2(0.0f, dot(c.n, wi));
    if (cosWo <= 0.0f || cosWi <= 0.0f) {
        return ColN<C>::splat(0.0f);
    }
    const V3 h = normalize(add(wi, c.wo));
    const float cosH = max2(0.0f, dot(c.n, h));
    const float cosHo = max2(0.0f, dot(h, c.wo));
    if (cosHo <= 0.0f) {
        return ColN<C>::splat(0.0f);
    }
    const float alpha = alpha_of(c.m.roughness);
    const float D = ggx_d(cosH, alpha);
    const float G = smith_g1(cosWo, alpha) * smith_g1(cosWi, alpha);
    // Schlick F with F0 from PrincipledBrdf.h:38-54
    const float metallic = clamp01(c.m.metallic);
    const float ior = material_ior(c);
    const float f0s = (ior - 1.0f) / (ior + 1.0f);
    const float f0v = f0s * f0s * clamp01(clamp01(c.m.specular));
    C f0;
#pragma unroll
    for (int k = 0; k < ColN<C>::n; ++k) {
        cset(f0, k, f0v + (cget(base, k) - f0v) * metallic);
    }
    const float t = max2(0.0f, min2(1.0f, 1.0f - cosHo));
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    const C F = add(scale(f0, 1.0f - t5), ColN<C>::splat(t5));
    const float ec = energy_comp(c.m.roughness, cosWo);
    const float denom = max2(4.0f * cosWo * cosWi, 1.0e-8f);
    const float sc = c.w.specular * ec * (D * G) / denom;
    return scale(F, sc);
}

template <class C>
PT_DEV C eval_transmit_t(const Ctx& c, V3 wi, const C& base)
{
    const float cosWoS = dot(c.n, c.wo);
    if (!transmit_valid(c, cosWoS)) {
        return ColN<C>::splat(0.0f);
    }
    const float thin = clamp01(c.m.thin);
    const float ior = material_ior(c);
    const float etaT = cosWoS > 0.0f ? ior : 1.0f;
    const float refl = dielectric_fresnel(cosWoS, c.eta, etaT);
    const float trans = 1.0f - refl;
    const V3 reflected = mirror(c.wo, c.n);
    if (dot(wi, reflected) > 0.999f) {
        const float cosWi = absf(dot(c.n, wi));
        return scale(base, c.w.transmit * refl / max2(cosWi, 1.0e-8f));
    }
    V3 tr;
    float next = 1.0f;
    if (thin > 0.5f) {
        tr = normalize(scale(c.wo, -1.0f));
    } else if (!refract_dir(c, c.n, tr, next)) {
        return ColN<C>::splat(0.0f);
    }
    if (dot(wi, tr) > 0.999f) {
        const float cosWi = absf(dot(c.n, tr));
        const C tint = thin > 0.5f ? base : ColN<C>::splat(1.0f);
        C out;
#pragma unroll
        for (int k = 0; k < ColN<C>::n; ++k) {
            cset(out, k, cget(tint, k) * c.w.transmit * trans / max2(cosWi, 1.0e-8f));
        }
        return out;
    }
    const float alpha = alpha_of(c.m.roughness);
    if (alpha > 0.04f) {
        const V3 h = normalize(add(wi, c.wo));
        const float cosH = max2(0.0f, absf(dot(c.n, h)));
        const float D = ggx_d(cosH, alpha);
        const float roughPdf = D * cosH;
        if (roughPdf > kMinPdf) {
            const C tint = thin > 0.5f ? base : ColN<C>::splat(1.0f);
            const float sc = c.w.transmit * trans * roughPdf;
            return scale(tint, sc);
        }
    }
    return ColN<C>::splat(0.0f);
}

// Subsurface lobe. RGB: the reference's per-channel Burley profiles. Hero wavelengths: each
// wavelength is evaluated as a grey material whose scatter radius is that wavelength's radius.
template <class C>
PT_DEV C eval_subsurface_t(const Ctx& c, V3 wi, const C& base, const C& rad)
{
    const float cosWi = max2(0.0f, dot(c.n, wi));
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    if (cosWi <= 0.0f || cosWo <= 0.0f) {
        return ColN<C>::splat(0.0f);
    }
    if constexpr (ColN<C>::n == 3) {
        const float mr = max_radius(c.m);
        if (mr <= kMinRadius) {
            const float on = on_factor(c, wi);
            return scale(base, c.w.subsurface * on * kInvPi);
        }
        V3 t, b;
        basis(c.n, t, b);
        const V3 wt = add(scale(t, dot(wi, t)), scale(b, dot(wi, b)));
        const float dist = sqrtf(dot(wt, wt));
        const float sum = max2(rad.x + rad.y + rad.z, kMinRadius);
        const V3 cw = mk(rad.x / sum, rad.y / sum, rad.z / sum);
        const float pr = burley_profile(dist, max2(rad.x, kMinRadius));
        const float pg = burley_profile(dist, max2(rad.y, kMinRadius));
        const float pb = burley_profile(dist, max2(rad.z, kMinRadius));
        const float dif = max2(cosWi, 0.0f) * max2(cosWo, 0.0f);
        const V3 bur = mk(cw.x * pr * dif, cw.y * pg * dif, cw.z * pb * dif);
        return mk(base.x * c.w.subsurface * bur.x, base.y * c.w.subsurface * bur.y,
                  base.z * c.w.subsurface * bur.z);
    } else {
        V3 t, b;
        basis(c.n, t, b);
        const V3 wt = add(scale(t, dot(wi, t)), scale(b, dot(wi, b)));
        const float dist = sqrtf(dot(wt, wt));
        const float dif = max2(cosWi, 0.0f) * max2(cosWo, 0.0f);
        float flat = 0.0f;
        bool haveFlat = false;
        C out;
#pragma unroll
        for (int k = 0; k < ColN<C>::n; ++k) {
            const float rk = cget(rad, k);
            if (rk <= kMinRadius) {
                if (!haveFlat) {
                    flat = c.w.subsurface * on_factor(c, wi) * kInvPi;
                    haveFlat = true;
                }
                cset(out, k, cget(base, k) * flat);
            } else {
                const float sum = max2(rk + rk + rk, kMinRadius);
                const float bur = rk / sum * burley_profile(dist, max2(rk, kMinRadius)) * dif;
                cset(out, k, cget(base, k) * c.w.subsurface * bur);
            }
        }
        return out;
    }
}

template <class C>
PT_DEV C brdf_eval_t(const Ctx& c, V3 wi, const C& base, const C& rad)
{
    const C d = eval_diffuse_t(c, wi, base);
    const C s = eval_specular_t(c, wi, base);
    const C t = c.w.transmit != 0.0f ? eval_transmit_t(c, wi, base) : ColN<C>::splat(0.0f);
    const C ss = c.w.subsurface != 0.0f ? eval_subsurface_t(c, wi, base, rad) : ColN<C>::splat(0.0f);
    return add(add(d, s), add(t, ss));
}

PT_DEV V3 mat_base(const Mat& m) { return mk(m.r, m.g, m.b); }
PT_DEV V3 mat_radii(const Mat& m) { return mk(max2(0.0f, m.srR), max2(0.0f, m.srG), max2(0.0f, m.srB)); }

PT_DEV V3 eval_diffuse(const Ctx& c, V3 wi) { return eval_diffuse_t(c, wi, mat_base(c.m)); }
PT_DEV V3 eval_specular(const Ctx& c, V3 wi) { return eval_specular_t(c, wi, mat_base(c.m)); }
PT_DEV V3 eval_transmit(const Ctx& c, V3 wi) { return eval_transmit_t(c, wi, mat_base(c.m)); }
PT_DEV V3 eval_subsurface(const Ctx& c, V3 wi) { return eval_subsurface_t(c, wi, mat_base(c.m), mat_radii(c.m)); }
PT_DEV V3 brdf_eval(const Ctx& c, V3 wi) { return brdf_eval_t(c, wi, mat_base(c.m), mat_radii(c.m)); }

PT_DEV float pdf_specular(const Ctx& c, V3 wi)
{
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    const float cosWi = max2(0.0f, dot(c.n, wi));
    if (cosWo <= 0.0f || cosWi <= 0.0f) {
        return 0.0f;
    }
    const V3 h = normalize(add(wi, c.wo));
    const float cosH = max2(0.0f, dot(c.n, h));
    const float cosHo = max2(0.0f, dot(h, c.wo));
    if (cosHo <= 0.0f) {
        return 0.0f;
    }
    const float alpha = alpha_of(c.m.roughness);
    const float hPdf = ggx_d(cosH, alpha) * cosH;
    return c.w.specular * hPdf / max2(4.0f * cosHo, 1.0e-8f);
}

PT_DEV float pdf_transmit(const Ctx& c, V3 wi)
{
    const float cosWoS = dot(c.n, c.wo);
    if (!transmit_valid(c, cosWoS)) {
        return 0.0f;
    }
    const float thin = clamp01(c.m.thin);
    const float ior = material_ior(c);
    const float etaT = cosWoS > 0.0f ? ior : 1.0f;
    const float refl = dielectric_fresnel(cosWoS, c.eta, etaT);
    const float trans = 1.0f - refl;
    float pdf = 0.0f;
    const V3 reflected = mirror(c.wo, c.n);
    if (dot(wi, reflected) > 0.999f) {
        const float cosWi = absf(dot(c.n, wi));
        pdf += c.w.transmit * refl / max2(cosWi, 1.0e-8f);
    }
    V3 tr;
    if (thin > 0.5f) {
        tr = normalize(scale(c.wo, -1.0f));
    } else {
        float next = 1.0f;
        if (!refract_dir(c, c.n, tr, next)) {
            return pdf;
        }
    }
    if (dot(wi, tr) > 0.999f) {
        const float cosWi = absf(dot(c.n, tr));
        pdf += c.w.transmit * trans / max2(cosWi, 1.0e-8f);
    }
    return pdf;
}

PT_DEV float pdf_subsurface(const Ctx& c, V3 wi)
{
    const float cosPdf = max2(0.0f, max2(0.0f, dot(c.n, wi))) * kInvPi;
    const float mr = max_radius(c.m);
    if (mr <= kMinRadius) {
        return c.w.subsurface * cosPdf;
    }
    V3 t, b;
    basis(c.n, t, b);
    const V3 wt = add(scale(t, dot(wi, t)), scale(b, dot(wi, b)));
    const float dist = sqrtf(dot(wt, wt));
    const float radial = burley_pdf_radius(dist, mr);
    return c.w.subsurface * max2(cosPdf, radial * 0.5f);
}

PT_DEV float brdf_pdf(const Ctx& c, V3 wi)
{
    const float pd = c.w.diffuse * (max2(0.0f, max2(0.0f, dot(c.n, wi))) * kInvPi);
    const float ps = pdf_specular(c, wi);
    const float pt = c.w.transmit != 0.0f ? pdf_transmit(c, wi) : 0.0f;
    const float pss = c.w.subsurface != 0.0f ? pdf_subsurface(c, wi) : 0.0f;
    return pd + ps + pt + pss;
}

PT_DEV Sample empty_sample()
{
    Sample s;
    s.dir = mk(0.0f, 0.0f, 0.0f);
    s.pdf = 0.0f;
    s.valid = false;
    s.transmitted = false;
    s.sss = false;
    s.nextEta = 1.0f;
    s.exitOffset = mk(0.0f, 0.0f, 0.0f);
    s.steps = 0;
    return s;
}

PT_DEV Sample sample_diffuse(const Ctx& c, float u1, float u2)
{
    Sample r = empty_sample();
    V3 t, b;
    basis(c.n, t, b);
    r.dir = to_world(cosine_local(u1, u2), c.n, t, b);
    r.pdf = brdf_pdf(c, r.dir);
    r.valid = dot(c.n, r.dir) > 0.0f && r.pdf > kMinPdf;
    r.nextEta = c.eta;
    return r;
}

PT_DEV Sample sample_specular(const Ctx& c, float u1, float u2)
{
    Sample r = empty_sample();
    const float cosWo = max2(0.0f, dot(c.n, c.wo));
    if (cosWo <= 0.0f) {
        return r;
    }
    const float alpha = alpha_of(c.m.roughness);
    V3 t, b;
    basis(c.n, t, b);
    float hPdf = 0.0f;
    const V3 hl = ggx_half_local(alpha, u1, u2, hPdf);
    const V3 h = to_world(hl, c.n, t, b);
    const float cosHo = max2(0.0f, dot(h, c.wo));
    if (cosHo <= 0.0f || hPdf <= kMinPdf) {
        return r;
    }
    r.dir = normalize(sub(scale(h, 2.0f * cosHo), c.wo));
    const float cosWi = max2(0.0f, dot(c.n, r.dir));
    if (cosWi <= 0.0f) {
        return r;
    }
    r.pdf = brdf_pdf(c, r.dir);
    r.valid = r.pdf > kMinPdf;
    r.nextEta = c.eta;
    return r;
}

PT_DEV Sample sample_transmit(const Ctx& c, float u1, float u2)
{
    Sample r = empty_sample();
    const float cosWoS = dot(c.n, c.wo);
    if (!transmit_valid(c, cosWoS)) {
        return r;
    }
    const float thin = clamp01(c.m.thin);
    const float ior = material_ior(c);
    const float etaT = cosWoS > 0.0f ? ior : 1.0f;
    const float refl = dielectric_fresnel(cosWoS, c.eta, etaT);
    const float trans = 1.0f - refl;
    bool refractBranch = u1 < trans;
    if ((c.flags & 2) != 0) {
        refractBranch = false;
    }
    if ((c.flags & 4) != 0) {
        refractBranch = true;
    }
    V3 rn = c.n;
    const float alpha = alpha_of(c.m.roughness);
    if (refractBranch && thin < 0.5f && alpha > 0.04f) {
        V3 t, b;
        basis(c.n, t, b);
        float hPdf = 0.0f;
        const V3 hl = ggx_half_local(alpha, u2, u1 * 0.413f + u2 * 0.587f, hPdf);
        rn = to_world(hl, c.n, t, b);
    }
    if (refractBranch) {
        if (thin > 0.5f) {
            r.dir = normalize(scale(c.wo, -1.0f));
            r.nextEta = c.eta;
        } else {
            if (!refract_dir(c, rn, r.dir, r.nextEta)) {
                if ((c.flags & 8) != 0) {
                    return r;
                }
                r.dir = mirror(c.wo, c.n);
                r.transmitted = false;
                r.nextEta = c.eta;
            } else {
                r.transmitted = true;
            }
        }
        if (thin > 0.5f) {
            r.transmitted = true;
        }
        if (absf(dot(c.n, r.dir)) <= 0.0f) {
            return r;
        }
    } else {
        r.dir = mirror(c.wo, c.n);
        if (absf(dot(c.n, r.dir)) <= 0.0f) {
            return r;
        }
        r.transmitted = false;
        r.nextEta = c.eta;
    }
    r.pdf = brdf_pdf(c, r.dir);
    r.valid = r.pdf > kMinPdf;
    return r;
}

PT_DEV Sample sample_subsurface(const Ctx& c, float u1, float u2)
{
    const float mr = max_radius(c.m);
    if (mr <= kMinRadius) {
        return sample_diffuse(c, u1, u2);
    }
    Sample r = empty_sample();
    const float u3 = u1 * 0.318f + u2 * 0.682f;
    V3 t, b;
    basis(c.n, t, b);
    // SubsurfaceCore.h:50-64, 84-106
    const V3 rad = mk(max2(0.0f, c.m.srR), max2(0.0f, c.m.srG), max2(0.0f, c.m.srB));
    float er;
    if (u3 < rad.x / (rad.x + rad.y + rad.z + kMinRadius)) {
        er = max2(rad.x, kMinRadius);
    } else if (u3 < (rad.x + rad.y) / (rad.x + rad.y + rad.z + kMinRadius)) {
        er = max2(rad.y, kMinRadius);
    } else {
        er = max2(rad.z, kMinRadius);
    }
    const float rr = burley_sample_radius(u1, er);
    const float phi = 2.0f * kPi * u2;
    float sp, cp;
    sincosf_pt(phi, sp, cp);
    r.exitOffset = add(scale(t, rr * cp), scale(b, rr * sp));
    const V3 l = cosine_local(max2(kMinPdf, u1 * 0.5f + 0.25f), max2(kMinPdf, u2 * 0.5f + 0.25f));
    r.dir = to_world(l, c.n, t, b);
    r.sss = true;
    r.steps = clamp01(c.m.thin) > 0.5f ? 0 : 2;
    r.pdf = brdf_pdf(c, r.dir);
    r.valid = dot(c.n, r.dir) > 0.0f && r.pdf > kMinPdf;
    r.nextEta = c.eta;
    return r;
}

// PrincipledBrdf.h:553-579
PT_DEV Sample brdf_sample(const Ctx& c, float u1, float u2)
{
    if ((c.flags & 1) != 0) {
        return sample_transmit(c, u1, u2);
    }
    const Lobes& w = c.w;
    float u = u1;
    if (u < w.diffuse) {
        return sample_diffuse(c, u1 / max2(w.diffuse, kMinPdf), u2);
    }
    u -= w.diffuse;
    if (u < w.specular) {
        return sample_specular(c, (u1 - w.diffuse) / max2(w.specular, kMinPdf), u2);
    }
    u -= w.specular;
    if (u < w.transmit) {
        return sample_transmit(c, (u1 - w.diffuse - w.specular) / max2(w.transmit, kMinPdf), u2);
    }
    return sample_subsurface(c, (u1 - w.diffuse - w.specular - w.transmit) / max2(w.subsurface, kMinPdf), u2);
}

// BrdfDispatch.h:42-76
template <class C>
PT_DEV C apply_throughput_t(C T, const Ctx& c, const Sample& s, C f, const C& base)
{
    if (!s.valid || s.pdf <= kMinPdf) {
        return ColN<C>::splat(0.0f);
    }
    const float tr = max2(0.0f, min2(1.0f, c.m.transmission));
    const float thin = max2(0.0f, min2(1.0f, c.m.thin));
    const bool pure = tr > 0.99f && c.m.metallic < 0.01f;
    C sc;
    if (pure && thin < 0.5f) {
        sc = scale(f, 1.0f / s.pdf);
    } else if (s.transmitted) {
        sc = scale(f, absf(dot(c.n, s.dir)) / s.pdf);
        if (thin > 0.5f) {
            sc = mul(sc, base);
        }
    } else if (s.sss) {
        sc = scale(f, max2(0.0f, dot(c.n, s.dir)) / s.pdf);
    } else {
        const float cs = c.eta > 1.0f + 1.0e-4f ? absf(dot(c.n, s.dir)) : max2(0.0f, dot(c.n, s.dir));
        sc = scale(f, cs / s.pdf);
    }
    return mul(T, sc);
}

PT_DEV V3 apply_throughput(V3 T, const Ctx& c, const Sample& s, V3 f)
{
    return apply_throughput_t(T, c, s, f, mat_base(c.m));
}

PT_DEV bool skips_env_nee(const Mat& m)
{
    const float tr = max2(0.0f, min2(1.0f, m.transmission));
    const float me = max2(0.0f, min2(1.0f, m.metallic));
    return tr > 0.99f && me < 0.01f;
}

PT_DEV float mis_balance(float a, float b)
{
    if (a <= 0.0f) {
        return 0.0f;
    }
    const float d = a + b;
    if (d <= 0.0f) {
        return 0.0f;
    }
    return a / d;
}

// ----------------------------------------------------------------------------------------
// Environment and emissive lights (Sampling/LightSamplingCore.h)
// ----------------------------------------------------------------------------------------
PT_DEV V3 lat_long_to_dir(float u, float v)
{
    const float phi = u * 2.0f * kPi - kPi;
    const float theta = v * kPi;
    float st, ct, sp, cp;
    sincosf_pt(theta, st, ct);
    sincosf_pt(phi, sp, cp);
    return normalize(mk(st * sp, ct, st * cp));
}

PT_DEV V3 dir_to_lat_long(V3 d)
{
    const V3 w = normalize(d);
    const float phi = atan2f_pt(w.x, w.z);
    const float theta = acosf_pt(max2(-1.0f, min2(w.y, 1.0f)));
    return mk((phi + kPi) / (2.0f * kPi), theta / kPi, 0.0f);
}

PT_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

PT_DEV V3 env_eval(const EnvironmentMapGpu* env, const RenderParamsGpu& p, V3 dir)
{
    if (env != nullptr && env->valid != 0) {
        if (env->pixels == nullptr || env->width <= 0 || env->height <= 0) {
            return mk(0.0f, 0.0f, 0.0f);
        }
        const V3 uv = dir_to_lat_long(dir);
        const int x = clampi(static_cast<int>(uv.x * static_cast<float>(env->width)), 0, env->width - 1);
        const int y = clampi(static_cast<int>(uv.y * static_cast<float>(env->height)), 0, env->height - 1);
        const int i = (y * env->width + x) * 3;
        return mk(env->pixels[i], env->pixels[i + 1], env->pixels[i + 2]);
    }
    const float k = p.environmentIntensity;
    return mk(p.backgroundR * k, p.backgroundG * k, p.backgroundB * k);
}

PT_DEV bool env_tables_ok(const EnvironmentMapGpu* env)
{
    return !(env->pixels == nullptr || env->marginalCdf == nullptr || env->rowCdf == nullptr || env->width <= 0 ||
             env->height <= 0);
}

PT_DEV int find_cdf(const float* cdf, int count, float u)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (cdf[mid] < u) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

PT_DEV V3 env_sample(const EnvironmentMapGpu* env, float u1, float u2, float& pdf)
{
    if (env != nullptr && env->valid != 0) {
        pdf = 0.0f;
        if (!env_tables_ok(env)) {
            return mk(0.0f, 1.0f, 0.0f);
        }
        const int y = find_cdf(env->marginalCdf, env->height, u1);
        const int ro = y * (env->width + 1);
        const int x = find_cdf(env->rowCdf + ro, env->width, u2);
        const float u = (static_cast<float>(x) + 0.5f) / static_cast<float>(env->width);
        const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(env->height);
        const V3 dir = lat_long_to_dir(u, v);
        const int pi = (y * env->width + x) * 3;
        const float lum = max2(max2(env->pixels[pi], env->pixels[pi + 1]), env->pixels[pi + 2]);
        const float rowPdf = env->rowCdf[ro + x + 1] - env->rowCdf[ro + x];
        const float margPdf = env->marginalCdf[y + 1] - env->marginalCdf[y];
        const float st = max2(1.0e-4f, sinf_pt(v * kPi));
        pdf = (margPdf * rowPdf) * static_cast<float>(env->width * env->height) / (2.0f * kPi * kPi * st);
        if (lum <= 0.0f) {
            pdf = 0.0f;
        }
        return dir;
    }
    pdf = 1.0f / (4.0f * kPi);
    return lat_long_to_dir(u1, u2);
}

PT_DEV float env_pdf(const EnvironmentMapGpu* env, V3 wi)
{
    if (env != nullptr && env->valid != 0) {
        if (!env_tables_ok(env)) {
            return 0.0f;
        }
        const V3 uv = dir_to_lat_long(wi);
        const int x = clampi(static_cast<int>(uv.x * static_cast<float>(env->width)), 0, env->width - 1);
        const int y = clampi(static_cast<int>(uv.y * static_cast<float>(env->height)), 0, env->height - 1);
        const int ro = y * (env->width + 1);
        const float rowPdf = env->rowCdf[ro + x + 1] - env->rowCdf[ro + x];
        const float margPdf = env->marginalCdf[y + 1] - env->marginalCdf[y];
        const float st = max2(1.0e-4f, sinf_pt(uv.y * kPi));
        return (margPdf * rowPdf) * static_cast<float>(env->width * env->height) / (2.0f * kPi * kPi * st);
    }
    return 1.0f / (4.0f * kPi);
}

// LightSamplingCore.h:257-313 on the emissive array (v0, v1, v2 with r*e, g*e, b*e in .w).
PT_DEV bool sample_emissive(const MeshAccelSceneGpu& s, float uTri, float u1, float u2, V3& pos, V3& ln, V3& L,
                            float& pdf, int* entry = nullptr)
{
    pdf = 0.0f;
    if (s.emissiveTriangleCount == 0 || s.emissiveTriangleCdf == nullptr || s.emissiveTris == nullptr) {
        return false;
    }
    const float total = s.emissiveTriangleCdf[s.emissiveTriangleCount];
    if (total <= kMinPdf) {
        return false;
    }
    const float target = uTri * total;
    int lo = 0, hi = static_cast<int>(s.emissiveTriangleCount) - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s.emissiveTriangleCdf[mid + 1] < target) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    const float4 a = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(lo) + 0u);
    const float4 b = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(lo) + 1u);
    const float4 c = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(lo) + 2u);
    const V3 v0 = mk(a.x, a.y, a.z), v1 = mk(b.x, b.y, b.z), v2 = mk(c.x, c.y, c.z);
    if (entry != nullptr) {
        *entry = lo;
    }
    const float su = sqrtf(max2(0.0f, u1));
    const float b0 = 1.0f - su;
    const float b1 = u2 * su;
    const float b2 = 1.0f - b0 - b1;
    pos = add(add(scale(v0, b0), scale(v1, b1)), scale(v2, b2));
    const V3 cr = cross(sub(v1, v0), sub(v2, v0));
    ln = normalize(cr);
    L = mk(a.w, b.w, c.w);
    const float area = 0.5f * length(cr);
    if (area <= kMinPdf) {
        return false;
    }
    const float sel = (s.emissiveTriangleCdf[lo + 1] - s.emissiveTriangleCdf[lo]) / total;
    pdf = sel / area;
    return pdf > kMinPdf;
}

// Whether an any-hit trace of r over [tMin, tMax) is certain to report a hit because of emissive
// list entry `entry` (its vertex floats are the scene triangle's). Every traversal here tests a
// triangle exactly when the slab test of its own box (its vertex floats as planes) admits the ray
// within [tMin, tMax * kCullSlack]: the boxes above it contain that box, and slab distances move
// monotonically with the planes, so no ancestor rejects what the triangle's box admits. When that
// box admits the ray and tri_test finds the triangle at t < tMax, every traversal reports the ray
// occluded. The reference's emissive next-event shadow ray runs to kRayTMax along the direction of
// a point sampled on the triangle, so it nearly always meets the triangle itself: one triangle test
// settles it, and only the rays this cannot settle are traced.
PT_DEV bool emissive_blocks(const MeshAccelSceneGpu& s, int entry, const Ray& r, float tMin, float tMax)
{
    const float4 a = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(entry) + 0u);
    const float4 b = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(entry) + 1u);
    const float4 c = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(entry) + 2u);
    float te;
    if (!slab(r, min2(min2(a.x, b.x), c.x), min2(min2(a.y, b.y), c.y), min2(min2(a.z, b.z), c.z),
              max2(max2(a.x, b.x), c.x), max2(max2(a.y, b.y), c.y), max2(max2(a.z, b.z), c.z), tMin,
              tMax * kCullSlack, te)) {
        return false;
    }
    const float4 e1 = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
    const float4 e2 = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
    float t, u, v;
    return tri_test(r, a, e1, e2, tMin, tMax, t, u, v) && t < tMax;
}

// Position of original triangle `tri` in the emissive list (ascending triangle indices), or -1.
PT_DEV int emissive_slot(const MeshAccelSceneGpu& s, uint32_t tri)
{
    int lo = 0, hi = static_cast<int>(s.emissiveTriangleCount) - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        const uint32_t v = s.emissiveTriangleIndices[mid];
        if (v == tri) {
            return mid;
        }
        if (v < tri) {
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    return -1;
}

// Solid-angle density with which emissive next-event estimation (sample_emissive, then
// areaPdf * d2 / cosL as the integrator forms it) picks the point at distance t along dir on
// triangle tri; 0 where it never does (not an emissive entry, zero weight, back face).
PT_DEV float emissive_pdf_hit(const MeshAccelSceneGpu& s, uint32_t tri, V3 dir, float t)
{
    if (s.emissiveTriangleCount == 0 || s.emissiveTriangleCdf == nullptr || s.emissiveTris == nullptr) {
        return 0.0f;
    }
    const float total = s.emissiveTriangleCdf[s.emissiveTriangleCount];
    if (total <= kMinPdf) {
        return 0.0f;
    }
    const int li = emissive_slot(s, tri);
    if (li < 0) {
        return 0.0f;
    }
    const float4 a = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(li) + 0u);
    const float4 b = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(li) + 1u);
    const float4 c = ld4(s.emissiveTris, 3u * static_cast<uint32_t>(li) + 2u);
    const V3 v0 = mk(a.x, a.y, a.z), v1 = mk(b.x, b.y, b.z), v2 = mk(c.x, c.y, c.z);
    const V3 cr = cross(sub(v1, v0), sub(v2, v0));
    const V3 ln = normalize(cr);
    const float area = 0.5f * length(cr);
    if (area <= kMinPdf) {
        return 0.0f;
    }
    const float cosL = max2(0.0f, dot(ln, scale(dir, -1.0f)));
    if (cosL <= 0.0f) {
        return 0.0f;
    }
    const float sel = (s.emissiveTriangleCdf[li + 1] - s.emissiveTriangleCdf[li]) / total;
    const float areaPdf = sel / area;
    return areaPdf * (t * t) / cosL;
}

// Balance weight of emission found by a BSDF-sampled ray (kLightMis): prevPdf is the pdf of the
// sampled direction, reach whether the previous vertex's next-event estimate covers it.
PT_DEV float emission_mis_weight(const MeshAccelSceneGpu& s, uint32_t tri, V3 dir, float t, float prevPdf,
                                 bool reach)
{
    if (!reach) {
        return 1.0f;
    }
    const float lightPdf = emissive_pdf_hit(s, tri, dir, t);
    if (lightPdf <= 0.0f) {
        return 1.0f;
    }
    return mis_balance(prevPdf, lightPdf);
}

// ----------------------------------------------------------------------------------------
// Camera (CameraDevice.cuh:8-139)
// ----------------------------------------------------------------------------------------
PT_DEV V3 rot_quat(const CameraGpu& cam, V3 v)
{
    const float s = cam.orientation.x;
    const V3 u = mk(cam.orientation.y, cam.orientation.z, cam.orientation.w);
    const V3 c1 = mk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
    const V3 c2 = mk(u.y * c1.z - u.z * c1.y, u.z * c1.x - u.x * c1.z, u.x * c1.y - u.y * c1.x);
    return mk(v.x + 2.0f * c1.x * s + 2.0f * c2.x, v.y + 2.0f * c1.y * s + 2.0f * c2.y,
              v.z + 2.0f * c1.z * s + 2.0f * c2.z);
}

PT_DEV V3 film_local(const CameraGpu& cam, float u, float v)
{
    const float ndcX = 2.0f * u - 1.0f;
    const float ndcY = 1.0f - 2.0f * v;
    const float tanHalf = tanf_pt(cam.fovY * 0.5f);
    const float top = cam.nearPlane * tanHalf;
    const float right = top * cam.aspect;
    return mk(ndcX * right, ndcY * top, -cam.nearPlane);
}

PT_DEV void camera_ray(const CameraGpu& cam, float u, float v, float l1, float l2, V3& ro, V3& rd)
{
    const V3 pos = mk(cam.position.x, cam.position.y, cam.position.z);
    if (cam.focusValid == 0 || cam.apertureRadius <= 0.0f || cam.focusDistance <= 0.0f || cam.nearPlane <= 0.0f) {
        ro = pos;
        rd = normalize(rot_quat(cam, film_local(cam, u, v)));
        return;
    }
    const V3 right = normalize(rot_quat(cam, mk(1.0f, 0.0f, 0.0f)));
    const V3 up = normalize(rot_quat(cam, mk(0.0f, 1.0f, 0.0f)));
    const V3 film = film_local(cam, u, v);
    const float fs = cam.focusDistance / cam.nearPlane;
    const V3 rf = rot_quat(cam, mk(film.x * fs, film.y * fs, -cam.focusDistance));
    const V3 focusWorld = mk(pos.x + rf.x, pos.y + rf.y, pos.z + rf.z);
    float dx = 0.0f, dy = 0.0f;
    const float a = 2.0f * l1 - 1.0f;
    const float b = 2.0f * l2 - 1.0f;
    if (!(a == 0.0f && b == 0.0f)) {
        float r, th;
        if (fabsf(a) > fabsf(b)) {
            r = a;
            th = (3.14159265358979323846f * 0.25f) * (b / a);
        } else {
            r = b;
            th = (3.14159265358979323846f * 0.5f) - (3.14159265358979323846f * 0.25f) * (a / b);
        }
        float sth, cth;
        sincosf_pt(th, sth, cth);
        dx = r * cth;
        dy = r * sth;
    }
    const float lx = dx * cam.apertureRadius;
    const float ly = dy * cam.apertureRadius;
    ro = mk(pos.x + right.x * lx + up.x * ly, pos.y + right.y * lx + up.y * ly, pos.z + right.z * lx + up.z * ly);
    rd = normalize(mk(focusWorld.x - ro.x, focusWorld.y - ro.y, focusWorld.z - ro.z));
}

// PathTracer.cu:49-67
PT_DEV V3 clamp_firefly(V3 r)
{
    if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) {
        return mk(0.0f, 0.0f, 0.0f);
    }
    r = mk(max2(0.0f, r.x), max2(0.0f, r.y), max2(0.0f, r.z));
    const float lum = r.x * 0.2126f + r.y * 0.7152f + r.z * 0.0722f;
    if (lum <= 100.0f) {
        return r;
    }
    const float k = 100.0f / lum;
    return mk(r.x * k, r.y * k, r.z * k);
}

} // namespace pt
