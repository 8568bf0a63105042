/*
 * orc_math.c -- oracle vector helpers, project transcendental functions and RNG.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Transcendentals: the reference calls CUDA's sinf/cosf/tanf/expf/logf/atan2f/acosf.
 * Those cannot be reproduced here, so the oracle and the HIP kernels share one
 * specification: evaluate in IEEE double (Cody-Waite argument reduction with the
 * split constants of pi/2 and ln 2, then the Taylor series of the function, which
 * converges to far below float resolution on the reduced interval) and round once
 * to float. Only IEEE add/sub/mul/div/sqrt/rint are used, so every conforming
 * implementation produces the same bits. tests/test_oracle_math.py bounds the
 * error against libm (<= 1 ulp) and checks identities.
 */
#include "oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <string.h>

/* pi/2 split into a 33-bit head and a tail; ln 2 likewise (digits of the constants). */
static const double kPio2Hi = 1.57079632673412561417e+00;
static const double kPio2Lo = 6.07710050650619224932e-11;
static const double kTwoOverPi = 6.36619772367581382433e-01;
static const double kLn2Hi = 6.93147180369123816490e-01;
static const double kLn2Lo = 1.90821492927058770002e-10;
static const double kInvLn2 = 1.44269504088896338700e+00;
static const double kLn2 = 6.93147180559945286227e-01;
static const double kPiD = 3.14159265358979311600e+00;
static const double kPio2D = 1.57079632679489655800e+00;

static double sin_series(double r)
{
    const double r2 = r * r;
    /* Taylor coefficients (-1)^k / (2k+1)!, k = 1..7 */
    double p = -1.0 / 1307674368000.0;
    p = p * r2 + 1.0 / 6227020800.0;
    p = p * r2 + -1.0 / 39916800.0;
    p = p * r2 + 1.0 / 362880.0;
    p = p * r2 + -1.0 / 5040.0;
    p = p * r2 + 1.0 / 120.0;
    p = p * r2 + -1.0 / 6.0;
    return r + r * (r2 * p);
}

static double cos_series(double r)
{
    const double r2 = r * r;
    /* Taylor coefficients (-1)^k / (2k)!, k = 1..8 */
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

/* Reduce x to r in [-pi/4, pi/4] with x = r + q*pi/2. */
static double reduce_pio2(double x, int* quadrant)
{
    const double k = rint(x * kTwoOverPi);
    const double r = (x - k * kPio2Hi) - k * kPio2Lo;
    const long long ki = (long long)k;
    *quadrant = (int)(ki & 3);
    return r;
}

double orc_sin_d(double x)
{
    int q = 0;
    const double r = reduce_pio2(x, &q);
    switch (q) {
    case 0: return sin_series(r);
    case 1: return cos_series(r);
    case 2: return -sin_series(r);
    default: return -cos_series(r);
    }
}

double orc_cos_d(double x)
{
    int q = 0;
    const double r = reduce_pio2(x, &q);
    switch (q) {
    case 0: return cos_series(r);
    case 1: return -sin_series(r);
    case 2: return -cos_series(r);
    default: return sin_series(r);
    }
}

float orc_sinf(float x)
{
    if (!isfinite(x)) {
        return x - x;
    }
    if (x == 0.0f) {
        return x; /* keeps the sign of zero */
    }
    return (float)orc_sin_d((double)x);
}

float orc_cosf(float x)
{
    if (!isfinite(x)) {
        return x - x;
    }
    return (float)orc_cos_d((double)x);
}

float orc_tanf(float x)
{
    if (!isfinite(x)) {
        return x - x;
    }
    if (x == 0.0f) {
        return x;
    }
    const double xd = (double)x;
    return (float)(orc_sin_d(xd) / orc_cos_d(xd));
}

float orc_expf(float x)
{
    if (x != x) {
        return x;
    }
    if (x > 100.0f) {
        return INFINITY;
    }
    if (x < -120.0f) {
        return 0.0f;
    }
    const double xd = (double)x;
    const double k = rint(xd * kInvLn2);
    const double r = (xd - k * kLn2Hi) - k * kLn2Lo;
    /* Taylor series of e^r, |r| <= ln2/2, through r^15/15! */
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
    return (float)ldexp(p, (int)k);
}

/* Natural log of a positive finite double m*2^e with the mantissa split at sqrt(2). */
static double log_positive(double x)
{
    int e = 0;
    double m = frexp(x, &e); /* m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    /* log(m) = 2 atanh(s) = 2 sum s^(2k+1)/(2k+1), k = 0..10 */
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
    return (double)e * kLn2 + 2.0 * s * p;
}

float orc_logf(float x)
{
    if (x != x) {
        return x;
    }
    if (x < 0.0f) {
        return NAN;
    }
    if (x == 0.0f) {
        return -INFINITY;
    }
    if (isinf(x)) {
        return x;
    }
    return (float)log_positive((double)x);
}

/* atan of z in [0, 1] in double: two half-angle reductions, then Taylor. */
static double atan_unit(double z)
{
    /* atan(z) = 2 atan(z / (1 + sqrt(1 + z^2))), applied twice: |z| <= tan(pi/16) */
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

double orc_atan2_d(double y, double x)
{
    const double ay = fabs(y);
    const double ax = fabs(x);
    double a = 0.0;
    if (ay == 0.0 && ax == 0.0) {
        a = 0.0;
    } else if (ay <= ax) {
        a = atan_unit(ay / ax);
    } else {
        a = kPio2D - atan_unit(ax / ay);
    }
    if (x < 0.0 || (x == 0.0 && signbit(x))) {
        a = kPiD - a;
    }
    return signbit(y) ? -a : a;
}

float orc_atan2f(float y, float x)
{
    if (x != x || y != y) {
        return x + y;
    }
    return (float)orc_atan2_d((double)y, (double)x);
}

float orc_acosf(float x)
{
    if (x != x || x > 1.0f || x < -1.0f) {
        return NAN;
    }
    const double xd = (double)x;
    const double s = sqrt((1.0 - xd) * (1.0 + xd));
    return (float)orc_atan2_d(s, xd);
}

/* ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -------------------- */
static void philox_round(uint32_t c[4], const uint32_t k[2])
{
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n1 = lo1;
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    const uint32_t n3 = lo0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof(c));
}

/* Sampling/RandCore.h:9-10: the reference's per-sample seed hash (u32 wrap). */
uint32_t orc_seed_hash(int pixelIndex, int sampleIndex, uint32_t globalSeed)
{
    return (uint32_t)pixelIndex * 747796405u + (uint32_t)sampleIndex + globalSeed * 2891336453u;
}

/* Stream for one pixel sample: key = (seed hash, global seed), counter = (block, pixel, sample, 0). */
void orc_rng_init(ORng* rng, int pixelIndex, int sampleIndex, uint32_t globalSeed)
{
    rng->key[0] = orc_seed_hash(pixelIndex, sampleIndex, globalSeed);
    rng->key[1] = globalSeed;
    rng->ctr[0] = 0u;
    rng->ctr[1] = (uint32_t)pixelIndex;
    rng->ctr[2] = (uint32_t)sampleIndex;
    rng->ctr[3] = 0u;
    rng->avail = 0;
}

/* u32 -> float in (0, 1]: ((x >> 8) + 1) * 2^-24, exact in float. */
float orc_u32_to_unit(uint32_t x)
{
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

float orc_rand01(ORng* rng)
{
    if (rng->avail == 0) {
        orc_philox4x32_10(rng->ctr, rng->key, rng->buf);
        rng->ctr[0] += 1u;
        rng->avail = 4;
    }
    const uint32_t x = rng->buf[4 - rng->avail];
    rng->avail -= 1;
    return orc_u32_to_unit(x);
}
