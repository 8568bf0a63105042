/*
 * orc_mesh.c -- oracle restatement of the reference's tube loft and root transform.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). The reference's Loft.cpp and ProceduralMeshBuilder.cpp
 * cannot be compiled here (they need the Manifold library), so the triangles they produce for a
 * lofted segment are restated from the source:
 *   - buildLoftHostMeshFromFrames, Application/src/Procedural/Loft.cpp:108-155, with ringVertex
 *     :24-32, planarCapUv :34-40, sideWallCorner :63-75, capRimCorner :77-83, capCenterCorner
 *     :85-89;
 *   - applyRootTransform / rotateYawPitchRoll / rotateAroundAxis, ProceduralMeshBuilder.cpp:198-247.
 * Host code: cosf/sinf are the C library's, as on the reference's host. Float order follows the
 * reference's expressions (MathCore.h helpers: vecNormalize3 scales by 1/length, (0, 0, -1) for a
 * zero vector).
 *
 * Layouts: a frame is 13 floats (position, tangent, normal, binormal, radius); a triangle is the
 * 25-word HostTriangle (v0 v1 v2 n0 n1 n2 as 3 floats each, uv0 uv1 uv2 as 2 floats each, then the
 * material index word, left untouched here).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static void unit3(const float v[3], float out[3])
{
    const float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len <= 0.0f) {
        out[0] = 0.0f;
        out[1] = 0.0f;
        out[2] = -1.0f;
        return;
    }
    const float inv = 1.0f / len;
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
}

/* Loft.cpp:24-32: frame.position + (normal * (c * r) + binormal * (s * r)), angle = k * 2pi / N. */
static void ring_point(const float* fr, int k, int n, float out[3])
{
    const float angle = (float)k * 6.283185307f / (float)n;
    const float c = cosf(angle);
    const float s = sinf(angle);
    const float r = fr[12];
    for (int a = 0; a < 3; ++a) {
        out[a] = fr[a] + (fr[6 + a] * (c * r) + fr[9 + a] * (s * r));
    }
}

typedef struct {
    float p[3], n[3], uv[2];
} Corner;

/* Loft.cpp:63-75: the column index wraps to 0 for the position; u = column / N. */
static Corner wall_corner(const float* fr, int column, int n, float v)
{
    Corner c;
    ring_point(fr, column < n ? column : 0, n, c.p);
    const float d[3] = {c.p[0] - fr[0], c.p[1] - fr[1], c.p[2] - fr[2]};
    unit3(d, c.n);
    c.uv[0] = (float)column / (float)n;
    c.uv[1] = v;
    return c;
}

/* Loft.cpp:77-89 */
static Corner cap_corner(const float* fr, int k, int n, float sign, int centre)
{
    Corner c;
    const float t[3] = {fr[3] * sign, fr[4] * sign, fr[5] * sign};
    unit3(t, c.n);
    if (centre) {
        memcpy(c.p, fr, sizeof c.p);
        c.uv[0] = 0.5f;
        c.uv[1] = 0.5f;
        return c;
    }
    ring_point(fr, k, n, c.p);
    const float angle = (float)k * 6.283185307f / (float)n;
    c.uv[0] = 0.5f + 0.5f * sinf(angle);
    c.uv[1] = 0.5f + 0.5f * cosf(angle);
    return c;
}

static void emit(float* out, const Corner* a, const Corner* b, const Corner* c)
{
    const Corner* k[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) {
        memcpy(out + 3 * i, k[i]->p, 3 * sizeof(float));
        memcpy(out + 9 + 3 * i, k[i]->n, 3 * sizeof(float));
        memcpy(out + 18 + 2 * i, k[i]->uv, 2 * sizeof(float));
    }
}

/* Loft.cpp:108-155. Writes (frames - 1) * 2N + 2N triangles (0 when frames < 2 or N < 3) and
 * returns the count; out may be NULL to query it. */
int orc_loft_triangles(const float* frames, int frameCount, int n, float* out)
{
    if (frameCount < 2 || n < 3) {
        return 0;
    }
    const int total = (frameCount - 1) * n * 2 + n * 2;
    if (out == NULL) {
        return total;
    }
    const float invSpan = 1.0f / (float)(frameCount - 1);
    int w = 0;
    for (int ring = 0; ring < frameCount - 1; ++ring) {
        const float* f0 = frames + 13 * ring;
        const float* f1 = frames + 13 * (ring + 1);
        const float v0 = (float)ring * invSpan;
        const float v1 = (float)(ring + 1) * invSpan;
        for (int k = 0; k < n; ++k) {
            const Corner a = wall_corner(f0, k, n, v0);
            const Corner b = wall_corner(f0, k + 1, n, v0);
            const Corner c = wall_corner(f1, k, n, v1);
            const Corner d = wall_corner(f1, k + 1, n, v1);
            emit(out + 25 * (size_t)w++, &a, &b, &c);
            emit(out + 25 * (size_t)w++, &b, &d, &c);
        }
    }
    const float* fs = frames;
    const float* fe = frames + 13 * (frameCount - 1);
    const Corner cs = cap_corner(fs, 0, n, -1.0f, 1);
    const Corner ce = cap_corner(fe, 0, n, 1.0f, 1);
    for (int k = 0; k < n; ++k) {
        const int k1 = (k + 1) % n;
        const Corner s0 = cap_corner(fs, k, n, -1.0f, 0);
        const Corner s1 = cap_corner(fs, k1, n, -1.0f, 0);
        emit(out + 25 * (size_t)w++, &cs, &s1, &s0);
        const Corner e0 = cap_corner(fe, k, n, 1.0f, 0);
        const Corner e1 = cap_corner(fe, k1, n, 1.0f, 0);
        emit(out + 25 * (size_t)w++, &ce, &e0, &e1);
    }
    return w;
}

/* ProceduralMeshBuilder.cpp:203-217: v c + (a x v) s + a (a.v)(1 - c), a the unit axis. */
static void rotate_axis(float v[3], const float axis[3], float rad)
{
    float a[3];
    unit3(axis, a);
    const float c = cosf(rad);
    const float s = sinf(rad);
    const float d = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
    const float x[3] = {a[1] * v[2] - a[2] * v[1], a[2] * v[0] - a[0] * v[2], a[0] * v[1] - a[1] * v[0]};
    float r[3];
    for (int i = 0; i < 3; ++i) {
        r[i] = (v[i] * c + x[i] * s) + a[i] * (d * (1.0f - c));
    }
    memcpy(v, r, sizeof r);
}

/* ProceduralMeshBuilder.cpp:219-232: yaw about +Y, pitch about +X, roll about +Z (each only when
 * |angle| > 1e-6 degrees), angles times 0.0174532925. */
static void rotate_ypr(float v[3], const float deg[3])
{
    static const float ax[3][3] = {{0.0f, 1.0f, 0.0f}, {1.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    for (int i = 0; i < 3; ++i) {
        if (fabsf(deg[i]) > 1.0e-6f) {
            rotate_axis(v, ax[i], deg[i] * 0.0174532925f);
        }
    }
}

/* ProceduralMeshBuilder.cpp:234-247: vertices rotated then translated, normals rotated. */
void orc_root_transform(float* tris, int count, const float translation[3], const float rotationDeg[3])
{
    for (int t = 0; t < count; ++t) {
        float* x = tris + 25 * (size_t)t;
        for (int k = 0; k < 6; ++k) {
            float* v = x + 3 * k;
            rotate_ypr(v, rotationDeg);
            if (k < 3) {
                for (int a = 0; a < 3; ++a) {
                    v[a] = v[a] + translation[a];
                }
            }
        }
    }
}
