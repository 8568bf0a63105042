/*
 * orc_light.c -- oracle lights, MIS, material resolve, textures and camera rays.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "orc_internal.h"

#include <string.h>

#define K_PI 3.14159265f
#define K_MIN_PDF 1.0e-8f

/* Sampling/MisCore.h:11-21 */
float orc_mis_balance(float a, float b)
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

/* LightSamplingCore.h:25-33 */
OVec3 orc_dir_to_lat_long(OVec3 d)
{
    const OVec3 w = vnorm(d);
    const float phi = orc_atan2f(w.x, w.z);
    const float theta = orc_acosf(fmax2(-1.0f, fmin2(w.y, 1.0f)));
    const float u = (phi + K_PI) / (2.0f * K_PI);
    const float v = theta / K_PI;
    return v3(u, v, 0.0f);
}

/* LightSamplingCore.h:35-41 */
OVec3 orc_lat_long_to_dir(float u, float v)
{
    const float phi = u * 2.0f * K_PI - K_PI;
    const float theta = v * K_PI;
    const float st = orc_sinf(theta);
    return vnorm(v3(st * orc_sinf(phi), orc_cosf(theta), st * orc_cosf(phi)));
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* LightSamplingCore.h:43-54 */
static OVec3 env_eval_map(const OEnv* env, OVec3 dir)
{
    if (env == NULL || env->valid == 0 || env->pixels == NULL || env->width <= 0 || env->height <= 0) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    const OVec3 uv = orc_dir_to_lat_long(dir);
    const int x = clampi((int)(uv.x * (float)env->width), 0, env->width - 1);
    const int y = clampi((int)(uv.y * (float)env->height), 0, env->height - 1);
    const int i = (y * env->width + x) * 3;
    return v3(env->pixels[i], env->pixels[i + 1], env->pixels[i + 2]);
}

/* LightSamplingCore.h:56-77 */
OVec3 orc_env_eval(const OEnv* env, const OParams* params, OVec3 dir)
{
    if (env != NULL && env->valid != 0) {
        return env_eval_map(env, dir);
    }
    if (params == NULL) {
        return v3(0.04f, 0.04f, 0.04f);
    }
    const float k = params->environmentIntensity;
    return v3(params->backgroundR * k, params->backgroundG * k, params->backgroundB * k);
}

/* LightSamplingCore.h:79-92 */
static int find_cdf_row(const float* cdf, int count, float u)
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

static int env_tables_ok(const OEnv* env)
{
    return !(env == NULL || env->valid == 0 || env->pixels == NULL || env->marginalCdf == NULL ||
             env->rowCdf == NULL || env->width <= 0 || env->height <= 0);
}

/* LightSamplingCore.h:94-129, 149-153, 161-173 */
OVec3 orc_env_sample(const OEnv* env, const OParams* params, float u1, float u2, float* pdf)
{
    (void)params;
    if (env != NULL && env->valid != 0) {
        *pdf = 0.0f;
        if (!env_tables_ok(env)) {
            return v3(0.0f, 1.0f, 0.0f);
        }
        const int y = find_cdf_row(env->marginalCdf, env->height, u1);
        const int rowOffset = y * (env->width + 1);
        const int x = find_cdf_row(env->rowCdf + rowOffset, env->width, u2);
        const float u = ((float)x + 0.5f) / (float)env->width;
        const float v = ((float)y + 0.5f) / (float)env->height;
        const OVec3 dir = orc_lat_long_to_dir(u, v);
        const int pi = (y * env->width + x) * 3;
        const OVec3 rad = v3(env->pixels[pi], env->pixels[pi + 1], env->pixels[pi + 2]);
        const float lum = fmax2(fmax2(rad.x, rad.y), rad.z);
        const float rowPdf = env->rowCdf[rowOffset + x + 1] - env->rowCdf[rowOffset + x];
        const float margPdf = env->marginalCdf[y + 1] - env->marginalCdf[y];
        const float st = fmax2(1.0e-4f, orc_sinf(v * K_PI));
        *pdf = (margPdf * rowPdf) * (float)(env->width * env->height) / (2.0f * K_PI * K_PI * st);
        if (lum <= 0.0f) {
            *pdf = 0.0f;
        }
        return dir;
    }
    *pdf = 1.0f / (4.0f * K_PI);
    return orc_lat_long_to_dir(u1, u2);
}

/* LightSamplingCore.h:131-147, 155-159, 175-185 */
float orc_env_pdf(const OEnv* env, const OParams* params, OVec3 wi)
{
    (void)params;
    if (env != NULL && env->valid != 0) {
        if (!env_tables_ok(env)) {
            return 0.0f;
        }
        const OVec3 uv = orc_dir_to_lat_long(wi);
        const int x = clampi((int)(uv.x * (float)env->width), 0, env->width - 1);
        const int y = clampi((int)(uv.y * (float)env->height), 0, env->height - 1);
        const int rowOffset = y * (env->width + 1);
        const float rowPdf = env->rowCdf[rowOffset + x + 1] - env->rowCdf[rowOffset + x];
        const float margPdf = env->marginalCdf[y + 1] - env->marginalCdf[y];
        const float st = fmax2(1.0e-4f, orc_sinf(uv.y * K_PI));
        return (margPdf * rowPdf) * (float)(env->width * env->height) / (2.0f * K_PI * K_PI * st);
    }
    return 1.0f / (4.0f * K_PI);
}

/* EnvironmentMap.cpp:76-127: importance tables from an RGB float image.
 * marginalCdf holds height+1 floats, rowCdf holds height*(width+1). */
int orc_env_build_tables(const float* px, int width, int height, float* marg, float* rowCdf)
{
    if (width <= 0 || height <= 0 || px == NULL) {
        return 0;
    }
    memset(marg, 0, sizeof(float) * (size_t)(height + 1));
    memset(rowCdf, 0, sizeof(float) * (size_t)height * (size_t)(width + 1));
    float total = 0.0f;
    for (int y = 0; y < height; ++y) {
        const float v = ((float)y + 0.5f) / (float)height;
        const float st = orc_sinf(v * 3.14159265f);
        float rowSum = 0.0f;
        const size_t ro = (size_t)y * (size_t)(width + 1);
        rowCdf[ro] = 0.0f;
        for (int x = 0; x < width; ++x) {
            const int i = (y * width + x) * 3;
            const float a = px[i], b = px[i + 1], c = px[i + 2];
            const float bc = b < c ? c : b; /* std::max(b, c) */
            const float lum = a < bc ? bc : a;
            const float w = (lum < 1.0e-6f ? 1.0e-6f : lum) * st;
            rowSum += w;
            rowCdf[ro + (size_t)x + 1] = rowSum;
        }
        if (rowSum > 0.0f) {
            for (int x = 0; x < width; ++x) {
                rowCdf[ro + (size_t)x + 1] /= rowSum;
            }
        }
        total += rowSum;
        marg[y + 1] = total;
    }
    if (total > 0.0f) {
        for (int y = 0; y < height; ++y) {
            marg[y + 1] /= total;
        }
    }
    return 1;
}

/* LightSamplingCore.h:199-213: fallback grey 0.8 when the index is out of range. */
OMaterial orc_fetch_material(const OScene* scene, uint32_t tri)
{
    OMaterial fb;
    memset(&fb, 0, sizeof(fb));
    fb.r = 0.8f;
    fb.g = 0.8f;
    fb.b = 0.8f;
    fb.roughness = 0.5f;
    fb.ior = 1.5f;
    fb.diffuseRoughness = -1.0f;
    fb.specular = 1.0f;
    if (scene == NULL || scene->materials == NULL || tri >= scene->triangleCount) {
        return fb;
    }
    const uint32_t mi = scene->triangles[tri].materialIndex;
    if (mi >= scene->materialCount) {
        return fb;
    }
    return scene->materials[mi];
}

/* Texture/ProceduralTexture.h:12-15 */
static float fract1(float v) { return v - floorf(v); }
static float clamp01_tex(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

/* ProceduralTexture.h:22-48 */
static float tex_scalar(const OTexDesc* d, OVec2 uv, float u1d)
{
    switch (d->kind) {
    case 0:
        return d->p0.x;
    case 3: {
        const float c = fract1(u1d * d->p0.x);
        return c < d->p0.y ? d->p0.z : d->p0.w;
    }
    case 2: {
        const float fx = fract1(uv.x * d->p0.x);
        const float fy = fract1(uv.y * d->p0.y);
        return (fx < d->p0.z || fy < d->p0.z) ? 1.0f : 0.0f;
    }
    default:
        return 0.0f;
    }
}

/* ProceduralTexture.h:50-74 */
static OVec3 tex_rgb(const OTexDesc* d, OVec2 uv)
{
    switch (d->kind) {
    case 1:
        return v3(d->p0.x, d->p0.y, d->p0.z);
    case 2: {
        const float fx = fract1(uv.x * d->p0.x);
        const float fy = fract1(uv.y * d->p0.y);
        const float g = (fx < d->p0.z || fy < d->p0.z) ? 1.0f : 0.0f;
        return v3(d->p1.x + (d->p2.x - d->p1.x) * g, d->p1.y + (d->p2.y - d->p1.y) * g,
                  d->p1.z + (d->p2.z - d->p1.z) * g);
    }
    default:
        return v3(0.0f, 0.0f, 0.0f);
    }
}

static float chan_scalar(float inl, uint32_t ti, const OScene* s, OVec2 uv)
{
    const OTexDesc* bank = s != NULL ? s->textures : NULL;
    const uint32_t n = s != NULL ? s->textureCount : 0u;
    if (ti == 0u || bank == NULL || ti >= n) {
        return inl;
    }
    return tex_scalar(&bank[ti], uv, uv.x);
}

/* ProceduralTexture.h:138-207; the resolved material carries no texture indices. */
OMaterial orc_resolve_material(const OMaterial* m, OVec2 uv, const OScene* s)
{
    OMaterial r;
    memset(&r, 0, sizeof(r));
    const OTexDesc* bank = s != NULL ? s->textures : NULL;
    const uint32_t n = s != NULL ? s->textureCount : 0u;
    if (m->albedoTex == 0u || bank == NULL || m->albedoTex >= n) {
        r.r = m->r;
        r.g = m->g;
        r.b = m->b;
    } else {
        const OVec3 c = tex_rgb(&bank[m->albedoTex], uv);
        r.r = c.x;
        r.g = c.y;
        r.b = c.z;
    }
    r.roughness = clamp01_tex(chan_scalar(m->roughness, m->roughnessTex, s, uv));
    r.metallic = clamp01_tex(chan_scalar(m->metallic, m->metallicTex, s, uv));
    r.transmission = clamp01_tex(chan_scalar(m->transmission, m->transmissionTex, s, uv));
    r.thin = clamp01_tex(chan_scalar(m->thin, m->thinTex, s, uv));
    r.ior = fmaxf(1.0e-3f, chan_scalar(m->ior, m->iorTex, s, uv));
    r.subsurface = clamp01_tex(chan_scalar(m->subsurface, m->subsurfaceTex, s, uv));
    r.emission = fmaxf(0.0f, chan_scalar(m->emission, m->emissionTex, s, uv));
    r.diffuseRoughness = m->diffuseRoughness;
    if (r.diffuseRoughness < 0.0f) {
        r.diffuseRoughness = r.roughness;
    } else {
        r.diffuseRoughness = clamp01_tex(r.diffuseRoughness);
    }
    r.scatterRadiusR = fmaxf(0.0f, m->scatterRadiusR);
    r.scatterRadiusG = fmaxf(0.0f, m->scatterRadiusG);
    r.scatterRadiusB = fmaxf(0.0f, m->scatterRadiusB);
    r.specular = clamp01_tex(m->specular);
    return r;
}

/* LightSamplingCore.h:215-228 */
static float tri_area(const OTriangle* t)
{
    return 0.5f * vlen(vcross(vsub(t->v1, t->v0), vsub(t->v2, t->v0)));
}
static OVec3 tri_normal(const OTriangle* t) { return vnorm(vcross(vsub(t->v1, t->v0), vsub(t->v2, t->v0))); }

/* LightSamplingCore.h:238-255 */
static float emissive_selection_pdf(const OScene* s, uint32_t li)
{
    if (s == NULL || s->emissiveCdf == NULL || s->emissiveCount == 0) {
        return 0.0f;
    }
    if (li >= s->emissiveCount) {
        return 0.0f;
    }
    const float total = s->emissiveCdf[s->emissiveCount];
    if (total <= K_MIN_PDF) {
        return 0.0f;
    }
    return (s->emissiveCdf[li + 1] - s->emissiveCdf[li]) / total;
}

/* LightSamplingCore.h:257-313 */
int orc_light_sample_emissive(const OScene* s, float uTri, float u1, float u2, OVec3* pos, OVec3* normal,
                              OVec3* radiance, float* pdf)
{
    return orc_light_sample_emissive_tri(s, uTri, u1, u2, pos, normal, radiance, pdf, NULL);
}

/* As orc_light_sample_emissive; *triOut (if not NULL) receives the chosen triangle's index. */
int orc_light_sample_emissive_tri(const OScene* s, float uTri, float u1, float u2, OVec3* pos, OVec3* normal,
                                  OVec3* radiance, float* pdf, uint32_t* triOut)
{
    *pdf = 0.0f;
    if (s == NULL || s->emissiveCount == 0 || s->emissiveIndices == NULL || s->emissiveCdf == NULL ||
        s->triangles == NULL) {
        return 0;
    }
    const float total = s->emissiveCdf[s->emissiveCount];
    if (total <= K_MIN_PDF) {
        return 0;
    }
    const float target = uTri * total;
    int lo = 0, hi = (int)s->emissiveCount - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s->emissiveCdf[mid + 1] < target) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    const uint32_t li = (uint32_t)lo;
    const uint32_t ti = s->emissiveIndices[li];
    if (triOut != NULL) {
        *triOut = ti;
    }
    const OTriangle* t = &s->triangles[ti];
    const OMaterial m = orc_fetch_material(s, ti);
    const float su = sqrtf(fmax2(0.0f, u1));
    const float b0 = 1.0f - su;
    const float b1 = u2 * su;
    const float b2 = 1.0f - b0 - b1;
    *pos = vadd(vadd(vscale(t->v0, b0), vscale(t->v1, b1)), vscale(t->v2, b2));
    *normal = tri_normal(t);
    *radiance = v3(m.r * m.emission, m.g * m.emission, m.b * m.emission);
    const float area = tri_area(t);
    if (area <= K_MIN_PDF) {
        return 0;
    }
    *pdf = emissive_selection_pdf(s, li) / area;
    return *pdf > K_MIN_PDF;
}

/* CameraDevice.cuh:8-25 */
static OVec3 rotate_by_quat(OFloat4 q, OVec3 v)
{
    const float s = q.x;
    const OVec3 u = v3(q.y, q.z, q.w);
    const OVec3 c1 = v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
    const OVec3 c2 = v3(u.y * c1.z - u.z * c1.y, u.z * c1.x - u.x * c1.z, u.x * c1.y - u.y * c1.x);
    return v3(v.x + 2.0f * c1.x * s + 2.0f * c2.x, v.y + 2.0f * c1.y * s + 2.0f * c2.y,
              v.z + 2.0f * c1.z * s + 2.0f * c2.z);
}

/* CameraDevice.cuh:37-55 */
static void concentric_disk(float u1, float u2, float* dx, float* dy)
{
    const float a = 2.0f * u1 - 1.0f;
    const float b = 2.0f * u2 - 1.0f;
    if (a == 0.0f && b == 0.0f) {
        *dx = 0.0f;
        *dy = 0.0f;
        return;
    }
    float r, th;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        th = (3.14159265358979323846f * 0.25f) * (b / a);
    } else {
        r = b;
        th = (3.14159265358979323846f * 0.5f) - (3.14159265358979323846f * 0.25f) * (a / b);
    }
    *dx = r * orc_cosf(th);
    *dy = r * orc_sinf(th);
}

/* CameraDevice.cuh:64-74 */
static OVec3 film_point_local(const OCamera* c, float u, float v)
{
    const float ndcX = 2.0f * u - 1.0f;
    const float ndcY = 1.0f - 2.0f * v;
    const float tanHalf = orc_tanf(c->fovY * 0.5f);
    const float top = c->nearPlane * tanHalf;
    const float right = top * c->aspect;
    return v3(ndcX * right, ndcY * top, -c->nearPlane);
}

/* CameraDevice.cuh:88-139 */
void orc_camera_ray(const OCamera* c, float u, float v, float lensU1, float lensU2, OVec3* ro, OVec3* rd)
{
    if (c == NULL) {
        *ro = v3(0.0f, 0.0f, 0.0f);
        *rd = v3(0.0f, 0.0f, -1.0f);
        return;
    }
    if (c->focusValid == 0 || c->apertureRadius <= 0.0f || c->focusDistance <= 0.0f || c->nearPlane <= 0.0f) {
        *ro = c->position;
        *rd = vnorm(rotate_by_quat(c->orientation, film_point_local(c, u, v)));
        return;
    }
    const OVec3 right = vnorm(rotate_by_quat(c->orientation, v3(1.0f, 0.0f, 0.0f)));
    const OVec3 up = vnorm(rotate_by_quat(c->orientation, v3(0.0f, 1.0f, 0.0f)));
    const OVec3 film = film_point_local(c, u, v);
    const float fs = c->focusDistance / c->nearPlane;
    const OVec3 focusLocal = v3(film.x * fs, film.y * fs, -c->focusDistance);
    const OVec3 rf = rotate_by_quat(c->orientation, focusLocal);
    const OVec3 focusWorld = v3(c->position.x + rf.x, c->position.y + rf.y, c->position.z + rf.z);
    float dx = 0.0f, dy = 0.0f;
    concentric_disk(lensU1, lensU2, &dx, &dy);
    const float lx = dx * c->apertureRadius;
    const float ly = dy * c->apertureRadius;
    *ro = v3(c->position.x + right.x * lx + up.x * ly, c->position.y + right.y * lx + up.y * ly,
             c->position.z + right.z * lx + up.z * ly);
    *rd = vnorm(v3(focusWorld.x - ro->x, focusWorld.y - ro->y, focusWorld.z - ro->z));
}
