// Exposes the reference's header-only host policies (compiled unmodified from the
// reference sources) through a C ABI, so tests can compare the oracle against them.
// Only headers that compile without CUDA, Qt or glm are included.
#include "Geometry/MathCore.h"
#include "PathTracerAdaptiveSampling.h"
#include "PathTracerPreviewLevels.h"
#include "PathTracerSampleBudget.h"
#include "Sampling/MisCore.h"
#include "Spectral/Rgb2Spec.h"

extern "C" {

int ref_is_converged(uint32_t n, float mean, float m2, int minSamples, float thr)
{
    return isAdaptivePixelConverged(n, mean, m2, minSamples, thr) ? 1 : 0;
}

float ref_sample_luminance(float r, float g, float b) { return sampleLuminance(r, g, b); }

int ref_budget_total(int preview, int maxSpp) { return sampleBudgetTotalIterations(preview, maxSpp); }

int ref_can_take_sample(int count, int preview, int maxSpp) { return canTakeSampleAtIteration(count, preview, maxSpp); }

int ref_budget_exhausted(int count, int preview, int maxSpp) { return isSampleBudgetExhausted(count, preview, maxSpp); }

int ref_adaptive_exhausted(int active, int preview, int count)
{
    return isAdaptiveSampleBudgetExhausted(active, preview, count);
}

void ref_preview_level(int w, int h, int levels, int index, int* out3)
{
    const PreviewLevelDimensions d = previewLevelDimensionsAt(w, h, levels, index);
    out3[0] = d.width;
    out3[1] = d.height;
    out3[2] = d.downscale;
}

float ref_mis_balance(float a, float b) { return misBalanceWeight(a, b); }

void ref_normalize(const float* in3, float* out3)
{
    const Vec3 r = vecNormalize3(vecMake3(in3[0], in3[1], in3[2]));
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
}

void ref_rotate_ypr(const float* in3, float yaw, float pitch, float roll, float* out3)
{
    const Vec3 r = vecRotateYawPitchRoll(vecMake3(in3[0], in3[1], in3[2]), yaw, pitch, roll);
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
}

static Rgb2SpecGpu ref_model(const float* scale, const float* data, int res)
{
    Rgb2SpecGpu m;
    m.scale = scale;
    m.data = data;
    m.res = res;
    return m;
}

void ref_rgb2spec_coeffs(const float* scale, const float* data, int res, const float* rgb3, float* out3)
{
    float rgb[3] = {rgb3[0], rgb3[1], rgb3[2]};
    rgb2specFetchCoefficients(ref_model(scale, data, res), rgb, out3);
}

float ref_rgb2spec_reflectance(const float* scale, const float* data, int res, float r, float g, float b, float lambda)
{
    return rgb2specEvalReflectance(ref_model(scale, data, res), r, g, b, lambda);
}

float ref_rgb2spec_radiance(const float* scale, const float* data, int res, float r, float g, float b, float lambda)
{
    return rgb2specEvalRadiance(ref_model(scale, data, res), r, g, b, lambda);
}

}
