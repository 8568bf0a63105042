// pt_spectral.hip -- spectral state and hero-wavelength settings behind the C ABI.
//
// The reference loads an rgb2spec coefficient file, computes the white normalisation on the host
// and uploads the tables (Spectral/SpectralState.cpp:15-141), then stores the device state with
// spectralUploadStateToSymbol (PathTracer.cu:27-33). PtSpectral plays the role of
// SpectralStateHost; the stored state is what the launchers pass to their kernels once hero
// wavelength sampling is switched on with pathTracerSetSpectralSampling.
#include "pt_internal.h"
#include "pt_spectral.h"

#include "../../include/pt_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

std::mutex gSpecMutex;
SpectralStateGpu gState{nullptr, nullptr, 0, 1.0f, 1.0f, 1.0f};
bool gEnabled = false;
float gLambdaMin = pt::kLambdaMinNm;
float gLambdaRange = pt::kLambdaRangeNm;

} // namespace

struct PtSpectral {
    std::vector<float> scale;
    std::vector<float> data;
    int res = 0;
    float white[3] = {1.0f, 1.0f, 1.0f};
    float* dScale = nullptr;
    float* dData = nullptr;
};

bool pt_spectral_enabled()
{
    std::lock_guard<std::mutex> lock(gSpecMutex);
    return gEnabled;
}

pt::SpecModel pt_spectral_model()
{
    std::lock_guard<std::mutex> lock(gSpecMutex);
    pt::SpecModel m;
    m.scale = gState.rgb2specScale;
    m.data = gState.rgb2specData;
    m.res = gState.rgb2specRes;
    m.whiteR = gState.whiteNormR;
    m.whiteG = gState.whiteNormG;
    m.whiteB = gState.whiteNormB;
    m.lambdaMin = gLambdaMin;
    m.lambdaRange = gLambdaRange;
    return m;
}

namespace {

bool table_ok(const float* scale, const float* data, int res, const char* who)
{
    if (scale == nullptr || data == nullptr || res < 2 || res > 1024) {
        ptSetError("%s: need scale and data tables and 2 <= res <= 1024 (got res %d)", who, res);
        return false;
    }
    return true;
}

void free_device(PtSpectral* s)
{
    if (s->dScale != nullptr) {
        (void)hipFree(s->dScale);
        s->dScale = nullptr;
    }
    if (s->dData != nullptr) {
        (void)hipFree(s->dData);
        s->dData = nullptr;
    }
}

} // namespace

extern "C" {

hipError_t spectralUploadStateToSymbol(const SpectralStateGpu* hostState)
{
    if (hostState == nullptr) {
        return hipErrorInvalidValue;
    }
    std::lock_guard<std::mutex> lock(gSpecMutex);
    gState = *hostState;
    return hipSuccess;
}

bool pathTracerSetSpectralSampling(int wavelengthsPerSample, float lambdaMinNm, float lambdaMaxNm)
{
    if (wavelengthsPerSample == 0) {
        std::lock_guard<std::mutex> lock(gSpecMutex);
        gEnabled = false;
        return true;
    }
    if (wavelengthsPerSample != 4) {
        ptSetError("pathTracerSetSpectralSampling: wavelengthsPerSample must be 0 (RGB) or 4 (got %d)",
                   wavelengthsPerSample);
        return false;
    }
    if (!(lambdaMinNm >= pt::kLambdaMinNm) || !(lambdaMaxNm <= pt::kLambdaMaxNm) || !(lambdaMinNm < lambdaMaxNm)) {
        ptSetError("pathTracerSetSpectralSampling: need %g <= lambdaMin < lambdaMax <= %g nm", pt::kLambdaMinNm,
                   pt::kLambdaMaxNm);
        return false;
    }
    std::lock_guard<std::mutex> lock(gSpecMutex);
    if (gState.rgb2specScale == nullptr || gState.rgb2specData == nullptr || gState.rgb2specRes < 2) {
        ptSetError("pathTracerSetSpectralSampling: no rgb2spec table stored (spectralUploadStateToSymbol)");
        return false;
    }
    gEnabled = true;
    gLambdaMin = lambdaMinNm;
    gLambdaRange = lambdaMaxNm - lambdaMinNm;
    return true;
}

int pathTracerSpectralSampling(void)
{
    std::lock_guard<std::mutex> lock(gSpecMutex);
    return gEnabled ? 4 : 0;
}

bool ptSpectralWhiteNorm(const float* scale, const float* data, int res, float* out3)
{
    if (!table_ok(scale, data, res, "ptSpectralWhiteNorm") || out3 == nullptr) {
        if (out3 == nullptr) {
            ptSetError("ptSpectralWhiteNorm: null output");
        }
        return false;
    }
    pt::SpecModel m{scale, data, res, 1.0f, 1.0f, 1.0f, pt::kLambdaMinNm, pt::kLambdaRangeNm};
    pt::spectral_white_norm(m, out3);
    return true;
}

PtSpectral* ptSpectralCreate(const float* scale, const float* data, int res)
{
    if (!table_ok(scale, data, res, "ptSpectralCreate")) {
        return nullptr;
    }
    PtSpectral* s = new PtSpectral();
    s->res = res;
    const size_t nData = static_cast<size_t>(9) * res * res * res;
    s->scale.assign(scale, scale + res);
    s->data.assign(data, data + nData);
    pt::SpecModel m{s->scale.data(), s->data.data(), res, 1.0f, 1.0f, 1.0f, pt::kLambdaMinNm, pt::kLambdaRangeNm};
    pt::spectral_white_norm(m, s->white);
    if (hipMalloc(&s->dScale, res * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->dData, nData * sizeof(float)) != hipSuccess ||
        hipMemcpy(s->dScale, s->scale.data(), res * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->dData, s->data.data(), nData * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        ptSetError("ptSpectralCreate: device allocation or copy failed");
        free_device(s);
        delete s;
        return nullptr;
    }
    return s;
}

// SpectralState.cpp:48-103: "SPEC", uint32 res, float scale[res], float data[3 * 3 * res^3]
PtSpectral* ptSpectralLoadCoeffFile(const char* path)
{
    if (path == nullptr) {
        ptSetError("ptSpectralLoadCoeffFile: null path");
        return nullptr;
    }
    FILE* f = std::fopen(path, "rb");
    if (f == nullptr) {
        ptSetError("ptSpectralLoadCoeffFile: failed to open rgb2spec coefficient file: %s", path);
        return nullptr;
    }
    char magic[4] = {};
    uint32_t res = 0;
    if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "SPEC", 4) != 0) {
        std::fclose(f);
        ptSetError("ptSpectralLoadCoeffFile: invalid rgb2spec coefficient header");
        return nullptr;
    }
    if (std::fread(&res, sizeof(uint32_t), 1, f) != 1 || res < 2 || res > 1024) {
        std::fclose(f);
        ptSetError("ptSpectralLoadCoeffFile: invalid rgb2spec coefficient resolution");
        return nullptr;
    }
    std::vector<float> scale(res);
    std::vector<float> data(static_cast<size_t>(9) * res * res * res);
    const bool ok = std::fread(scale.data(), sizeof(float), res, f) == res &&
                    std::fread(data.data(), sizeof(float), data.size(), f) == data.size();
    std::fclose(f);
    if (!ok) {
        ptSetError("ptSpectralLoadCoeffFile: truncated rgb2spec coefficient file");
        return nullptr;
    }
    return ptSpectralCreate(scale.data(), data.data(), static_cast<int>(res));
}

bool ptSpectralGetState(const PtSpectral* s, SpectralStateGpu* out)
{
    if (s == nullptr || out == nullptr) {
        ptSetError("ptSpectralGetState: null argument");
        return false;
    }
    out->rgb2specScale = s->dScale;
    out->rgb2specData = s->dData;
    out->rgb2specRes = s->res;
    out->whiteNormR = s->white[0];
    out->whiteNormG = s->white[1];
    out->whiteNormB = s->white[2];
    return true;
}

void ptSpectralDestroy(PtSpectral* s)
{
    if (s == nullptr) {
        return;
    }
    {
        std::lock_guard<std::mutex> lock(gSpecMutex);
        if (gState.rgb2specData == s->dData) {
            // the stored state would dangle: back to RGB with no table
            gState = SpectralStateGpu{nullptr, nullptr, 0, 1.0f, 1.0f, 1.0f};
            gEnabled = false;
        }
    }
    free_device(s);
    delete s;
}

} // extern "C"
