// pt_internal.h -- host-side helpers shared by the library's translation units.
#pragma once

#include <cstdarg>
#include <vector>

// Records the message returned by pathTracerLastError() (thread-local).
void ptSetError(const char* fmt, ...);

// Deepest leaf level the binned-SAH build produces; traversal stacks hold 64 entries.
constexpr int kMaxSahDepth = 48;

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
// Binned-SAH BVH over device triangles (25-word records): internal nodes (n - 1 x 64 B), leaf
// slots (3 x float4 each) and shading records (4 x float4 per triangle). rootOut receives the root
// box and the root reference (bits of an int in rootOut[6]); levelsOut the number of levels.
// levelStarts (optional) receives the first internal node of every level.
bool pt_build_bvh_sah(const float* dTris, int n, float4* nodes, float4* triPos, float4* triShade, hipStream_t stream,
                      float rootOut[7], int* levelsOut, std::vector<int>* levelStarts);
// Device workspace of the binned-SAH build for scenes of up to `capacity` triangles.
struct PtSahWork;
PtSahWork* pt_sah_work_create(int capacity);
void pt_sah_work_destroy(PtSahWork* w);
// Enqueues a whole binned-SAH build of n <= capacity triangles into nodes/triPos/triShade (as
// pt_build_bvh_sah), and when wide is not null the 4-wide collapse into wide (room for
// wideCapacity nodes): a fixed sequence of launches with no allocation and no host
// synchronisation, so a caller can capture it into a graph. It enqueues `levels` levels (0: as
// many as the depth cap allows; a rebuild of unchanged triangles needs the levels of their first
// build, and pt_sah_result reports a build left unfinished). The root box and reference, the
// level starts and any error stay in the workspace until pt_sah_result reads them.
bool pt_sah_enqueue(PtSahWork* w, const float* dTris, int n, float4* nodes, float4* triPos, float4* triShade,
                    float4* wide, int wideCapacity, int levels, hipStream_t stream);
// Waits for the stream and checks the last build enqueued with w (n triangles): false with the
// error set if it did not produce a complete tree; otherwise the outputs of pt_build_bvh_sah and
// the 4-wide node count (each optional).
bool pt_sah_result(PtSahWork* w, int n, hipStream_t stream, float rootOut[7], int* levelsOut,
                   std::vector<int>* levelStarts, int* wideCount);
// Status of a build as a flat record (the control block, then the root box and reference as 8
// words): pt_sah_snapshot enqueues a copy of the last build's status into dStatus (a kernel, so
// it can be captured), pt_sah_status_result checks such a copy as pt_sah_result does.
size_t pt_sah_status_words();
bool pt_sah_snapshot(PtSahWork* w, int* dStatus, hipStream_t stream);
// Marks a status copy as read (synchronous): until a rebuild's snapshot overwrites it,
// pt_sah_status_result reports that no rebuild ran.
bool pt_sah_status_consume(int* dStatus);
bool pt_sah_status_result(const int* dStatus, int n, hipStream_t stream, float rootOut[7], int* levelsOut,
                          int* wideCount);
// 4-wide tree collapsed from a level-numbered binary tree of internalCount internal nodes
// (pt_sah.hip): wideOut (device, 8 float4 per node) and its node count; the wide root is node 0
// where the binary root is internal.
bool pt_build_wide(const float4* nodes, int internalCount, const std::vector<int>& levelStarts, float4** wideOut,
                   int* wideCount, hipStream_t stream);

#include "pt_spectral.h"
// Whether the scene description at d was built with a 4-wide tree (ptSceneBuild registers it).
bool pt_scene_is_wide(const MeshAccelSceneGpu* d);
// Spectral sampling state set through the C ABI (pt_spectral.hip): whether the launchers
// render hero wavelengths, and the model (coefficient table, white normalisation, wavelength
// range) the kernels then read.
bool pt_spectral_enabled();
// Light sampling mode set through pathTracerSetLightSampling (pt_render.hip): 0 the reference's,
// 1 multiple importance sampling on area lights.
int pt_light_mode();
pt::SpecModel pt_spectral_model();

// Adjoint of the fixed-spp render (pathTracerAdjointRender after argument checks), scheduled on
// the wavefront machinery (pt_wavefront.hip).
bool pt_adjoint_wavefront(const float* d_gradImage, int width, int height, int spp, const CameraGpu* d_camera,
                          const MeshAccelSceneGpu* d_scene, const EnvironmentMapGpu* d_env,
                          const RenderParamsGpu* d_params, int materialCount, const float* d_sampleRadiance,
                          double* d_grad, unsigned int firstSeed, hipStream_t stream);
#endif
