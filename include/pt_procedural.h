/*
 * pt_procedural.h -- procedural scene input: L-system definitions -> HostMesh -> PtScene.
 *
 * The reference builds its scenes from ProceduralInstance lists (Application/src/Procedural/
 * ProceduralTypes.h:80-86) through meshSceneBuild (PathTracerLib/src/MeshAccel/
 * MeshSceneContent.cpp:5-52): each instance's definition is evaluated by LSystemEvaluator
 * (LSystemLib/src/LSystemEvaluator.cpp), interpreted by the turtle (Turtle.cpp:168-283), every
 * segment is fitted with a Hermite spline and parallel-transport frames (Spline.cpp:102-254),
 * lofted into a capped tube (Loft.cpp:108-155, 319-346), given its `Mat(id)` material
 * (ProceduralMeshBuilder.cpp:139-196, 266-330) and moved by the instance's root transform
 * (ProceduralMeshBuilder.cpp:198-247); the instances' meshes are concatenated with material and
 * texture indices offset (HostMesh.h:77-120) and uploaded (MeshAccelScene::build).
 *
 * Everything here runs on the host (C++), deterministically: the same definition gives the same
 * triangles bit for bit on every machine with IEEE-754 single precision and the C library's
 * cosf/sinf/asinf.
 *
 * Differences from the reference:
 *   - a segment of zero arc length (e.g. `F(0)`) becomes an octahedral sphere of this library's
 *     own construction (the reference uses manifold::Manifold::Sphere and CalculateNormals): same
 *     centre, radius and segment count, smooth radial normals, zero UVs;
 *   - the turtle extensions of PtProceduralBuildParams (turn symbols, radius decay, jitter, leaf
 *     quads) are off by default; with them off the output is the reference's.
 * Errors: functions return false (or NULL) and set the message read by pathTracerLastError().
 */
#ifndef PT_PROCEDURAL_H
#define PT_PROCEDURAL_H

#include "pt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ProceduralInstance (ProceduralTypes.h:80-86): a definition, its derivation count and its root
 * transform (rotationDeg.x = yaw about +Y, .y = pitch about +X, .z = roll about +Z, applied in
 * that order, then the translation). */
typedef struct {
    const char* commandString; /* NUL-terminated definition text */
    uint64_t iterations;
    Vec3 translation;
    Vec3 rotationDeg;
} PtProceduralInstance;

/* MeshSceneBuildParams (MeshSceneContent.h:9-13). */
typedef struct {
    int circularSegments; /* 64 */
    float creaseAngleDeg; /* 50 */
} PtMeshSceneBuildParams;

/* Turtle extension flags (no reference counterpart). */
#define PT_TURTLE_TURN_SYMBOLS 1u /* parse and interpret + - & ^ / \ | */
#define PT_TURTLE_LEAVES 2u       /* interpret Leaf / Leaf(length) / Leaf(length, width) */

/* ProceduralBuildParams (ProceduralTypes.h:67-76) followed by the turtle extensions. */
typedef struct {
    int circularSegments;         /* 32 */
    int samplesPerSpan;           /* 4 */
    float hermiteTension;         /* 1 (<= 0 means 1) */
    float segmentRefineTolerance; /* 0.01: the reference refines its Manifold copy only; unused here */
    float globalRefineTolerance;  /* 0.01: unused by the reference too */
    float creaseAngleDeg;         /* 50: sphere normals only */
    float defaultStepLength;      /* 500 (TurtleParams) */
    float defaultRadius;          /* 100 (TurtleParams) */
    /* extensions */
    uint32_t turtleFlags;         /* PT_TURTLE_*; 0 = the reference turtle */
    float turnAngleDeg;           /* + - (yaw about up), & ^ (pitch about right); `+(a)` turns by a */
    float rollAngleDeg;           /* / \ (roll about the tangent) */
    float radiusDecay;            /* F radii scaled by radiusDecay^(bracket depth); <= 0 or 1: off */
    float jitterDeg;              /* before each F: yaw, pitch, roll by uniform(-j, j) degrees; 0: off */
    uint64_t jitterSeed;          /* SplitMix64 seed of the jitter draws */
    float leafLength;             /* Leaf quad along the tangent (default size) */
    float leafWidth;              /* Leaf quad across, along the turtle's right axis */
    uint64_t ruleSeed;            /* rule-choice seed (LSystemEvaluator default 0xC0FFEE) */
} PtProceduralBuildParams;

/* The reference's defaults (ProceduralTypes.h:67-76, LSystemEvaluator.h:26-29), extensions off. */
void ptProceduralDefaultParams(PtProceduralBuildParams* out);

/* Host mesh owned by the library (HostMesh, Application/src/MeshBuilder/HostMesh.h:23-28). */
typedef struct PtHostMeshBuf PtHostMeshBuf;
void ptHostMeshDestroy(PtHostMeshBuf* mesh);
void ptHostMeshCounts(const PtHostMeshBuf* mesh, uint32_t* triangles, uint32_t* materials, uint32_t* textures);
const HostTriangle* ptHostMeshTriangles(const PtHostMeshBuf* mesh);
const MaterialGpu* ptHostMeshMaterials(const PtHostMeshBuf* mesh);
const TextureDescGpu* ptHostMeshTextures(const PtHostMeshBuf* mesh);

/* ProceduralMeshBuilder::buildHostMesh (ProceduralMeshBuilder.cpp:266-330): NULL when the
 * definition does not parse (the parser's message), when the turtle draws no segment, or when no
 * segment yields triangles. params NULL = ptProceduralDefaultParams. */
PtHostMeshBuf* ptProceduralBuildHostMesh(const char* definition, uint64_t iterations, Vec3 translation,
                                         Vec3 rotationDeg, const PtProceduralBuildParams* params);

/* The HostMesh half of meshSceneBuild (MeshSceneContent.cpp:5-52) with full build parameters:
 * every instance built, then concatenated in order; a failing instance fails the whole call with
 * "Procedural instance <i>: <message>". count 0 gives an empty mesh. */
PtHostMeshBuf* ptProceduralBuildScene(const PtProceduralInstance* instances, uint32_t count,
                                      const PtProceduralBuildParams* params);

/* meshSceneBuild (MeshSceneContent.cpp:5-52): MeshSceneBuildParams select circularSegments and
 * creaseAngleDeg, the rest are ProceduralBuildParams defaults; the concatenated mesh is built into
 * `scene` (ptSceneBuild on `stream`, current build mode). params NULL = {64, 50}. */
bool ptMeshSceneBuild(const PtProceduralInstance* instances, uint32_t count, PtScene* scene,
                      const PtMeshSceneBuildParams* params, hipStream_t stream);

/* ---- inspection hooks (used by the parity tests; all outputs caller-owned) ---- */

/* LSystemEvaluator::evaluate_steps (LSystemEvaluator.cpp:431-470) as text: generation g is
 * written as its modules separated by single spaces (arguments as %.17g, material identifiers
 * verbatim, subsystem references as ref.Name), generations separated by '\n' (allSteps) or only
 * the last one. Returns the length needed (excluding the NUL), or -1 on a parse error. The text is
 * truncated to capacity - 1 characters. flags: PT_TURTLE_TURN_SYMBOLS accepts the turn symbols. */
int64_t ptLSystemEvaluateText(const char* definition, uint64_t iterations, uint64_t seed, uint32_t flags,
                              int allSteps, char* out, uint64_t capacity);

/* Structure of a parse (ModuleStringParser::parse / parse_rule / parse_rules_block,
 * LSystem::parse; ModuleStringParser.cpp, LSystemImpl.cpp) as text. kind 0: a module string,
 * dumped as `[m1 m2 ...]` (a module is `name`, `name(e1,e2)` or `ref:Name`; expressions in prefix
 * form `(op a b)` with op one of + - * / < <= > >= ==, numbers %.17g); kind 1: one rule; kind 2: a
 * rules block, one line per rule, `left=[..] pred=P right=[..] cond=E|null prob=p succ=[..]`;
 * kind 3: a whole definition, `axiom=[..]`, then `rule ...` lines, `const name=value` lines in
 * name order and `mat id` lines. Same return convention as ptLSystemEvaluateText; the parser's
 * message (with line and column) is in pathTracerLastError() on -1. */
int64_t ptLSystemParseDump(int kind, const char* text, uint32_t flags, char* out, uint64_t capacity);

/* The parsed `Mat(id)` definitions (LSystemMaterialParse.cpp) as text, one line per material:
 * id, then per channel (albedo, roughness, metallic, transmission, thin, ior, subsurface,
 * emission, diffuseRoughness, scatterRadiusR, scatterRadiusG, scatterRadiusB, specular)
 * " i:<r>,<g>,<b>" for the albedo or " i:<value>" inline, " t:<kind>:<p0>,<p1>,..." for a texture,
 * floats as %.9g. Same return convention as ptLSystemEvaluateText. */
int64_t ptLSystemMaterialsText(const char* definition, char* out, uint64_t capacity);

/* Turtle output of the evaluated definition (turtleExecute, Turtle.cpp:168-283). states receives
 * 7 floats per state (position, tangent, radius) for every segment in order, stateCounts one count
 * per segment, materialIds the segments' material ids separated by '\n', groups the branch-group
 * sizes (BranchGroup, ProceduralTypes.h:21-24), leaves 9 floats per leaf (position, tangent,
 * right) and leafLengths 2 floats per leaf (length, width). Capacities are in those units; counts
 * are always written; NULL outputs are skipped. */
typedef struct {
    uint32_t segments, states, groups, leaves, materialIdChars;
} PtTurtleCounts;
bool ptTurtleExecute(const char* definition, uint64_t iterations, const PtProceduralBuildParams* params,
                     PtTurtleCounts* counts, float* states, uint32_t stateCapacity, uint32_t* stateCounts,
                     uint32_t segmentCapacity, char* materialIds, uint32_t materialIdCapacity, uint32_t* groups,
                     uint32_t groupCapacity, float* leaves, float* leafSizes, uint32_t leafCapacity);

/* SplinePath over one segment's states (7 floats each) with the given tension:
 * mode 0 = computeLoftFrames(samplesPerSpan) (Spline.cpp:222-254): 13 floats per frame (position,
 * tangent, normal, binormal, radius); mode 1 = sampleUniform(samplesPerSpan) (Spline.cpp:159-183):
 * 7 floats per sample. Writes the count to *outCount (and the frames when capacity allows);
 * *arcLength receives totalArcLength(). false when fewer than two states. */
bool ptSplineFrames(const float* states, uint32_t stateCount, float tension, int samplesPerSpan, int mode,
                    float* out, uint32_t capacity, uint32_t* outCount, float* arcLength);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PT_PROCEDURAL_H */
