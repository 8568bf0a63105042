// pt_mesher.h -- procedural meshes from L-system definitions: turtle, Hermite spline with
// parallel-transport frames, lofted tubes with caps, the material table of `Mat(id)` lines, and
// the concatenation of several instances into one scene mesh. Host C++.
//
// The pipeline is the Application's (Application/src/Procedural/Turtle.cpp:168-283,
// Spline.cpp:102-254, Loft.cpp:108-155, ProceduralMeshBuilder.cpp:266-330) and
// MeshSceneContent.cpp:5-52 (meshSceneBuild), with the turtle extended for the renderer's
// workloads (turn symbols, radius decay per bracket depth, seeded angle jitter, leaf quads), each
// off by default.
#pragma once

#include "../../include/pt_procedural.h"
#include "pt_lsystem.h"

#include <string>
#include <vector>

namespace mesh {

struct State {
    Vec3 position, tangent;
    float radius;
};

struct Segment {
    std::vector<State> states;
    std::string material = "0";
};

struct Frame {
    Vec3 position, tangent, normal, binormal;
    float radius;
};

struct Leaf {
    Vec3 position, tangent, right;
    std::string material;
};

struct TurtleResult {
    std::vector<Segment> segments;
    std::vector<Leaf> leaves;
};

struct HostMesh {
    std::vector<HostTriangle> triangles;
    std::vector<MaterialGpu> materials;
    std::vector<TextureDescGpu> textures;
};

TurtleResult turtle(const std::vector<lsys::Module>& modules, const PtProceduralBuildParams& p);
// Loft frames of one segment (SplinePath::computeLoftFrames); empty when the segment has fewer
// than two states; zeroLength is set when its chords sum to at most 1e-6.
std::vector<Frame> loft_frames(const Segment& s, const PtProceduralBuildParams& p, bool* zeroLength);
// Side walls and caps of a tube through the frames (circularSegments around).
void loft_triangles(const std::vector<Frame>& frames, int circularSegments, std::vector<HostTriangle>& out);
// The mesh of one definition (ProceduralMeshBuilder::buildHostMesh); throws std::runtime_error.
HostMesh build(const std::string& definition, std::size_t iterations, Vec3 translation, Vec3 rotationDeg,
               const PtProceduralBuildParams& p);
// Every instance's mesh, concatenated with material and texture indices offset (meshSceneBuild).
HostMesh build_scene(const PtProceduralInstance* instances, uint32_t count, const PtProceduralBuildParams& p);

} // namespace mesh
