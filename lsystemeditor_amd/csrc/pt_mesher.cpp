// pt_mesher.cpp -- turtle, spline frames, lofted tubes, spheres, leaves, material tables and the
// procedural C ABI of pt_procedural.h.
//
// Floating-point order: the mesh must come out of this file bit for bit as the reference's
// pipeline makes it (Turtle.cpp, Spline.cpp, Loft.cpp:108-155, ProceduralMeshBuilder.cpp:198-247),
// so each expression below keeps the reference's operation order and its float literals, and the
// file is compiled without FP contraction. The vector helpers are the MathCore ones
// (PathTracerLib/src/Geometry/MathCore.h): dot = (x*x + y*y) + z*z, length = sqrtf(dot),
// normalize = scale by 1/length ((0, 0, -1) for a zero vector), lerp = a + (b - a) * t.
#include "pt_mesher.h"

#include "pt_internal.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace mesh {

namespace {

// ---- vector arithmetic ----------------------------------------------------------------------

inline Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
inline Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(Vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 cross(Vec3 a, Vec3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline Vec3 unit(Vec3 a)
{
    const float len = length(a);
    if (len <= 0.0f) {
        return v3(0.0f, 0.0f, -1.0f);
    }
    return a * (1.0f / len);
}
inline float lerp(float a, float b, float t) { return a + (b - a) * t; }

// Rotation by `angle` about the unit axis n as the axis-angle (Rodrigues) matrix; apply() sums
// each row left to right.
struct AxisRotation {
    float m[3][3];
    AxisRotation(Vec3 n, float c, float s)
    {
        const float k = 1.0f - c;
        m[0][0] = c + n.x * n.x * k;
        m[0][1] = n.x * n.y * k - n.z * s;
        m[0][2] = n.x * n.z * k + n.y * s;
        m[1][0] = n.y * n.x * k + n.z * s;
        m[1][1] = c + n.y * n.y * k;
        m[1][2] = n.y * n.z * k - n.x * s;
        m[2][0] = n.z * n.x * k - n.y * s;
        m[2][1] = n.z * n.y * k + n.x * s;
        m[2][2] = c + n.z * n.z * k;
    }
    Vec3 apply(Vec3 v) const
    {
        return v3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                  m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
    }
};

// ---- turtle (Turtle.cpp:168-283) ------------------------------------------------------------

struct Pose {
    Vec3 position{0.0f, 0.0f, 0.0f};
    Vec3 tangent{0.0f, 0.0f, -1.0f};
    Vec3 up{0.0f, 1.0f, 0.0f};

    Vec3 right() const { return unit(cross(up, tangent)); }
    void turn(Vec3 axis, float degrees)
    {
        const float rad = degrees * 3.14159265f / 180.0f;
        const AxisRotation r(unit(axis), std::cos(rad), std::sin(rad));
        tangent = unit(r.apply(tangent));
        up = unit(r.apply(up));
    }
    void yaw(float deg) { turn(up, deg); }
    void pitch(float deg) { turn(right(), deg); }
    void roll(float deg) { turn(tangent, deg); }
    void advance(float step) { position = position + tangent * step; }
};

bool same_name_ci(const std::string& a, const char* b)
{
    const size_t n = std::strlen(b);
    if (a.size() != n) {
        return false;
    }
    for (size_t i = 0; i < n; ++i) {
        if (std::tolower(static_cast<unsigned char>(a[i])) != std::tolower(static_cast<unsigned char>(b[i]))) {
            return false;
        }
    }
    return true;
}

bool number_arg(const lsys::Module& m, size_t i, float& out)
{
    if (i >= m.args.size() || !m.args[i] || m.args[i]->op != lsys::Expr::Num) {
        return false;
    }
    out = static_cast<float>(m.args[i]->num);
    return true;
}

// Mat(id): an identifier, or a non-negative integral number written in decimal.
bool material_arg(const lsys::Module& m, std::string& out)
{
    if (m.args.empty() || !m.args[0]) {
        return false;
    }
    const lsys::Expr& e = *m.args[0];
    if (e.op == lsys::Expr::Ident) {
        out = e.ident;
        return true;
    }
    if (e.op == lsys::Expr::Num) {
        if (e.num < 0.0 || e.num != std::floor(e.num)) {
            return false;
        }
        out = std::to_string(static_cast<unsigned long long>(e.num));
        return true;
    }
    return false;
}

class SplitMix64 {
public:
    explicit SplitMix64(uint64_t seed) : s_(seed) {}
    double uniform()
    {
        s_ += 0x9E3779B97F4A7C15ull;
        uint64_t z = s_;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    }

private:
    uint64_t s_;
};

class Turtle {
public:
    explicit Turtle(const PtProceduralBuildParams& p) : p_(p), jitter_(p.jitterSeed) {}

    TurtleResult run(const std::vector<lsys::Module>& modules)
    {
        const bool turns = (p_.turtleFlags & PT_TURTLE_TURN_SYMBOLS) != 0;
        const bool leaves = (p_.turtleFlags & PT_TURTLE_LEAVES) != 0;
        for (const lsys::Module& m : modules) {
            if (m.subsystemRef) {
                continue;
            }
            const std::string& n = m.name;
            if (same_name_ci(n, "[")) {
                flush();
                stack_.push_back(Saved{pose_, out_.segments.size(), radiusScale_});
                if (decay()) {
                    radiusScale_ *= p_.radiusDecay;
                }
                continue;
            }
            if (same_name_ci(n, "]")) {
                flush();
                if (!stack_.empty()) {
                    const Saved s = stack_.back();
                    stack_.pop_back();
                    pose_ = s.pose;
                    radiusScale_ = s.radiusScale;
                    const size_t made = out_.segments.size() - s.segmentsAtPush;
                    if (made > 1) {
                        groups_.push_back(static_cast<uint32_t>(made));
                    }
                }
                continue;
            }
            if (n == "f") {
                flush();
                float step, radius;
                forward_args(m, step, radius);
                pose_.advance(step);
                continue;
            }
            if (n == "F") {
                float step, radius;
                forward_args(m, step, radius);
                if (p_.jitterDeg > 0.0f) {
                    pose_.yaw(draw());
                    pose_.pitch(draw());
                    pose_.roll(draw());
                }
                if (decay()) {
                    radius *= radiusScale_;
                }
                states_.push_back(State{pose_.position, pose_.tangent, radius});
                pose_.advance(step);
                continue;
            }
            float a = 0.0f;
            const bool hasArg = number_arg(m, 0, a);
            std::string mat;
            const bool hasMat = material_arg(m, mat);
            if (same_name_ci(n, "Yaw") && hasArg) {
                pose_.yaw(a);
                continue;
            }
            if ((same_name_ci(n, "Pitch") || same_name_ci(n, "Pich")) && hasArg) {
                pose_.pitch(a);
                continue;
            }
            if (same_name_ci(n, "Roll") && hasArg) {
                pose_.roll(a);
                continue;
            }
            if (same_name_ci(n, "Mat") && hasMat) {
                material_ = std::move(mat);
                continue;
            }
            if (turns && n.size() == 1) {
                const char c = n[0];
                const float turn = hasArg ? a : p_.turnAngleDeg;
                const float roll = hasArg ? a : p_.rollAngleDeg;
                switch (c) {
                case '+': pose_.yaw(turn); continue;
                case '-': pose_.yaw(-turn); continue;
                case '&': pose_.pitch(turn); continue;
                case '^': pose_.pitch(-turn); continue;
                case '/': pose_.roll(roll); continue;
                case '\\': pose_.roll(-roll); continue;
                case '|': pose_.yaw(180.0f); continue;
                default: break;
                }
            }
            if (leaves && n == "Leaf") {
                float len = p_.leafLength, wid = p_.leafWidth;
                (void)number_arg(m, 0, len);
                if (!number_arg(m, 1, wid) && hasArg) {
                    wid = len * (p_.leafLength > 0.0f ? p_.leafWidth / p_.leafLength : 1.0f);
                }
                out_.leaves.push_back(Leaf{pose_.position, pose_.tangent, pose_.right(), material_});
                leafSizes_.push_back(len);
                leafSizes_.push_back(wid);
                continue;
            }
        }
        flush();
        return std::move(out_);
    }

    std::vector<uint32_t> groups_;
    std::vector<float> leafSizes_;

private:
    struct Saved {
        Pose pose;
        size_t segmentsAtPush;
        float radiusScale;
    };

    bool decay() const { return p_.radiusDecay > 0.0f && p_.radiusDecay != 1.0f; }

    float draw() { return static_cast<float>((2.0 * jitter_.uniform() - 1.0) * static_cast<double>(p_.jitterDeg)); }

    // F / f arguments: (step, radius), (x) meaning step = radius = x, or the defaults.
    void forward_args(const lsys::Module& m, float& step, float& radius) const
    {
        float a0 = 0.0f, a1 = 0.0f;
        const bool h0 = number_arg(m, 0, a0);
        const bool h1 = number_arg(m, 1, a1);
        step = p_.defaultStepLength;
        radius = p_.defaultRadius;
        if (h0 && h1) {
            step = a0;
            radius = a1;
        } else if (h0) {
            step = a0;
            radius = a0;
        }
    }

    // Ends the open segment: its last state is the current pose with the last drawn radius.
    void flush()
    {
        if (states_.empty()) {
            return;
        }
        const float r = states_.back().radius;
        states_.push_back(State{pose_.position, pose_.tangent, r});
        Segment s;
        s.states = std::move(states_);
        s.material = material_;
        out_.segments.push_back(std::move(s));
        states_.clear();
    }

    PtProceduralBuildParams p_;
    SplitMix64 jitter_;
    TurtleResult out_;
    Pose pose_;
    std::vector<State> states_;
    std::vector<Saved> stack_;
    std::string material_ = "0";
    float radiusScale_ = 1.0f;
};

// ---- Hermite spans and loft frames (Spline.cpp:102-254) -------------------------------------

struct Span {
    Vec3 p0, p1, m0, m1;
    float r0, r1, chord;

    Vec3 point(float t) const
    {
        const float t2 = t * t;
        const float t3 = t2 * t;
        const float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f;
        const float h10 = t3 - 2.0f * t2 + t;
        const float h01 = -2.0f * t3 + 3.0f * t2;
        const float h11 = t3 - t2;
        return (p0 * h00 + m0 * h10) + (p1 * h01 + m1 * h11);
    }
    Vec3 direction(float t) const
    {
        const float t2 = t * t;
        const float d00 = 6.0f * t2 - 6.0f * t;
        const float d10 = 3.0f * t2 - 4.0f * t + 1.0f;
        const float d01 = -6.0f * t2 + 6.0f * t;
        const float d11 = 3.0f * t2 - 2.0f * t;
        return unit((p0 * d00 + m0 * d10) + (p1 * d01 + m1 * d11));
    }
    // A span is subdivided unless both end tangents lie along the chord, agree with each other and
    // the curve's midpoint is within 1e-4 chord lengths of the chord's.
    bool curved(int perSpan) const
    {
        if (perSpan <= 2 || chord <= 1e-6f) {
            return false;
        }
        const Vec3 along = (p1 - p0) * (1.0f / chord);
        const Vec3 t0 = unit(m0);
        const Vec3 t1 = unit(m1);
        if (dot(t0, along) < 0.999f || dot(t1, along) < 0.999f || dot(t0, t1) < 0.999f) {
            return true;
        }
        const Vec3 mid = point(0.5f);
        const Vec3 straight = (p0 + p1) * 0.5f;
        return length(mid - straight) > 1e-4f * chord;
    }
};

std::vector<Span> spans_of(const std::vector<State>& states, float tension)
{
    std::vector<Span> out;
    if (states.size() < 2) {
        return out;
    }
    const float k = tension > 0.0f ? tension : 1.0f;
    out.reserve(states.size() - 1);
    for (size_t i = 0; i + 1 < states.size(); ++i) {
        const State& a = states[i];
        const State& b = states[i + 1];
        const float c = length(b.position - a.position);
        out.push_back(Span{a.position, b.position, unit(a.tangent) * (c * k), unit(b.tangent) * (c * k), a.radius,
                           b.radius, c});
    }
    return out;
}

float arc_length(const std::vector<Span>& spans)
{
    float s = 0.0f;
    for (const Span& sp : spans) {
        s += sp.chord;
    }
    return s;
}

// Parallel transport of the normal from tangent a to tangent b: a rotation about a x b by the angle
// whose sine is |a x b|.
Vec3 transport(Vec3 a, Vec3 b, Vec3 normal)
{
    const Vec3 axis = cross(unit(a), unit(b));
    const float s = length(axis);
    if (s <= 1e-6f) {
        return normal;
    }
    const Vec3 n = axis * (1.0f / s);
    const float angle = std::asin(std::max(-1.0f, std::min(s, 1.0f)));
    const AxisRotation r(n, std::cos(angle), std::sin(angle));
    return unit(r.apply(normal));
}

Vec3 first_normal(Vec3 tangent)
{
    const Vec3 t = unit(tangent);
    const Vec3 up = std::fabs(dot(t, v3(0.0f, 1.0f, 0.0f))) > 0.95f ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 1.0f, 0.0f);
    return unit(cross(up, t));
}

struct Knot {
    Vec3 position, tangent;
    float radius;
};

std::vector<Frame> frames_of(const std::vector<Knot>& knots)
{
    std::vector<Frame> out;
    out.reserve(knots.size());
    if (knots.empty()) {
        return out;
    }
    Vec3 n = first_normal(knots.front().tangent);
    for (size_t i = 0; i < knots.size(); ++i) {
        if (i > 0) {
            n = transport(knots[i - 1].tangent, knots[i].tangent, n);
        }
        const Vec3 t = unit(knots[i].tangent);
        const Vec3 b = unit(cross(t, n));
        n = unit(cross(b, t));
        out.push_back(Frame{knots[i].position, t, n, b, knots[i].radius});
    }
    return out;
}

void add_knot(std::vector<Knot>& knots, const Knot& k)
{
    if (!knots.empty() && length(k.position - knots.back().position) <= 1e-6f) {
        return;
    }
    knots.push_back(k);
}

std::vector<Frame> loft_frames_of(const std::vector<Span>& spans, int samplesPerSpan)
{
    if (spans.empty()) {
        return {};
    }
    const int per = std::max(samplesPerSpan, 2);
    std::vector<Knot> knots;
    knots.reserve(spans.size() + 1);
    for (size_t i = 0; i < spans.size(); ++i) {
        const Span& s = spans[i];
        if (i == 0) {
            add_knot(knots, Knot{s.p0, unit(s.m0), s.r0});
        }
        if (s.curved(per)) {
            for (int k = 1; k < per - 1; ++k) {
                const float t = static_cast<float>(k) / static_cast<float>(per - 1);
                add_knot(knots, Knot{s.point(t), s.direction(t), lerp(s.r0, s.r1, t)});
            }
        }
        add_knot(knots, Knot{s.p1, unit(s.m1), s.r1});
    }
    return frames_of(knots);
}

// SplinePath::sampleUniform: per span `per` samples at t = k / (per - 1), the first sample of
// every span after the first dropped (it repeats the previous span's last).
std::vector<Knot> uniform_samples(const std::vector<Span>& spans, int samplesPerSpan)
{
    std::vector<Knot> out;
    const int per = std::max(samplesPerSpan, 2);
    for (size_t i = 0; i < spans.size(); ++i) {
        for (int k = i == 0 ? 0 : 1; k < per; ++k) {
            const float t = static_cast<float>(k) / static_cast<float>(per - 1);
            out.push_back(Knot{spans[i].point(t), spans[i].direction(t), lerp(spans[i].r0, spans[i].r1, t)});
        }
    }
    return out;
}

// ---- tube walls and caps (Loft.cpp:108-155) -------------------------------------------------

struct Corner {
    Vec3 p, n;
    Vec2 uv;
};

struct Ring {
    // cos and sin of seg * 2 pi / segments, for seg = 0 .. segments - 1
    std::vector<float> c, s;
    explicit Ring(int segments) : c(segments), s(segments)
    {
        for (int k = 0; k < segments; ++k) {
            const float a = static_cast<float>(k) * 6.283185307f / static_cast<float>(segments);
            c[k] = std::cos(a);
            s[k] = std::sin(a);
        }
    }
    Vec3 at(const Frame& f, int k) const { return f.position + (f.normal * (c[k] * f.radius) + f.binormal * (s[k] * f.radius)); }
};

void push(std::vector<HostTriangle>& out, const Corner& a, const Corner& b, const Corner& c)
{
    HostTriangle t{};
    t.v0 = a.p;
    t.v1 = b.p;
    t.v2 = c.p;
    t.n0 = a.n;
    t.n1 = b.n;
    t.n2 = c.n;
    t.uv0 = a.uv;
    t.uv1 = b.uv;
    t.uv2 = c.uv;
    out.push_back(t);
}

// ---- spheres for zero-length segments ------------------------------------------------------

// Octahedron with each face cut into n^2 triangles (n = (segments + 3) / 4, so that 4n vertices
// run round the equator), vertices pushed out to the sphere; normals radial, UVs zero.
void sphere(Vec3 centre, float radius, int segments, std::vector<HostTriangle>& out)
{
    const int n = std::max(1, (std::max(segments, 3) + 3) / 4);
    auto corner = [&](const double d[3]) {
        const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const Vec3 u = v3(static_cast<float>(d[0] / len), static_cast<float>(d[1] / len), static_cast<float>(d[2] / len));
        return Corner{v3(static_cast<float>(centre.x + radius * d[0] / len),
                         static_cast<float>(centre.y + radius * d[1] / len),
                         static_cast<float>(centre.z + radius * d[2] / len)),
                      u, Vec2{0.0f, 0.0f}};
    };
    for (int face = 0; face < 8; ++face) {
        const double sx = (face & 1) ? -1.0 : 1.0, sy = (face & 2) ? -1.0 : 1.0, sz = (face & 4) ? -1.0 : 1.0;
        double A[3] = {sx, 0.0, 0.0}, B[3] = {0.0, sy, 0.0}, Cc[3] = {0.0, 0.0, sz};
        if (sx * sy * sz < 0.0) {
            std::swap(B, Cc);   // keep the winding counter-clockwise seen from outside
        }
        auto grid = [&](int i, int j) {
            double d[3];
            for (int k = 0; k < 3; ++k) {
                d[k] = A[k] + (B[k] - A[k]) * i / n + (Cc[k] - A[k]) * j / n;
            }
            return corner(d);
        };
        for (int i = 0; i < n; ++i) {
            for (int j = 0; i + j < n; ++j) {
                push(out, grid(i, j), grid(i + 1, j), grid(i, j + 1));
                if (i + j + 1 < n) {
                    push(out, grid(i + 1, j), grid(i + 1, j + 1), grid(i, j + 1));
                }
            }
        }
    }
}

// ---- leaves ---------------------------------------------------------------------------------

void leaf_quad(const Leaf& l, float len, float wid, std::vector<HostTriangle>& out)
{
    const float hw = wid * 0.5f;
    const Vec3 a = l.position - l.right * hw;
    const Vec3 b = l.position + l.right * hw;
    const Vec3 c = b + l.tangent * len;
    const Vec3 d = a + l.tangent * len;
    const Vec3 n = unit(cross(l.right, l.tangent));
    push(out, Corner{a, n, Vec2{0.0f, 0.0f}}, Corner{b, n, Vec2{1.0f, 0.0f}}, Corner{c, n, Vec2{1.0f, 1.0f}});
    push(out, Corner{a, n, Vec2{0.0f, 0.0f}}, Corner{c, n, Vec2{1.0f, 1.0f}}, Corner{d, n, Vec2{0.0f, 1.0f}});
}

// ---- root transform (ProceduralMeshBuilder.cpp:198-247) -------------------------------------

Vec3 rotate_about(Vec3 v, Vec3 axis, float rad)
{
    const Vec3 a = unit(axis);
    const float c = std::cos(rad);
    const float s = std::sin(rad);
    const float along = dot(a, v);
    return (v * c + cross(a, v) * s) + a * (along * (1.0f - c));
}

struct Root {
    Vec3 t, deg;
    Vec3 rotate(Vec3 v) const
    {
        const float toRad = 0.0174532925f;
        if (std::fabs(deg.x) > 1.0e-6f) {
            v = rotate_about(v, v3(0.0f, 1.0f, 0.0f), deg.x * toRad);
        }
        if (std::fabs(deg.y) > 1.0e-6f) {
            v = rotate_about(v, v3(1.0f, 0.0f, 0.0f), deg.y * toRad);
        }
        if (std::fabs(deg.z) > 1.0e-6f) {
            v = rotate_about(v, v3(0.0f, 0.0f, 1.0f), deg.z * toRad);
        }
        return v;
    }
    void apply(std::vector<HostTriangle>& tris, size_t from) const
    {
        for (size_t i = from; i < tris.size(); ++i) {
            HostTriangle& x = tris[i];
            x.v0 = rotate(x.v0) + t;
            x.v1 = rotate(x.v1) + t;
            x.v2 = rotate(x.v2) + t;
            x.n0 = rotate(x.n0);
            x.n1 = rotate(x.n1);
            x.n2 = rotate(x.n2);
        }
    }
};

// ---- material table (ProceduralMeshBuilder.cpp:39-196) --------------------------------------

TextureDescGpu pack_texture(const lsys::Channel& ch)
{
    TextureDescGpu d{};
    const std::vector<float>& p = ch.textureParams;
    if (ch.textureKind == "Grid") {
        d.kind = 2;
        float fu = 8.0f, fv = 8.0f, th = 0.05f;
        switch (p.size()) {
        case 6: break;
        case 7: fu = fv = p[6]; break;
        case 8: fu = fv = p[6]; th = p[7]; break;
        default: fu = p[6]; fv = p[7]; th = p[8]; break;   // 9 or 10 parameters
        }
        d.p0 = PtFloat4{fu, fv, th, 0.0f};
        d.p1 = PtFloat4{p[0], p[1], p[2], 0.0f};
        d.p2 = PtFloat4{p[3], p[4], p[5], 0.0f};
    } else {
        d.kind = 3;
        d.p0 = PtFloat4{p[0], p[1], p.size() > 2 ? p[2] : 1.0f, p.size() > 3 ? p[3] : 0.0f};
    }
    return d;
}

// Texture slot of a channel: 0 for an inline value; otherwise a new bank entry (index 0 of a
// non-empty bank is a constant placeholder).
uint32_t texture_slot(const lsys::Channel& ch, std::vector<TextureDescGpu>& bank)
{
    if (!ch.texture) {
        return 0u;
    }
    if (bank.empty()) {
        bank.push_back(TextureDescGpu{});   // kind 0 = ConstantScalar
    }
    bank.push_back(pack_texture(ch));
    return static_cast<uint32_t>(bank.size() - 1);
}

MaterialGpu default_material()
{
    MaterialGpu m{};
    m.r = m.g = m.b = 0.8f;
    m.roughness = 0.5f;
    m.ior = 1.5f;
    m.diffuseRoughness = -1.0f;
    m.specular = 1.0f;
    return m;
}

MaterialGpu pack_material(const lsys::MaterialDef& d, std::vector<TextureDescGpu>& bank)
{
    auto scalar = [](const lsys::Channel& c, float dflt) { return c.texture ? dflt : c.value; };
    MaterialGpu m{};
    m.r = d.albedo.texture ? 0.8f : d.albedo.r;
    m.g = d.albedo.texture ? 0.8f : d.albedo.g;
    m.b = d.albedo.texture ? 0.8f : d.albedo.b;
    m.roughness = scalar(d.roughness, 0.5f);
    m.metallic = scalar(d.metallic, 0.0f);
    m.emission = scalar(d.emission, 0.0f);
    m.ior = scalar(d.ior, 1.5f);
    m.transmission = scalar(d.transmission, 0.0f);
    m.thin = scalar(d.thin, 0.0f);
    m.subsurface = scalar(d.subsurface, 0.0f);
    m.diffuseRoughness = scalar(d.diffuseRoughness, -1.0f);
    m.scatterRadiusR = scalar(d.scatterR, 0.0f);
    m.scatterRadiusG = scalar(d.scatterG, 0.0f);
    m.scatterRadiusB = scalar(d.scatterB, 0.0f);
    m.specular = scalar(d.specular, 1.0f);
    // bank order: albedo, roughness, metallic, transmission, thin, ior, subsurface, emission
    m.albedoTex = texture_slot(d.albedo, bank);
    m.roughnessTex = texture_slot(d.roughness, bank);
    m.metallicTex = texture_slot(d.metallic, bank);
    m.transmissionTex = texture_slot(d.transmission, bank);
    m.thinTex = texture_slot(d.thin, bank);
    m.iorTex = texture_slot(d.ior, bank);
    m.subsurfaceTex = texture_slot(d.subsurface, bank);
    m.emissionTex = texture_slot(d.emission, bank);
    return m;
}

// Materials in definition order (first of a repeated id wins), then the default grey as "0"
// unless defined; returns the id -> index map.
std::unordered_map<std::string, uint32_t> material_table(const std::vector<lsys::MaterialDef>& defs, HostMesh& out)
{
    std::unordered_map<std::string, uint32_t> index;
    for (const lsys::MaterialDef& d : defs) {
        if (index.count(d.id) != 0) {
            continue;
        }
        index.emplace(d.id, static_cast<uint32_t>(out.materials.size()));
        out.materials.push_back(pack_material(d, out.textures));
    }
    if (index.count("0") == 0) {
        index.emplace("0", static_cast<uint32_t>(out.materials.size()));
        out.materials.push_back(default_material());
    }
    return index;
}

uint32_t material_index(const std::unordered_map<std::string, uint32_t>& table, const std::string& id)
{
    auto it = table.find(id);
    if (it == table.end()) {
        it = table.find("0");
    }
    return it == table.end() ? 0u : it->second;
}

// Appends src to dst: material indices offset by `materialOffset`, non-zero texture slots by the
// texture count of dst (HostMesh.h:77-120).
void append(HostMesh& dst, const HostMesh& src, uint32_t materialOffset)
{
    const size_t t0 = dst.triangles.size();
    dst.triangles.insert(dst.triangles.end(), src.triangles.begin(), src.triangles.end());
    for (size_t i = t0; i < dst.triangles.size(); ++i) {
        dst.triangles[i].materialIndex += materialOffset;
    }
    const uint32_t texOffset = static_cast<uint32_t>(dst.textures.size());
    dst.textures.insert(dst.textures.end(), src.textures.begin(), src.textures.end());
    const size_t m0 = dst.materials.size();
    dst.materials.insert(dst.materials.end(), src.materials.begin(), src.materials.end());
    for (size_t i = m0; i < dst.materials.size(); ++i) {
        MaterialGpu& m = dst.materials[i];
        for (uint32_t* slot : {&m.albedoTex, &m.roughnessTex, &m.metallicTex, &m.transmissionTex, &m.thinTex,
                               &m.iorTex, &m.subsurfaceTex, &m.emissionTex}) {
            if (*slot != 0u) {
                *slot += texOffset;
            }
        }
    }
}

lsys::System parse_definition(const std::string& def, const PtProceduralBuildParams& p)
{
    return lsys::parse(def, (p.turtleFlags & PT_TURTLE_TURN_SYMBOLS) != 0);
}

} // namespace

TurtleResult turtle(const std::vector<lsys::Module>& modules, const PtProceduralBuildParams& p)
{
    Turtle t(p);
    return t.run(modules);
}

std::vector<Frame> loft_frames(const Segment& s, const PtProceduralBuildParams& p, bool* zeroLength)
{
    const std::vector<Span> spans = spans_of(s.states, p.hermiteTension);
    if (zeroLength != nullptr) {
        *zeroLength = !spans.empty() && arc_length(spans) <= 1e-6f;
    }
    if (spans.empty()) {
        return {};
    }
    return loft_frames_of(spans, p.samplesPerSpan);
}

void loft_triangles(const std::vector<Frame>& frames, int segments, std::vector<HostTriangle>& out)
{
    if (frames.size() < 2 || segments < 3) {
        return;
    }
    const Ring ring(segments);
    const int rings = static_cast<int>(frames.size());
    const float vStep = 1.0f / static_cast<float>(rings - 1);
    const size_t need = out.size() + static_cast<size_t>((rings - 1) * segments * 2 + segments * 2);
    if (need > out.capacity()) {
        out.reserve(std::max(need, 2 * out.capacity()));   // geometric growth: appends stay linear
    }
    // side wall: column k in 0..segments (column `segments` repeats column 0's position with u = 1)
    auto wall = [&](const Frame& f, int k, float v) {
        const Vec3 p = ring.at(f, k < segments ? k : 0);
        return Corner{p, unit(p - f.position), Vec2{static_cast<float>(k) / static_cast<float>(segments), v}};
    };
    for (int r = 0; r + 1 < rings; ++r) {
        const float va = static_cast<float>(r) * vStep;
        const float vb = static_cast<float>(r + 1) * vStep;
        const Frame& fa = frames[static_cast<size_t>(r)];
        const Frame& fb = frames[static_cast<size_t>(r + 1)];
        for (int k = 0; k < segments; ++k) {
            const Corner a0 = wall(fa, k, va), a1 = wall(fa, k + 1, va);
            const Corner b0 = wall(fb, k, vb), b1 = wall(fb, k + 1, vb);
            push(out, a0, a1, b0);
            push(out, a1, b1, b0);
        }
    }
    // caps: flat normals along -tangent (start) and +tangent (end), planar UVs (0.5 + 0.5 sin, 0.5 + 0.5 cos)
    const Frame& fs = frames.front();
    const Frame& fe = frames.back();
    const Vec3 ns = unit(fs.tangent * -1.0f);
    const Vec3 ne = unit(fe.tangent * 1.0f);
    const Corner cs{fs.position, ns, Vec2{0.5f, 0.5f}};
    const Corner ce{fe.position, ne, Vec2{0.5f, 0.5f}};
    auto rim = [&](const Frame& f, int k, Vec3 n) {
        return Corner{ring.at(f, k), n, Vec2{0.5f + 0.5f * ring.s[k], 0.5f + 0.5f * ring.c[k]}};
    };
    for (int k = 0; k < segments; ++k) {
        const int k1 = (k + 1) % segments;
        push(out, cs, rim(fs, k1, ns), rim(fs, k, ns));
        push(out, ce, rim(fe, k, ne), rim(fe, k1, ne));
    }
}

HostMesh build(const std::string& definition, std::size_t iterations, Vec3 translation, Vec3 rotationDeg,
               const PtProceduralBuildParams& p)
{
    const lsys::System sys = parse_definition(definition, p);
    const std::vector<lsys::Module> gen = lsys::evaluate(sys, iterations, p.ruleSeed);
    Turtle t(p);
    const TurtleResult tr = t.run(gen);
    if (tr.segments.empty()) {
        throw std::runtime_error("");   // no message: the turtle drew nothing
    }
    HostMesh out;
    const auto table = material_table(sys.materials, out);
    const Root root{translation, rotationDeg};
    for (const Segment& s : tr.segments) {
        const size_t first = out.triangles.size();
        bool zero = false;
        const std::vector<Frame> frames = loft_frames(s, p, &zero);
        if (zero) {
            float r = s.states.front().radius;
            if (r <= 1e-6f) {
                r = p.defaultRadius;
            }
            if (r > 1e-6f) {
                sphere(s.states.front().position, r, p.circularSegments, out.triangles);
            }
        } else {
            // a tube whose rings all have zero radius has no area (the reference's manifold check
            // drops it)
            const bool flat = std::all_of(frames.begin(), frames.end(), [](const Frame& f) { return f.radius == 0.0f; });
            if (!flat) {
                loft_triangles(frames, p.circularSegments, out.triangles);
            }
        }
        root.apply(out.triangles, first);
        const uint32_t mi = material_index(table, s.material);
        for (size_t i = first; i < out.triangles.size(); ++i) {
            out.triangles[i].materialIndex = mi;
        }
    }
    for (size_t i = 0; i < tr.leaves.size(); ++i) {
        const size_t first = out.triangles.size();
        leaf_quad(tr.leaves[i], t.leafSizes_[2 * i], t.leafSizes_[2 * i + 1], out.triangles);
        root.apply(out.triangles, first);
        const uint32_t mi = material_index(table, tr.leaves[i].material);
        for (size_t k = first; k < out.triangles.size(); ++k) {
            out.triangles[k].materialIndex = mi;
        }
    }
    if (out.triangles.empty()) {
        throw std::runtime_error("");
    }
    return out;
}

HostMesh build_scene(const PtProceduralInstance* instances, uint32_t count, const PtProceduralBuildParams& p)
{
    HostMesh scene;
    for (uint32_t i = 0; i < count; ++i) {
        const PtProceduralInstance& in = instances[i];
        HostMesh piece;
        try {
            piece = build(in.commandString != nullptr ? in.commandString : "", static_cast<std::size_t>(in.iterations),
                          in.translation, in.rotationDeg, p);
        } catch (const std::exception& e) {
            throw std::runtime_error("Procedural instance " + std::to_string(i) + ": " + e.what());
        }
        append(scene, piece, static_cast<uint32_t>(scene.materials.size()));
    }
    return scene;
}

} // namespace mesh

// ---- C ABI ----------------------------------------------------------------------------------

struct PtHostMeshBuf {
    mesh::HostMesh m;
};

namespace {

PtProceduralBuildParams resolved(const PtProceduralBuildParams* p)
{
    PtProceduralBuildParams d;
    ptProceduralDefaultParams(&d);
    return p != nullptr ? *p : d;
}

bool valid_params(const PtProceduralBuildParams& p, const char* who)
{
    if (p.circularSegments < 3 || p.circularSegments > 4096 || p.samplesPerSpan < 0 || p.samplesPerSpan > 1024) {
        ptSetError("%s: circularSegments must be 3..4096 and samplesPerSpan 0..1024", who);
        return false;
    }
    return true;
}

int64_t copy_text(const std::string& s, char* out, uint64_t capacity)
{
    if (out != nullptr && capacity > 0) {
        const size_t n = std::min<uint64_t>(s.size(), capacity - 1);
        std::memcpy(out, s.data(), n);
        out[n] = '\0';
    }
    return static_cast<int64_t>(s.size());
}

void append_float(std::string& s, float v)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%.9g", static_cast<double>(v));
    s += buf;
}

void channel_text(std::string& s, const lsys::Channel& c, bool rgb)
{
    if (c.texture) {
        s += " t:" + c.textureKind + ":";
        for (size_t i = 0; i < c.textureParams.size(); ++i) {
            if (i != 0) {
                s.push_back(',');
            }
            append_float(s, c.textureParams[i]);
        }
        return;
    }
    s += " i:";
    if (rgb) {
        append_float(s, c.r);
        s.push_back(',');
        append_float(s, c.g);
        s.push_back(',');
        append_float(s, c.b);
    } else {
        append_float(s, c.value);
    }
}

} // namespace

extern "C" {

void ptProceduralDefaultParams(PtProceduralBuildParams* out)
{
    if (out == nullptr) {
        return;
    }
    *out = PtProceduralBuildParams{};
    out->circularSegments = 32;
    out->samplesPerSpan = 4;
    out->hermiteTension = 1.0f;
    out->segmentRefineTolerance = 0.01f;
    out->globalRefineTolerance = 0.01f;
    out->creaseAngleDeg = 50.0f;
    out->defaultStepLength = 500.0f;
    out->defaultRadius = 100.0f;
    out->turtleFlags = 0u;
    out->turnAngleDeg = 25.0f;
    out->rollAngleDeg = 25.0f;
    out->radiusDecay = 1.0f;
    out->jitterDeg = 0.0f;
    out->jitterSeed = 0u;
    out->leafLength = 0.0f;
    out->leafWidth = 0.0f;
    out->ruleSeed = 0xC0FFEEull;
}

void ptHostMeshDestroy(PtHostMeshBuf* mesh) { delete mesh; }

void ptHostMeshCounts(const PtHostMeshBuf* mesh, uint32_t* triangles, uint32_t* materials, uint32_t* textures)
{
    if (triangles != nullptr) {
        *triangles = mesh != nullptr ? static_cast<uint32_t>(mesh->m.triangles.size()) : 0u;
    }
    if (materials != nullptr) {
        *materials = mesh != nullptr ? static_cast<uint32_t>(mesh->m.materials.size()) : 0u;
    }
    if (textures != nullptr) {
        *textures = mesh != nullptr ? static_cast<uint32_t>(mesh->m.textures.size()) : 0u;
    }
}

const HostTriangle* ptHostMeshTriangles(const PtHostMeshBuf* mesh)
{
    return mesh != nullptr && !mesh->m.triangles.empty() ? mesh->m.triangles.data() : nullptr;
}

const MaterialGpu* ptHostMeshMaterials(const PtHostMeshBuf* mesh)
{
    return mesh != nullptr && !mesh->m.materials.empty() ? mesh->m.materials.data() : nullptr;
}

const TextureDescGpu* ptHostMeshTextures(const PtHostMeshBuf* mesh)
{
    return mesh != nullptr && !mesh->m.textures.empty() ? mesh->m.textures.data() : nullptr;
}

PtHostMeshBuf* ptProceduralBuildHostMesh(const char* definition, uint64_t iterations, Vec3 translation,
                                         Vec3 rotationDeg, const PtProceduralBuildParams* params)
{
    const PtProceduralBuildParams p = resolved(params);
    if (definition == nullptr || !valid_params(p, "ptProceduralBuildHostMesh")) {
        if (definition == nullptr) {
            ptSetError("ptProceduralBuildHostMesh: definition is NULL");
        }
        return nullptr;
    }
    try {
        auto* out = new PtHostMeshBuf;
        try {
            out->m = mesh::build(definition, static_cast<std::size_t>(iterations), translation, rotationDeg, p);
        } catch (...) {
            delete out;
            throw;
        }
        return out;
    } catch (const std::exception& e) {
        ptSetError("ptProceduralBuildHostMesh: %s", e.what()[0] != '\0' ? e.what() : "the definition draws no geometry");
        return nullptr;
    }
}

PtHostMeshBuf* ptProceduralBuildScene(const PtProceduralInstance* instances, uint32_t count,
                                      const PtProceduralBuildParams* params)
{
    const PtProceduralBuildParams p = resolved(params);
    if ((instances == nullptr && count != 0) || !valid_params(p, "ptProceduralBuildScene")) {
        if (instances == nullptr && count != 0) {
            ptSetError("ptProceduralBuildScene: instances is NULL");
        }
        return nullptr;
    }
    try {
        auto* out = new PtHostMeshBuf;
        try {
            out->m = mesh::build_scene(instances, count, p);
        } catch (...) {
            delete out;
            throw;
        }
        return out;
    } catch (const std::exception& e) {
        ptSetError("ptProceduralBuildScene: %s", e.what());
        return nullptr;
    }
}

bool ptMeshSceneBuild(const PtProceduralInstance* instances, uint32_t count, PtScene* scene,
                      const PtMeshSceneBuildParams* params, hipStream_t stream)
{
    if (scene == nullptr) {
        ptSetError("ptMeshSceneBuild: scene is NULL");
        return false;
    }
    PtProceduralBuildParams p;
    ptProceduralDefaultParams(&p);
    p.circularSegments = params != nullptr ? params->circularSegments : 64;
    p.creaseAngleDeg = params != nullptr ? params->creaseAngleDeg : 50.0f;
    PtHostMeshBuf* m = ptProceduralBuildScene(instances, count, &p);
    if (m == nullptr) {
        return false;
    }
    const bool ok = ptSceneBuild(scene, m->m.triangles.data(), static_cast<uint32_t>(m->m.triangles.size()),
                                 m->m.materials.data(), static_cast<uint32_t>(m->m.materials.size()),
                                 m->m.textures.data(), static_cast<uint32_t>(m->m.textures.size()), stream);
    ptHostMeshDestroy(m);
    return ok;
}

int64_t ptLSystemEvaluateText(const char* definition, uint64_t iterations, uint64_t seed, uint32_t flags,
                              int allSteps, char* out, uint64_t capacity)
{
    if (definition == nullptr) {
        ptSetError("ptLSystemEvaluateText: definition is NULL");
        return -1;
    }
    try {
        const lsys::System sys = lsys::parse(definition, (flags & PT_TURTLE_TURN_SYMBOLS) != 0);
        std::string text;
        if (allSteps != 0) {
            const auto steps = lsys::evaluate_steps(sys, static_cast<std::size_t>(iterations), seed);
            for (size_t g = 0; g < steps.size(); ++g) {
                if (g != 0) {
                    text.push_back('\n');
                }
                text += lsys::to_text(steps[g]);
            }
        } else {
            text = lsys::to_text(lsys::evaluate(sys, static_cast<std::size_t>(iterations), seed));
        }
        return copy_text(text, out, capacity);
    } catch (const std::exception& e) {
        ptSetError("%s", e.what());
        return -1;
    }
}

int64_t ptLSystemParseDump(int kind, const char* text, uint32_t flags, char* out, uint64_t capacity)
{
    if (text == nullptr || kind < 0 || kind > 3) {
        ptSetError("ptLSystemParseDump: invalid arguments");
        return -1;
    }
    const bool ext = (flags & PT_TURTLE_TURN_SYMBOLS) != 0;
    try {
        std::string d;
        switch (kind) {
        case 0: d = lsys::dump_modules(lsys::parse_modules(text, ext)); break;
        case 1: d = lsys::dump_rules({lsys::parse_rule(text, ext)}); break;
        case 2: d = lsys::dump_rules(lsys::parse_rules(text, ext)); break;
        default: d = lsys::dump_system(lsys::parse(text, ext)); break;
        }
        return copy_text(d, out, capacity);
    } catch (const std::exception& e) {
        ptSetError("%s", e.what());
        return -1;
    }
}

int64_t ptLSystemMaterialsText(const char* definition, char* out, uint64_t capacity)
{
    if (definition == nullptr) {
        ptSetError("ptLSystemMaterialsText: definition is NULL");
        return -1;
    }
    try {
        const lsys::System sys = lsys::parse(definition);
        std::string text;
        for (size_t i = 0; i < sys.materials.size(); ++i) {
            const lsys::MaterialDef& d = sys.materials[i];
            if (i != 0) {
                text.push_back('\n');
            }
            text += d.id;
            channel_text(text, d.albedo, true);
            for (const lsys::Channel* c : {&d.roughness, &d.metallic, &d.transmission, &d.thin, &d.ior, &d.subsurface,
                                           &d.emission, &d.diffuseRoughness, &d.scatterR, &d.scatterG, &d.scatterB,
                                           &d.specular}) {
                channel_text(text, *c, false);
            }
        }
        return copy_text(text, out, capacity);
    } catch (const std::exception& e) {
        ptSetError("%s", e.what());
        return -1;
    }
}

bool ptTurtleExecute(const char* definition, uint64_t iterations, const PtProceduralBuildParams* params,
                     PtTurtleCounts* counts, float* states, uint32_t stateCapacity, uint32_t* stateCounts,
                     uint32_t segmentCapacity, char* materialIds, uint32_t materialIdCapacity, uint32_t* groups,
                     uint32_t groupCapacity, float* leaves, float* leafSizes, uint32_t leafCapacity)
{
    if (definition == nullptr || counts == nullptr) {
        ptSetError("ptTurtleExecute: invalid arguments");
        return false;
    }
    const PtProceduralBuildParams p = resolved(params);
    try {
        const lsys::System sys = mesh::parse_definition(definition, p);
        mesh::Turtle t(p);
        const mesh::TurtleResult r = t.run(lsys::evaluate(sys, static_cast<std::size_t>(iterations), p.ruleSeed));
        PtTurtleCounts c{};
        std::string ids;
        c.segments = static_cast<uint32_t>(r.segments.size());
        uint32_t st = 0;
        for (size_t i = 0; i < r.segments.size(); ++i) {
            const mesh::Segment& s = r.segments[i];
            if (stateCounts != nullptr && i < segmentCapacity) {
                stateCounts[i] = static_cast<uint32_t>(s.states.size());
            }
            for (const mesh::State& x : s.states) {
                if (states != nullptr && st < stateCapacity) {
                    const float v[7] = {x.position.x, x.position.y, x.position.z, x.tangent.x,
                                        x.tangent.y, x.tangent.z, x.radius};
                    std::memcpy(states + 7 * static_cast<size_t>(st), v, sizeof v);
                }
                ++st;
            }
            if (i != 0) {
                ids.push_back('\n');
            }
            ids += s.material;
        }
        c.states = st;
        c.groups = static_cast<uint32_t>(t.groups_.size());
        for (size_t i = 0; groups != nullptr && i < t.groups_.size() && i < groupCapacity; ++i) {
            groups[i] = t.groups_[i];
        }
        c.leaves = static_cast<uint32_t>(r.leaves.size());
        for (size_t i = 0; i < r.leaves.size() && i < leafCapacity; ++i) {
            const mesh::Leaf& l = r.leaves[i];
            if (leaves != nullptr) {
                const float v[9] = {l.position.x, l.position.y, l.position.z, l.tangent.x, l.tangent.y,
                                    l.tangent.z, l.right.x, l.right.y, l.right.z};
                std::memcpy(leaves + 9 * i, v, sizeof v);
            }
            if (leafSizes != nullptr) {
                leafSizes[2 * i] = t.leafSizes_[2 * i];
                leafSizes[2 * i + 1] = t.leafSizes_[2 * i + 1];
            }
        }
        c.materialIdChars = static_cast<uint32_t>(ids.size());
        (void)copy_text(ids, materialIds, materialIdCapacity);
        *counts = c;
        return true;
    } catch (const std::exception& e) {
        ptSetError("%s", e.what());
        return false;
    }
}

bool ptSplineFrames(const float* states, uint32_t stateCount, float tension, int samplesPerSpan, int mode,
                    float* out, uint32_t capacity, uint32_t* outCount, float* arcLength)
{
    if (states == nullptr || outCount == nullptr || (mode != 0 && mode != 1)) {
        ptSetError("ptSplineFrames: invalid arguments");
        return false;
    }
    if (stateCount < 2) {
        *outCount = 0;
        ptSetError("ptSplineFrames: a spline needs at least two states");
        return false;
    }
    std::vector<mesh::State> st(stateCount);
    for (uint32_t i = 0; i < stateCount; ++i) {
        const float* s = states + 7 * static_cast<size_t>(i);
        st[i] = mesh::State{Vec3{s[0], s[1], s[2]}, Vec3{s[3], s[4], s[5]}, s[6]};
    }
    const std::vector<mesh::Span> spans = mesh::spans_of(st, tension);
    if (arcLength != nullptr) {
        *arcLength = mesh::arc_length(spans);
    }
    if (mode == 0) {
        const std::vector<mesh::Frame> f = mesh::loft_frames_of(spans, samplesPerSpan);
        *outCount = static_cast<uint32_t>(f.size());
        for (size_t i = 0; out != nullptr && i < f.size() && i < capacity; ++i) {
            const mesh::Frame& x = f[i];
            const float v[13] = {x.position.x, x.position.y, x.position.z, x.tangent.x, x.tangent.y,
                                 x.tangent.z, x.normal.x, x.normal.y, x.normal.z, x.binormal.x,
                                 x.binormal.y, x.binormal.z, x.radius};
            std::memcpy(out + 13 * i, v, sizeof v);
        }
    } else {
        const std::vector<mesh::Knot> k = mesh::uniform_samples(spans, samplesPerSpan);
        *outCount = static_cast<uint32_t>(k.size());
        for (size_t i = 0; out != nullptr && i < k.size() && i < capacity; ++i) {
            const float v[7] = {k[i].position.x, k[i].position.y, k[i].position.z, k[i].tangent.x,
                                k[i].tangent.y, k[i].tangent.z, k[i].radius};
            std::memcpy(out + 7 * i, v, sizeof v);
        }
    }
    return true;
}

} // extern "C"
