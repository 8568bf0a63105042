// Exposes the reference's L-system library (LSystemLib/src, compiled unmodified) and its turtle
// and spline (Application/src/Procedural/Turtle.cpp, Spline.cpp, compiled unmodified) through a C
// ABI whose outputs use the same text and float layouts as the product's inspection hooks
// (ptLSystemParseDump, ptLSystemEvaluateText, ptLSystemMaterialsText, ptTurtleExecute,
// ptSplineFrames), so the tests can compare the two bit for bit. Test infrastructure only.
// Loft.cpp and ProceduralMeshBuilder.cpp need the Manifold library, which is not in this image,
// so the tube triangles are checked against the oracle's restatement instead.
#include "LSystem.h"
#include "LSystemEvaluator.h"
#include "LSystemMaterials.h"
#include "ModuleStringParser.h"
#include "Spline.h"
#include "Turtle.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

void num(std::string& s, double v)
{
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    s += b;
}

void flt(std::string& s, float v)
{
    char b[32];
    std::snprintf(b, sizeof b, "%.9g", static_cast<double>(v));
    s += b;
}

void expr(std::string& s, const Expr* e)
{
    if (e == nullptr) {
        s += "null";
        return;
    }
    const char* op = nullptr;
    switch (e->kind) {
    case Expr::Kind::Number: num(s, e->number_value); return;
    case Expr::Kind::Ident: s += e->ident; return;
    case Expr::Kind::Add: op = "+"; break;
    case Expr::Kind::Sub: op = "-"; break;
    case Expr::Kind::Mul: op = "*"; break;
    case Expr::Kind::Div: op = "/"; break;
    case Expr::Kind::Lt: op = "<"; break;
    case Expr::Kind::Lte: op = "<="; break;
    case Expr::Kind::Gt: op = ">"; break;
    case Expr::Kind::Gte: op = ">="; break;
    case Expr::Kind::Eq: op = "=="; break;
    }
    s += "(";
    s += op;
    s += " ";
    expr(s, e->left.get());
    s += " ";
    expr(s, e->right.get());
    s += ")";
}

void symbol(std::string& s, const Symbol& m)
{
    if (m.kind == SymbolKind::SubsystemRef) {
        s += "ref:" + m.subsystem;
        return;
    }
    s += m.name;
    if (!m.args.empty()) {
        s += "(";
        for (size_t i = 0; i < m.args.size(); ++i) {
            if (i) {
                s += ",";
            }
            expr(s, m.args[i].get());
        }
        s += ")";
    }
}

void list(std::string& s, const std::vector<Symbol>& v)
{
    s += "[";
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) {
            s += " ";
        }
        symbol(s, v[i]);
    }
    s += "]";
}

void rule(std::string& s, const ParsedRule& r)
{
    s += "left=";
    list(s, r.left_context);
    s += " pred=";
    symbol(s, r.predecessor);
    s += " right=";
    list(s, r.right_context);
    s += " cond=";
    expr(s, r.condition.get());
    s += " prob=";
    num(s, r.probability);
    s += " succ=";
    list(s, r.successor);
}

std::string generation_text(const std::vector<Symbol>& g)
{
    std::string s;
    for (size_t i = 0; i < g.size(); ++i) {
        if (i) {
            s += " ";
        }
        const Symbol& m = g[i];
        if (m.kind == SymbolKind::SubsystemRef) {
            s += "ref." + m.subsystem;
            continue;
        }
        s += m.name;
        if (!m.args.empty()) {
            s += "(";
            for (size_t k = 0; k < m.args.size(); ++k) {
                if (k) {
                    s += ",";
                }
                const Expr* e = m.args[k].get();
                if (e != nullptr && e->kind == Expr::Kind::Number) {
                    num(s, e->number_value);
                } else if (e != nullptr && e->kind == Expr::Kind::Ident) {
                    s += e->ident;
                } else {
                    s += "?";
                }
            }
            s += ")";
        }
    }
    return s;
}

void channel(std::string& s, const MaterialChannel& c, bool rgb)
{
    if (c.mode == MaterialChannel::Mode::Texture) {
        s += " t:" + c.texture.kind + ":";
        for (size_t i = 0; i < c.texture.params.size(); ++i) {
            if (i) {
                s += ",";
            }
            flt(s, c.texture.params[i]);
        }
        return;
    }
    s += " i:";
    if (rgb) {
        flt(s, c.r);
        s += ",";
        flt(s, c.g);
        s += ",";
        flt(s, c.b);
    } else {
        flt(s, c.scalar);
    }
}

int64_t put(const std::string& s, char* out, uint64_t cap)
{
    if (out != nullptr && cap > 0) {
        const size_t n = std::min<uint64_t>(s.size(), cap - 1);
        std::memcpy(out, s.data(), n);
        out[n] = '\0';
    }
    return static_cast<int64_t>(s.size());
}

} // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

int64_t ref_lsystem_parse_dump(int kind, const char* text, char* out, uint64_t cap)
{
    try {
        std::string d;
        if (kind == 0) {
            list(d, ModuleStringParser::parse(text));
        } else if (kind == 1) {
            rule(d, ModuleStringParser::parse_rule(text));
        } else if (kind == 2) {
            const auto rules = ModuleStringParser::parse_rules_block(text);
            for (size_t i = 0; i < rules.size(); ++i) {
                if (i) {
                    d += "\n";
                }
                rule(d, rules[i]);
            }
        } else {
            LSystem sys;
            sys.parse(text);
            d = "axiom=";
            list(d, sys.axiom_modules());
            for (const ParsedRule& r : sys.rules()) {
                d += "\nrule ";
                rule(d, r);
            }
            std::vector<std::string> names;
            for (const auto& kv : sys.constants()) {
                names.push_back(kv.first);
            }
            std::sort(names.begin(), names.end());
            for (const std::string& n : names) {
                d += "\nconst " + n + "=";
                num(d, sys.constants().at(n));
            }
            for (const MaterialDefinition& m : sys.material_definitions()) {
                d += "\nmat " + m.id;
            }
        }
        return put(d, out, cap);
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int64_t ref_lsystem_evaluate_text(const char* definition, uint64_t iterations, uint64_t seed, int allSteps, char* out,
                                  uint64_t cap)
{
    try {
        LSystem sys;
        sys.parse(definition);
        const auto steps = LSystemEvaluator::evaluate_steps(sys, static_cast<std::size_t>(iterations), seed);
        std::string s;
        if (allSteps) {
            for (size_t g = 0; g < steps.size(); ++g) {
                if (g) {
                    s += "\n";
                }
                s += generation_text(steps[g]);
            }
        } else if (!steps.empty()) {
            s = generation_text(steps.back());
        }
        return put(s, out, cap);
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int64_t ref_lsystem_materials_text(const char* definition, char* out, uint64_t cap)
{
    try {
        LSystem sys;
        sys.parse(definition);
        std::string s;
        const auto& defs = sys.material_definitions();
        for (size_t i = 0; i < defs.size(); ++i) {
            const MaterialEntry& e = defs[i].entry;
            if (i) {
                s += "\n";
            }
            s += defs[i].id;
            channel(s, e.albedo, true);
            for (const MaterialChannel* c : {&e.roughness, &e.metallic, &e.transmission, &e.thin, &e.ior,
                                             &e.subsurface, &e.emission, &e.diffuseRoughness, &e.scatterRadiusR,
                                             &e.scatterRadiusG, &e.scatterRadiusB, &e.specular}) {
                channel(s, *c, false);
            }
        }
        return put(s, out, cap);
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

// counts: segments, states, groups, materialIdChars. Group sizes are the BranchGroup sizes.
int ref_turtle_execute(const char* definition, uint64_t iterations, float step, float radius, uint32_t* counts,
                       float* states, uint32_t stateCap, uint32_t* stateCounts, uint32_t segCap, char* ids,
                       uint32_t idCap, uint32_t* groups, uint32_t groupCap)
{
    try {
        const LSystemEvaluationResult r = LSystemEvaluator::evaluate(definition, static_cast<std::size_t>(iterations));
        TurtleParams p{};
        p.defaultStepLength = step;
        p.defaultRadius = radius;
        const TurtleOutput t = turtleExecute(r.generation, p);
        uint32_t n = 0;
        std::string idText;
        for (size_t i = 0; i < t.segments.size(); ++i) {
            if (stateCounts != nullptr && i < segCap) {
                stateCounts[i] = static_cast<uint32_t>(t.segments[i].states.size());
            }
            for (const TurtleState& x : t.segments[i].states) {
                if (states != nullptr && n < stateCap) {
                    const float v[7] = {x.position.x, x.position.y, x.position.z, x.tangent.x,
                                        x.tangent.y, x.tangent.z, x.radius};
                    std::memcpy(states + 7 * static_cast<size_t>(n), v, sizeof v);
                }
                ++n;
            }
            if (i) {
                idText += "\n";
            }
            idText += t.segments[i].materialId;
        }
        for (size_t i = 0; groups != nullptr && i < t.branchGroups.size() && i < groupCap; ++i) {
            groups[i] = static_cast<uint32_t>(t.branchGroups[i].segmentIndices.size());
        }
        counts[0] = static_cast<uint32_t>(t.segments.size());
        counts[1] = n;
        counts[2] = static_cast<uint32_t>(t.branchGroups.size());
        counts[3] = static_cast<uint32_t>(idText.size());
        put(idText, ids, idCap);
        return 1;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 0;
    }
}

// mode 0: computeLoftFrames (13 floats per frame); mode 1: sampleUniform (7 floats per sample).
int ref_spline_frames(const float* states, uint32_t n, float tension, int samplesPerSpan, int mode, float* out,
                      uint32_t cap, uint32_t* outCount, float* arc)
{
    TurtleSegment seg{};
    for (uint32_t i = 0; i < n; ++i) {
        const float* s = states + 7 * static_cast<size_t>(i);
        seg.states.push_back(TurtleState{Vec3{s[0], s[1], s[2]}, Vec3{s[3], s[4], s[5]}, s[6]});
    }
    SplinePath path;
    if (!path.buildFromSegment(seg, tension)) {
        *outCount = 0;
        return 0;
    }
    *arc = path.totalArcLength();
    if (mode == 0) {
        const std::vector<PathFrame> f = path.computeLoftFrames(samplesPerSpan);
        *outCount = static_cast<uint32_t>(f.size());
        for (size_t i = 0; i < f.size() && i < cap; ++i) {
            const PathFrame& x = f[i];
            const float v[13] = {x.position.x, x.position.y, x.position.z, x.tangent.x, x.tangent.y,
                                 x.tangent.z, x.normal.x, x.normal.y, x.normal.z, x.binormal.x,
                                 x.binormal.y, x.binormal.z, x.radius};
            std::memcpy(out + 13 * i, v, sizeof v);
        }
    } else {
        const std::vector<SplineSample> k = path.sampleUniform(samplesPerSpan);
        *outCount = static_cast<uint32_t>(k.size());
        for (size_t i = 0; i < k.size() && i < cap; ++i) {
            const float v[7] = {k[i].position.x, k[i].position.y, k[i].position.z, k[i].tangent.x,
                                k[i].tangent.y, k[i].tangent.z, k[i].radius};
            std::memcpy(out + 7 * i, v, sizeof v);
        }
    }
    return 1;
}

} // extern "C"
