// pt_lsystem.h -- L-system definitions: parser, simultaneous rewriting and the material table of
// `Mat(id) = {...}` lines. Host C++ (no device code); the mesher (pt_mesher.cpp) and the
// renderer facade consume it.
//
// The language is LSystemLib's (LSystemLib/src/details/ModuleStringParser.cpp,
// LSystemConstantsParse.cpp, LSystemMaterialParse.cpp, LSystemImpl.cpp; evaluation semantics of
// LSystemEvaluator.cpp:17-470):
//   - lines before the first line holding "->" form the axiom (joined with spaces), except
//     `Mat(id) = {...}` material lines and `name = number` constant lines; `#` starts a comment;
//   - a rule is `[left <] pred [> right] [: cond] [: prob] -> successor`, or `pred :: prob -> ...`;
//     an indented line continues the rule above it;
//   - modules are identifiers (letters first, then letters, digits, `_`, `.`), `^`, `&`, `[`, `]`,
//     optionally with a parenthesised list of arithmetic / comparison expressions;
//   - `ref.Name` is a subsystem reference (no arguments);
//   - rewriting is simultaneous; identifiers in predecessor and context arguments bind to the
//     matched values (a repeated identifier must match the same value), constants are bound
//     first, a condition is true when non-zero, and among several matching rules one is drawn
//     with the probabilities as weights (one draw per rewritten module).
// Extension (turtle-style modules the reference rejects): with `extended` parsing, `+`, `-`, `/`,
// `\` and `|` are accepted as modules too.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace lsys {

// Expression tree (immutable; shared between generations).
struct Expr {
    enum Op : uint8_t { Num, Ident, Add, Sub, Mul, Div, Lt, Le, Gt, Ge, Eq };
    Op op = Num;
    double num = 0.0;
    std::string ident;
    std::shared_ptr<const Expr> a, b;
};
using ExprP = std::shared_ptr<const Expr>;

struct Module {
    bool subsystemRef = false;
    std::string name;       // module name (or the subsystem's name for a reference)
    std::vector<ExprP> args;
};

struct Rule {
    std::vector<Module> left, right;
    Module pred;
    ExprP condition;        // null: always
    double probability = 1.0;
    std::vector<Module> successor;
};

// One material channel of a `Mat(id)` line: an inline value or a procedural texture.
struct Channel {
    bool texture = false;
    float value = 0.0f;
    float r = 0.8f, g = 0.8f, b = 0.8f;  // albedo
    std::string textureKind;             // "Grid" or "Stripe"
    std::vector<float> textureParams;
};

struct MaterialDef {
    std::string id;
    // albedo, roughness, metallic, transmission, thin, ior, subsurface, emission,
    // diffuseRoughness, scatterRadiusR/G/B, specular
    Channel albedo, roughness, metallic, transmission, thin, ior, subsurface, emission, diffuseRoughness,
        scatterR, scatterG, scatterB, specular;
};

struct System {
    std::vector<Module> axiom;
    std::vector<Rule> rules;
    std::vector<MaterialDef> materials;
    std::unordered_map<std::string, double> constants;
};

// Parses a whole definition; throws std::runtime_error with a message naming line and column.
System parse(const std::string& definition, bool extended = false);
// Parses a module string (axiom syntax) or a single rule.
std::vector<Module> parse_modules(const std::string& text, bool extended = false);
Rule parse_rule(const std::string& text, bool extended = false);
// Parses a rules block (one rule per line, indented lines continue the rule above).
std::vector<Rule> parse_rules(const std::string& block, bool extended = false);
// Parses one `Mat(id) = {...}` line into defs; false if the line is not a material line; throws
// on a malformed or duplicate one.
bool parse_material_line(const std::string& trimmed, std::vector<MaterialDef>& defs);

// Generations 0..iterations (generation 0 is the axiom). The rule choice uses std::mt19937_64
// seeded with `seed` (one 64-bit draw per rewritten module whose matches have a positive total
// weight), mapped to [0, total weight) as the GNU C++ library's uniform_real_distribution does.
std::vector<std::vector<Module>> evaluate_steps(const System& sys, std::size_t iterations, uint64_t seed);
std::vector<Module> evaluate(const System& sys, std::size_t iterations, uint64_t seed);

// Numeric value of an expression under the constants alone (false if it does not evaluate).
bool eval_number(const ExprP& e, const std::unordered_map<std::string, double>& constants, double& out);

// Canonical text of a generation: modules separated by spaces, arguments with %.17g.
std::string to_text(const std::vector<Module>& gen);

// Structural dumps (expressions in prefix form, `(op a b)`; numbers %.17g; subsystem references
// `ref:Name`): a module list as `[m1 m2 ...]`; a rule as `left=[..] pred=P right=[..] cond=E
// prob=p succ=[..]`; a system as `axiom=[..]`, then `rule ...` lines, `const name=value` lines in
// name order and `mat id` lines in definition order.
std::string dump_modules(const std::vector<Module>& ms);
std::string dump_rules(const std::vector<Rule>& rules);
std::string dump_system(const System& sys);

} // namespace lsys
