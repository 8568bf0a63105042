// pt_lsystem.cpp -- L-system parser, rewriting and material lines (see pt_lsystem.h for the
// language; the reference's parser is LSystemLib/src/details/ModuleStringParser.cpp and its
// evaluator LSystemLib/src/LSystemEvaluator.cpp).
#include "pt_lsystem.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string_view>

namespace lsys {

namespace {

// ---- text helpers ---------------------------------------------------------------------------

bool space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }
bool digit(char c) { return c >= '0' && c <= '9'; }
bool letter(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); }
bool ident_char(char c) { return letter(c) || digit(c) || c == '_' || c == '.'; }

std::string_view trim(std::string_view s)
{
    while (!s.empty() && space(s.front())) {
        s.remove_prefix(1);
    }
    while (!s.empty() && space(s.back())) {
        s.remove_suffix(1);
    }
    return s;
}

std::string_view before_hash(std::string_view s)
{
    const size_t h = s.find('#');
    return h == std::string_view::npos ? s : s.substr(0, h);
}

// A line whose first non-blank character is '#'.
bool comment_line(std::string_view raw)
{
    for (char c : raw) {
        if (!space(c)) {
            return c == '#';
        }
    }
    return false;
}

std::vector<std::string_view> split_lines(const std::string& s)
{
    std::vector<std::string_view> out;
    size_t i = 0;
    while (i < s.size()) {
        const size_t nl = s.find('\n', i);
        const size_t end = nl == std::string::npos ? s.size() : nl;
        out.push_back(std::string_view(s).substr(i, end - i));
        if (nl == std::string::npos) {
            break;
        }
        i = nl + 1;
    }
    return out;
}

[[noreturn]] void fail_at(std::string_view text, size_t at, const char* what)
{
    size_t line = 1, col = 1;
    for (size_t i = 0; i < at && i < text.size(); ++i) {
        if (text[i] == '\n') {
            ++line;
            col = 1;
        } else {
            ++col;
        }
    }
    throw std::runtime_error(std::string(what) + " at line " + std::to_string(line) + " column " +
                             std::to_string(col));
}

// ---- lexer ------------------------------------------------------------------------------------

enum class T { End, Ident, Number, LPar, RPar, Comma, Plus, Minus, Star, Slash, Lt, Le, Gt, Ge, Eq, Colon,
               Arrow, LBr, RBr, Other };

struct Tok {
    T kind = T::End;
    size_t at = 0;
    std::string_view text;
    double value = 0.0;
};

class Lexer {
public:
    explicit Lexer(std::string_view s) : src_(s) { next(); }

    const Tok& tok() const { return tok_; }
    std::string_view src() const { return src_; }

    void next()
    {
        while (pos_ < src_.size() && space(src_[pos_])) {
            ++pos_;
        }
        const size_t start = pos_;
        if (pos_ >= src_.size()) {
            tok_ = Tok{T::End, pos_, {}, 0.0};
            return;
        }
        const char c = src_[pos_];
        auto single = [&](T k) {
            ++pos_;
            tok_ = Tok{k, start, src_.substr(start, 1), 0.0};
        };
        switch (c) {
        case '(': return single(T::LPar);
        case ')': return single(T::RPar);
        case '[': return single(T::LBr);
        case ']': return single(T::RBr);
        case ',': return single(T::Comma);
        case '^':
        case '&': return single(T::Ident);
        case '+': return single(T::Plus);
        case '*': return single(T::Star);
        case '/': return single(T::Slash);
        case ':': return single(T::Colon);
        case '\\':
        case '|': return single(T::Other);
        default: break;
        }
        if (c == '-') {
            if (pos_ + 1 < src_.size() && src_[pos_ + 1] == '>') {
                pos_ += 2;
                tok_ = Tok{T::Arrow, start, src_.substr(start, 2), 0.0};
                return;
            }
            if (unary_position(start)) {
                size_t p = pos_ + 1;
                bool dot = false;
                if (mantissa(p, dot)) {
                    pos_ = p;
                    number(start, dot);
                    return;
                }
            }
            return single(T::Minus);
        }
        if (c == '<' || c == '>' || c == '=') {
            const bool eq2 = pos_ + 1 < src_.size() && src_[pos_ + 1] == '=';
            T k = c == '<' ? (eq2 ? T::Le : T::Lt) : c == '>' ? (eq2 ? T::Ge : T::Gt) : T::Eq;
            pos_ += eq2 ? 2 : 1;
            tok_ = Tok{k, start, src_.substr(start, pos_ - start), 0.0};
            return;
        }
        if (digit(c)) {
            size_t p = pos_;
            bool dot = false;
            if (!mantissa(p, dot)) {
                fail_at(src_, start, "invalid number");
            }
            pos_ = p;
            number(start, dot);
            return;
        }
        if (letter(c)) {
            ++pos_;
            while (pos_ < src_.size() && ident_char(src_[pos_])) {
                ++pos_;
            }
            tok_ = Tok{T::Ident, start, src_.substr(start, pos_ - start), 0.0};
            return;
        }
        fail_at(src_, start, "unexpected character");
    }

private:
    // '-' at p starts a signed literal where an operand is expected: at the start, after an
    // opening bracket, a separator or an operator.
    bool unary_position(size_t p) const
    {
        while (p > 0 && space(src_[p - 1])) {
            --p;
        }
        if (p == 0) {
            return true;
        }
        switch (src_[p - 1]) {
        case '(': case '[': case ',': case ':': case '+': case '-': case '*': case '/': case '<': case '>':
        case '=':
            return true;
        default:
            return false;
        }
    }

    // Digits with at most one '.', or '.' followed by digits. Advances p past it.
    bool mantissa(size_t& p, bool& dot) const
    {
        dot = false;
        if (p >= src_.size()) {
            return false;
        }
        if (src_[p] == '.') {
            if (p + 1 >= src_.size() || !digit(src_[p + 1])) {
                return false;
            }
            dot = true;
            ++p;
            while (p < src_.size() && digit(src_[p])) {
                ++p;
            }
            return true;
        }
        if (!digit(src_[p])) {
            return false;
        }
        while (p < src_.size() && (digit(src_[p]) || (src_[p] == '.' && !dot))) {
            dot = dot || src_[p] == '.';
            ++p;
        }
        return true;
    }

    void number(size_t start, bool dot)
    {
        const std::string_view s = src_.substr(start, pos_ - start);
        if (dot && (s.back() == '.')) {
            fail_at(src_, start, "invalid number");
        }
        const std::string tmp(s);
        char* end = nullptr;
        const double v = std::strtod(tmp.c_str(), &end);
        if (end == tmp.c_str()) {
            fail_at(src_, start, "invalid number");
        }
        tok_ = Tok{T::Number, start, s, v};
    }

    std::string_view src_;
    size_t pos_ = 0;
    Tok tok_;
};

// ---- parser -----------------------------------------------------------------------------------

ExprP node(Expr::Op op, ExprP a = nullptr, ExprP b = nullptr)
{
    auto e = std::make_shared<Expr>();
    e->op = op;
    e->a = std::move(a);
    e->b = std::move(b);
    return e;
}

bool is_ref_name(std::string_view s) { return s == "ref" || (s.size() >= 4 && s.substr(0, 4) == "ref."); }

class Parser {
public:
    Parser(std::string_view s, bool extended) : lx_(s), ext_(extended) {}

    std::vector<Module> modules()
    {
        std::vector<Module> out;
        while (lx_.tok().kind != T::End) {
            out.push_back(module_or_bracket(true));
        }
        return out;
    }

    Rule rule()
    {
        Rule r;
        std::vector<Module> lead;
        while (starts_module(false)) {
            lead.push_back(module(false));
        }
        const T k = lx_.tok().kind;
        if (k == T::Lt) {
            r.left = std::move(lead);
            lx_.next();
            r.pred = module(false);
            if (lx_.tok().kind == T::Gt) {
                lx_.next();
                while (starts_module(false)) {
                    r.right.push_back(module(false));
                }
            }
        } else if (k == T::Arrow || k == T::Colon) {
            if (lead.empty()) {
                fail_at(lx_.src(), lx_.tok().at, "expected predecessor module");
            }
            if (lead.size() > 1) {
                fail_at(lx_.src(), lx_.tok().at, "ambiguous rule: use '<' to separate context from predecessor");
            }
            r.pred = std::move(lead[0]);
        } else {
            fail_at(lx_.src(), lx_.tok().at,
                    lead.empty() ? "expected module name or '<'" : "expected '<', ':', or '->' after module(s)");
        }
        if (lx_.tok().kind == T::Arrow) {
            lx_.next();
        } else if (lx_.tok().kind == T::Colon) {
            lx_.next();
            if (lx_.tok().kind == T::Colon) {
                lx_.next();
                r.probability = probability("expected probability after '::'");
            } else if (lx_.tok().kind != T::Arrow) {
                r.condition = expr();
                if (lx_.tok().kind == T::Colon) {
                    lx_.next();
                    r.probability = probability("expected probability after second ':'");
                }
            }
            expect(T::Arrow, "expected '->' before successor");
        } else {
            fail_at(lx_.src(), lx_.tok().at, "expected ':', or '->' after context");
        }
        while (lx_.tok().kind != T::End) {
            r.successor.push_back(module_or_bracket(true));
        }
        return r;
    }

private:
    double probability(const char* what)
    {
        if (lx_.tok().kind != T::Number) {
            fail_at(lx_.src(), lx_.tok().at, what);
        }
        const double p = lx_.tok().value;
        lx_.next();
        return p;
    }

    void expect(T k, const char* what)
    {
        if (lx_.tok().kind != k) {
            fail_at(lx_.src(), lx_.tok().at, what);
        }
        lx_.next();
    }

    // Single-character turtle modules of the extended language.
    bool turtle_symbol() const
    {
        const T k = lx_.tok().kind;
        return ext_ && (k == T::Plus || k == T::Minus || k == T::Slash || k == T::Other);
    }

    bool starts_module(bool) const { return lx_.tok().kind == T::Ident || turtle_symbol(); }

    Module module_or_bracket(bool allowRef)
    {
        const T k = lx_.tok().kind;
        if (k == T::LBr || k == T::RBr) {
            Module m;
            m.name = std::string(lx_.tok().text);
            lx_.next();
            return m;
        }
        return module(allowRef);
    }

    Module module(bool allowRef)
    {
        const Tok t = lx_.tok();
        Module m;
        if (turtle_symbol()) {
            m.name = std::string(t.text);
            lx_.next();
        } else {
            if (t.kind != T::Ident) {
                fail_at(lx_.src(), t.at, "expected identifier (module name)");
            }
            lx_.next();
            if (is_ref_name(t.text)) {
                if (!allowRef) {
                    fail_at(lx_.src(), t.at, "subsystem reference is not allowed here");
                }
                if (t.text == "ref") {
                    fail_at(lx_.src(), t.at, "invalid subsystem reference (expected ref.Name)");
                }
                if (t.text.size() == 4) {
                    fail_at(lx_.src(), t.at, "invalid subsystem reference (empty name)");
                }
                const std::string target(t.text.substr(4));
                if (target.front() == '.' || target.find("..") != std::string::npos) {
                    fail_at(lx_.src(), t.at, "invalid subsystem reference");
                }
                if (lx_.tok().kind == T::LPar) {
                    fail_at(lx_.src(), lx_.tok().at, "subsystem reference does not support arguments");
                }
                m.subsystemRef = true;
                m.name = target;
                return m;
            }
            m.name = std::string(t.text);
        }
        if (lx_.tok().kind == T::LPar) {
            lx_.next();
            if (lx_.tok().kind != T::RPar) {
                while (true) {
                    m.args.push_back(expr());
                    if (lx_.tok().kind != T::Comma) {
                        break;
                    }
                    lx_.next();
                }
            }
            expect(T::RPar, "expected ')' after argument list");
        }
        return m;
    }

    ExprP expr()
    {
        ExprP l = additive();
        const T k = lx_.tok().kind;
        if (k == T::Lt || k == T::Le || k == T::Gt || k == T::Ge || k == T::Eq) {
            lx_.next();
            const Expr::Op op = k == T::Lt ? Expr::Lt : k == T::Le ? Expr::Le : k == T::Gt ? Expr::Gt
                                : k == T::Ge ? Expr::Ge : Expr::Eq;
            l = node(op, l, additive());
        }
        return l;
    }

    ExprP additive()
    {
        ExprP l = term();
        while (lx_.tok().kind == T::Plus || lx_.tok().kind == T::Minus) {
            const Expr::Op op = lx_.tok().kind == T::Plus ? Expr::Add : Expr::Sub;
            lx_.next();
            l = node(op, l, term());
        }
        return l;
    }

    ExprP term()
    {
        ExprP l = factor();
        while (lx_.tok().kind == T::Star || lx_.tok().kind == T::Slash) {
            const Expr::Op op = lx_.tok().kind == T::Star ? Expr::Mul : Expr::Div;
            lx_.next();
            l = node(op, l, factor());
        }
        return l;
    }

    ExprP factor()
    {
        const Tok t = lx_.tok();
        if (t.kind == T::Number) {
            auto e = std::make_shared<Expr>();
            e->op = Expr::Num;
            e->num = t.value;
            lx_.next();
            return e;
        }
        if (t.kind == T::Ident) {
            if (is_ref_name(t.text)) {
                fail_at(lx_.src(), t.at, "subsystem reference is not allowed in expressions");
            }
            auto e = std::make_shared<Expr>();
            e->op = Expr::Ident;
            e->ident = std::string(t.text);
            lx_.next();
            return e;
        }
        if (t.kind == T::LPar) {
            lx_.next();
            ExprP inner = expr();
            expect(T::RPar, "expected ')' after expression");
            return inner;
        }
        fail_at(lx_.src(), t.at, "expected number, identifier, or '('");
    }

    Lexer lx_;
    bool ext_;
};

// Rules block: a rule per line; an indented line continues the rule above it; blank lines and
// comment lines do not end a rule that a continuation line follows.
std::vector<Rule> parse_rules_block(const std::vector<std::string_view>& lines, bool extended)
{
    std::vector<Rule> rules;
    std::string buf;
    auto flush = [&]() {
        if (!buf.empty()) {
            rules.push_back(parse_rule(buf, extended));
            buf.clear();
        }
    };
    for (std::string_view raw : lines) {
        if (comment_line(raw)) {
            continue;
        }
        const std::string_view t = trim(before_hash(raw));
        if (t.empty()) {
            continue;
        }
        if (!buf.empty() && (raw.front() == ' ' || raw.front() == '\t')) {
            buf.push_back(' ');
            buf.append(t.data(), t.size());
        } else {
            flush();
            buf.assign(t.data(), t.size());
        }
    }
    flush();
    return rules;
}

// ---- constants: `name = number` lines ---------------------------------------------------------

bool parse_name(std::string_view s, size_t& p, std::string& out)
{
    while (p < s.size() && space(s[p])) {
        ++p;
    }
    if (p >= s.size() || !(std::isalpha(static_cast<unsigned char>(s[p])) || s[p] == '_')) {
        return false;
    }
    const size_t b = p++;
    while (p < s.size() && (std::isalnum(static_cast<unsigned char>(s[p])) || s[p] == '_')) {
        ++p;
    }
    out.assign(s.data() + b, p - b);
    return true;
}

bool parse_double(std::string_view s, size_t& p, double& out)
{
    while (p < s.size() && space(s[p])) {
        ++p;
    }
    if (p >= s.size()) {
        return false;
    }
    const std::string tmp(s.substr(p));
    char* end = nullptr;
    out = std::strtod(tmp.c_str(), &end);
    if (end == tmp.c_str()) {
        return false;
    }
    p += static_cast<size_t>(end - tmp.c_str());
    return true;
}

bool is_material_line(std::string_view t);

// Fills name/value when t is `name = number`; the reference's LSystemConstantsParse.cpp form.
bool constant_line(std::string_view t, std::string& name, double& value)
{
    if (t.empty() || is_material_line(t)) {
        return false;
    }
    size_t p = 0;
    if (!parse_name(t, p, name)) {
        return false;
    }
    while (p < t.size() && space(t[p])) {
        ++p;
    }
    if (p >= t.size() || t[p] != '=') {
        return false;
    }
    ++p;
    if (!parse_double(t, p, value)) {
        return false;
    }
    while (p < t.size() && space(t[p])) {
        ++p;
    }
    return p == t.size();
}

// ---- materials: `Mat(id) = {...}` lines -------------------------------------------------------

[[noreturn]] void material_fail(std::string_view line, const char* what)
{
    throw std::runtime_error(std::string(what) + ": " + std::string(line));
}

void skip_spaces(std::string_view s, size_t& p)
{
    while (p < s.size() && space(s[p])) {
        ++p;
    }
}

// Material id: digits, or a letter or '_' followed by letters, digits and '_'.
bool material_id(std::string_view s, size_t& p, std::string& id)
{
    skip_spaces(s, p);
    if (p >= s.size()) {
        return false;
    }
    const size_t b = p;
    if (digit(s[p])) {
        while (p < s.size() && digit(s[p])) {
            ++p;
        }
    } else if (std::isalpha(static_cast<unsigned char>(s[p])) || s[p] == '_') {
        ++p;
        while (p < s.size() && (std::isalnum(static_cast<unsigned char>(s[p])) || s[p] == '_')) {
            ++p;
        }
    } else {
        return false;
    }
    id.assign(s.data() + b, p - b);
    return true;
}

bool mat_head(std::string_view s, size_t& p, std::string& id)
{
    if (s.size() < 3 || s.substr(0, 3) != "Mat") {
        return false;
    }
    p = 3;
    skip_spaces(s, p);
    if (p >= s.size() || s[p] != '(') {
        return false;
    }
    ++p;
    if (!material_id(s, p, id)) {
        return false;
    }
    skip_spaces(s, p);
    if (p >= s.size() || s[p] != ')') {
        return false;
    }
    ++p;
    return true;
}

bool is_material_line(std::string_view t)
{
    size_t p = 0;
    std::string id;
    if (!mat_head(t, p, id)) {
        return false;
    }
    skip_spaces(t, p);
    if (p >= t.size() || t[p] != '=') {
        return false;
    }
    ++p;
    skip_spaces(t, p);
    return p < t.size() && t[p] == '{';
}

bool parse_float_at(std::string_view s, size_t& p, float& out)
{
    double d = 0.0;
    if (!parse_double(s, p, d)) {
        return false;
    }
    out = static_cast<float>(d);
    return true;
}

struct Component {
    bool texture = false;
    float value = 0.0f;
    std::string kind;
    std::vector<float> params;
};

// `{Kind, p0, p1, ...}`
void texture_block(std::string_view line, size_t& p, Component& c)
{
    ++p;  // '{'
    if (!material_id(line, p, c.kind)) {
        material_fail(line, "texture block expected kind");
    }
    skip_spaces(line, p);
    if (p >= line.size() || line[p] != ',') {
        material_fail(line, "texture block expected ',' after kind");
    }
    ++p;
    while (true) {
        skip_spaces(line, p);
        if (p < line.size() && line[p] == '}') {
            ++p;
            break;
        }
        float v = 0.0f;
        if (!parse_float_at(line, p, v)) {
            material_fail(line, "texture block expected number");
        }
        c.params.push_back(v);
        skip_spaces(line, p);
        if (p < line.size() && line[p] == ',') {
            ++p;
            continue;
        }
        if (p < line.size() && line[p] == '}') {
            ++p;
            break;
        }
        material_fail(line, "texture block expected ',' or '}'");
    }
    if (c.kind == "Grid") {
        if (c.params.size() < 6 || c.params.size() > 10) {
            material_fail(line, "Grid texture requires 6 to 10 params (ar, ag, ab, br, bg, bb, [freq or freqU], "
                                "[freqV], [lineThickness])");
        }
    } else if (c.kind == "Stripe") {
        if (c.params.size() < 2 || c.params.size() > 4) {
            material_fail(line, "Stripe texture requires 2 to 4 params (frequency, lineThickness, [onValue], "
                                "[offValue])");
        }
    } else {
        material_fail(line, "unknown texture kind");
    }
    c.texture = true;
}

float clamp01(float v) { return std::min(std::max(v, 0.0f), 1.0f); }

void set_value(Channel& ch, float v)
{
    ch.texture = false;
    ch.value = v;
}

// Channel values of a component list. All scalars: r, g, b then the 12 scalar channels by
// position with their clamps; otherwise the albedo is a texture or three floats, and the next
// eleven components fill roughness .. scatterRadiusB (the reference's list stops before specular).
void assign(std::string_view line, MaterialDef& d, const std::vector<Component>& comps)
{
    Channel* slots[12] = {&d.roughness, &d.metallic, &d.transmission, &d.thin, &d.ior, &d.subsurface, &d.emission,
                          &d.diffuseRoughness, &d.scatterR, &d.scatterG, &d.scatterB, &d.specular};
    // slot clamps: ior >= 1e-3, emission and scatter radii >= 0, the rest to [0, 1]
    auto clamp_slot = [](int slot, float v) {
        if (slot == 4) {
            return std::max(1.0e-3f, v);
        }
        if (slot == 6 || (slot >= 8 && slot <= 10)) {
            return std::max(0.0f, v);
        }
        return clamp01(v);
    };
    const float defaults[12] = {0.5f, 0.0f, 0.0f, 0.0f, 1.5f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    d.albedo = Channel{};
    for (int k = 0; k < 12; ++k) {
        set_value(*slots[k], defaults[k]);
    }
    const bool scalars = std::none_of(comps.begin(), comps.end(), [](const Component& c) { return c.texture; });
    if (scalars) {
        if (comps.size() >= 3) {
            d.albedo.r = clamp01(comps[0].value);
            d.albedo.g = clamp01(comps[1].value);
            d.albedo.b = clamp01(comps[2].value);
        }
        for (int k = 0; k < 12 && 3 + static_cast<size_t>(k) < comps.size(); ++k) {
            set_value(*slots[k], clamp_slot(k, comps[3 + k].value));
        }
        // diffuseRoughness: -1 (take the roughness) unless given
        if (comps.size() <= 10) {
            set_value(d.diffuseRoughness, -1.0f);
        }
        return;
    }
    size_t i = 0;
    if (comps[0].texture) {
        d.albedo.texture = true;
        d.albedo.textureKind = comps[0].kind;
        d.albedo.textureParams = comps[0].params;
        i = 1;
    } else {
        if (comps.size() < 3) {
            material_fail(line, "material albedo requires 3 floats or a texture block");
        }
        d.albedo.r = clamp01(comps[0].value);
        d.albedo.g = clamp01(comps[1].value);
        d.albedo.b = clamp01(comps[2].value);
        i = 3;
    }
    for (int k = 0; k < 11 && i < comps.size(); ++k, ++i) {
        if (comps[i].texture) {
            slots[k]->texture = true;
            slots[k]->textureKind = comps[i].kind;
            slots[k]->textureParams = comps[i].params;
        } else {
            set_value(*slots[k], clamp_slot(k, comps[i].value));
        }
    }
}

void parse_materials(const std::vector<std::string_view>& lines, std::vector<MaterialDef>& defs)
{
    for (std::string_view raw : lines) {
        const std::string_view t = trim(before_hash(raw));
        if (!t.empty() && is_material_line(t)) {
            (void)parse_material_line(std::string(t), defs);
        }
    }
}

// ---- evaluation ---------------------------------------------------------------------------------

constexpr double kEps = 1e-9;
bool near(double a, double b) { return std::fabs(a - b) <= kEps; }

using Bindings = std::vector<std::pair<std::string, double>>;

const double* lookup(const Bindings& b, const std::string& k)
{
    for (const auto& e : b) {
        if (e.first == k) {
            return &e.second;
        }
    }
    return nullptr;
}

bool eval(const Expr* e, const Bindings& b, const std::unordered_map<std::string, double>& consts, double& out)
{
    if (e == nullptr) {
        return false;
    }
    switch (e->op) {
    case Expr::Num:
        out = e->num;
        return true;
    case Expr::Ident: {
        if (const double* v = lookup(b, e->ident)) {
            out = *v;
            return true;
        }
        const auto it = consts.find(e->ident);
        if (it == consts.end()) {
            return false;
        }
        out = it->second;
        return true;
    }
    default:
        break;
    }
    double l = 0.0, r = 0.0;
    if (!eval(e->a.get(), b, consts, l) || !eval(e->b.get(), b, consts, r)) {
        return false;
    }
    switch (e->op) {
    case Expr::Add: out = l + r; return true;
    case Expr::Sub: out = l - r; return true;
    case Expr::Mul: out = l * r; return true;
    case Expr::Div:
        if (near(r, 0.0)) {
            return false;
        }
        out = l / r;
        return true;
    case Expr::Lt: out = l < r ? 1.0 : 0.0; return true;
    case Expr::Le: out = l <= r ? 1.0 : 0.0; return true;
    case Expr::Gt: out = l > r ? 1.0 : 0.0; return true;
    case Expr::Ge: out = l >= r ? 1.0 : 0.0; return true;
    case Expr::Eq: out = near(l, r) ? 1.0 : 0.0; return true;
    default: return false;
    }
}

// Binds a pattern argument to an actual value: an identifier binds (or must repeat its binding,
// constants included, which bind before any module); other patterns must evaluate to the value.
bool bind(const Expr* pat, double actual, Bindings& b, const std::unordered_map<std::string, double>& consts)
{
    if (pat == nullptr) {
        return false;
    }
    if (pat->op == Expr::Ident) {
        if (const double* v = lookup(b, pat->ident)) {
            return near(*v, actual);
        }
        const auto it = consts.find(pat->ident);
        if (it != consts.end()) {
            return near(it->second, actual);
        }
        b.emplace_back(pat->ident, actual);
        return true;
    }
    double want = 0.0;
    return eval(pat, b, consts, want) && near(want, actual);
}

bool match(const Module& pat, const Module& act, Bindings& b, const std::unordered_map<std::string, double>& consts)
{
    if (pat.subsystemRef != act.subsystemRef || pat.name != act.name) {
        return false;
    }
    if (pat.subsystemRef) {
        return true;
    }
    if (pat.args.size() != act.args.size()) {
        return false;
    }
    for (size_t k = 0; k < pat.args.size(); ++k) {
        double v = 0.0;
        if (!eval(act.args[k].get(), b, consts, v) || !bind(pat.args[k].get(), v, b, consts)) {
            return false;
        }
    }
    return true;
}

bool rule_matches(const Rule& r, const std::vector<Module>& gen, size_t i,
                  const std::unordered_map<std::string, double>& consts, Bindings& b)
{
    b.clear();
    if (!r.left.empty()) {
        if (r.left.size() > i) {
            return false;
        }
        const size_t s = i - r.left.size();
        for (size_t k = 0; k < r.left.size(); ++k) {
            if (!match(r.left[k], gen[s + k], b, consts)) {
                return false;
            }
        }
    }
    if (!match(r.pred, gen[i], b, consts)) {
        return false;
    }
    if (!r.right.empty()) {
        if (i + 1 + r.right.size() > gen.size()) {
            return false;
        }
        for (size_t k = 0; k < r.right.size(); ++k) {
            if (!match(r.right[k], gen[i + 1 + k], b, consts)) {
                return false;
            }
        }
    }
    if (r.condition) {
        double c = 0.0;
        if (!eval(r.condition.get(), b, consts, c) || near(c, 0.0)) {
            return false;
        }
    }
    return true;
}

ExprP number(double v)
{
    auto e = std::make_shared<Expr>();
    e->op = Expr::Num;
    e->num = v;
    return e;
}

Module instantiate(const Module& tpl, const Bindings& b, const std::unordered_map<std::string, double>& consts)
{
    Module out;
    out.subsystemRef = tpl.subsystemRef;
    out.name = tpl.name;
    out.args.reserve(tpl.args.size());
    for (const ExprP& a : tpl.args) {
        double v = 0.0;
        if (eval(a.get(), b, consts, v)) {
            out.args.push_back(number(v));
        } else if (tpl.name == "Mat" && a && a->op == Expr::Ident) {
            out.args.push_back(a);   // a named material id stays a name
        } else {
            throw std::runtime_error("unable to evaluate successor argument expression");
        }
    }
    return out;
}

std::vector<Module> derive(const std::vector<Module>& gen, const System& sys, std::mt19937_64& rng,
                           const std::unordered_map<std::string, std::vector<size_t>>& byName)
{
    std::vector<Module> next;
    next.reserve(gen.size() * 2);
    std::vector<size_t> hits;
    std::vector<Bindings> hitBindings;
    Bindings b;
    for (size_t i = 0; i < gen.size(); ++i) {
        hits.clear();
        hitBindings.clear();
        const Module& m = gen[i];
        const auto it = byName.find(m.subsystemRef ? "ref." + m.name : m.name);
        if (it != byName.end()) {
            for (size_t ri : it->second) {
                if (rule_matches(sys.rules[ri], gen, i, sys.constants, b)) {
                    hits.push_back(ri);
                    hitBindings.push_back(b);
                }
            }
        }
        if (hits.empty()) {
            next.push_back(m);
            continue;
        }
        // weighted choice among the matches (non-positive or non-finite weights count 0); one
        // draw per rewritten module whenever some weight is positive, else the first match
        double total = 0.0;
        for (size_t ri : hits) {
            const double p = sys.rules[ri].probability;
            total += (std::isfinite(p) && p > 0.0) ? p : 0.0;
        }
        size_t pick = 0;
        if (total > 0.0) {
            // u = x / 2^64 rounded to double, kept below 1, then scaled by the total weight: the
            // uniform_real_distribution<double>(0, total) of the GNU C++ library on one 64-bit
            // draw (the reference's own build with MSVC may map the draw differently)
            double u = static_cast<double>(rng()) * 0x1p-64;
            if (u >= 1.0) {
                u = std::nextafter(1.0, 0.0);
            }
            const double sel = u * total + 0.0;
            double run = 0.0;
            pick = hits.size() - 1;
            for (size_t k = 0; k < hits.size(); ++k) {
                const double p = sys.rules[hits[k]].probability;
                run += (std::isfinite(p) && p > 0.0) ? p : 0.0;
                if (sel <= run) {
                    pick = k;
                    break;
                }
            }
        }
        for (const Module& s : sys.rules[hits[pick]].successor) {
            next.push_back(instantiate(s, hitBindings[pick], sys.constants));
        }
    }
    return next;
}

void append_number(std::string& s, double v)
{
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    s += buf;
}

} // namespace

// ---- public --------------------------------------------------------------------------------------

std::vector<Module> parse_modules(const std::string& text, bool extended)
{
    return Parser(text, extended).modules();
}

Rule parse_rule(const std::string& text, bool extended) { return Parser(text, extended).rule(); }

std::vector<Rule> parse_rules(const std::string& block, bool extended)
{
    return parse_rules_block(split_lines(block), extended);
}

namespace {

void dump_expr(std::string& s, const Expr* e)
{
    static const char* const ops[] = {"", "", "+", "-", "*", "/", "<", "<=", ">", ">=", "=="};
    if (e == nullptr) {
        s += "null";
    } else if (e->op == Expr::Num) {
        append_number(s, e->num);
    } else if (e->op == Expr::Ident) {
        s += e->ident;
    } else {
        s += "(";
        s += ops[e->op];
        s.push_back(' ');
        dump_expr(s, e->a.get());
        s.push_back(' ');
        dump_expr(s, e->b.get());
        s += ")";
    }
}

void dump_module(std::string& s, const Module& m)
{
    if (m.subsystemRef) {
        s += "ref:" + m.name;
        return;
    }
    s += m.name;
    if (!m.args.empty()) {
        s.push_back('(');
        for (size_t i = 0; i < m.args.size(); ++i) {
            if (i != 0) {
                s.push_back(',');
            }
            dump_expr(s, m.args[i].get());
        }
        s.push_back(')');
    }
}

void dump_list(std::string& s, const std::vector<Module>& ms)
{
    s.push_back('[');
    for (size_t i = 0; i < ms.size(); ++i) {
        if (i != 0) {
            s.push_back(' ');
        }
        dump_module(s, ms[i]);
    }
    s.push_back(']');
}

void dump_rule(std::string& s, const Rule& r)
{
    s += "left=";
    dump_list(s, r.left);
    s += " pred=";
    dump_module(s, r.pred);
    s += " right=";
    dump_list(s, r.right);
    s += " cond=";
    dump_expr(s, r.condition.get());
    s += " prob=";
    append_number(s, r.probability);
    s += " succ=";
    dump_list(s, r.successor);
}

} // namespace

std::string dump_modules(const std::vector<Module>& ms)
{
    std::string s;
    dump_list(s, ms);
    return s;
}

std::string dump_rules(const std::vector<Rule>& rules)
{
    std::string s;
    for (size_t i = 0; i < rules.size(); ++i) {
        if (i != 0) {
            s.push_back('\n');
        }
        dump_rule(s, rules[i]);
    }
    return s;
}

std::string dump_system(const System& sys)
{
    std::string s = "axiom=" + dump_modules(sys.axiom);
    for (const Rule& r : sys.rules) {
        s += "\nrule ";
        dump_rule(s, r);
    }
    std::vector<std::string> names;
    for (const auto& kv : sys.constants) {
        names.push_back(kv.first);
    }
    std::sort(names.begin(), names.end());
    for (const std::string& n : names) {
        s += "\nconst " + n + "=";
        append_number(s, sys.constants.at(n));
    }
    for (const MaterialDef& m : sys.materials) {
        s += "\nmat " + m.id;
    }
    return s;
}

bool parse_material_line(const std::string& trimmedIn, std::vector<MaterialDef>& defs)
{
    const std::string_view line(trimmedIn);
    if (line.empty()) {
        return false;
    }
    size_t p = 0;
    std::string id;
    if (!mat_head(line, p, id)) {
        return false;
    }
    skip_spaces(line, p);
    if (p >= line.size() || line[p] != '=') {
        return false;
    }
    ++p;
    skip_spaces(line, p);
    if (p >= line.size() || line[p] != '{') {
        material_fail(line, "material declaration expected '{' (r, g, b, [roughness], [metallic], [transmission], "
                            "[thin], [ior], [subsurface], [emission], [diffuseRoughness], [scatterRadiusR], "
                            "[scatterRadiusG], [scatterRadiusB], [specular])");
    }
    ++p;
    std::vector<Component> comps;
    while (true) {
        skip_spaces(line, p);
        if (p < line.size() && line[p] == '}') {
            ++p;
            break;
        }
        Component c;
        size_t q = p + 1;
        skip_spaces(line, q);
        if (p < line.size() && line[p] == '{' && q < line.size() &&
            (std::isalpha(static_cast<unsigned char>(line[q])) || line[q] == '_')) {
            texture_block(line, p, c);
        } else if (!parse_float_at(line, p, c.value)) {
            material_fail(line, "material declaration expected component");
        }
        comps.push_back(std::move(c));
        skip_spaces(line, p);
        if (p < line.size() && line[p] == ',') {
            ++p;
            continue;
        }
        if (p < line.size() && line[p] == '}') {
            ++p;
            break;
        }
        material_fail(line, "material declaration expected ',' or '}'");
    }
    skip_spaces(line, p);
    if (p != line.size()) {
        material_fail(line, "unexpected text after material declaration");
    }
    const bool scalars = std::none_of(comps.begin(), comps.end(), [](const Component& c) { return c.texture; });
    if (comps.empty() || comps.size() > 15) {
        material_fail(line, "material requires 1 to 15 components (albedo texture or r,g,b, then optional channel "
                            "values)");
    }
    if (scalars && comps.size() < 3) {
        material_fail(line, "material requires 3 to 15 scalar components (r, g, b, [roughness], [metallic], "
                            "[transmission], [thin], [ior], [subsurface], [emission], [diffuseRoughness], "
                            "[scatterRadiusR], [scatterRadiusG], [scatterRadiusB], [specular])");
    }
    if (!scalars && !comps[0].texture && comps.size() < 3) {
        material_fail(line, "material albedo requires 3 floats or a texture block");
    }
    for (const MaterialDef& d : defs) {
        if (d.id == id) {
            material_fail(line, "duplicate material id");
        }
    }
    MaterialDef d;
    d.id = id;
    assign(line, d, comps);
    defs.push_back(std::move(d));
    return true;
}

System parse(const std::string& definition, bool extended)
{
    System sys;
    const std::vector<std::string_view> lines = split_lines(definition);
    parse_materials(lines, sys.materials);
    for (std::string_view raw : lines) {
        const std::string_view t = trim(before_hash(raw));
        std::string name;
        double v = 0.0;
        if (constant_line(t, name, v)) {
            if (sys.constants.count(name) != 0) {
                throw std::runtime_error("duplicate constant name: " + std::string(t));
            }
            sys.constants.emplace(name, v);
        }
    }
    size_t firstRule = lines.size();
    for (size_t i = 0; i < lines.size(); ++i) {
        if (trim(before_hash(lines[i])).find("->") != std::string_view::npos) {
            firstRule = i;
            break;
        }
    }
    std::string axiom;
    for (size_t i = 0; i < firstRule; ++i) {
        const std::string_view t = trim(before_hash(lines[i]));
        std::string name;
        double v = 0.0;
        if (t.empty() || is_material_line(t) || constant_line(t, name, v)) {
            continue;
        }
        if (!axiom.empty()) {
            axiom.push_back(' ');
        }
        axiom.append(t.data(), t.size());
    }
    sys.axiom = parse_modules(axiom, extended);
    if (firstRule < lines.size()) {
        sys.rules = parse_rules_block(std::vector<std::string_view>(lines.begin() + static_cast<long>(firstRule),
                                                                    lines.end()),
                                      extended);
    }
    return sys;
}

std::vector<std::vector<Module>> evaluate_steps(const System& sys, std::size_t iterations, uint64_t seed)
{
    std::unordered_map<std::string, std::vector<size_t>> byName;
    for (size_t r = 0; r < sys.rules.size(); ++r) {
        const Module& p = sys.rules[r].pred;
        byName[p.subsystemRef ? "ref." + p.name : p.name].push_back(r);
    }
    std::mt19937_64 rng(seed);
    std::vector<std::vector<Module>> steps;
    steps.reserve(iterations + 1);
    steps.push_back(sys.axiom);
    for (std::size_t it = 0; it < iterations; ++it) {
        steps.push_back(derive(steps.back(), sys, rng, byName));
    }
    return steps;
}

std::vector<Module> evaluate(const System& sys, std::size_t iterations, uint64_t seed)
{
    std::unordered_map<std::string, std::vector<size_t>> byName;
    for (size_t r = 0; r < sys.rules.size(); ++r) {
        const Module& p = sys.rules[r].pred;
        byName[p.subsystemRef ? "ref." + p.name : p.name].push_back(r);
    }
    std::mt19937_64 rng(seed);
    std::vector<Module> gen = sys.axiom;
    for (std::size_t it = 0; it < iterations; ++it) {
        gen = derive(gen, sys, rng, byName);
    }
    return gen;
}

bool eval_number(const ExprP& e, const std::unordered_map<std::string, double>& constants, double& out)
{
    return eval(e.get(), Bindings{}, constants, out);
}

std::string to_text(const std::vector<Module>& gen)
{
    std::string s;
    for (size_t i = 0; i < gen.size(); ++i) {
        const Module& m = gen[i];
        if (i != 0) {
            s.push_back(' ');
        }
        if (m.subsystemRef) {
            s += "ref.";
        }
        s += m.name;
        if (!m.args.empty()) {
            s.push_back('(');
            for (size_t k = 0; k < m.args.size(); ++k) {
                if (k != 0) {
                    s.push_back(',');
                }
                const Expr* e = m.args[k].get();
                if (e != nullptr && e->op == Expr::Num) {
                    append_number(s, e->num);
                } else if (e != nullptr && e->op == Expr::Ident) {
                    s += e->ident;
                } else {
                    s += "?";
                }
            }
            s.push_back(')');
        }
    }
    return s;
}

} // namespace lsys
