// Model readers (API: include/sovereign/io.hpp).
//
// CPLEX LP subset (the forms the reference's LP reader documents,
// src/io/lp_reader.cpp:7-21): "\" comments; Maximize/Minimize with an
// optional objective name and a constant; Subject To / such that / st with
// optional row names and the two-sided form lo <= expr <= hi; Bounds
// (x >= v, x <= v, lo <= x <= hi, x = v, x free, free x, -inf <= x);
// General/Integer; Binary; End.  Anything else is rejected with an Error.
//
// MPS: free-format tokens (names without blanks), sections NAME, OBJSENSE,
// ROWS, COLUMNS (with MARKER INTORG/INTEND), RHS, RANGES, BOUNDS (UP LO FX FR
// MI PL BV LI UI), ENDATA; rules below, at the MPS reader.  Gzip input is not
// supported.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <string_view>
#include <unordered_set>

#include "sovereign/io.hpp"

namespace sov {
namespace {

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

std::string trim(const std::string& s) {
    std::size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

// ----------------------------------------------------------------- LP ----

enum class Tk { Num, Id, Op, Sign, Colon, End };

struct Token {
    Tk kind = Tk::End;
    std::string text;
    double num = 0.0;
    int dir = 0;   // Op: -1 '<=', +1 '>=', 0 '='
};

bool name_char(char c) {
    return std::isalnum(static_cast<unsigned char>(c)) || std::string("_.[]!\"#$%&()/,;?@`'{}|~^").find(c) != std::string::npos;
}

std::vector<Token> tokenize(const std::string& s, const std::string& origin) {
    std::vector<Token> out;
    std::size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (std::isspace(static_cast<unsigned char>(c))) { ++i; continue; }
        Token t;
        if (c == '<' || c == '>' || c == '=') {
            std::size_t j = i + 1;
            if (j < s.size() && (s[j] == '=' || s[j] == '<' || s[j] == '>')) ++j;
            const std::string op = s.substr(i, j - i);
            t.kind = Tk::Op;
            t.text = op;
            t.dir = (op.find('<') != std::string::npos) ? -1 : (op.find('>') != std::string::npos) ? 1 : 0;
            i = j;
        } else if (c == '+' || c == '-') {
            t.kind = Tk::Sign;
            t.text = std::string(1, c);
            ++i;
        } else if (c == ':') {
            t.kind = Tk::Colon;
            ++i;
        } else if (std::isdigit(static_cast<unsigned char>(c)) || (c == '.' && i + 1 < s.size() && std::isdigit(static_cast<unsigned char>(s[i + 1])))) {
            std::size_t j = i;
            while (j < s.size() && (std::isdigit(static_cast<unsigned char>(s[j])) || s[j] == '.')) ++j;
            if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
                std::size_t k = j + 1;
                if (k < s.size() && (s[k] == '+' || s[k] == '-')) ++k;
                if (k < s.size() && std::isdigit(static_cast<unsigned char>(s[k]))) {
                    j = k;
                    while (j < s.size() && std::isdigit(static_cast<unsigned char>(s[j]))) ++j;
                }
            }
            t.kind = Tk::Num;
            t.text = s.substr(i, j - i);
            t.num = std::strtod(t.text.c_str(), nullptr);
            i = j;
        } else if (name_char(c)) {
            std::size_t j = i;
            while (j < s.size() && name_char(s[j])) ++j;
            t.kind = Tk::Id;
            t.text = s.substr(i, j - i);
            const std::string l = lower(t.text);
            if (l == "inf" || l == "infinity") {
                t.kind = Tk::Num;
                t.num = kInfinity;
            }
            i = j;
        } else {
            throw Error(origin + ": unexpected character '" + std::string(1, c) + "'");
        }
        out.push_back(t);
    }
    out.push_back(Token{});
    return out;
}

class LpParser {
public:
    explicit LpParser(std::string origin) : origin_(std::move(origin)) {}

    ReadResult parse(const std::string& text) {
        enum class Sec { None, Obj, Cons, Bounds, Gen, Bin, Done };
        Sec sec = Sec::None;
        std::string obj, cons;
        std::vector<std::string> bounds;
        std::string gen, bin;
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            const auto bs = line.find('\\');
            if (bs != std::string::npos) line = line.substr(0, bs);
            const std::string l = lower(trim(line));
            if (l.empty()) continue;
            if (l == "maximize" || l == "maximise" || l == "max" || l == "maximum") {
                b_.set_sense(ObjSense::Maximize);
                sec = Sec::Obj;
                continue;
            }
            if (l == "minimize" || l == "minimise" || l == "min" || l == "minimum") {
                b_.set_sense(ObjSense::Minimize);
                sec = Sec::Obj;
                continue;
            }
            if (l == "subject to" || l == "such that" || l == "st" || l == "s.t." || l == "st." || l == "subjectto") {
                sec = Sec::Cons;
                continue;
            }
            if (l == "bounds" || l == "bound") { sec = Sec::Bounds; continue; }
            if (l == "general" || l == "generals" || l == "gen" || l == "integer" || l == "integers") { sec = Sec::Gen; continue; }
            if (l == "binary" || l == "binaries" || l == "bin") { sec = Sec::Bin; continue; }
            if (l == "end") { sec = Sec::Done; break; }
            switch (sec) {
                case Sec::Obj: obj += line + "\n"; break;
                case Sec::Cons: cons += line + "\n"; break;
                case Sec::Bounds: bounds.push_back(line); break;
                case Sec::Gen: gen += line + "\n"; break;
                case Sec::Bin: bin += line + "\n"; break;
                default: throw Error(origin_ + ": text outside any section: '" + trim(line) + "'");
            }
        }
        parse_objective(obj);
        parse_constraints(cons);
        for (const auto& bl : bounds) parse_bound(bl);
        declare(gen, false);
        declare(bin, true);
        ReadResult r;
        r.problem = b_.finish();
        r.format = "lp";
        std::string why;
        if (r.problem.validate(&why) != Status::Ok) throw Error(origin_ + ": invalid model: " + why);
        return r;
    }

private:
    Idx column(const std::string& name) {
        const Idx c = b_.find_column(name);
        return c >= 0 ? c : b_.add_column(name);
    }

    // expression: sequence of [sign] [num] [id] terms; stops at Op / End /
    // a "name :" start.  Returns the constant part.
    double expression(const std::vector<Token>& t, std::size_t& p, std::vector<std::pair<Idx, double>>& terms) {
        double constant = 0.0;
        bool any = false;
        while (t[p].kind == Tk::Sign || t[p].kind == Tk::Num || t[p].kind == Tk::Id) {
            if (t[p].kind == Tk::Id && t[p + 1].kind == Tk::Colon) break;   // next statement
            if (any && t[p].kind != Tk::Sign) break;   // two terms without an operator: a new statement
            double sgn = 1.0;
            while (t[p].kind == Tk::Sign) {
                if (t[p].text == "-") sgn = -sgn;
                ++p;
            }
            double coef = 1.0;
            bool has_num = false;
            if (t[p].kind == Tk::Num) {
                coef = t[p].num;
                has_num = true;
                ++p;
            }
            if (t[p].kind == Tk::Id && !(t[p + 1].kind == Tk::Colon)) {
                terms.emplace_back(column(t[p].text), sgn * coef);
                ++p;
            } else if (has_num) {
                constant += sgn * coef;
            } else {
                throw Error(origin_ + ": malformed expression");
            }
            any = true;
        }
        return constant;
    }

    void parse_objective(const std::string& s) {
        const auto t = tokenize(s, origin_);
        std::size_t p = 0;
        if (t[p].kind == Tk::Id && t[p + 1].kind == Tk::Colon) {
            b_.set_objective_name(t[p].text);
            p += 2;
        }
        std::vector<std::pair<Idx, double>> terms;
        const double k = expression(t, p, terms);
        if (t[p].kind != Tk::End) throw Error(origin_ + ": unexpected token in objective");
        std::map<Idx, double> acc;
        for (auto& [c, v] : terms) acc[c] += v;
        for (auto& [c, v] : acc) b_.set_objective_coefficient(c, v);
        b_.set_objective_constant(k);
    }

    void parse_constraints(const std::string& s) {
        const auto t = tokenize(s, origin_);
        std::size_t p = 0;
        int auto_id = 0;
        while (t[p].kind != Tk::End) {
            std::string name;
            if (t[p].kind == Tk::Id && t[p + 1].kind == Tk::Colon) {
                name = t[p].text;
                p += 2;
            } else {
                name = "R" + std::to_string(++auto_id);
            }
            // two-sided form: [sign] num op expr op [sign] num
            double lo = -kInfinity, hi = kInfinity;
            std::size_t q = p;
            double s1 = 1.0;
            while (t[q].kind == Tk::Sign) {
                if (t[q].text == "-") s1 = -s1;
                ++q;
            }
            bool two_sided = false;
            std::vector<std::pair<Idx, double>> terms;
            if (t[q].kind == Tk::Num && t[q + 1].kind == Tk::Op) {
                const double first = s1 * t[q].num;
                const int d1 = t[q + 1].dir;
                p = q + 2;
                const double k = expression(t, p, terms);
                if (t[p].kind != Tk::Op) throw Error(origin_ + ": incomplete two-sided constraint " + name);
                const int d2 = t[p].dir;
                ++p;
                const double second = rhs(t, p) - k;
                const double f = first - k;
                if (d1 < 0 && d2 < 0) { lo = f; hi = second; }
                else if (d1 > 0 && d2 > 0) { lo = second; hi = f; }
                else throw Error(origin_ + ": inconsistent two-sided constraint " + name);
                two_sided = true;
            }
            if (!two_sided) {
                const double k = expression(t, p, terms);
                if (t[p].kind != Tk::Op) throw Error(origin_ + ": constraint " + name + " has no relation");
                const int d = t[p].dir;
                ++p;
                const double r = rhs(t, p) - k;
                if (d < 0) hi = r;
                else if (d > 0) lo = r;
                else lo = hi = r;
            }
            const Idx row = b_.add_row(name, lo, hi);
            for (auto& [c, v] : terms) b_.add_coefficient(row, c, v);
        }
    }

    double rhs(const std::vector<Token>& t, std::size_t& p) {
        double sg = 1.0;
        while (t[p].kind == Tk::Sign) {
            if (t[p].text == "-") sg = -sg;
            ++p;
        }
        if (t[p].kind != Tk::Num) throw Error(origin_ + ": expected a number");
        return sg * t[p++].num;
    }

    void parse_bound(const std::string& line) {
        const auto t = tokenize(line, origin_);
        const std::string l0 = t[0].kind == Tk::Id ? lower(t[0].text) : "";
        if (t.size() == 3 && t[0].kind == Tk::Id && lower(t[1].text) == "free") {
            const Idx c = column(t[0].text);
            b_.set_column_bounds(c, -kInfinity, kInfinity);
            return;
        }
        if (t.size() == 3 && l0 == "free" && t[1].kind == Tk::Id) {
            b_.set_column_bounds(column(t[1].text), -kInfinity, kInfinity);
            return;
        }
        std::size_t p = 0;
        auto num = [&](double& out) {
            double sg = 1.0;
            std::size_t q = p;
            while (t[q].kind == Tk::Sign) {
                if (t[q].text == "-") sg = -sg;
                ++q;
            }
            if (t[q].kind != Tk::Num) return false;
            out = sg * t[q].num;
            p = q + 1;
            return true;
        };
        double a = 0.0;
        if (num(a)) {   // a op x [op b]
            if (t[p].kind != Tk::Op || t[p + 1].kind != Tk::Id) throw Error(origin_ + ": malformed bound '" + line + "'");
            const int d1 = t[p].dir;
            const Idx c = column(t[p + 1].text);
            p += 2;
            apply(c, a, -d1);   // a <= x  is  x >= a
            if (t[p].kind == Tk::Op) {
                const int d2 = t[p].dir;
                ++p;
                double b = 0.0;
                if (!num(b)) throw Error(origin_ + ": malformed bound '" + line + "'");
                apply(c, b, d2);
            }
        } else if (t[0].kind == Tk::Id && t[1].kind == Tk::Op) {   // x op b
            const Idx c = column(t[0].text);
            const int d = t[1].dir;
            p = 2;
            double b = 0.0;
            if (!num(b)) throw Error(origin_ + ": malformed bound '" + line + "'");
            apply(c, b, d);
        } else {
            throw Error(origin_ + ": malformed bound '" + line + "'");
        }
        if (t[p].kind != Tk::End) throw Error(origin_ + ": trailing text in bound '" + line + "'");
    }

    // x (dir) v : dir -1 means x <= v, +1 x >= v, 0 x = v
    void apply(Idx c, double v, int dir) {
        if (dir < 0) b_.set_column_bounds(c, b_.column_lower(c), v);
        else if (dir > 0) b_.set_column_bounds(c, v, b_.column_upper(c));
        else b_.set_column_bounds(c, v, v);
    }

    void declare(const std::string& s, bool binary) {
        for (const Token& t : tokenize(s, origin_)) {
            if (t.kind == Tk::End) break;
            if (t.kind != Tk::Id) throw Error(origin_ + ": expected a variable name in an integer section");
            const Idx c = column(t.text);
            b_.set_column_kind(c, VarKind::Integer);
            if (binary) b_.set_column_bounds(c, 0.0, 1.0);
        }
    }

    std::string origin_;
    Problem::Builder b_;
};

std::string slurp(const std::string& path) {
    if (path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0)
        throw Error(path + ": gzip-compressed models are not supported by this build");
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open " + path);
    std::ostringstream os;
    os << f.rdbuf();
    return os.str();
}

}  // namespace

ReadResult read_lp_string(const std::string& text, const std::string& origin) { return LpParser(origin).parse(text); }

ReadResult read_lp(const std::string& path) { return read_lp_string(slurp(path), path); }

// ----------------------------------------------------------------- MPS ---
//
// One pass over the text, tokens as views into it (no per-line streams), so
// a million-column file reads at memory speed.  Semantics follow the
// reference reader (src/io/mps_reader.cpp): numbers must parse completely
// ("expected a number"); bound values accept INF / INFINITY / |v| >= 1e30 as
// infinite; a valued bound card without a value is an error; the column of a
// BOUNDS card is its last token (valueless types) or the one before the
// value, and an unknown one is added with a warning; only UP moves a default
// lower bound of 0 to -inf when its value is negative (MpsOptions); a second
// N row is kept as a free row (keep_free_rows) or dropped together with every
// coefficient, right-hand side and range naming it; an RHS entry on the
// objective row sets the objective constant to minus its value; RANGES
// widens L and G rows by |R| and E rows by the signed R.

namespace {

using Tokens = std::vector<std::string_view>;

bool iequal(std::string_view a, std::string_view b) {
    if (a.size() != b.size()) return false;
    for (std::size_t i = 0; i < a.size(); ++i)
        if (std::toupper(static_cast<unsigned char>(a[i])) != std::toupper(static_cast<unsigned char>(b[i])))
            return false;
    return true;
}

/// A complete decimal number, or nothing.
bool to_number(std::string_view t, double* out) {
    if (t.empty()) return false;
    const char* b = t.data();
    const char* e = b + t.size();
    if (*b == '+') ++b;   // from_chars takes no leading '+'
    const auto r = std::from_chars(b, e, *out);
    return r.ec == std::errc() && r.ptr == e;
}

class MpsText {
public:
    MpsText(const std::string& text, const MpsOptions& opt, const std::string& origin)
        : text_(text), opt_(opt), origin_(origin) {}

    ReadResult read() {
        std::size_t pos = 0;
        Tokens tok;
        while (pos < text_.size()) {
            std::size_t eol = text_.find('\n', pos);
            if (eol == std::string::npos) eol = text_.size();
            const std::string_view line(text_.data() + pos, eol - pos);
            pos = eol + 1;
            ++line_;
            if (line.empty() || line[0] == '*') continue;
            split(line, tok);
            if (tok.empty()) continue;
            if (!std::isspace(static_cast<unsigned char>(line[0]))) {
                if (!header(tok)) break;   // ENDATA
                continue;
            }
            card(tok);
        }
        res_.problem = b_.finish();
        res_.format = "mps";
        // a model that cannot be solved is not handed on (the reference's
        // readers validate too)
        std::string why;
        if (res_.problem.validate(&why) != Status::Ok) throw Error(origin_ + ": invalid model: " + why);
        return std::move(res_);
    }

private:
    enum class Sec { None, Name, ObjSense, Rows, Columns, Rhs, Ranges, Bounds };

    [[noreturn]] void fail(const std::string& msg) const {
        throw Error(origin_ + ":" + std::to_string(line_) + ": " + msg);
    }
    void warn(const char* section, std::string msg) { res_.warnings.push_back({line_, section, std::move(msg)}); }

    static void split(std::string_view line, Tokens& tok) {
        tok.clear();
        std::size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
            const std::size_t s = i;
            while (i < line.size() && !std::isspace(static_cast<unsigned char>(line[i]))) ++i;
            if (i > s) tok.push_back(line.substr(s, i - s));
        }
    }

    double number(std::string_view t) const {
        double v = 0.0;
        if (!to_number(t, &v)) fail("expected a number, got '" + std::string(t) + "'");
        return v;
    }

    double bound_number(std::string_view t) const {
        std::string_view u = t;
        bool negative = false;
        if (!u.empty() && (u[0] == '+' || u[0] == '-')) {
            negative = u[0] == '-';
            u.remove_prefix(1);
        }
        if (iequal(u, "INF") || iequal(u, "INFINITY")) return negative ? -kInfinity : kInfinity;
        const double v = number(t);
        if (std::abs(v) >= 1e30) return v > 0 ? kInfinity : -kInfinity;
        return v;
    }

    void objsense(std::string_view t) {
        if (iequal(t, "MAX") || iequal(t, "MAXIMIZE")) b_.set_sense(ObjSense::Maximize);
        else if (iequal(t, "MIN") || iequal(t, "MINIMIZE")) b_.set_sense(ObjSense::Minimize);
        else fail("unknown OBJSENSE '" + std::string(t) + "'");
    }

    /// A section header; false at ENDATA.
    bool header(const Tokens& tok) {
        const std::string_view h = tok[0];
        if (iequal(h, "ENDATA")) return false;
        if (iequal(h, "NAME")) {
            sec_ = Sec::Name;
            b_.set_name(tok.size() > 1 ? std::string(tok[1]) : std::string());
        } else if (iequal(h, "OBJSENSE")) {
            sec_ = Sec::ObjSense;
            if (tok.size() > 1) objsense(tok[1]);
        } else if (iequal(h, "ROWS")) {
            sec_ = Sec::Rows;
        } else if (iequal(h, "COLUMNS")) {
            sec_ = Sec::Columns;
        } else if (iequal(h, "RHS")) {
            sec_ = Sec::Rhs;
        } else if (iequal(h, "RANGES")) {
            sec_ = Sec::Ranges;
        } else if (iequal(h, "BOUNDS")) {
            sec_ = Sec::Bounds;
        } else {
            fail("unknown section '" + std::string(h) + "'");
        }
        return true;
    }

    void card(const Tokens& tok) {
        switch (sec_) {
            case Sec::ObjSense: objsense(tok[0]); break;
            case Sec::Rows: row(tok); break;
            case Sec::Columns: column(tok); break;
            case Sec::Rhs: rhs(tok); break;
            case Sec::Ranges: range(tok); break;
            case Sec::Bounds: bound(tok); break;
            default: fail("data outside a known section");
        }
    }

    void row(const Tokens& tok) {
        if (tok.size() < 2) fail("ROWS card needs a type and a name");
        const std::string name(tok[1]);
        char sense = static_cast<char>(std::toupper(static_cast<unsigned char>(tok[0][0])));
        if (tok[0].size() != 1 || std::string_view("NLGE").find(sense) == std::string_view::npos)
            fail("unknown row type '" + std::string(tok[0]) + "'");
        if (sense == 'N') {
            if (obj_.empty()) {
                obj_ = name;
                b_.set_objective_name(name);
                return;
            }
            if (!opt_.keep_free_rows) {
                dropped_.insert(name);
                return;
            }
        }
        const double lo = sense == 'G' || sense == 'E' ? 0.0 : -kInfinity;
        const double hi = sense == 'L' || sense == 'E' ? 0.0 : kInfinity;
        b_.add_row(name, lo, hi);
        sense_.push_back(sense);
        rhs_.push_back(0.0);
    }

    /// The row a card names, or -1 for a dropped free row.
    Idx row_of(std::string_view name, const char* section) {
        const std::string key(name);
        const Idx r = b_.find_row(key);
        if (r >= 0) return r;
        if (dropped_.count(key)) return -1;
        fail("unknown row '" + key + "' in " + section);
    }

    void column(const Tokens& tok) {
        bool marker = false, start = false, end = false;
        for (std::string_view t : tok) {
            std::string u;
            for (char ch : t)
                if (ch != '\'') u.push_back(static_cast<char>(std::toupper(static_cast<unsigned char>(ch))));
            marker = marker || u == "MARKER";
            start = start || u == "INTORG";
            end = end || u == "INTEND";
        }
        if (marker && (start || end)) {
            integer_ = start;
            return;
        }
        if (tok.size() < 3) fail("COLUMNS card needs a column, a row and a value");
        if (tok[0] != last_col_name_) {   // cards of one column come together
            last_col_name_ = std::string(tok[0]);
            last_col_ = b_.find_column(last_col_name_);
            if (last_col_ < 0) last_col_ = b_.add_column(last_col_name_);
        }
        if (integer_) b_.set_column_kind(last_col_, VarKind::Integer);
        for (std::size_t k = 1; k + 1 < tok.size(); k += 2) {
            const double v = number(tok[k + 1]);
            if (tok[k] == obj_) {
                b_.set_objective_coefficient(last_col_, v);
                continue;
            }
            const Idx r = row_of(tok[k], "COLUMNS");
            if (r >= 0) b_.add_coefficient(r, last_col_, v);
        }
    }

    void rhs(const Tokens& tok) {
        if (tok.size() < 2) fail("RHS card needs a row and a value");
        for (std::size_t k = tok.size() % 2; k + 1 < tok.size(); k += 2) {   // odd count: a set name first
            const double v = number(tok[k + 1]);
            if (tok[k] == obj_) {
                b_.set_objective_constant(-v);
                continue;
            }
            const Idx r = row_of(tok[k], "RHS");
            if (r < 0) continue;
            const auto i = static_cast<std::size_t>(r);
            rhs_[i] = v;
            switch (sense_[i]) {
                case 'L': b_.set_row_bounds(r, -kInfinity, v); break;
                case 'G': b_.set_row_bounds(r, v, kInfinity); break;
                case 'E': b_.set_row_bounds(r, v, v); break;
                default: warn("RHS", "right-hand side of free row '" + std::string(tok[k]) + "' ignored");
            }
        }
    }

    void range(const Tokens& tok) {
        if (tok.size() < 2) fail("RANGES card needs a row and a value");
        for (std::size_t k = tok.size() % 2; k + 1 < tok.size(); k += 2) {
            const Idx r = row_of(tok[k], "RANGES");
            const double R = number(tok[k + 1]);
            if (r < 0) continue;
            const auto i = static_cast<std::size_t>(r);
            const double v = rhs_[i];
            switch (sense_[i]) {
                case 'L': b_.set_row_bounds(r, v - std::abs(R), v); break;
                case 'G': b_.set_row_bounds(r, v, v + std::abs(R)); break;
                case 'E':
                    if (R >= 0) b_.set_row_bounds(r, v, v + R);
                    else b_.set_row_bounds(r, v + R, v);
                    break;
                default: warn("RANGES", "range of free row '" + std::string(tok[k]) + "' ignored");
            }
        }
    }

    void bound(const Tokens& tok) {
        if (tok.size() < 2) fail("BOUNDS card needs a type and a column");
        std::string type;
        for (char ch : tok[0]) type.push_back(static_cast<char>(std::toupper(static_cast<unsigned char>(ch))));
        const bool valueless = type == "FR" || type == "MI" || type == "PL" || type == "BV";
        if (!valueless && tok.size() < 3) fail("bound type " + type + " needs a value");
        const std::string name(tok[valueless ? tok.size() - 1 : tok.size() - 2]);
        Idx c = b_.find_column(name);
        if (c < 0) {
            warn("BOUNDS", "bound on column '" + name + "', which has no coefficients");
            c = b_.add_column(name);
        }
        const double lo = b_.column_lower(c), hi = b_.column_upper(c);
        const double v = valueless ? 0.0 : bound_number(tok.back());
        if (type == "UP") {
            if (v < 0.0 && lo == 0.0 && opt_.negative_upper_bound == NegativeUpperBoundRule::LowerBecomesMinusInfinity) {
                b_.set_column_bounds(c, -kInfinity, v);
                warn("BOUNDS", "negative UP bound on '" + name + "' with the default lower bound 0: lower bound "
                               "set to -inf");
            } else {
                b_.set_column_bounds(c, lo, v);
                if (v < 0.0 && lo == 0.0) warn("BOUNDS", "negative UP bound on '" + name + "' leaves crossed bounds");
            }
        } else if (type == "LO") {
            b_.set_column_bounds(c, v, hi);
        } else if (type == "FX") {
            b_.set_column_bounds(c, v, v);
        } else if (type == "FR") {
            b_.set_column_bounds(c, -kInfinity, kInfinity);
        } else if (type == "MI") {
            b_.set_column_bounds(c, -kInfinity, hi);
        } else if (type == "PL") {
            b_.set_column_bounds(c, lo, kInfinity);
        } else if (type == "BV") {
            b_.set_column_bounds(c, 0.0, 1.0);
            b_.set_column_kind(c, VarKind::Integer);
        } else if (type == "LI") {
            b_.set_column_bounds(c, v, hi);
            b_.set_column_kind(c, VarKind::Integer);
        } else if (type == "UI") {
            b_.set_column_bounds(c, lo, v);
            b_.set_column_kind(c, VarKind::Integer);
        } else {
            fail("unsupported bound type '" + std::string(tok[0]) + "'");
        }
    }

    const std::string& text_;
    const MpsOptions& opt_;
    const std::string& origin_;
    int line_ = 0;
    Sec sec_ = Sec::None;
    Problem::Builder b_;
    ReadResult res_;
    std::string obj_;
    std::vector<char> sense_;
    std::vector<double> rhs_;
    std::unordered_set<std::string> dropped_;
    bool integer_ = false;
    std::string last_col_name_;
    Idx last_col_ = -1;
};

}  // namespace

ReadResult read_mps_string(const std::string& text, const MpsOptions& opt, const std::string& origin) {
    return MpsText(text, opt, origin).read();
}

// --------------------------------------------------------- MPS writer ---

namespace {

/// Appends `v` in its shortest round-trip decimal form.
void put_number(std::string& out, double v) {
    char buf[32];
    const auto r = std::to_chars(buf, buf + sizeof buf, v);
    out.append(buf, r.ptr);
}

}  // namespace

std::string write_mps_string(const Problem& p) {
    const std::span<const std::string> rn = p.row_names(), cn = p.col_names();
    const Idx m = p.num_rows(), n = p.num_cols();
    std::string out;
    out.reserve(static_cast<std::size_t>(p.num_nonzeros()) * 40 + static_cast<std::size_t>(m + n) * 24 + 256);
    std::string obj = "OBJ";   // the objective row: a name no row has
    {
        const std::unordered_set<std::string_view> used(rn.begin(), rn.end());
        while (used.count(obj)) obj += '_';
    }
    out += "NAME ";
    out += p.name().empty() ? std::string("model") : p.name();
    out += "\n";
    if (p.sense() == ObjSense::Maximize) out += "OBJSENSE\n    MAX\n";
    out += "ROWS\n N  " + obj + "\n";
    std::vector<char> sense(static_cast<std::size_t>(m));
    for (Idx i = 0; i < m; ++i) {
        const Real lo = p.row_lower()[i], hi = p.row_upper()[i];
        char t = 'N';
        if (is_finite_bound(lo) && is_finite_bound(hi)) t = lo == hi ? 'E' : 'G';
        else if (is_finite_bound(lo)) t = 'G';
        else if (is_finite_bound(hi)) t = 'L';
        sense[static_cast<std::size_t>(i)] = t;
        out += ' ';
        out += t;
        out += "  ";
        out += rn[static_cast<std::size_t>(i)];
        out += '\n';
    }
    out += "COLUMNS\n";
    const CscMatrix A = p.matrix().to_csc();
    bool in_integer = false;
    int marker = 0;
    for (Idx j = 0; j < n; ++j) {
        const bool integer = p.col_kind()[static_cast<std::size_t>(j)] == VarKind::Integer;
        if (integer != in_integer) {
            out += "    M" + std::to_string(marker++) + "  'MARKER'  " + (integer ? "'INTORG'\n" : "'INTEND'\n");
            in_integer = integer;
        }
        const std::string& name = cn[static_cast<std::size_t>(j)];
        const Real c = p.objective()[j];
        const Idx b = A.col_ptr()[j], e = A.col_ptr()[j + 1];
        if (c != 0.0 || b == e) {   // every column appears at least once
            out += "    " + name + "  " + obj + "  ";
            put_number(out, c);
            out += '\n';
        }
        for (Idx k = b; k < e; ++k) {
            out += "    " + name + "  ";
            out += rn[static_cast<std::size_t>(A.row_idx()[k])];
            out += "  ";
            put_number(out, A.values()[k]);
            out += '\n';
        }
    }
    if (in_integer) out += "    M" + std::to_string(marker) + "  'MARKER'  'INTEND'\n";
    out += "RHS\n";
    if (p.objective_constant() != 0.0) {
        out += "    RHS  " + obj + "  ";
        put_number(out, -p.objective_constant());
        out += '\n';
    }
    std::string ranges;
    for (Idx i = 0; i < m; ++i) {
        const char t = sense[static_cast<std::size_t>(i)];
        if (t == 'N') continue;
        const Real lo = p.row_lower()[i], hi = p.row_upper()[i];
        const Real v = t == 'L' ? hi : lo;
        if (v != 0.0) {
            out += "    RHS  " + rn[static_cast<std::size_t>(i)] + "  ";
            put_number(out, v);
            out += '\n';
        }
        if (t == 'G' && is_finite_bound(hi)) {
            // G rows read a range R as [lo, lo + |R|]; hi - lo must give hi back
            const Real R = hi - lo;
            if (lo + R != hi) throw Error("write_mps: row " + rn[static_cast<std::size_t>(i)] +
                                          " has a range that does not round-trip");
            ranges += "    RNG  " + rn[static_cast<std::size_t>(i)] + "  ";
            put_number(ranges, R);
            ranges += '\n';
        }
    }
    if (!ranges.empty()) out += "RANGES\n" + ranges;
    out += "BOUNDS\n";
    for (Idx j = 0; j < n; ++j) {
        const Real lo = p.col_lower()[j], hi = p.col_upper()[j];
        const std::string& name = cn[static_cast<std::size_t>(j)];
        auto card = [&](const char* type, const Real* v) {
            out += ' ';
            out += type;
            out += " BND  " + name;
            if (v != nullptr) {
                out += "  ";
                put_number(out, *v);
            }
            out += '\n';
        };
        const bool lf = is_finite_bound(lo), hf = is_finite_bound(hi);
        if (lf && hf && lo == hi) {
            card("FX", &lo);
            continue;
        }
        if (!lf && !hf) {
            card("FR", nullptr);
            continue;
        }
        if (!lf) card("MI", nullptr);
        else if (lo != 0.0) card("LO", &lo);
        if (hf) card("UP", &hi);   // after MI / LO, so a negative UP keeps the lower bound written
    }
    out += "ENDATA\n";
    return out;
}

void write_mps(const Problem& p, const std::string& path) {
    const std::string text = write_mps_string(p);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error("cannot write " + path);
    f.write(text.data(), static_cast<std::streamsize>(text.size()));
    if (!f) throw Error("write failed: " + path);
}

ReadResult read_mps(const std::string& path, const MpsOptions& opt) { return read_mps_string(slurp(path), opt, path); }

ReadResult read_model(const std::string& path) {
    std::string p = lower(path);
    if (p.size() > 4 && p.compare(p.size() - 4, 4, ".mps") == 0) return read_mps(path);
    if (p.size() > 3 && p.compare(p.size() - 3, 3, ".lp") == 0) return read_lp(path);
    const std::string text = slurp(path);
    const std::string head = lower(text.substr(0, 4096));
    if (head.find("rows") != std::string::npos && head.find("columns") != std::string::npos)
        return read_mps_string(text, {}, path);
    return read_lp_string(text, path);
}

}  // namespace sov
