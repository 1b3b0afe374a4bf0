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
// ROWS, COLUMNS (with MARKER INTORG/INTEND), RHS (an entry on the objective
// row sets the constant to minus that value), RANGES, BOUNDS (UP LO FX FR MI
// PL BV LI UI; UP < 0 on a column with lower bound 0 moves the lower bound to
// -inf under the default rule), ENDATA.  Gzip input is not supported.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>

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

ReadResult read_mps_string(const std::string& text, const MpsOptions& opt, const std::string& origin) {
    Problem::Builder b;
    ReadResult res;
    std::string obj_row;
    std::istringstream in(text);
    std::string line, section;
    int ln = 0;
    bool integer_block = false;
    std::map<Idx, double> rhs, ranges;
    auto err = [&](const std::string& m) { return Error(origin + ":" + std::to_string(ln) + ": " + m); };
    std::vector<std::string> tok;
    while (std::getline(in, line)) {
        ++ln;
        if (line.empty() || line[0] == '*') continue;
        tok.clear();
        {
            std::istringstream ls(line);
            std::string w;
            while (ls >> w) tok.push_back(w);
        }
        if (tok.empty()) continue;
        const bool header = !std::isspace(static_cast<unsigned char>(line[0]));
        if (header) {
            section = tok[0];
            for (char& ch : section) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
            if (section == "NAME") b.set_name(tok.size() > 1 ? tok[1] : "");
            if (section == "OBJSENSE" && tok.size() > 1) {
                const std::string s = lower(tok[1]);
                b.set_sense(s.rfind("max", 0) == 0 ? ObjSense::Maximize : ObjSense::Minimize);
            }
            if (section == "ENDATA") break;
            continue;
        }
        if (section == "OBJSENSE") {
            const std::string s = lower(tok[0]);
            b.set_sense(s.rfind("max", 0) == 0 ? ObjSense::Maximize : ObjSense::Minimize);
        } else if (section == "ROWS") {
            if (tok.size() < 2) throw err("ROWS entry needs a type and a name");
            const char ty = static_cast<char>(std::toupper(static_cast<unsigned char>(tok[0][0])));
            if (ty == 'N') {
                if (obj_row.empty()) {
                    obj_row = tok[1];
                    b.set_objective_name(tok[1]);
                    continue;
                }
                if (!opt.keep_free_rows) continue;
                b.add_row(tok[1], -kInfinity, kInfinity);
            } else if (ty == 'L') {
                b.add_row(tok[1], -kInfinity, 0.0);
            } else if (ty == 'G') {
                b.add_row(tok[1], 0.0, kInfinity);
            } else if (ty == 'E') {
                b.add_row(tok[1], 0.0, 0.0);
            } else {
                throw err("unknown row type " + tok[0]);
            }
        } else if (section == "COLUMNS") {
            if (tok.size() >= 3 && lower(tok[1]) == "'marker'") {
                const std::string mk = lower(tok[2]);
                if (mk == "'intorg'") integer_block = true;
                else if (mk == "'intend'") integer_block = false;
                continue;
            }
            if (tok.size() < 3 || tok.size() % 2 == 0) throw err("COLUMNS entry needs name/value pairs");
            Idx c = b.find_column(tok[0]);
            if (c < 0) {
                c = b.add_column(tok[0]);
                if (integer_block) {
                    b.set_column_kind(c, VarKind::Integer);
                    b.set_column_bounds(c, 0.0, kInfinity);
                }
            }
            for (std::size_t k = 1; k + 1 < tok.size(); k += 2) {
                const double v = std::strtod(tok[k + 1].c_str(), nullptr);
                if (tok[k] == obj_row) {
                    b.set_objective_coefficient(c, v);
                    continue;
                }
                const Idx r = b.find_row(tok[k]);
                if (r < 0) throw err("unknown row " + tok[k]);
                b.add_coefficient(r, c, v);
            }
        } else if (section == "RHS" || section == "RANGES") {
            const std::size_t start = (tok.size() % 2 == 1) ? 1 : 0;   // optional set name
            for (std::size_t k = start; k + 1 < tok.size(); k += 2) {
                const double v = std::strtod(tok[k + 1].c_str(), nullptr);
                if (section == "RHS" && tok[k] == obj_row) {
                    b.set_objective_constant(-v);
                    continue;
                }
                const Idx r = b.find_row(tok[k]);
                if (r < 0) throw err("unknown row " + tok[k]);
                (section == "RHS" ? rhs : ranges)[r] = v;
            }
        } else if (section == "BOUNDS") {
            if (tok.size() < 3) throw err("BOUNDS entry too short");
            const std::string ty = lower(tok[0]);
            const bool no_value = ty == "fr" || ty == "mi" || ty == "pl" || ty == "bv";
            const std::string& cname = tok[2];
            (void)no_value;
            Idx c = b.find_column(cname);
            if (c < 0) throw err("unknown column " + cname);
            const double v = tok.size() > 3 ? std::strtod(tok[3].c_str(), nullptr) : 0.0;
            const double lo = b.column_lower(c), hi = b.column_upper(c);
            if (ty == "up" || ty == "ui") {
                double nlo = lo;
                if (v < 0.0 && lo == 0.0 && opt.negative_upper_bound == NegativeUpperBoundRule::LowerBecomesMinusInfinity) {
                    nlo = -kInfinity;
                    res.warnings.push_back({ln, "BOUNDS", "negative UP bound moved the lower bound of " + cname + " to -inf"});
                }
                b.set_column_bounds(c, nlo, v);
                if (ty == "ui") b.set_column_kind(c, VarKind::Integer);
            } else if (ty == "lo" || ty == "li") {
                b.set_column_bounds(c, v, hi);
                if (ty == "li") b.set_column_kind(c, VarKind::Integer);
            } else if (ty == "fx") {
                b.set_column_bounds(c, v, v);
            } else if (ty == "fr") {
                b.set_column_bounds(c, -kInfinity, kInfinity);
            } else if (ty == "mi") {
                b.set_column_bounds(c, -kInfinity, hi);
            } else if (ty == "pl") {
                b.set_column_bounds(c, lo, kInfinity);
            } else if (ty == "bv") {
                b.set_column_bounds(c, 0.0, 1.0);
                b.set_column_kind(c, VarKind::Integer);
            } else {
                throw err("unknown bound type " + tok[0]);
            }
        } else {
            throw err("data outside a known section");
        }
    }
    // apply RHS and RANGES to the row bounds
    std::vector<char> types;
    {
        // recover the row order's types
        std::istringstream in2(text);
        std::string l2, sec2;
        while (std::getline(in2, l2)) {
            if (l2.empty() || l2[0] == '*') continue;
            std::istringstream ls(l2);
            std::vector<std::string> tk;
            std::string w;
            while (ls >> w) tk.push_back(w);
            if (tk.empty()) continue;
            if (!std::isspace(static_cast<unsigned char>(l2[0]))) {
                sec2 = tk[0];
                for (char& ch : sec2) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
                continue;
            }
            if (sec2 == "ROWS" && tk.size() >= 2 && tk[1] != obj_row) {
                const char ty = static_cast<char>(std::toupper(static_cast<unsigned char>(tk[0][0])));
                if (ty == 'N' && !opt.keep_free_rows) continue;
                types.push_back(ty);
            }
        }
    }
    for (Idx r = 0; r < b.num_rows(); ++r) {
        const char ty = types[static_cast<std::size_t>(r)];
        const double v = rhs.count(r) ? rhs[r] : 0.0;
        double lo = -kInfinity, hi = kInfinity;
        if (ty == 'L') hi = v;
        else if (ty == 'G') lo = v;
        else if (ty == 'E') lo = hi = v;
        if (ranges.count(r) && ty != 'N') {
            const double R = ranges[r];
            if (ty == 'E') {
                if (R > 0) hi = v + std::abs(R);
                else lo = v - std::abs(R);
            } else if (ty == 'L') {
                lo = v - std::abs(R);
            } else if (ty == 'G') {
                hi = v + std::abs(R);
            }
        }
        b.set_row_bounds(r, lo, hi);
    }
    res.problem = b.finish();
    res.format = "mps";
    return res;
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
