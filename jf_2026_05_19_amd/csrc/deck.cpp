// deck.cpp -- native Nastran deck reader of the SOL 105 path: deck text to
// the structure-of-arrays model (sol105_model_desc), host code only.
//
// It replaces, for the card subset this path supports, the reference's
// parse and model build (JFEM/src/parsing/card_processor.jl:51-212,
// JFEM/src/parsing/utilities.jl field rules, JFEM/src/ModelBuilder.jl:
// 421-823) followed by the per-element array derivation of
// assemble_stiffness (JFEM/src/solver/assembly.jl:1953-2530):
//   * sections: executive control up to CEND (SOL kept), case control up to
//     BEGIN BULK, bulk data up to ENDDATA; a deck without CEND is all bulk;
//   * cards: '$' comments, free field (commas), small (8) and large (16,
//     '*') fixed fields, '+' / '*' / blank-name continuations, numbers with
//     the Nastran exponent shorthand (1.-5);
//   * GRID, CQUAD4, CTRIA3, CBAR, CBEAM, PSHELL (blank MID2: membrane only),
//     PBARL TYPE=BAR (section constants, Cowper shear factor at nu = 0.3),
//     MAT1 (completion of E, G, nu), SPC1 (THRU), SPC, FORCE, EIGRL, PARAM,
//     RBE1, RBE2, RBE3, RBAR (as an RBE2), RSPLINE and MPC;
//   * case control SOL / SUBCASE / key = value, global keys inherited by the
//     subcases, the buckling and static subcase pairing (METHOD / STATSUB /
//     LOAD / SPC);
//   * node ranks = ascending GRID id, shells and bars in ascending element
//     id, the SPC set and GRID PS of the static subcase, the FORCE vector of
//     its LOAD (card order), EIGRL ND / V1 / V2 of the buckling METHOD,
//     PARAM K6ROT / AUTOSPC;
//   * the same refusals as the host model builder (cards outside the subset,
//     CP / CD != 0, PBARL other than BAR, pin flags, offsets, enforced
//     displacements, FORCE CID != 0, SNORM > 0, warped CQUAD4) as
//     SOL105_E_UNSUPPORTED with the same message.
#include "../../include/sol105.h"

#include <algorithm>
#include <charconv>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

struct sol105_deck {
    std::vector<int64_t> node_ids, q4_eid, q4_pid, t3_eid, t3_pid, bar_eid;
    std::vector<int32_t> rbe2, rbe3, rbe3_grp, rbe3_node;  // constraint arrays (sol105_model_desc)
    std::vector<double> rbe3_wt;
    std::vector<int32_t> mpc_eq, mpc_term;  // explicit MPC equations (sol105_model_desc)
    std::vector<double> mpc_coef;
    std::vector<int32_t> rspline, rspline_indep, rspline_dep;  // RSPLINE elements (sol105_model_desc)
    std::vector<int32_t> rbe1, rbe1_indep, rbe1_dep;           // RBE1 elements (sol105_model_desc)
    std::vector<double> xyz, q4prop, t3prop, bar_vref, barprop, bar_points, force;
    std::vector<int32_t> q4, t3, bar;
    std::vector<uint8_t> spc_mask, has_shell, has_bar;
    double k6rot = 100.0, v1 = 0.0, v2 = 0.0;
    int32_t autospc = 1, nd = 3, shell_cards = 0;
    int64_t buckling_subcase = 1, static_subcase = 1, sol = 101;
};

namespace {

struct Unsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Invalid : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// numbers and strings
// ---------------------------------------------------------------------------

// A parsed field: absent, an integer, a real, or (case control) a word.
struct Num {
    enum Kind { NONE, INT, REAL } kind = NONE;
    int64_t i = 0;
    double d = 0.0;
    bool ok() const { return kind != NONE; }
    double val() const { return kind == INT ? (double)i : d; }
    bool truthy() const { return ok() && val() != 0.0; }
};

using sv = std::string_view;

sv strip(sv s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}
std::string strip(const std::string &s) { return std::string(strip(sv(s))); }

std::string upper(sv s) {
    std::string u(s);
    for (char &c : u) c = (char)std::toupper((unsigned char)c);
    return u;
}

bool all_digits(sv s, size_t from) {
    if (from >= s.size()) return false;
    for (size_t k = from; k < s.size(); k++)
        if (!std::isdigit((unsigned char)s[k])) return false;
    return true;
}

bool is_int_text(sv s) { return all_digits(s, (!s.empty() && (s[0] == '+' || s[0] == '-')) ? 1 : 0); }

// a real over the whole of s (a leading '+' allowed, no hexadecimal forms)
bool full_real(sv s, double &out) {
    if (!s.empty() && s[0] == '+') {
        s.remove_prefix(1);
        if (!s.empty() && (s[0] == '+' || s[0] == '-')) return false;
    }
    if (s.empty()) return false;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), out, std::chars_format::general);
    return r.ec == std::errc() && r.ptr == s.data() + s.size();
}

// Nastran number (JFEM/src/parsing/utilities.jl): integers, reals with an E
// or D exponent, and the shorthand mantissa[+-]exponent (1.5+3, 1.-5).
Num parse_number(sv tok) {
    Num n;
    const sv s = strip(tok);
    if (s.empty()) return n;
    if (is_int_text(s)) {
        const char *b = s.data() + (s[0] == '+' ? 1 : 0);
        const auto r = std::from_chars(b, s.data() + s.size(), n.i);
        if (r.ec == std::errc()) {
            n.kind = Num::INT;
            return n;
        }
    }
    bool dexp = false;
    for (char c : s) dexp = dexp || c == 'D' || c == 'd';
    if (dexp) {
        std::string t(s);
        for (char &c : t) {
            if (c == 'D') c = 'E';
            else if (c == 'd') c = 'e';
        }
        if (full_real(t, n.d)) {
            n.kind = Num::REAL;
            return n;
        }
    } else if (full_real(s, n.d)) {
        n.kind = Num::REAL;
        return n;
    }
    // shorthand: [+-]?[0-9.]+ followed by [+-][0-9]+
    size_t k = (s[0] == '+' || s[0] == '-') ? 1 : 0, m0 = k;
    while (k < s.size() && (std::isdigit((unsigned char)s[k]) || s[k] == '.')) k++;
    if (k > m0 && k < s.size() && (s[k] == '+' || s[k] == '-') && all_digits(s, k + 1)) {
        const std::string e = std::string(s.substr(0, k)) + "e" + std::string(s.substr(k));
        if (full_real(e, n.d)) {
            n.kind = Num::REAL;
            return n;
        }
    }
    return n;
}

// ---------------------------------------------------------------------------
// lines, sections and logical cards
// ---------------------------------------------------------------------------

std::vector<sv> split_lines(const char *text, size_t len) {
    std::vector<sv> lines;
    size_t a = 0;
    for (size_t k = 0; k < len; k++) {
        const char c = text[k];
        if (c == '\n' || c == '\r' || c == '\v' || c == '\f' || c == '\x1c' || c == '\x1d' || c == '\x1e') {
            lines.emplace_back(text + a, k - a);
            if (c == '\r' && k + 1 < len && text[k + 1] == '\n') k++;
            a = k + 1;
        }
    }
    if (a < len) lines.emplace_back(text + a, len - a);
    return lines;
}

// A card: its data fields (stripped views into the deck text, the name
// excluded) are pool[off, off + n).
struct Card {
    uint32_t off = 0, n = 0;
};

struct Cards {
    std::vector<sv> pool;
    std::map<std::string, std::vector<Card>> by_name;  // sorted names
    sv at(const Card &c, size_t i) const { return i < c.n ? pool[c.off + i] : sv(); }
    size_t size(const Card &c) const { return c.n; }
};

// fixed fields of a line padded to 80 columns: columns 9-72 in 8 (or 16)
void fixed_fields(sv line, bool large, std::vector<sv> &pool) {
    const size_t w = large ? 16 : 8;
    for (size_t a = 8; a < 72; a += w) {
        const size_t b = std::min(a + w, line.size());
        pool.push_back(a < line.size() ? strip(line.substr(a, b - a)) : sv());
    }
}

void read_cards(const std::vector<sv> &lines, Cards &cards) {
    std::string cur_name;
    Card cur;
    bool have = false;
    auto flush = [&]() {
        if (have) cards.by_name[cur_name].push_back(cur);
        have = false;
    };
    auto start = [&](sv name) {
        flush();
        while (!name.empty() && name.back() == '*') name.remove_suffix(1);
        cur_name = upper(name);
        cur.off = (uint32_t)cards.pool.size();
        cur.n = 0;
        have = true;
    };
    auto grow = [&]() { cur.n = (uint32_t)(cards.pool.size() - cur.off); };
    for (sv line : lines) {
        const size_t dollar = line.find('$');
        if (dollar != sv::npos) line = line.substr(0, dollar);
        if (strip(line).empty()) continue;
        if (line.find(',') != sv::npos) {
            size_t a = 0;
            size_t b = line.find(',');
            const sv name = strip(line.substr(0, b));
            const bool cont = name.empty() || name[0] == '+' || name[0] == '*';
            if (cont && !have) continue;
            if (!cont) start(name);
            while (b != sv::npos) {
                a = b + 1;
                b = line.find(',', a);
                cards.pool.push_back(strip(line.substr(a, b == sv::npos ? sv::npos : b - a)));
            }
            grow();
            continue;
        }
        const sv head = line.substr(0, std::min<size_t>(8, line.size()));
        const sv hs = strip(head);
        const bool large = (!hs.empty() && hs.back() == '*') || (!head.empty() && head[0] == '*');
        if ((!head.empty() && (head[0] == '+' || head[0] == '*')) || (hs.empty() && have)) {
            if (have) {
                fixed_fields(line, large, cards.pool);
                grow();
            }
            continue;
        }
        start(hs);
        fixed_fields(line, large, cards.pool);
        grow();
    }
    flush();
}

// a card with its fields
struct CardView {
    const Cards *C;
    Card c;
    size_t size() const { return c.n; }
    sv operator[](size_t i) const { return C->at(c, i); }
};

Num field(const CardView &c, size_t i) { return i < c.size() ? parse_number(c[i]) : Num(); }
double real_or(const CardView &c, size_t i, double dflt) {
    const Num n = field(c, i);
    return n.ok() ? n.val() : dflt;
}
// value "or" default as the host builder's `x or d`: a zero value takes the default too
double real_or0(const CardView &c, size_t i) {
    const Num n = field(c, i);
    return n.truthy() ? n.val() : 0.0;
}

// ---------------------------------------------------------------------------
// case control
// ---------------------------------------------------------------------------

struct CCValue {
    Num num;
    std::string word;  // when num is absent
};
using Subcase = std::map<std::string, CCValue>;

struct CaseControl {
    int64_t sol = 101;               // SOL of the executive control (101 when absent)
    std::map<double, Subcase> subs;  // by subcase id
    std::vector<std::pair<std::string, CCValue>> global;
};

std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    size_t k = 0;
    while (k < s.size()) {
        while (k < s.size() && std::isspace((unsigned char)s[k])) k++;
        size_t a = k;
        while (k < s.size() && !std::isspace((unsigned char)s[k])) k++;
        if (k > a) out.push_back(s.substr(a, k - a));
    }
    return out;
}

CaseControl read_case_control(const std::vector<sv> &lines) {
    CaseControl cc;
    Subcase global;
    std::vector<std::string> global_order;
    Subcase *cur = &global;
    auto set = [&](Subcase *s, const std::string &key, const CCValue &v) {
        if (s == &global && !global.count(key)) global_order.push_back(key);
        (*s)[key] = v;
    };
    for (sv ln : lines) {
        const size_t dollar = ln.find('$');
        const std::string s(strip(dollar == sv::npos ? ln : ln.substr(0, dollar)));
        if (s.empty()) continue;
        const std::string u = upper(s);
        if (u.compare(0, 3, "SOL") == 0) {
            const std::vector<std::string> tok = split_ws(u);
            if (tok.size() > 1) {
                const Num n = parse_number(tok[1]);
                cc.sol = n.ok() ? (int64_t)std::trunc(n.val()) : 101;
            }
            continue;
        }
        if (u.compare(0, 7, "SUBCASE") == 0) {
            const std::vector<std::string> tok = split_ws(u);
            if (tok.size() < 2) throw Invalid("SUBCASE without an id");
            const Num sid = parse_number(tok[1]);
            cur = &cc.subs[sid.ok() ? sid.val() : 1.0];
            continue;
        }
        const size_t eq = u.find('=');
        if (eq != std::string::npos) {
            std::string key = strip(u.substr(0, eq));
            const std::string val = strip(u.substr(eq + 1));
            key = strip(key.substr(0, key.find('(')));
            const std::vector<std::string> tok = split_ws(val);
            CCValue v;
            v.num = parse_number(tok.empty() ? std::string() : tok[0]);
            if (!v.num.ok()) v.word = val;
            set(cur, key, v);
        }
    }
    if (cc.subs.empty()) cc.subs[1.0] = Subcase();
    for (auto &kv : cc.subs)
        for (const std::string &k : global_order)
            if (!kv.second.count(k)) kv.second[k] = global[k];
    return cc;
}

int64_t as_int(const CCValue &v, const char *what) {
    if (!v.num.ok()) throw Invalid(std::string("case control ") + what + " is not a number");
    return (int64_t)std::trunc(v.num.val());
}

// buckling / static subcase pairs (METHOD and STATSUB, else the first
// subcase with a LOAD that is not a buckling one, else the first subcase)
std::vector<std::pair<double, double>> subcase_pairing(const CaseControl &cc) {
    std::vector<double> sids;
    for (const auto &kv : cc.subs) sids.push_back(kv.first);
    std::vector<std::pair<double, int64_t>> buck;
    for (double sid : sids) {
        const Subcase &sub = cc.subs.at(sid);
        if (sub.count("METHOD")) {
            auto st = sub.find("STATSUB");
            buck.push_back({sid, st != sub.end() ? as_int(st->second, "STATSUB") : 0});
        }
    }
    std::vector<std::pair<double, double>> out;
    bool all_zero = !buck.empty();
    for (auto &p : buck) all_zero = all_zero && p.second == 0;
    if (all_zero) {
        double dflt = sids[0];
        bool found = false;
        for (double sid : sids) {
            bool is_buck = false;
            for (auto &p : buck) is_buck = is_buck || p.first == sid;
            if (cc.subs.at(sid).count("LOAD") && !is_buck) {
                dflt = sid;
                found = true;
                break;
            }
        }
        (void)found;
        for (auto &p : buck) out.push_back({p.first, dflt});
        return out;
    }
    for (auto &p : buck) out.push_back({p.first, (double)p.second});
    if (out.empty()) out.push_back({sids.size() >= 2 ? sids[1] : sids[0], sids[0]});
    return out;
}

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------

// Dictionary key of an id-like field: absent, integer or real (an integer
// and a real of equal value are different keys, as the host builder's str()).
struct Key {
    int kind = 0;  // 0 absent, 1 integer, 2 real
    int64_t i = 0;
    double d = 0.0;
    bool operator<(const Key &o) const {
        if (kind != o.kind) return kind < o.kind;
        if (kind == 1) return i < o.i;
        if (kind == 2) return d < o.d;
        return false;
    }
};
Key key_of(const Num &n) {
    Key k;
    if (n.kind == Num::INT) { k.kind = 1; k.i = n.i; }
    else if (n.kind == Num::REAL) { k.kind = 2; k.d = n.d; }
    return k;
}

int64_t element_id(const Num &n, const char *card) {
    if (n.kind != Num::INT) throw Invalid(std::string(card) + " without an integer id");
    return n.i;
}

struct Mat {
    double E, G, nu;
};
struct PShell {
    Key mid;
    double T, br, tst;
    Key mid3;  // absent when 0 or blank
};
struct PBarl {
    Num pid;
    Key mid;
    bool bar = false;
    double A = 0, I1 = 0, I2 = 0, J = 0, K = 0;
    double pts[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // stress points C1 C2 D1 D2 E1 E2 F1 F2
};
struct Shell {
    Num pid;
    Num nodes[4];
    int nn = 0;
};
struct Bar {
    Num pid, ga, gb;
    double pa, pb;
    double w[6];
    bool g0_set;
    int64_t g0;
    double v[3];
};
struct Spc {
    Num sid;
    std::string comps;  // sorted distinct characters
    std::vector<Num> nodes;
    double d;
};
struct Force {
    Num sid, gid;
    double cid, mag, dir[3];
};

// id-keyed entries in card order -> ascending id, the last card of an id
// kept (a later card of the same id replaces an earlier one)
template <class T>
void sort_last_wins(std::vector<std::pair<int64_t, T>> &v) {
    std::stable_sort(v.begin(), v.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    size_t w = 0;
    for (size_t i = 0; i < v.size(); i++) {
        if (i + 1 < v.size() && v[i + 1].first == v[i].first) continue;
        if (w != i) v[w] = std::move(v[i]);  // no self-move (it would empty a vector member)
        w++;
    }
    v.resize(w);
}

std::string sorted_chars(const std::string &s) {
    std::string t = s;
    std::sort(t.begin(), t.end());
    t.erase(std::unique(t.begin(), t.end()), t.end());
    return t;
}

std::string fmt_g(double x) {
    char b[64];
    std::snprintf(b, sizeof(b), "%g", x);
    return b;
}

void build(const Cards &cards, const CaseControl &cc, int buckling_index, sol105_deck &D) {
    static const char *supported[] = {"GRID", "CQUAD4", "CTRIA3", "CBAR", "CBEAM", "PSHELL", "PBARL",
                                      "MAT1", "SPC1", "SPC", "FORCE", "EIGRL", "PARAM", "RBE2", "RBE3", "RBAR",
                                      "MPC", "RSPLINE", "RBE1"};
    {
        std::string list;
        for (const auto &kv : cards.by_name) {
            bool ok = false;
            for (const char *s : supported) ok = ok || kv.first == s;
            if (!ok) list += (list.empty() ? "'" : ", '") + kv.first + "'";
        }
        if (!list.empty()) throw Unsupported("unsupported cards: [" + list + "]");
    }
    auto get = [&](const char *name) {
        std::vector<CardView> out;
        auto it = cards.by_name.find(name);
        if (it != cards.by_name.end()) {
            out.reserve(it->second.size());
            for (const Card &c : it->second) out.push_back(CardView{&cards, c});
        }
        return out;
    };
    // GRID (a later card of the same id replaces an earlier one)
    struct Grid {
        double x[3];
        double cp, cd;
        sv ps;
    };
    std::vector<std::pair<int64_t, Grid>> grids;
    for (const CardView &c : get("GRID")) {
        const int64_t gid = element_id(field(c, 0), "GRID");
        Grid g;
        g.cp = real_or0(c, 1);
        g.x[0] = real_or(c, 2, 0.0);
        g.x[1] = real_or(c, 3, 0.0);
        g.x[2] = real_or(c, 4, 0.0);
        g.cd = real_or0(c, 5);
        g.ps = c.size() > 6 ? strip(c[6]) : sv();
        grids.push_back({gid, g});
    }
    sort_last_wins(grids);
    const int n = (int)grids.size();
    D.node_ids.clear();
    D.xyz.assign(3 * (size_t)n, 0.0);
    for (const auto &kv : grids) {
        const int r = (int)D.node_ids.size();
        if (kv.second.cd != 0.0 || kv.second.cp != 0.0)
            throw Unsupported("GRID " + std::to_string(kv.first) + ": CP/CD coordinate systems are not supported");
        D.node_ids.push_back(kv.first);
        for (int k = 0; k < 3; k++) D.xyz[3 * (size_t)r + k] = kv.second.x[k];
    }
    // a node reference: present and numerically equal to a GRID id
    auto rank_of = [&](const Num &g, int &r) {
        if (!g.ok()) return false;
        const double v = g.val();
        if (v != std::floor(v) || std::fabs(v) > 9.0e18) return false;
        auto it = std::lower_bound(D.node_ids.begin(), D.node_ids.end(), (int64_t)v);
        if (it == D.node_ids.end() || *it != (int64_t)v) return false;
        r = (int)(it - D.node_ids.begin());
        return true;
    };
    // shells (CQUAD4, then CTRIA3 replacing a CQUAD4 of the same id)
    std::vector<std::pair<int64_t, Shell>> shells;
    for (const char *name : {"CQUAD4", "CTRIA3"}) {
        const size_t nn = name[1] == 'Q' ? 4 : 3;
        for (const CardView &c : get(name)) {
            Shell s;
            const int64_t eid = element_id(field(c, 0), name);
            s.pid = field(c, 1);
            s.nn = (int)nn;
            for (size_t k = 0; k < nn; k++) s.nodes[k] = field(c, 2 + k);
            shells.push_back({eid, s});
        }
    }
    sort_last_wins(shells);
    D.shell_cards = (int32_t)shells.size();
    // bars: CBAR and CBEAM groups
    std::vector<std::pair<int64_t, Bar>> bars_by_group[2];
    for (int gi = 0; gi < 2; gi++) {
        for (const CardView &c : get(gi == 0 ? "CBAR" : "CBEAM")) {
            Bar b;
            const int64_t eid = element_id(field(c, 0), gi == 0 ? "CBAR" : "CBEAM");
            b.pid = field(c, 1);
            b.ga = field(c, 2);
            b.gb = field(c, 3);
            b.pa = real_or0(c, 8);
            b.pb = real_or0(c, 9);
            for (int k = 0; k < 6; k++) b.w[k] = real_or0(c, 10 + k);
            const sv x1 = c.size() > 4 ? strip(c[4]) : sv();
            const bool x2 = c.size() > 5 && !strip(c[5]).empty();
            b.g0_set = !x1.empty() && is_int_text(x1) && !x2;
            b.g0 = b.g0_set ? std::strtoll(std::string(x1).c_str(), nullptr, 10) : 0;
            for (int k = 0; k < 3; k++) b.v[k] = b.g0_set ? 0.0 : real_or0(c, 4 + k);
            bars_by_group[gi].push_back({eid, b});
        }
        sort_last_wins(bars_by_group[gi]);
    }
    std::map<Key, PShell> pshells;
    for (const CardView &c : get("PSHELL")) {
        PShell p;
        const Key pid = key_of(field(c, 0));
        p.mid = key_of(field(c, 1));
        p.T = real_or0(c, 2);
        const Num mid2 = field(c, 3);
        const bool blank = !mid2.ok() || mid2.val() == 0.0;
        const Num br = field(c, 4);
        p.br = blank ? 0.0 : (br.ok() ? br.val() : 1.0);
        const Num mid3 = field(c, 5);
        p.mid3 = (!mid3.ok() || mid3.val() == 0.0) ? Key() : key_of(mid3);
        const Num ts = field(c, 6);
        p.tst = ts.ok() ? ts.val() : 5.0 / 6.0;
        pshells[pid] = p;
    }
    std::map<Key, PBarl> pbarls;
    for (const CardView &c : get("PBARL")) {
        PBarl p;
        p.pid = field(c, 0);
        p.mid = key_of(field(c, 1));
        std::string typ = "ROD";
        size_t dim_start = 0;
        bool have_type = false;
        for (size_t k = 2; k < std::min<size_t>(10, c.size()); k++) {
            const sv v = strip(c[k]);
            if (!v.empty() && std::isalpha((unsigned char)v[0])) {
                typ = upper(v);
                dim_start = k + 1;
                have_type = true;
                break;
            }
        }
        std::vector<double> nums;
        if (have_type)
            for (size_t k = dim_start; k < c.size(); k++) {
                const Num x = parse_number(c[k]);
                if (x.ok()) nums.push_back(x.val());
            }
        if (typ == "BAR" && nums.size() >= 2) {
            // section constants of a solid rectangle w x h
            const double w = nums[0], h = nums[1], nu = 0.3;
            p.bar = true;
            p.A = w * h;
            p.I1 = w * std::pow(h, 3) / 12;
            p.I2 = h * std::pow(w, 3) / 12;
            if (w >= h && w > 0) p.J = w * std::pow(h, 3) / 3 * (1 - 0.63 * h / w);
            else if (h > 0) p.J = h * std::pow(w, 3) / 3 * (1 - 0.63 * w / h);
            else p.J = p.I1 + p.I2;
            p.K = 10.0 * (1.0 + nu) / (12.0 + 11.0 * nu);
            // (y, z) = C (h/2, w/2), D (-h/2, w/2), E (-h/2, -w/2), F (h/2, -w/2)
            // (extract_properties.jl:176-178)
            const double w2 = w / 2, h2 = h / 2;
            const double pts[8] = {h2, w2, -h2, w2, -h2, -w2, h2, -w2};
            std::copy(pts, pts + 8, p.pts);
        }
        pbarls[key_of(p.pid)] = p;
    }
    std::map<Key, Mat> mats;
    for (const CardView &c : get("MAT1")) {
        const Key mid = key_of(field(c, 0));
        const Num e = field(c, 1), g = field(c, 2), v = field(c, 3);
        double E = e.ok() ? e.val() : 0.0, G = g.ok() ? g.val() : 0.0, nu = v.ok() ? v.val() : -1.0;
        if (E > 0 && G > 0 && nu < 0) nu = E / (2 * G) - 1.0;
        else if (E > 0 && nu >= 0 && G <= 0) G = E / (2 * (1 + nu));
        else if (G > 0 && nu >= 0 && E <= 0) E = 2 * G * (1 + nu);
        if (nu < 0) nu = 0.3;
        mats[mid] = Mat{E, G, nu};
    }
    // RBE2, RBAR (as RBE2) and RBE3 (extract_constraints.jl:46-141,
    // ModelBuilder.jl:708-725); ids as to_id: integers, reals rounded, else 0
    auto to_id = [](const Num &x) -> int64_t {
        if (x.kind == Num::INT) return x.i;
        if (x.kind == Num::REAL && std::isfinite(x.d) && std::fabs(x.d) < 9.0e18) return (int64_t)std::nearbyint(x.d);
        return 0;
    };
    auto id_at = [&](const CardView &c, size_t i) { return to_id(field(c, i)); };
    struct Rbe2 {
        int64_t gn, cm;
        std::vector<int64_t> gm;
    };
    std::vector<std::pair<int64_t, Rbe2>> rbe2s;
    for (const CardView &c : get("RBE2")) {
        const int64_t eid = id_at(c, 0), gn = id_at(c, 1), cm = id_at(c, 2);
        Rbe2 r{gn, cm, {}};
        for (size_t k = 3; k < c.size(); k++) {
            const int64_t g = id_at(c, k);
            if (g > 0) r.gm.push_back(g);
        }
        if (eid > 0 && gn > 0) rbe2s.push_back({eid, r});
    }
    for (const CardView &c : get("RBAR")) {
        const int64_t eid = id_at(c, 0), ga = id_at(c, 1), gb = id_at(c, 2), cna = id_at(c, 3), cnb = id_at(c, 4);
        if (eid > 0 && ga > 0 && gb > 0) {
            if (cnb > 0 && cna == 0) rbe2s.push_back({eid, Rbe2{gb, cnb, {ga}}});
            else if (cna > 0 && cnb == 0) rbe2s.push_back({eid, Rbe2{ga, cna, {gb}}});
            else rbe2s.push_back({eid, Rbe2{ga, 123456, {gb}}});
        }
    }
    sort_last_wins(rbe2s);
    D.rbe2.clear();
    for (const auto &kv : rbe2s) {
        int m, r;
        Num gn;
        gn.kind = Num::INT;
        gn.i = kv.second.gn;
        if (!rank_of(gn, m)) continue;
        for (int64_t g : kv.second.gm) {
            Num x;
            x.kind = Num::INT;
            x.i = g;
            if (!rank_of(x, r)) continue;
            D.rbe2.insert(D.rbe2.end(), {m, (int32_t)kv.second.cm, r});
        }
    }
    auto is_cont = [](sv t) {
        t = strip(t);
        return !t.empty() && (t[0] == '+' || t[0] == '*');
    };
    auto is_real_field = [&](sv t) {
        t = strip(t);
        if (t.empty() || is_cont(t)) return false;
        for (char ch : t)
            if (ch == '.' || ch == 'E' || ch == 'e' || ch == 'D' || ch == 'd') return true;
        for (size_t k = 1; k < t.size(); k++)
            if (t[k] == '+' || t[k] == '-') return true;
        return false;
    };
    struct Rbe3Group {
        double wt;
        int64_t comps;
        std::vector<int64_t> grids;
    };
    struct Rbe3 {
        int64_t ref, refc;
        std::vector<Rbe3Group> groups;
    };
    std::vector<std::pair<int64_t, Rbe3>> rbe3s;
    for (const CardView &c : get("RBE3")) {
        Rbe3 r{id_at(c, 2), id_at(c, 3), {}};
        size_t pos = 4;
        const size_t nf = c.size();
        while (pos < nf) {
            const sv w = strip(c[pos]);
            if (w.empty() || is_cont(w)) {
                pos++;
                continue;
            }
            const Num wt = parse_number(w);
            if (!wt.ok()) break;
            pos++;
            while (pos < nf) {
                const sv t = strip(c[pos]);
                if (!t.empty() && !is_cont(t)) break;
                pos++;
            }
            if (pos >= nf) break;
            Rbe3Group g{wt.val(), id_at(c, pos), {}};
            pos++;
            while (pos < nf) {
                const sv t = strip(c[pos]);
                if (t.empty() || is_cont(t)) {
                    pos++;
                    continue;
                }
                if (is_real_field(t)) break;
                const Num x = parse_number(t);
                if (!x.ok()) break;
                g.grids.push_back((int64_t)std::trunc(x.val()));
                pos++;
            }
            if (!g.grids.empty()) r.groups.push_back(std::move(g));
        }
        rbe3s.push_back({id_at(c, 0), std::move(r)});
    }
    sort_last_wins(rbe3s);
    D.rbe3.clear();
    D.rbe3_grp.clear();
    D.rbe3_wt.clear();
    D.rbe3_node.clear();
    for (const auto &kv : rbe3s) {
        int ref;
        Num x;
        x.kind = Num::INT;
        x.i = kv.second.ref;
        if (!rank_of(x, ref)) continue;
        const int32_t g0 = (int32_t)D.rbe3_wt.size();
        for (const Rbe3Group &g : kv.second.groups) {
            const int32_t first = (int32_t)D.rbe3_node.size();
            for (int64_t gid : g.grids) {
                Num y;
                y.kind = Num::INT;
                y.i = gid;
                int r;
                if (rank_of(y, r)) D.rbe3_node.push_back(r);
            }
            D.rbe3_grp.insert(D.rbe3_grp.end(), {(int32_t)g.comps, first, (int32_t)D.rbe3_node.size() - first});
            D.rbe3_wt.push_back(g.wt);
        }
        D.rbe3.insert(D.rbe3.end(), {ref, (int32_t)kv.second.refc, g0, (int32_t)D.rbe3_wt.size() - g0});
    }
    // RBE1 (extract_constraints.jl:147-200): EID, (GRID, components) pairs
    // up to the "UM" field, then the dependent pairs; every digit of a
    // component field is one DOF; kept with both lists non-empty
    struct Rbe1 {
        std::vector<std::pair<int64_t, int>> indep, dep;
    };
    std::vector<std::pair<int64_t, Rbe1>> rbe1s;
    for (const CardView &c : get("RBE1")) {
        const int64_t eid = id_at(c, 0);
        size_t um = c.size();
        for (size_t k = 0; k < c.size(); k++) {
            std::string t(strip(c[k]));
            for (char &ch : t) ch = (char)std::toupper((unsigned char)ch);
            if (t == "UM") {
                um = k;
                break;
            }
        }
        if (eid <= 0 || um == c.size()) continue;
        auto pairs = [&](size_t k, size_t stop, std::vector<std::pair<int64_t, int>> &out) {
            for (; k + 1 < c.size() && k < stop; k += 2) {
                const int64_t g = id_at(c, k), cn = id_at(c, k + 1);
                if (g > 0 && cn > 0)
                    for (char ch : std::to_string(cn)) out.push_back({g, ch - '0'});
            }
        };
        Rbe1 r;
        pairs(1, um, r.indep);
        pairs(um + 1, c.size(), r.dep);
        if (!r.indep.empty() && !r.dep.empty()) rbe1s.push_back({eid, std::move(r)});
    }
    sort_last_wins(rbe1s);
    D.rbe1.clear();
    D.rbe1_indep.clear();
    D.rbe1_dep.clear();
    for (const auto &kv : rbe1s) {
        for (const auto *lst : {&kv.second.indep, &kv.second.dep})
            for (const auto &gc : *lst)
                if (gc.second < 1 || gc.second > 6)
                    throw Unsupported("RBE1 " + std::to_string(kv.first) + ": component " +
                                      std::to_string(gc.second) + " of GRID " + std::to_string(gc.first));
        auto rank_id = [&](int64_t g, int &rr) {
            Num x;
            x.kind = Num::INT;
            x.i = g;
            return rank_of(x, rr);
        };
        const auto &in = kv.second.indep;
        if (in.size() > 6) continue;
        std::vector<int32_t> ind;
        bool all = true;
        for (const auto &gc : in) {
            int rr;
            if (!rank_id(gc.first, rr)) {
                all = false;
                break;
            }
            ind.insert(ind.end(), {rr, (int32_t)gc.second});
        }
        if (!all) continue;
        const int32_t i0 = (int32_t)(D.rbe1_indep.size() / 2), d0 = (int32_t)(D.rbe1_dep.size() / 2);
        D.rbe1_indep.insert(D.rbe1_indep.end(), ind.begin(), ind.end());
        for (const auto &gc : kv.second.dep) {
            int rr;
            if (rank_id(gc.first, rr)) D.rbe1_dep.insert(D.rbe1_dep.end(), {rr, (int32_t)gc.second});
        }
        D.rbe1.insert(D.rbe1.end(), {i0, (int32_t)in.size(), d0, (int32_t)(D.rbe1_dep.size() / 2) - d0});
    }
    // RSPLINE (extract_constraints.jl:202-250): EID, D/L, then GRIDs in
    // polyline order; a GRID followed by a component code (digits 0-6, at
    // most 123456) is dependent in each non-zero digit, any other GRID is
    // independent; kept with a dependent and two independent GRIDs
    struct Rspline {
        std::vector<int64_t> indep;
        std::vector<std::pair<int64_t, int>> dep;
    };
    std::vector<std::pair<int64_t, Rspline>> rsplines;
    for (const CardView &c : get("RSPLINE")) {
        const int64_t eid = id_at(c, 0), g1 = id_at(c, 2);
        if (eid <= 0 || g1 <= 0) continue;
        Rspline r;
        r.indep.push_back(g1);
        size_t k = 3;
        while (k < c.size()) {
            const int64_t g = id_at(c, k);
            if (g <= 0) {
                k++;
                continue;
            }
            if (k + 1 < c.size()) {
                const Num nx = field(c, k + 1);
                const int64_t nxt = nx.ok() ? to_id(nx) : -1;
                if (nxt > 0 && nxt <= 123456) {
                    const std::string sd = std::to_string(nxt);
                    if (sd.find_first_not_of("0123456") == std::string::npos) {
                        for (char ch : sd)
                            if (ch != '0') r.dep.push_back({g, ch - '0'});
                        k += 2;
                        continue;
                    }
                }
            }
            r.indep.push_back(g);
            k++;
        }
        if (!r.dep.empty() && r.indep.size() >= 2) rsplines.push_back({eid, std::move(r)});
    }
    sort_last_wins(rsplines);
    D.rspline.clear();
    D.rspline_indep.clear();
    D.rspline_dep.clear();
    for (const auto &kv : rsplines) {
        auto rank_id = [&](int64_t g, int &rr) {
            Num x;
            x.kind = Num::INT;
            x.i = g;
            return rank_of(x, rr);
        };
        std::vector<int32_t> ind;
        for (int64_t g : kv.second.indep) {
            int rr;
            if (rank_id(g, rr)) ind.push_back(rr);
        }
        if (ind.size() < 2) continue;
        const int32_t i0 = (int32_t)D.rspline_indep.size(), d0 = (int32_t)(D.rspline_dep.size() / 2);
        D.rspline_indep.insert(D.rspline_indep.end(), ind.begin(), ind.end());
        for (const auto &gc : kv.second.dep) {
            int rr;
            if (rank_id(gc.first, rr)) D.rspline_dep.insert(D.rspline_dep.end(), {rr, (int32_t)gc.second});
        }
        D.rspline.insert(D.rspline.end(), {i0, (int32_t)ind.size(), d0, (int32_t)(D.rspline_dep.size() / 2) - d0});
    }
    // MPC (extract_constraints.jl:251-271): (G, C, A) triples after the SID
    // while a whole triple remains, terms with G > 0 and C > 0, equations of
    // two or more terms; in card order, filtered as constraints.jl:294-312
    D.mpc_eq.clear();
    D.mpc_term.clear();
    D.mpc_coef.clear();
    for (const CardView &c : get("MPC")) {
        struct Term { int64_t g, comp; double a; };
        std::vector<Term> terms;
        for (size_t k = 1; k + 2 < c.size(); k += 3) {
            const int64_t g = id_at(c, k), comp = id_at(c, k + 1);
            const double a = real_or(c, k + 2, 0.0);
            if (g > 0 && comp > 0) terms.push_back({g, comp, a});
        }
        if (terms.size() < 2) continue;
        auto rank_id = [&](int64_t g, int &r) {
            Num x;
            x.kind = Num::INT;
            x.i = g;
            return rank_of(x, r);
        };
        int r0;
        if (!rank_id(terms[0].g, r0) || terms[0].comp < 1 || terms[0].comp > 6 || !(std::fabs(terms[0].a) >= 1e-30))
            continue;
        std::vector<std::pair<int, const Term *>> rest;
        for (size_t t = 1; t < terms.size(); t++) {
            int r;
            if (rank_id(terms[t].g, r) && terms[t].comp >= 1 && terms[t].comp <= 6) rest.push_back({r, &terms[t]});
        }
        if (rest.empty()) continue;
        D.mpc_eq.insert(D.mpc_eq.end(), {(int32_t)(D.mpc_term.size() / 2), (int32_t)(1 + rest.size())});
        D.mpc_term.insert(D.mpc_term.end(), {r0, (int32_t)terms[0].comp});
        D.mpc_coef.push_back(terms[0].a);
        for (const auto &rt : rest) {
            D.mpc_term.insert(D.mpc_term.end(), {rt.first, (int32_t)rt.second->comp});
            D.mpc_coef.push_back(rt.second->a);
        }
    }
    std::vector<Spc> spcs;
    for (const CardView &c : get("SPC1")) {
        Spc s;
        s.sid = field(c, 0);
        s.comps = sorted_chars(c.size() > 1 ? std::string(strip(c[1])) : std::string());
        std::vector<std::string> rest;
        for (size_t k = 2; k < c.size(); k++)
            if (!strip(c[k]).empty()) rest.push_back(std::string(c[k]));
        auto node_num = [&](const std::string &t) {
            const Num x = parse_number(t);
            if (!x.ok()) throw Invalid("SPC1: node field '" + t + "' is not a number");
            Num r;
            r.kind = Num::INT;
            r.i = (int64_t)std::trunc(x.val());
            return r;
        };
        if (rest.size() >= 3 && upper(strip(rest[1])) == "THRU") {
            const int64_t a = node_num(rest[0]).i, b = node_num(rest[2]).i;
            for (int64_t g = a; g <= b; g++) {
                Num x;
                x.kind = Num::INT;
                x.i = g;
                s.nodes.push_back(x);
            }
        } else {
            for (const std::string &t : rest) s.nodes.push_back(node_num(t));
        }
        s.d = 0.0;
        spcs.push_back(s);
    }
    for (const CardView &c : get("SPC")) {
        const Num sid = field(c, 0);
        for (size_t k = 1; k + 1 < c.size(); k += 3) {
            const Num g = field(c, k);
            if (!g.ok()) continue;
            Spc s;
            s.sid = sid;
            s.comps = sorted_chars(std::string(strip(c[k + 1])));
            s.nodes.push_back(g);
            s.d = real_or0(c, k + 2);
            spcs.push_back(s);
        }
    }
    std::vector<Force> forces;
    for (const CardView &c : get("FORCE")) {
        Force f;
        f.sid = field(c, 0);
        f.gid = field(c, 1);
        f.cid = real_or0(c, 2);
        f.mag = real_or0(c, 3);
        for (int k = 0; k < 3; k++) f.dir[k] = real_or0(c, 4 + k);
        forces.push_back(f);
    }
    struct Eigrl {
        double v1, v2;
        int nd;
    };
    std::map<Key, Eigrl> eigrls;
    for (const CardView &c : get("EIGRL")) {
        const Num nd = field(c, 3);
        eigrls[key_of(field(c, 0))] = Eigrl{real_or0(c, 1), real_or0(c, 2), nd.truthy() ? (int)std::trunc(nd.val()) : 3};
    }
    std::map<std::string, CCValue> params;
    for (const CardView &c : get("PARAM")) {
        const std::string key = upper(c.size() == 0 ? sv() : strip(c[0]));
        const std::string val = c.size() > 1 ? std::string(strip(c[1])) : std::string();
        CCValue v;
        v.num = parse_number(val);
        if (!v.num.ok()) v.word = val;
        params[key] = v;
    }

    // ---- arrays ----
    D.q4.clear(); D.q4_eid.clear(); D.q4_pid.clear(); D.q4prop.clear();
    D.t3.clear(); D.t3_eid.clear(); D.t3_pid.clear(); D.t3prop.clear();
    for (const auto &kv : shells) {
        const Shell &el = kv.second;
        auto pit = pshells.find(key_of(el.pid));
        if (pit == pshells.end()) continue;
        const PShell &prop = pit->second;
        auto mit = mats.find(prop.mid);
        if (mit == mats.end()) continue;
        int r[4];
        bool ok = true;
        for (int k = 0; k < el.nn; k++) ok = ok && rank_of(el.nodes[k], r[k]);
        if (!ok) continue;
        const Mat &mat = mit->second;
        const Mat *smat = &mat;
        if (prop.mid3.kind != 0) {
            auto s3 = mats.find(prop.mid3);
            if (s3 != mats.end()) smat = &s3->second;
        }
        double Gs = smat->G;
        if (Gs <= 0.0) Gs = smat->E > 0 ? smat->E / (2 * (1 + smat->nu)) : 0.0;
        const double row[6] = {mat.E, mat.nu, prop.T, prop.br, prop.tst, Gs};
        const int64_t pid = (int64_t)std::trunc(el.pid.val());
        if (el.nn == 3) {
            for (int k = 0; k < 3; k++) D.t3.push_back(r[k]);
            D.t3_eid.push_back(kv.first);
            D.t3_pid.push_back(pid);
            D.t3prop.insert(D.t3prop.end(), row, row + 6);
        } else {
            for (int k = 0; k < 4; k++) D.q4.push_back(r[k]);
            D.q4_eid.push_back(kv.first);
            D.q4_pid.push_back(pid);
            D.q4prop.insert(D.q4prop.end(), row, row + 6);
        }
    }
    D.bar.clear(); D.bar_eid.clear(); D.bar_vref.clear(); D.barprop.clear(); D.bar_points.clear();
    for (int gi = 0; gi < 2; gi++)
        for (const auto &kv : bars_by_group[gi]) {
            const Bar &b = kv.second;
            auto pit = pbarls.find(key_of(b.pid));
            if (pit == pbarls.end()) continue;
            const PBarl &prop = pit->second;
            if (!prop.bar) {
                const Num &p = prop.pid;
                const std::string ps = p.kind == Num::INT ? std::to_string(p.i) : (p.ok() ? fmt_g(p.d) : "None");
                throw Unsupported("PBARL " + ps + ": only TYPE=BAR is supported");
            }
            auto mit = mats.find(prop.mid);
            int ra, rb;
            if (mit == mats.end() || !rank_of(b.ga, ra) || !rank_of(b.gb, rb)) continue;
            if (b.pa != 0.0 || b.pb != 0.0)
                throw Unsupported("CBAR " + std::to_string(kv.first) + ": pin flags are not supported");
            for (int k = 0; k < 6; k++)
                if (std::fabs(b.w[k]) > 1e-10)
                    throw Unsupported("CBAR " + std::to_string(kv.first) + ": offsets are not supported");
            double v[3] = {b.v[0], b.v[1], b.v[2]};
            Num g0n;
            g0n.kind = Num::INT;
            g0n.i = b.g0;
            int rg0;
            if (b.g0 > 0 && rank_of(g0n, rg0))
                for (int k = 0; k < 3; k++) v[k] = D.xyz[3 * (size_t)rg0 + k] - D.xyz[3 * (size_t)ra + k];
            D.bar.push_back(ra);
            D.bar.push_back(rb);
            D.bar_eid.push_back(kv.first);
            D.bar_vref.insert(D.bar_vref.end(), v, v + 3);
            const double row[8] = {prop.A, prop.I1, prop.I2, prop.J, prop.K, prop.K, mit->second.E, mit->second.G};
            D.barprop.insert(D.barprop.end(), row, row + 8);
            D.bar_points.insert(D.bar_points.end(), prop.pts, prop.pts + 8);
        }
    const size_t ndof = 6 * (size_t)n;
    const auto pairs = subcase_pairing(cc);
    if (buckling_index < 0 || buckling_index >= (int)pairs.size())
        throw Invalid("buckling subcase index " + std::to_string(buckling_index) + " out of range");
    const double buck_sid = pairs[buckling_index].first, stat_sid = pairs[buckling_index].second;
    auto sit = cc.subs.find(stat_sid);
    const Subcase &sub_static = sit != cc.subs.end() ? sit->second : cc.subs.begin()->second;
    const Subcase &sub_buck = cc.subs.at(buck_sid);
    D.buckling_subcase = (int64_t)buck_sid;
    D.static_subcase = (int64_t)stat_sid;
    const CCValue *load = sub_static.count("LOAD") ? &sub_static.at("LOAD") : nullptr;
    const CCValue *spc = sub_buck.count("SPC") ? &sub_buck.at("SPC")
                         : (sub_static.count("SPC") ? &sub_static.at("SPC") : nullptr);
    D.spc_mask.assign(ndof, 0);
    auto set_dof = [&](int r, char comp) {
        if (!std::isdigit((unsigned char)comp)) throw Invalid(std::string("SPC component '") + comp + "'");
        int64_t i = 6 * (int64_t)r + (comp - '0') - 1;
        if (i < 0) i += (int64_t)ndof;  // component 0 of the first node wraps to the last DOF
        D.spc_mask[(size_t)i] = 1;
    };
    if (spc) {
        const int64_t set_id = as_int(*spc, "SPC");
        for (const Spc &s : spcs) {
            if (!s.sid.ok()) throw Invalid("SPC card without a set id");
            if ((int64_t)std::trunc(s.sid.val()) != set_id) continue;
            if (std::fabs(s.d) > 0) throw Unsupported("enforced displacements are not supported");
            for (const Num &g : s.nodes) {
                int r;
                if (!rank_of(g, r)) continue;
                for (char c : s.comps) set_dof(r, c);
            }
        }
    }
    {
        int r = 0;
        for (const auto &kv : grids) {
            for (char c : kv.second.ps)
                if (std::isdigit((unsigned char)c) && c != '0') set_dof(r, c);
            r++;
        }
    }
    D.force.assign(ndof, 0.0);
    if (load) {
        const int64_t load_id = as_int(*load, "LOAD");
        for (const Force &f : forces) {
            if (!f.sid.ok()) throw Invalid("FORCE without a set id");
            if ((int64_t)std::trunc(f.sid.val()) != load_id) continue;
            if (f.cid != 0.0) throw Unsupported("FORCE with CID != 0 is not supported");
            int r;
            if (!rank_of(f.gid, r)) throw Invalid("FORCE on an unknown GRID");
            for (int k = 0; k < 3; k++) D.force[6 * (size_t)r + k] += f.dir[k] * f.mag;
        }
    }
    D.nd = 3;
    D.v1 = D.v2 = 0.0;
    auto mth = sub_buck.find("METHOD");
    if (mth != sub_buck.end()) {
        Num m;
        m.kind = Num::INT;
        m.i = as_int(mth->second, "METHOD");
        auto e = eigrls.find(key_of(m));
        if (e != eigrls.end()) {
            D.nd = e->second.nd;
            D.v1 = e->second.v1;
            D.v2 = e->second.v2;
        }
    }
    D.k6rot = 100.0;
    if (params.count("K6ROT") && params["K6ROT"].num.ok()) D.k6rot = params["K6ROT"].num.val();
    D.autospc = 1;
    if (params.count("AUTOSPC")) {
        const CCValue &a = params["AUTOSPC"];
        if (a.num.ok()) D.autospc = std::fabs(a.num.val()) > 1e-12;
        else {
            const std::string w = upper(strip(a.word));
            D.autospc = !(w.empty() || w == "NO" || w == "N" || w == "FALSE" || w == "F" || w == "OFF" || w == "0");
        }
    }
    if (params.count("SNORM") && params["SNORM"].num.ok() && params["SNORM"].num.val() > 0)
        throw Unsupported("PARAM,SNORM > 0 is not supported");
    // warped CQUAD4: corner deviation from the mean plane >= 1e-6 x the
    // longer diagonal (assembly.jl:2815-2826)
    const size_t nq = D.q4_eid.size();
    for (size_t e = 0; e < nq; e++) {
        double p[4][3];
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < 3; i++) p[k][i] = D.xyz[3 * (size_t)D.q4[4 * e + k] + i];
        double d13[3], d24[3], c[3];
        for (int i = 0; i < 3; i++) {
            d13[i] = p[2][i] - p[0][i];
            d24[i] = p[3][i] - p[1][i];
            c[i] = (((p[0][i] + p[1][i]) + p[2][i]) + p[3][i]) / 4;
        }
        const double nv[3] = {d13[1] * d24[2] - d13[2] * d24[1], d13[2] * d24[0] - d13[0] * d24[2],
                              d13[0] * d24[1] - d13[1] * d24[0]};
        const double ln = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        const double L = std::max(std::sqrt(d13[0] * d13[0] + d13[1] * d13[1] + d13[2] * d13[2]),
                                  std::sqrt(d24[0] * d24[0] + d24[1] * d24[1] + d24[2] * d24[2]));
        double dev = 0.0;
        if (ln > 1e-12) {
            const double nn[3] = {nv[0] / ln, nv[1] / ln, nv[2] / ln};
            for (int k = 0; k < 4; k++) {
                const double q = (p[k][0] - c[0]) * nn[0] + (p[k][1] - c[1]) * nn[1] + (p[k][2] - c[2]) * nn[2];
                dev = std::max(dev, std::fabs(q));
            }
        }
        if (dev >= 1e-6 * std::max(L, 1e-12))
            throw Unsupported("element " + std::to_string(D.q4_eid[e]) +
                              ": warped CQUAD4 (curved-shell gates) not supported");
    }
    D.has_shell.assign(n, 0);
    D.has_bar.assign(n, 0);
    for (int32_t r : D.q4) D.has_shell[r] = 1;
    for (int32_t r : D.t3) D.has_shell[r] = 1;
    for (int32_t r : D.bar) D.has_bar[r] = 1;
}

void set_msg(char *err, int32_t err_len, const std::string &m) {
    if (!err || err_len <= 0) return;
    std::snprintf(err, (size_t)err_len, "%s", m.c_str());
}

}  // namespace

extern "C" {

int sol105_deck_read(const char *text, int64_t len, int32_t buckling_index, sol105_deck **out, char *err,
                     int32_t err_len) {
    if (!text || len < 0 || !out) {
        set_msg(err, err_len, "null argument");
        return SOL105_E_INVALID;
    }
    *out = nullptr;
    try {
        const std::vector<sv> lines = split_lines(text, (size_t)len);
        std::vector<sv> case_lines, bulk_lines;
        int section = 0;  // 0 executive, 1 case control, 2 bulk
        for (sv ln : lines) {
            const sv st = strip(ln);
            const std::string u = upper(st.substr(0, std::min<size_t>(st.size(), 8)));
            if (section == 0) {
                if (u.compare(0, 4, "CEND") == 0) section = 1;
                else if (u.compare(0, 3, "SOL") == 0) case_lines.push_back(ln);
                continue;
            }
            if (section == 1) {
                if (u.compare(0, 5, "BEGIN") == 0 && upper(st).find("BULK") != std::string::npos) section = 2;
                else case_lines.push_back(ln);
                continue;
            }
            if (u.compare(0, 7, "ENDDATA") == 0) break;
            bulk_lines.push_back(ln);
        }
        if (section == 0) {  // a bulk-data fragment
            bulk_lines.clear();
            bulk_lines = lines;
        }
        Cards cards;
        read_cards(bulk_lines, cards);
        const CaseControl cc = read_case_control(case_lines);
        auto D = std::make_unique<sol105_deck>();
        build(cards, cc, buckling_index, *D);
        D->sol = cc.sol;
        *out = D.release();
        return SOL105_OK;
    } catch (const Unsupported &e) {
        set_msg(err, err_len, e.what());
        return SOL105_E_UNSUPPORTED;
    } catch (const std::exception &e) {
        set_msg(err, err_len, e.what());
        return SOL105_E_INVALID;
    }
}

int sol105_deck_free(sol105_deck *deck) {
    delete deck;
    return SOL105_OK;
}

int sol105_deck_sizes(const sol105_deck *D, int64_t *out) {
    if (!D || !out) return SOL105_E_INVALID;
    out[0] = (int64_t)D->node_ids.size();
    out[1] = (int64_t)D->q4_eid.size();
    out[2] = (int64_t)D->t3_eid.size();
    out[3] = (int64_t)D->bar_eid.size();
    out[4] = D->nd;
    out[5] = D->buckling_subcase;
    out[6] = D->static_subcase;
    out[7] = D->shell_cards;
    out[8] = D->sol;
    return SOL105_OK;
}

int sol105_deck_desc(const sol105_deck *D, sol105_model_desc *d) {
    if (!D || !d) return SOL105_E_INVALID;
    *d = sol105_model_desc{};
    d->n_nodes = (int32_t)D->node_ids.size();
    d->xyz = D->xyz.data();
    d->n_q4 = (int32_t)D->q4_eid.size();
    d->q4_nodes = D->q4.data();
    d->q4_prop = D->q4prop.data();
    d->n_t3 = (int32_t)D->t3_eid.size();
    d->t3_nodes = D->t3.data();
    d->t3_prop = D->t3prop.data();
    d->n_bar = (int32_t)D->bar_eid.size();
    d->bar_nodes = D->bar.data();
    d->bar_vref = D->bar_vref.data();
    d->bar_prop = D->barprop.data();
    d->spc_mask = D->spc_mask.data();
    d->force = D->force.data();
    d->node_has_shell = D->has_shell.data();
    d->node_has_bar = D->has_bar.data();
    d->k6rot = D->k6rot;
    d->autospc = D->autospc;
    d->nd = D->nd;
    d->eigrl_v1 = D->v1;
    d->eigrl_v2 = D->v2;
    d->n_shell_cards = D->shell_cards;
    d->n_rbe2 = (int32_t)(D->rbe2.size() / 3);
    d->rbe2 = D->rbe2.data();
    d->n_rbe3 = (int32_t)(D->rbe3.size() / 4);
    d->rbe3 = D->rbe3.data();
    d->n_rbe3_grp = (int32_t)D->rbe3_wt.size();
    d->rbe3_grp = D->rbe3_grp.data();
    d->rbe3_wt = D->rbe3_wt.data();
    d->n_rbe3_node = (int32_t)D->rbe3_node.size();
    d->rbe3_node = D->rbe3_node.data();
    d->n_mpc = (int32_t)(D->mpc_eq.size() / 2);
    d->mpc_eq = D->mpc_eq.data();
    d->n_mpc_term = (int32_t)D->mpc_coef.size();
    d->mpc_term = D->mpc_term.data();
    d->mpc_coef = D->mpc_coef.data();
    d->n_rspline = (int32_t)(D->rspline.size() / 4);
    d->rspline = D->rspline.data();
    d->n_rspline_indep = (int32_t)D->rspline_indep.size();
    d->rspline_indep = D->rspline_indep.data();
    d->n_rspline_dep = (int32_t)(D->rspline_dep.size() / 2);
    d->rspline_dep = D->rspline_dep.data();
    d->n_rbe1 = (int32_t)(D->rbe1.size() / 4);
    d->rbe1 = D->rbe1.data();
    d->n_rbe1_indep = (int32_t)(D->rbe1_indep.size() / 2);
    d->rbe1_indep = D->rbe1_indep.data();
    d->n_rbe1_dep = (int32_t)(D->rbe1_dep.size() / 2);
    d->rbe1_dep = D->rbe1_dep.data();
    return SOL105_OK;
}

int sol105_deck_ids(const sol105_deck *D, int64_t *node_ids, int64_t *q4_eid, int64_t *q4_pid, int64_t *t3_eid,
                    int64_t *t3_pid, int64_t *bar_eid) {
    if (!D) return SOL105_E_INVALID;
    auto cp = [](const std::vector<int64_t> &v, int64_t *o) {
        if (o && !v.empty()) std::memcpy(o, v.data(), v.size() * sizeof(int64_t));
    };
    cp(D->node_ids, node_ids);
    cp(D->q4_eid, q4_eid);
    cp(D->q4_pid, q4_pid);
    cp(D->t3_eid, t3_eid);
    cp(D->t3_pid, t3_pid);
    cp(D->bar_eid, bar_eid);
    return SOL105_OK;
}

int sol105_deck_bar_points(const sol105_deck *D, double *out) {
    if (!D || (!out && !D->bar_points.empty())) return SOL105_E_INVALID;
    if (!D->bar_points.empty()) std::memcpy(out, D->bar_points.data(), D->bar_points.size() * sizeof(double));
    return SOL105_OK;
}

}  // extern "C"
