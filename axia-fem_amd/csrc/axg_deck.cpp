// INP deck ingestion and the linear-static driver for all-C3D8 / all-C3D20
// isotropic decks: the keyword subset of inp.rs (parse_with_base, :302-1061),
// the model compaction of model.rs (:1234-1309), and solve_linear's setup,
// reduction, solve, reconstruction, reactions and stresses (analysis.rs:548-1162).
// The setup runs on the host (it is cheap and sequential by definition); every
// per-element or per-DOF loop of the solve runs on the device through the
// C ABI of this library. Compiled with -ffp-contract=off: the load vector is
// summed exactly as the reference sums it.

#include <algorithm>
#include <array>
#include <functional>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <string_view>
#include <system_error>
#include <unordered_map>
#include <vector>

#include "axg_internal.h"
#include "axg_writers.h"

using namespace axg;

namespace {

struct DeckElem {
    int32_t id;
    int nn;
    std::array<int32_t, 20> nodes;  // node ids (nn of them)
    std::string elset;
};

struct Deferred {
    enum Kind { NESTED_NSET, NESTED_ELSET, BC, CL, DL } kind;
    std::string name, target;
    int dof = 0;  // BC / CL: 0-based dof; DL: face
    double val = 0.0;
};

struct NodeLoad {
    int32_t id;  // node id (BC, CLOAD) or element id (DLOAD)
    int k;       // 0-based dof, or the face for DLOAD
    double v;
};

using sv = std::string_view;

sv trim_v(sv s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string trim(const std::string &s) { return std::string(trim_v(s)); }

std::string upper(sv s) {
    std::string r(s);
    for (char &c : r) c = (char)std::toupper((unsigned char)c);
    return r;
}

// Tokens of a data line (separated by commas and blanks), appended to out as
// views into the line.
void split_tokens(sv line, std::vector<sv> &out) {
    size_t a = 0;
    const size_t n = line.size();
    while (a < n) {
        while (a < n && (line[a] == ',' || std::isspace((unsigned char)line[a]))) a++;
        size_t b = a;
        while (b < n && line[b] != ',' && !std::isspace((unsigned char)line[b])) b++;
        if (b > a) out.push_back(line.substr(a, b - a));
        a = b;
    }
}

std::vector<sv> split_tokens(sv line) {
    std::vector<sv> out;
    split_tokens(line, out);
    return out;
}

bool is_int_token(sv s) {
    sv t = trim_v(s);
    size_t i = 0;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) i++;
    if (i == t.size()) return false;
    for (; i < t.size(); i++)
        if (!std::isdigit((unsigned char)t[i])) return false;
    return true;
}

bool full_strtod(const std::string &s, double &out) {
    if (s.empty()) return false;
    const char *p = s.c_str();
    char *end = nullptr;
    double v = std::strtod(p, &end);
    if (end != p + s.size()) return false;
    // hexadecimal and "inf"/"nan" spellings are not numbers in a deck
    for (char c : s)
        if (c == 'x' || c == 'X' || c == 'n' || c == 'N' || c == 'i' || c == 'I') return false;
    out = v;
    return true;
}

// inp.rs:100-141: Fortran reals (1.d0, 1.0D-3, 1., .5, 1.0+3). A token made of
// digits, '.', 'e', 'E' and '-' only that from_chars reads to its end has the
// value strtod gives it (both round correctly; from_chars takes no '+' sign and
// no inf / nan / hex spellings): that is the common case and needs no copy.
// Everything else takes the reference's full rules.
bool parse_f64(sv tok, double &out) {
    const sv raw = trim_v(tok);
    if (raw.empty()) return false;
    bool plain = raw.size() < 64;
    for (char c : raw)
        if (!(std::isdigit((unsigned char)c) || c == '.' || c == 'e' || c == 'E' || c == '-')) {
            plain = false;
            break;
        }
    if (plain) {
        double v;
        const auto r = std::from_chars(raw.data(), raw.data() + raw.size(), v,
                                       std::chars_format::general);
        if (r.ec == std::errc() && r.ptr == raw.data() + raw.size()) {
            out = v;
            return true;
        }
    }
    std::string stripped(raw);
    while (!stripped.empty() && stripped.back() == '.') stripped.pop_back();
    if (full_strtod(stripped, out) || full_strtod(std::string(raw), out)) return true;
    std::string b(raw);
    for (char &c : b)
        if (c == 'd' || c == 'D') c = 'e';
    while (!b.empty() && b.back() == '.') b.pop_back();
    if (full_strtod(b, out)) return true;
    for (size_t i = 1; i < b.size(); i++) {
        if ((b[i] == '+' || b[i] == '-') &&
            (std::isdigit((unsigned char)b[i - 1]) || b[i - 1] == '.')) {
            return full_strtod(b.substr(0, i) + "e" + b.substr(i), out);
        }
    }
    return false;
}

double f64_or0(sv tok) {
    double v = 0.0;
    if (!parse_f64(tok, v)) return 0.0;
    return v == 0.0 ? 0.0 : v;
}

// An integer token: [-]digits (at most 9, so no overflow) directly; otherwise
// the reference's rules (an optional sign and digits read as a 64-bit integer
// and narrowed, or a real rounded to the nearest integer).
bool parse_i32(sv tok, int32_t &out) {
    const sv t = trim_v(tok);
    size_t i = t.size() && t[0] == '-' ? 1 : 0;
    if (t.size() > i && t.size() - i <= 9) {
        int64_t v = 0;
        size_t k = i;
        for (; k < t.size() && std::isdigit((unsigned char)t[k]); k++) v = 10 * v + (t[k] - '0');
        if (k == t.size()) {
            out = (int32_t)(i ? -v : v);
            return true;
        }
    }
    if (is_int_token(t)) {
        out = (int32_t)std::strtoll(std::string(t).c_str(), nullptr, 10);
        return true;
    }
    double v;
    if (!parse_f64(t, v) || !std::isfinite(v)) return false;
    out = (int32_t)std::llround(v);
    return true;
}

int32_t need_i32(sv tok) {
    int32_t v;
    AXG_REQUIRE(parse_i32(tok, v), AXG_E_INVALID, "Keine ganze Zahl: " + std::string(tok));
    return v;
}

bool is_comment_or_blank(sv raw) {
    const sv t = trim_v(raw);
    return t.empty() || t.compare(0, 2, "**") == 0;
}

bool is_keyword(sv raw) {
    const sv t = trim_v(raw);
    return !t.empty() && t[0] == '*' && t.compare(0, 2, "**") != 0;
}

// A keyword line's parameters (NAME=VALUE or bare NAME), upper case; a later
// repetition of a name replaces the earlier value. A handful per line, so a
// flat list.
struct KeywordParams {
    std::vector<std::pair<std::string, std::string>> kv;
    void clear() { kv.clear(); }
    const std::string *get(const std::string &k) const {
        for (auto &p : kv)
            if (p.first == k) return &p.second;
        return nullptr;
    }
    size_t count(const std::string &k) const { return get(k) ? 1 : 0; }
    std::string &operator[](const std::string &k) {
        for (auto &p : kv)
            if (p.first == k) return p.second;
        kv.emplace_back(k, std::string());
        return kv.back().second;
    }
};

void parse_keyword(sv raw, std::string &kw, KeywordParams &params) {
    sv r = trim_v(raw);
    if (r.compare(0, 3, "\xEF\xBB\xBF") == 0) r.remove_prefix(3);
    thread_local std::string t;  // reused: no allocation per keyword line
    t.assign(r);
    for (char &c : t) c = (char)std::toupper((unsigned char)c);
    params.clear();
    const sv tv(t);
    size_t start = 0;
    bool first = true;
    while (start <= tv.size()) {
        size_t comma = tv.find(',', start);
        const sv part = trim_v(tv.substr(start, comma == sv::npos ? sv::npos : comma - start));
        if (first) {
            kw.assign(part);
            first = false;
        } else if (!part.empty()) {
            size_t eq = part.find('=');
            if (eq == sv::npos)
                params[std::string(part)] = "";
            else
                params[std::string(trim_v(part.substr(0, eq)))] = std::string(trim_v(part.substr(eq + 1)));
        }
        if (comma == sv::npos) break;
        start = comma + 1;
    }
    // unglue_keyword (inp.rs:41-69), the aliases on this path
    static const char *const kAlias[][2] = {
        {"*SOLIDSECTION", "*SOLID SECTION"}, {"*ENDSTEP", "*END STEP"},
        {"*ELFILE", "*EL FILE"},             {"*NODEFILE", "*NODE FILE"},
        {"*NODEOUTPUT", "*NODE OUTPUT"},     {"*ELOUTPUT", "*EL OUTPUT"}};
    for (auto &a : kAlias)
        if (kw == a[0]) kw = a[1];
}

// Id -> position for node and element ids: a dense table when the ids are
// compact (their span at most 4 n + 1024), else a hash map. The first position
// of a repeated id wins. find returns -1 for an unknown id.
struct IdIndex {
    int32_t lo = 0;
    bool is_dense = true;
    std::vector<int64_t> dense;
    std::unordered_map<int32_t, int64_t> sparse;

    template <class Id>
    void build(size_t n, Id id_at) {
        dense.clear();
        sparse.clear();
        is_dense = true;
        if (n == 0) return;
        int32_t mn = id_at(0), mx = mn;
        for (size_t k = 1; k < n; k++) {
            mn = std::min(mn, id_at(k));
            mx = std::max(mx, id_at(k));
        }
        lo = mn;
        const int64_t span = (int64_t)mx - mn + 1;
        is_dense = span <= 4 * (int64_t)n + 1024;
        if (is_dense) {
            dense.assign((size_t)span, -1);
            for (size_t k = n; k-- > 0;) dense[(size_t)((int64_t)id_at(k) - lo)] = (int64_t)k;
        } else {
            sparse.reserve(n);
            for (size_t k = 0; k < n; k++) sparse.emplace(id_at(k), (int64_t)k);
        }
    }
    int64_t find(int32_t id) const {
        if (is_dense) {
            const int64_t o = (int64_t)id - lo;
            return o >= 0 && o < (int64_t)dense.size() ? dense[(size_t)o] : -1;
        }
        auto it = sparse.find(id);
        return it == sparse.end() ? -1 : it->second;
    }
};

}  // namespace

struct axg_deck {
    std::vector<int32_t> node_ids;
    std::vector<double> coords;  // 3 per node
    std::vector<DeckElem> elements;
    // hash maps: decks with one set, material and section per element have
    // hundreds of thousands of names; the two places that need an order
    // (section rebinding in compact, the strain output's material) sort
    std::unordered_map<std::string, std::vector<int32_t>> nsets, elsets;
    std::unordered_map<std::string, std::pair<double, double>> materials;
    std::unordered_map<std::string, double> densities;  // *DENSITY per material
    int32_t nmodes = 0;                       // *FREQUENCY n (0: static step)
    int32_t nbuckle = 0;                      // *BUCKLE n (0: no buckling step)
    bool nlgeom = false;                      // *STEP, NLGEOM
    std::string heading;                      // *HEADING's first line
    // *NODE FILE / *EL FILE selections (model.rs:1176-1180 defaults)
    bool output_u = true, output_s = true, output_rf = true, output_e = false, output_nt = false;
    int32_t max_inc = 100;                    // *STEP, INC=
    int32_t ninc = 1;                         // increments of the static step
    std::unordered_map<std::string, std::string> elset_material;
    std::vector<NodeLoad> bcs, cloads, dloads;
    std::vector<Deferred> deferred;
    IdIndex id_to_index;  // node id -> node index (after compact)
    // linear-static setup (full DOF space)
    int nn = 0;
    int64_t nnode = 0, nelem = 0, ndof = 0, n_free = 0;
    std::vector<double> xyz, e, nu, rho, u0, f_full;
    std::vector<int32_t> conn;
    std::vector<int64_t> ind_of;
    std::vector<uint8_t> prescribed;
    // the prescribed set and values before autospc (the NLGEOM driver has none)
    std::vector<uint8_t> prescribed0;
    std::vector<double> u00;
    // the text a writer's size query formatted, kept for the fill call with the
    // same arguments and unchanged inputs, so the two-call protocol formats once
    struct WriterCache {
        uint64_t key = 0;
        bool valid = false;
        std::string text;
    } wcache;

    std::vector<int32_t> expand_nset(const std::string &name) const {
        if (is_int_token(name)) return {(int32_t)std::strtoll(trim(name).c_str(), nullptr, 10)};
        const std::string up = upper(name);
        auto it = nsets.find(up);
        if (it != nsets.end()) return it->second;
        if (up == "NALL" || up == "ALL") return node_ids;
        throw Error{AXG_E_INVALID, "Unbekanntes Knotenset " + name};
    }

    // model.rs:1391-1407: the section's material, else the only one, else STEEL
    using MaterialIt = std::unordered_map<std::string, std::pair<double, double>>::const_iterator;
    MaterialIt material_entry_for(const std::string &elset) const {
        auto m = elset_material.find(elset);
        if (m != elset_material.end()) {
            auto it = materials.find(m->second);
            if (it != materials.end()) return it;
        }
        if (materials.size() == 1) return materials.begin();
        auto st = materials.find("STEEL");
        if (st != materials.end()) return st;
        throw Error{AXG_E_INVALID, "Kein Material für ELSET=" + elset};
    }
    const std::string &material_name_for(const std::string &elset) const {
        return material_entry_for(elset)->first;
    }
    std::pair<double, double> material_for(const std::string &elset) const {
        return materials.at(material_name_for(elset));
    }
    // the first material by name (the strain output's E and nu)
    const std::pair<double, double> &first_material() const {
        auto best = materials.begin();
        for (auto it = materials.begin(); it != materials.end(); ++it)
            if (it->first < best->first) best = it;
        return best->second;
    }
    double density_for(const std::string &elset) const {
        auto it = densities.find(material_name_for(elset));
        return it == densities.end() ? 0.0 : it->second;
    }
};

namespace {

// inp.rs:204-230: "set, first[, last][, value]"
bool boundary_card(const std::vector<sv> &t, std::string &name, int &first, int &last,
                   double &val) {
    if (t.size() < 2) return false;
    name = upper(t[0]);
    first = need_i32(t[1]);
    auto dof_id = [](sv s, int &d) {
        if (!is_int_token(s)) return false;
        long v = std::strtol(std::string(trim_v(s)).c_str(), nullptr, 10);
        if ((v >= 1 && v <= 6) || v == 11) {
            d = (int)v;
            return true;
        }
        return false;
    };
    val = 0.0;
    if (t.size() == 2) {
        last = first;
        return true;
    }
    int d2;
    if (t.size() == 3) {
        if (dof_id(t[2], d2)) {
            last = d2;
        } else {
            last = first;
            val = f64_or0(t[2]);
        }
        return true;
    }
    last = dof_id(t[2], d2) ? d2 : first;
    val = f64_or0(t[3]);
    return true;
}

// inp.rs:232-268
void push_boundary(axg_deck &m, const std::string &name, int first, int last, double val) {
    first = std::max(first, 1);
    last = std::max(last, first);
    std::vector<int32_t> ids;
    if (is_int_token(name)) {
        ids.push_back((int32_t)std::strtoll(trim(name).c_str(), nullptr, 10));
    } else {
        auto it = m.nsets.find(name);
        if (it != m.nsets.end() && !it->second.empty()) {
            ids = it->second;
        } else {
            for (int d = first; d <= last; d++)
                if (d >= 1 && d <= 6) m.deferred.push_back({Deferred::BC, name, "", d - 1, val});
            return;
        }
    }
    for (int32_t nid : ids)
        for (int d = first; d <= last; d++)
            if (d >= 1 && d <= 6) m.bcs.push_back({nid, d - 1, val});
}

// model.rs:1234-1309: nodes sorted by id (first definition wins), elements by
// id, NALL / EALL, elements outside any sectioned set rebound to the first
// sectioned set (by name) that lists them.
void compact(axg_deck &m) {
    const size_t n = m.node_ids.size();
    bool strictly_sorted = true;  // then sorting and de-duplication change nothing
    for (size_t i = 1; i < n && strictly_sorted; i++) strictly_sorted = m.node_ids[i - 1] < m.node_ids[i];
    if (!strictly_sorted) {
        std::vector<size_t> order(n);
        for (size_t i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return m.node_ids[a] < m.node_ids[b]; });
        std::vector<int32_t> ids;
        std::vector<double> coords;
        ids.reserve(n);
        coords.reserve(3 * n);
        for (size_t k = 0; k < n; k++) {
            size_t i = order[k];
            if (!ids.empty() && ids.back() == m.node_ids[i]) continue;
            ids.push_back(m.node_ids[i]);
            for (int q = 0; q < 3; q++) coords.push_back(m.coords[3 * i + q]);
        }
        m.node_ids.swap(ids);
        m.coords.swap(coords);
    }
    m.id_to_index.build(m.node_ids.size(), [&](size_t k) { return m.node_ids[k]; });
    auto by_id = [](const DeckElem &a, const DeckElem &b) { return a.id < b.id; };
    if (!std::is_sorted(m.elements.begin(), m.elements.end(), by_id))
        std::stable_sort(m.elements.begin(), m.elements.end(), by_id);
    m.nsets["NALL"] = m.node_ids;
    std::vector<int32_t> eall;
    eall.reserve(m.elements.size());
    for (auto &el : m.elements) eall.push_back(el.id);
    m.elsets["EALL"] = eall;
    if (!m.elset_material.empty()) {
        // each element's first sectioned set by name that lists it: the smallest
        // name, by element position
        const size_t ne = m.elements.size();
        IdIndex epos;
        epos.build(ne, [&](size_t k) { return m.elements[k].id; });
        std::vector<const std::string *> first_set(ne, nullptr);
        for (auto &kv : m.elset_material) {
            auto it = m.elsets.find(kv.first);
            if (it == m.elsets.end()) continue;
            for (int32_t eid : it->second) {
                const int64_t k = epos.find(eid);
                if (k < 0) continue;
                // every element of a repeated id takes the set (first position found here)
                for (int64_t q = k; q < (int64_t)ne && m.elements[q].id == eid; q++)
                    if (!first_set[q] || kv.first < *first_set[q]) first_set[q] = &kv.first;
            }
        }
        // an element whose own set has a section keeps it; consecutive elements
        // usually share their set, so its lookup is reused
        std::string last;
        bool have_last = false, last_sectioned = false;
        for (size_t k = 0; k < ne; k++) {
            DeckElem &el = m.elements[k];
            if (!have_last || el.elset != last) {
                last_sectioned = m.elset_material.count(el.elset) != 0;
                last = el.elset;
                have_last = true;
            }
            if (last_sectioned || !first_set[k]) continue;
            el.elset = *first_set[k];
        }
    }
}

// inp.rs:2256-2330 (the subset: nested sets, then BC, CLOAD, DLOAD)
void expand_deferred(axg_deck &m) {
    for (auto &d : m.deferred) {
        if (d.kind == Deferred::NESTED_NSET) {
            auto add = m.expand_nset(d.target);
            auto &dst = m.nsets[d.name];
            dst.insert(dst.end(), add.begin(), add.end());
        } else if (d.kind == Deferred::NESTED_ELSET) {
            auto it = m.elsets.find(d.target);
            std::vector<int32_t> add = it == m.elsets.end() ? std::vector<int32_t>() : it->second;
            auto &dst = m.elsets[d.name];
            dst.insert(dst.end(), add.begin(), add.end());
        }
    }
    for (auto &d : m.deferred) {
        if (d.kind == Deferred::BC) {
            for (int32_t nid : m.expand_nset(d.name)) m.bcs.push_back({nid, d.dof, d.val});
        } else if (d.kind == Deferred::CL) {
            for (int32_t nid : m.expand_nset(d.name)) m.cloads.push_back({nid, d.dof, d.val});
        } else if (d.kind == Deferred::DL) {
            auto it = m.elsets.find(d.name);
            if (it != m.elsets.end())
                for (int32_t eid : it->second) m.dloads.push_back({eid, d.dof, d.val});
        }
    }
    m.deferred.clear();
}

void parse_deck(axg_deck &m, const std::string &text) {
    // lines and tokens are views into text, which outlives the parse
    std::vector<sv> lines;
    {
        const sv all(text);
        size_t a = 0;
        while (a <= all.size()) {
            size_t b = all.find('\n', a);
            sv line = all.substr(a, b == sv::npos ? sv::npos : b - a);
            if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
            lines.push_back(line);
            if (b == sv::npos) break;
            a = b + 1;
        }
        if (!lines.empty() && lines.back().empty()) lines.pop_back();
    }
    const size_t n = lines.size();
    {
        size_t nkw = 0;  // decks with a set, material and section per element have many
        for (const sv l : lines) nkw += !l.empty() && l[0] == '*';
        m.elsets.reserve(nkw / 4 + 16);
        m.materials.reserve(nkw / 4 + 16);
        m.elset_material.reserve(nkw / 4 + 16);
    }
    size_t i = 0;
    std::string cur_mat;
    bool have_mat = false;
    std::pair<double, double> *cur_mat_val = nullptr;  // node-based map: stays valid
    auto collect = [&](size_t j, std::vector<sv> &toks) {
        toks.clear();
        while (j < n) {
            const sv raw = lines[j];
            if (is_comment_or_blank(raw)) {
                j++;
                continue;
            }
            if (is_keyword(raw)) break;
            split_tokens(raw, toks);
            j++;
        }
        return j;
    };
    // the first line at or after j that ends a keyword's data lines
    auto block_end = [&](size_t j) {
        while (j < n && (is_comment_or_blank(lines[j]) || !is_keyword(lines[j]))) j++;
        return j;
    };
    std::vector<sv> toks;
    KeywordParams params;
    std::string kw;
    while (i < n) {
        const sv raw = lines[i];
        if (is_comment_or_blank(raw) || !is_keyword(raw)) {
            i++;
            continue;
        }
        parse_keyword(raw, kw, params);
        auto param = [&](const char *key, const std::string &dflt) {
            const std::string *v = params.get(key);
            return v ? *v : dflt;
        };
        if (kw == "*NODE") {
            const std::string nset = param("NSET", "");
            const size_t j = block_end(i + 1);
            // data lines in parallel chunks, appended in line order
            const int64_t nl = (int64_t)(j - i - 1);
            const int nt = host_threads(nl, 16384);
            std::vector<std::vector<int32_t>> cid(nt);
            std::vector<std::vector<double>> cxyz(nt);
            std::vector<std::exception_ptr> cerr(nt);
            parallel_chunks(nl, nt, [&](int t, int64_t lo, int64_t hi) {
                try {
                    std::vector<sv> tk;
                    for (int64_t q = lo; q < hi; q++) {
                        const sv l = lines[i + 1 + q];
                        if (is_comment_or_blank(l)) continue;
                        tk.clear();
                        split_tokens(l, tk);
                        if (tk.empty()) continue;
                        cid[t].push_back(need_i32(tk[0]));
                        for (size_t c = 1; c <= 3; c++)
                            cxyz[t].push_back(tk.size() > c ? f64_or0(tk[c]) : 0.0);
                    }
                } catch (...) {
                    cerr[t] = std::current_exception();
                }
            });
            for (auto &e : cerr)
                if (e) std::rethrow_exception(e);
            size_t added = 0;
            for (int t = 0; t < nt; t++) {
                m.node_ids.insert(m.node_ids.end(), cid[t].begin(), cid[t].end());
                m.coords.insert(m.coords.end(), cxyz[t].begin(), cxyz[t].end());
                added += cid[t].size();
            }
            if (!nset.empty()) {
                auto &dst = m.nsets[nset];
                dst.insert(dst.end(), m.node_ids.end() - added, m.node_ids.end());
            }
            i = j;
        } else if (kw == "*ELEMENT") {
            const std::string typ = param("TYPE", "");
            int nn = 0;
            if (typ == "C3D8") nn = 8;
            else if (typ == "C3D20") nn = 20;
            AXG_REQUIRE(nn > 0, AXG_E_INVALID,
                        "Nicht unterstützter Elementtyp " + typ + " (this path: C3D8, C3D20).");
            const std::string elset = param("ELSET", "EALL");
            const size_t j = block_end(i + 1);
            // tokens of the data lines (an element may continue on the next
            // line), tokenised in parallel chunks and joined in line order; then
            // every complete group of 1 + nn tokens is an element
            const int64_t nl = (int64_t)(j - i - 1);
            const int nt = host_threads(nl, 16384);
            std::vector<std::vector<sv>> ctok(nt);
            parallel_chunks(nl, nt, [&](int t, int64_t lo, int64_t hi) {
                for (int64_t q = lo; q < hi; q++) {
                    const sv l = lines[i + 1 + q];
                    if (!is_comment_or_blank(l)) split_tokens(l, ctok[t]);
                }
            });
            toks.clear();
            for (auto &c : ctok) toks.insert(toks.end(), c.begin(), c.end());
            const size_t ne = toks.size() / (size_t)(nn + 1), e0 = m.elements.size();
            m.elements.resize(e0 + ne);
            const int ntp = host_threads((int64_t)ne, 8192);
            std::vector<std::exception_ptr> cerr(ntp);
            parallel_chunks((int64_t)ne, ntp, [&](int t, int64_t lo, int64_t hi) {
                try {
                    for (int64_t q = lo; q < hi; q++) {
                        const size_t k = (size_t)q * (nn + 1);
                        DeckElem &el = m.elements[e0 + q];
                        el.id = need_i32(toks[k]);
                        el.nn = nn;
                        for (int a = 0; a < nn; a++) el.nodes[a] = need_i32(toks[k + 1 + a]);
                        el.elset = elset;
                    }
                } catch (...) {
                    cerr[t] = std::current_exception();
                }
            });
            for (auto &e : cerr)
                if (e) std::rethrow_exception(e);
            auto &dst = m.elsets[elset];
            for (size_t q = 0; q < ne; q++) dst.push_back(m.elements[e0 + q].id);
            i = j;
        } else if (kw == "*NSET" || kw == "*ELSET") {
            const bool nsetkw = kw == "*NSET";
            const char *key = nsetkw ? "NSET" : "ELSET";
            AXG_REQUIRE(params.count(key), AXG_E_INVALID, kw + " ohne " + key + "=");
            const std::string name = params[key];
            i = collect(i + 1, toks);
            std::vector<int32_t> ids;
            if (params.count("GENERATE")) {
                size_t k = 0;
                while (k + 1 < toks.size()) {
                    int32_t a = need_i32(toks[k]), b = need_i32(toks[k + 1]), st = 1;
                    if (k + 2 < toks.size()) {
                        st = need_i32(toks[k + 2]);
                        if (st == 0) st = 1;
                        k += 3;
                    } else {
                        k += 2;
                    }
                    if (st > 0)
                        for (int64_t v = a; v <= b; v += st) ids.push_back((int32_t)v);
                    else
                        for (int64_t v = a; v >= b; v += st) ids.push_back((int32_t)v);
                }
            } else {
                for (auto &t : toks) {
                    int32_t v;
                    if (parse_i32(t, v))
                        ids.push_back(v);
                    else
                        m.deferred.push_back({nsetkw ? Deferred::NESTED_NSET : Deferred::NESTED_ELSET,
                                              name, upper(t), 0, 0.0});
                }
            }
            auto &dst = nsetkw ? m.nsets[name] : m.elsets[name];
            dst.insert(dst.end(), ids.begin(), ids.end());
        } else if (kw == "*MATERIAL") {
            cur_mat = param("NAME", "MATERIAL-1");
            have_mat = true;
            cur_mat_val = &m.materials.emplace(cur_mat, std::make_pair(0.0, 0.0)).first->second;
            i++;
        } else if (kw == "*ELASTIC") {
            i = collect(i + 1, toks);
            AXG_REQUIRE(toks.size() >= 2, AXG_E_INVALID, "*ELASTIC benötigt E und nu");
            double e = 0.0, nu = 0.0;
            parse_f64(toks[0], e);
            parse_f64(toks[1], nu);
            if (std::fabs(e) <= 1e-18) {
                e = 1.0;
                for (auto &t : toks) {
                    double v;
                    if (parse_f64(t, v) && std::fabs(v) > 1e-18) {
                        e = v;
                        break;
                    }
                }
            }
            nu = std::min(std::max(nu, -0.49), 0.49);
            if (have_mat && !cur_mat.empty())
                *cur_mat_val = {e, nu};
            else
                m.materials["MATERIAL-1"] = {e, nu};
        } else if (kw == "*DENSITY") {
            // inp.rs:675-686
            i = collect(i + 1, toks);
            AXG_REQUIRE(!toks.empty(), AXG_E_INVALID, "*DENSITY ohne Wert");
            double rho = 0.0;
            AXG_REQUIRE(parse_f64(toks[0], rho), AXG_E_INVALID, "Keine Zahl: " + std::string(toks[0]));
            const std::string name = have_mat && !cur_mat.empty() ? cur_mat : "MATERIAL-1";
            m.materials.emplace(name, std::make_pair(0.0, 0.0));
            m.densities[name] = rho;
        } else if (kw == "*FREQUENCY") {
            // inp.rs:1062-1071: n modes, 10 when absent or unreadable, at least 1
            i = collect(i + 1, toks);
            int32_t nm = 10;
            if (!toks.empty() && !parse_i32(toks[0], nm)) nm = 10;
            m.nmodes = std::max(nm, 1);
        } else if (kw == "*HEADING") {
            // inp.rs:410-417: the next non-blank line, trimmed, unless a keyword
            i++;
            while (i < n && is_comment_or_blank(lines[i])) i++;
            if (i < n && !is_keyword(lines[i])) {
                m.heading = std::string(trim_v(lines[i]));
                i++;
            }
        } else if (kw == "*NODE FILE" || kw == "*NODE OUTPUT") {
            // inp.rs:1989-2004
            i = collect(i + 1, toks);
            m.output_u = m.output_rf = false;
            for (auto &t : toks) {
                const std::string u = upper(trim_v(t));
                if (u == "U") m.output_u = true;
                else if (u == "RF") m.output_rf = true;
                else if (u == "NT") m.output_nt = true;
            }
            if (!m.output_u && !m.output_rf) m.output_u = true;
        } else if (kw == "*EL FILE" || kw == "*ELEMENT OUTPUT") {
            // inp.rs:2006-2021
            i = collect(i + 1, toks);
            m.output_s = m.output_e = false;
            for (auto &t : toks) {
                const std::string u = upper(trim_v(t));
                if (u == "S") m.output_s = true;
                else if (u == "E") m.output_e = true;
            }
            if (!m.output_s && !m.output_e) m.output_s = true;
        } else if (kw == "*STEP") {
            // inp.rs:953-982: NLGEOM (with any value) makes the step's static
            // procedure geometrically nonlinear; INC= caps the increments
            if (params.count("NLGEOM")) {
                m.nlgeom = true;
                m.ninc = 1;
            }
            if (params.count("INC")) {
                int32_t v;
                if (parse_i32(params["INC"], v) && v > 0) m.max_inc = v;
            }
            i++;
        } else if (kw == "*STATIC") {
            // inp.rs:983-1061 without RIKS: initial increment, period
            i = collect(i + 1, toks);
            auto positive = [&](size_t k) {
                double v;
                return (toks.size() > k && parse_f64(toks[k], v) && v > 0.0) ? v : 1.0;
            };
            const double dt0 = positive(0), period = positive(1);
            const long long n_from_time = std::max<long long>(std::llround(period / dt0), 1);
            m.ninc = (int32_t)std::max<long long>(
                std::min<long long>(n_from_time, std::max<int32_t>(m.max_inc, 1)), 1);
        } else if (kw == "*BUCKLE") {
            // inp.rs:1072-1081: n factors, 1 when absent or unreadable, at least 1
            i = collect(i + 1, toks);
            int32_t nb = 1;
            if (!toks.empty() && !parse_i32(toks[0], nb)) nb = 1;
            m.nbuckle = std::max(nb, 1);
        } else if (kw == "*SOLID SECTION") {
            const std::string elset = param("ELSET", "EALL");
            if (params.count("MATERIAL")) m.elset_material[elset] = params["MATERIAL"];
            i = collect(i + 1, toks);
        } else if (kw == "*BOUNDARY") {
            i++;
            while (i < n) {
                const sv l = lines[i];
                if (is_comment_or_blank(l)) {
                    i++;
                    continue;
                }
                if (is_keyword(l)) break;
                std::string name;
                int first, last;
                double val;
                if (boundary_card(split_tokens(l), name, first, last, val))
                    push_boundary(m, name, first, last, val);
                i++;
            }
        } else if (kw == "*CLOAD") {
            i = collect(i + 1, toks);
            size_t k = 0;
            while (k + 2 < toks.size()) {
                const std::string name = upper(toks[k]);
                const int32_t dof = need_i32(toks[k + 1]);
                double mag = 0.0;
                parse_f64(toks[k + 2], mag);
                k += 3;
                if (dof < 1 || dof > 6) continue;
                auto it = m.nsets.find(name);
                if (is_int_token(name)) {
                    m.cloads.push_back({(int32_t)std::strtoll(name.c_str(), nullptr, 10), dof - 1, mag});
                } else if (it != m.nsets.end() && !it->second.empty()) {
                    for (int32_t nid : it->second) m.cloads.push_back({nid, dof - 1, mag});
                } else {
                    m.deferred.push_back({Deferred::CL, name, "", dof - 1, mag});
                }
            }
        } else if (kw == "*DLOAD") {
            i = collect(i + 1, toks);
            size_t k = 0;
            while (k + 1 < toks.size()) {
                const std::string name = upper(toks[k]), typ = upper(toks[k + 1]);
                k += 2;
                AXG_REQUIRE(typ.size() >= 2 && typ[0] == 'P' &&
                                std::all_of(typ.begin() + 1, typ.end(),
                                            [](char c) { return std::isdigit((unsigned char)c); }),
                            AXG_E_INVALID, "DLOAD-Typ " + typ + " is not on this path");
                const int face = std::atoi(typ.c_str() + 1);
                AXG_REQUIRE(k < toks.size(), AXG_E_INVALID, "*DLOAD ohne Betrag");
                double mag = 0.0;
                parse_f64(toks[k], mag);
                k += 1;
                auto it = m.elsets.find(name);
                if (is_int_token(name)) {
                    m.dloads.push_back({(int32_t)std::strtoll(name.c_str(), nullptr, 10), face, mag});
                } else if (it != m.elsets.end() && !it->second.empty()) {
                    for (int32_t eid : it->second) m.dloads.push_back({eid, face, mag});
                } else {
                    m.deferred.push_back({Deferred::DL, name, "", face, mag});
                }
            }
        } else {
            i++;
        }
    }
    compact(m);
    expand_deferred(m);
}

// solve_linear's setup (analysis.rs:548-946) for an all-hex isotropic model,
// before assembly: arrays, prescribed values (last BC wins), isolated nodes
// pinned, the load vector (pressure faces in element order, then nodal loads
// summed per node and DOF).
void linear_setup(axg_deck &m) {
    AXG_REQUIRE(!m.elements.empty(), AXG_E_INVALID, "Modell ohne Elemente.");
    const int nn = m.elements[0].nn;
    for (auto &el : m.elements)
        AXG_REQUIRE(el.nn == nn, AXG_E_INVALID,
                    "only all-C3D8 or all-C3D20 decks are on this path");
    m.nn = nn;
    m.nnode = (int64_t)m.node_ids.size();
    m.nelem = (int64_t)m.elements.size();
    m.ndof = 3 * m.nnode;
    m.xyz = m.coords;
    m.conn.assign((size_t)m.nelem * nn, 0);
    m.e.assign(m.nelem, 0.0);
    m.nu.assign(m.nelem, 0.0);
    m.rho.assign(m.nelem, 0.0);
    IdIndex eidx;
    eidx.build(m.elements.size(), [&](size_t k) { return m.elements[k].id; });
    const std::string *last_set = nullptr;
    std::pair<double, double> mat{0.0, 0.0};
    double rho = 0.0;
    for (int64_t k = 0; k < m.nelem; k++) {
        const DeckElem &el = m.elements[k];
        for (int a = 0; a < nn; a++) {
            const int64_t idx = m.id_to_index.find(el.nodes[a]);
            AXG_REQUIRE(idx >= 0, AXG_E_INVALID,
                        "Element " + std::to_string(el.id) + ": unbekannter Knoten " +
                            std::to_string(el.nodes[a]));
            m.conn[k * nn + a] = (int32_t)idx;
        }
        if (!last_set || el.elset != *last_set) {  // consecutive elements often share a set
            auto mit = m.material_entry_for(el.elset);
            mat = mit->second;
            auto dit = m.densities.empty() ? m.densities.end() : m.densities.find(mit->first);
            rho = dit == m.densities.end() ? 0.0 : dit->second;
            last_set = &el.elset;
        }
        m.e[k] = mat.first;
        m.nu[k] = mat.second;
        m.rho[k] = rho;
    }
    auto node_index = [&](int32_t nid) {
        const int64_t idx = m.id_to_index.find(nid);
        AXG_REQUIRE(idx >= 0, AXG_E_INVALID, "Unbekannter Knoten " + std::to_string(nid));
        return idx;
    };
    // prescribed values, last BC wins (analysis.rs:555-563)
    m.prescribed.assign(m.ndof, 0);
    m.u0.assign(m.ndof, 0.0);
    for (auto &b : m.bcs) {
        if (b.k >= 3) continue;
        const int64_t d = 3 * node_index(b.id) + b.k;
        m.prescribed[d] = 1;
        m.u0[d] = b.v;
    }
    // pin_unused_dofs (analysis.rs:2230-2355): nodes no element uses
    std::vector<uint8_t> attached(m.nnode, 0);
    for (int32_t c : m.conn) attached[c] = 1;
    for (int64_t i = 0; i < m.nnode; i++)
        if (!attached[i])
            for (int d = 0; d < 3; d++) m.prescribed[3 * i + d] = 1;  // u0 keeps its value (0)
    m.prescribed0 = m.prescribed;
    m.u00 = m.u0;
    // pressure faces: elements ascending, each element's faces in deck order
    m.f_full.assign(m.ndof, 0.0);
    std::map<int64_t, std::vector<std::pair<int, double>>> by_elem;
    for (auto &dl : m.dloads) {
        const int64_t k = eidx.find(dl.id);
        AXG_REQUIRE(k >= 0, AXG_E_INVALID, "DLOAD auf unbekanntes Element " + std::to_string(dl.id));
        by_elem[k].push_back({dl.k, dl.v});
    }
    double exyz[60], fe[60];
    for (auto &kv : by_elem) {
        const int32_t *en = &m.conn[kv.first * nn];
        for (int a = 0; a < nn; a++)
            for (int q = 0; q < 3; q++) exyz[3 * a + q] = m.xyz[3 * (int64_t)en[a] + q];
        for (auto &fm : kv.second) {
            int st = face_pressure_host(nn, exyz, fm.first, fm.second, fe);
            AXG_REQUIRE(st == AXG_OK, AXG_E_INVALID,
                        std::string("Ungültige ") + (nn == 8 ? "C3D8" : "C3D20") + "-Fläche P" +
                            std::to_string(fm.first));
            for (int a = 0; a < nn; a++)
                for (int q = 0; q < 3; q++) m.f_full[3 * (int64_t)en[a] + q] += fe[3 * a + q];
        }
    }
    // add_cloads (analysis.rs:2709-2764): summed per (node, dof), then added once
    std::map<int64_t, double> acc;
    for (auto &cl : m.cloads) {
        if (cl.k >= 3) continue;
        const int64_t d = 3 * node_index(cl.id) + cl.k;
        auto it = acc.find(d);
        if (it == acc.end())
            acc.emplace(d, 0.0 + cl.v * 1.0);
        else
            it->second += cl.v * 1.0;
    }
    for (auto &kv : acc) m.f_full[kv.first] += kv.second;
}

// DofMap::build (constraint.rs:26-136) without MPCs
void number_dofs(axg_deck &m) {
    m.ind_of.assign(m.ndof, -1);
    int64_t k = 0;
    bool any = false;
    for (int64_t d = 0; d < m.ndof; d++) {
        if (m.prescribed[d]) {
            any = true;
            continue;
        }
        m.ind_of[d] = k++;
    }
    m.n_free = k;
    AXG_REQUIRE(k > 0 || any, AXG_E_INVALID, "Modell ohne freie Freiheitsgrade.");
}

}  // namespace

extern "C" {

int axg_deck_parse(const char *text, int64_t len, axg_deck **out) {
    API_BEGIN
    AXG_REQUIRE(text && out && len >= 0, AXG_E_INVALID, "axg_deck_parse: invalid argument");
    *out = nullptr;
    auto d = std::make_unique<axg_deck>();
    parse_deck(*d, std::string(text, (size_t)len));
    linear_setup(*d);
    number_dofs(*d);
    *out = d.release();
    API_END
}

int axg_deck_destroy(axg_deck *deck) {
    delete deck;
    return AXG_OK;
}

int axg_deck_info(axg_deck *deck, int32_t *nn, int64_t *nnode, int64_t *nelem, int64_t *n_free) {
    API_BEGIN
    AXG_REQUIRE(deck, AXG_E_INVALID, "axg_deck_info: NULL deck");
    if (nn) *nn = deck->nn;
    if (nnode) *nnode = deck->nnode;
    if (nelem) *nelem = deck->nelem;
    if (n_free) *n_free = deck->n_free;
    API_END
}

int axg_deck_arrays(axg_deck *deck, int32_t *node_ids, int32_t *elem_ids, double *xyz,
                    int32_t *conn, double *e, double *nu, int64_t *ind_of, double *u0,
                    double *f_full) {
    API_BEGIN
    AXG_REQUIRE(deck, AXG_E_INVALID, "axg_deck_arrays: NULL deck");
    const axg_deck &d = *deck;
    auto put = [](auto *dst, const auto &src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    put(node_ids, d.node_ids);
    if (elem_ids)
        for (int64_t k = 0; k < d.nelem; k++) elem_ids[k] = d.elements[k].id;
    put(xyz, d.xyz);
    put(conn, d.conn);
    put(e, d.e);
    put(nu, d.nu);
    put(ind_of, d.ind_of);
    put(u0, d.u0);
    put(f_full, d.f_full);
    API_END
}

// Assembly of the deck's K_ff on the device with autospc_null_diag
// (analysis.rs:2164-2228): DOFs whose full diagonal is at most 1e-14 of the
// largest are pinned at zero and the DOF map rebuilt before the final assembly.
struct DeckMesh {
    axg_mesh *m = nullptr;
    ~DeckMesh() {
        if (m) axg_mesh_destroy(m);
    }
};

static void check_status(int st) {
    if (st != AXG_OK) throw Error{st, axg_last_error()};
}

static void deck_mesh(axg_ctx *ctx, axg_deck &d, DeckMesh &g) {
    auto build = [&]() {
        if (g.m) axg_mesh_destroy(g.m), g.m = nullptr;
        check_status(axg_mesh_create(ctx, d.nn, d.nnode, d.xyz.data(), d.nelem, d.conn.data(),
                                     d.e.data(), d.nu.data(), d.ind_of.data(), &g.m));
        check_status(axg_assemble(g.m, AXG_PREC_FP64, nullptr));
    };
    build();
    std::vector<double> diag(d.ndof);
    check_status(axg_full_diagonal(g.m, diag.data()));
    double max_d = 0.0;
    for (double v : diag) max_d = std::max(max_d, std::fabs(v));
    bool more = false;
    if (max_d >= 1e-30) {
        const double tol = max_d * 1e-14;
        for (int64_t i = 0; i < d.ndof; i++)
            if (std::fabs(diag[i]) <= tol && !d.prescribed[i]) {
                d.prescribed[i] = 1;
                d.u0[i] = 0.0;
                more = true;
            }
    }
    if (more) {
        number_dofs(d);
        build();
    }
}

// DofMap::reconstruct (constraint.rs:207-218)
static void reconstruct(const axg_deck &d, const double *x, double *u_full) {
    for (int64_t i = 0; i < d.ndof; i++) {
        u_full[i] = d.u0[i];
        if (d.ind_of[i] >= 0) u_full[i] = u_full[i] + 1.0 * x[d.ind_of[i]];
    }
}

// Reduced right-hand side, PCG and DofMap::reconstruct of the step's loads.
static void static_solve(axg_deck &d, DeckMesh &g, const axg_solve_opts *opts, double *u_full,
                         axg_solve_info *si) {
    std::vector<double> rhs(std::max<int64_t>(d.n_free, 1)), x(std::max<int64_t>(d.n_free, 1));
    check_status(axg_reduce_rhs(g.m, d.f_full.data(), d.u0.data(), rhs.data()));
    std::memset(si, 0, sizeof *si);
    if (d.n_free > 0) check_status(axg_mesh_solve(g.m, rhs.data(), 1, opts, x.data(), si));
    reconstruct(d, x.data(), u_full);
}

int axg_deck_procedure(axg_deck *deck, int32_t *nmodes, int32_t *nbuckle, int32_t *nlgeom,
                       int32_t *ninc) {
    API_BEGIN
    AXG_REQUIRE(deck, AXG_E_INVALID, "axg_deck_procedure: NULL deck");
    if (nmodes) *nmodes = deck->nmodes;
    if (nbuckle) *nbuckle = deck->nbuckle;
    if (nlgeom) *nlgeom = deck->nlgeom ? 1 : 0;
    if (ninc) *ninc = deck->ninc;
    API_END
}

// solve_linear for the deck (analysis.rs:548-1162): assembly on the device,
// autospc, reduced right-hand side, GPU PCG, reconstruction, reactions and
// nodal stresses.
int axg_deck_solve(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *u_full,
                   double *rf, double *stress, double *vm, double *stress_gp,
                   axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && deck && u_full, AXG_E_INVALID, "axg_deck_solve: invalid argument");
    axg_deck &d = *deck;
    const int precision = opts ? opts->precision : AXG_PREC_FP64;
    DeckMesh g;
    deck_mesh(ctx, d, g);
    if (precision == AXG_PREC_MIXED) check_status(axg_assemble(g.m, AXG_PREC_MIXED, nullptr));
    axg_solve_info si;
    static_solve(d, g, opts, u_full, &si);
    if (rf) check_status(axg_reactions(g.m, u_full, d.f_full.data(), rf));
    if (stress || stress_gp) {
        std::vector<double> st(stress ? 0 : (size_t)d.nnode * 6), en;
        if (stress_gp) en.resize((size_t)d.nelem * d.nn * 6);
        check_status(axg_nodal_stress(g.m, u_full, stress ? stress : st.data(), vm,
                                      stress_gp ? en.data() : nullptr));
        // stress_gp (analysis.rs:1145-1151): nodal values / nn summed in node order
        if (stress_gp)
            for (int64_t e = 0; e < d.nelem; e++)
                for (int c = 0; c < 6; c++) {
                    double acc = 0.0;
                    for (int a = 0; a < d.nn; a++)
                        acc += en[((size_t)e * d.nn + a) * 6 + c] / (double)d.nn;
                    stress_gp[e * 6 + c] = acc;
                }
    }
    if (info) *info = si;
    API_END
}

// The *FREQUENCY branch of solve_linear (analysis.rs:955-1005) for a deck
// without consistent-mass terms: lumped mass (analysis.rs:669-717) restricted
// to the free DOFs, subspace iteration on the device, f = sqrt(lambda) / 2 pi
// (0 for lambda <= 0), mode shapes reconstructed over the full DOF space
// (mass-normalised, the reference's default EigNorm::Mass).
int axg_deck_frequency(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *freqs,
                       double *values, double *modes_full, int32_t *nfound, int32_t *sweeps,
                       axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && deck && freqs && nfound, AXG_E_INVALID,
                "axg_deck_frequency: invalid argument");
    axg_deck &d = *deck;
    AXG_REQUIRE(d.nmodes > 0, AXG_E_INVALID, "axg_deck_frequency: the deck has no *FREQUENCY step");
    DeckMesh g;
    deck_mesh(ctx, d, g);
    AXG_REQUIRE(d.n_free > 0, AXG_E_INVALID, "Eigenproblem ohne freie DOF / ohne Moden.");
    std::vector<double> m_full(d.ndof), m_ind(d.n_free, 0.0);
    check_status(axg_lumped_mass(g.m, d.rho.data(), m_full.data()));
    for (int64_t i = 0; i < d.ndof; i++)
        if (d.ind_of[i] >= 0) m_ind[d.ind_of[i]] += 1.0 * 1.0 * m_full[i];
    bool any = false;
    for (double v : m_ind) any = any || v > 0.0;
    AXG_REQUIRE(any, AXG_E_INVALID, "*FREQUENCY: *DENSITY fehlt oder Masse ist null.");
    const int32_t nm = d.nmodes;
    std::vector<double> vals(nm), vecs(modes_full ? (size_t)nm * d.n_free : 0);
    int32_t k = 0;
    check_status(axg_eigen(g.m, m_ind.data(), nm, opts, vals.data(),
                           modes_full ? vecs.data() : nullptr, &k, sweeps, info));
    for (int32_t j = 0; j < k; j++) {
        freqs[j] = vals[j] > 0.0 ? std::sqrt(vals[j]) / (2.0 * M_PI) : 0.0;
        if (values) values[j] = vals[j];
        if (modes_full) reconstruct(d, vecs.data() + (size_t)j * d.n_free, modes_full + (size_t)j * d.ndof);
    }
    *nfound = k;
    API_END
}

// assemble_fext_u (analysis.rs:3222-3300) for nodal loads and face pressure:
// the nodal loads summed per (node, dof) and added first, then each element's
// pressure faces in element order, on the deformed coordinates x + u when u is
// given (follower pressure).
static void nl_fext(const axg_deck &d, const double *u, std::vector<double> &f) {
    f.assign(d.ndof, 0.0);
    std::map<int64_t, double> acc;
    for (auto &cl : d.cloads) {
        if (cl.k >= 3) continue;
        const int64_t idx = d.id_to_index.find(cl.id);
        AXG_REQUIRE(idx >= 0, AXG_E_INVALID, "Unbekannter Knoten " + std::to_string(cl.id));
        const int64_t dof = 3 * idx + cl.k;
        auto a = acc.find(dof);
        if (a == acc.end())
            acc.emplace(dof, 0.0 + cl.v * 1.0);
        else
            a->second += cl.v * 1.0;
    }
    for (auto &kv : acc) f[kv.first] += kv.second;
    if (d.dloads.empty()) return;
    std::unordered_map<int32_t, int64_t> eidx;
    for (int64_t k = 0; k < d.nelem; k++) eidx.emplace(d.elements[k].id, k);
    std::map<int64_t, std::vector<std::pair<int, double>>> by_elem;
    for (auto &dl : d.dloads) {
        auto it = eidx.find(dl.id);
        AXG_REQUIRE(it != eidx.end(), AXG_E_INVALID,
                    "DLOAD auf unbekanntes Element " + std::to_string(dl.id));
        by_elem[it->second].push_back({dl.k, dl.v});
    }
    const int nn = d.nn;
    double exyz[60], fe[60];
    for (auto &kv : by_elem) {
        const int32_t *en = &d.conn[kv.first * nn];
        for (int a = 0; a < nn; a++)
            for (int q = 0; q < 3; q++) {
                const int64_t g = 3 * (int64_t)en[a] + q;
                exyz[3 * a + q] = u ? d.xyz[g] + u[g] : d.xyz[g];
            }
        for (auto &fm : kv.second) {
            int st = face_pressure_host(nn, exyz, fm.first, fm.second, fe);
            AXG_REQUIRE(st == AXG_OK, AXG_E_INVALID,
                        std::string("Ungültige ") + (nn == 8 ? "C3D8" : "C3D20") + "-Fläche P" +
                            std::to_string(fm.first));
            for (int a = 0; a < nn; a++)
                for (int q = 0; q < 3; q++) f[3 * (int64_t)en[a] + q] += fe[3 * a + q];
        }
    }
}

static double norm_seq(const std::vector<double> &v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

// solve_continuum_newton (analysis.rs:4313-4860) for an all-hex St. Venant-
// Kirchhoff deck without contact, pretension, rigid bodies, plasticity or
// amplitudes: load factor steps of 1 / ninc (x1.5 after a success, halved
// on failure, at most 16 retries, 1e-6 minimum), prescribed values and loads
// ramped with the factor, follower pressure re-assembled at every iteration,
// full Newton with the consistent tangent assembled on the device and the
// GPU PCG of opts, convergence at ||r_free|| < 1e-8 (1 + ||f||) within 25
// iterations. Results: u_full, rf = f_int - f_ext, nodal Cauchy stress, von
// Mises and Green-Lagrange strain (each may be NULL).
int axg_deck_nlgeom(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *u_full,
                    double *rf, double *stress, double *vm, double *strain, double *stress_gp,
                    int32_t *ninc_done, int32_t *newton_iters, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && deck && u_full, AXG_E_INVALID, "axg_deck_nlgeom: invalid argument");
    const axg_deck &d = *deck;
    AXG_REQUIRE(d.nlgeom, AXG_E_INVALID, "axg_deck_nlgeom: the deck has no *STEP, NLGEOM");
    // the tangent is re-assembled in FP64 for every Newton iteration: no FP32 operator
    AXG_REQUIRE(!opts || opts->precision == AXG_PREC_FP64, AXG_E_INVALID,
                "axg_deck_nlgeom: precision must be AXG_PREC_FP64 (the Newton solves run on "
                "the FP64 tangent)");
    const double newton_tol = 1e-8;  // model.rs:1173
    const int max_newton = 25;       // model.rs:1172
    const int64_t ndof = d.ndof;
    // DofMap::build over the step's prescribed DOFs (no autospc on this path)
    std::vector<int64_t> ind_of(ndof, -1);
    std::vector<int64_t> pres;
    int64_t n_free = 0;
    for (int64_t i = 0; i < ndof; i++) {
        if (d.prescribed0[i])
            pres.push_back(i);
        else
            ind_of[i] = n_free++;
    }
    AXG_REQUIRE(n_free > 0, AXG_E_INVALID, "NLGEOM Kontinuum: keine freien DOF.");
    DeckMesh g;
    check_status(axg_mesh_create(ctx, d.nn, d.nnode, d.xyz.data(), d.nelem, d.conn.data(),
                                 d.e.data(), d.nu.data(), ind_of.data(), &g.m));
    axg_mesh *mesh = g.m;
    hipStream_t s = ctx->stream;
    const int64_t ne = std::max<int64_t>(d.nelem, 1), m = 3 * d.nn;
    DevBuf<double> du(std::max<int64_t>(ndof, 1)), fe((size_t)ne * m), cau(6 * ne), glb(6 * ne),
        fint(std::max<int64_t>(ndof, 1));
    std::vector<double> f_tgt, f_0(ndof, 0.0), f_inc(ndof), fd, f_int(ndof), last_fint(ndof, 0.0);
    nl_fext(d, nullptr, f_tgt);
    const bool follower = !d.dloads.empty();
    const int ninc = std::max<int32_t>(d.ninc, 1);
    std::vector<double> u(ndof, 0.0), u_work, last_du(ndof), r(ndof), rhs(n_free), x(n_free),
        dsel_h(n_free);
    std::vector<double> f_now(f_0);
    double lam = 0.0, dt = 1.0 / ninc, residual = 0.0;
    int done_inc = 0, retries = 0, iters = 0;
    int64_t pcg_iters = 0;
    double pcg_ms = 0.0;
    std::string solver_name;
    auto upload = [&](const std::vector<double> &v) {
        AXG_HIP(hipMemcpyAsync(du.p, v.data(), 8 * ndof, hipMemcpyHostToDevice, s));
    };
    while (lam < 1.0 - 1e-14) {
        const double lam_new = std::min(lam + dt, 1.0);
        u_work = u;
        for (int64_t p : pres) u_work[p] = (1.0 - lam_new) * 0.0 + lam_new * d.u00[p];
        for (int64_t i = 0; i < ndof; i++) f_inc[i] = (1.0 - lam_new) * f_0[i] + lam_new * f_tgt[i];
        std::fill(last_du.begin(), last_du.end(), 0.0);
        bool have_du = false, inc_ok = false;
        std::string inc_err;
        for (int it = 0; it < max_newton; it++) {
            iters = it + 1;
            if (follower) {
                nl_fext(d, u_work.data(), fd);
                for (int64_t i = 0; i < ndof; i++) f_inc[i] = (1.0 - lam_new) * f_0[i] + lam_new * fd[i];
            }
            upload(u_work);
            try {
                nl_assemble_device(mesh, du.p, fe.p, cau.p, glb.p, fint.p);
            } catch (const Error &ex) {
                inc_err = ex.msg;
                if (ex.code == AXG_E_INVERSION && have_du) {
                    double nrm = 0.0;
                    for (int64_t i = 0; i < ndof; i++) {
                        u_work[i] -= 0.5 * last_du[i];
                        last_du[i] *= 0.5;
                    }
                    for (int64_t p : pres) u_work[p] = (1.0 - lam_new) * 0.0 + lam_new * d.u00[p];
                    nrm = norm_seq(last_du);
                    if (nrm > 1e-18) {
                        inc_err.clear();
                        continue;
                    }
                }
                break;
            }
            AXG_HIP(hipMemcpyAsync(f_int.data(), fint.p, 8 * ndof, hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
            last_fint = f_int;
            // DofMap::reduce_inc (constraint.rs:175-205): free rows of f_inc - f_int
            for (int64_t i = 0; i < ndof; i++) r[i] = f_inc[i] - f_int[i];
            std::fill(rhs.begin(), rhs.end(), 0.0);
            for (int64_t i = 0; i < ndof; i++)
                if (ind_of[i] >= 0 && std::fabs(r[i]) != 0.0) rhs[ind_of[i]] += 1.0 * r[i];
            residual = norm_seq(rhs);
            const double fref = norm_seq(f_inc);
            if (residual < newton_tol * (1.0 + fref)) {
                inc_ok = true;
                break;
            }
            if (it + 1 == max_newton) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "Newton konvergierte nicht (r=%.3e nach %d Iterationen).",
                              residual, iters);
                inc_err = buf;
                break;
            }
            // The reference solves K_T du = r directly; the PCG's absolute
            // breakdown guard (|p^T K p| < 1e-30, linalg.rs:231-233) would stop it
            // near convergence, where r is small. The right-hand side is divided
            // by the power of two nearest sqrt(r^T D^-1 r) (D: the Jacobi divisor)
            // and the solution multiplied back: exact, as a power of two changes no
            // rounding.
            AXG_HIP(hipMemcpyAsync(dsel_h.data(), mesh->dsel.p, 8 * n_free, hipMemcpyDeviceToHost, s));
            AXG_HIP(hipStreamSynchronize(s));
            double beta = 0.0;
            for (int64_t k = 0; k < n_free; k++) beta += rhs[k] * rhs[k] / dsel_h[k];
            double scale = 1.0;
            if (beta > 0.0 && std::isfinite(beta)) {
                scale = std::ldexp(1.0, (int)std::lround(0.5 * std::log2(beta)));
                for (double &v : rhs) v /= scale;
            }
            axg_solve_info si;
            const int st = axg_mesh_solve(mesh, rhs.data(), 1, opts, x.data(), &si);
            if (st != AXG_OK) {
                inc_err = axg_last_error();
                break;
            }
            for (double &v : x) v *= scale;
            solver_name = si.name;
            pcg_iters += si.iters;
            pcg_ms += si.solve_ms;
            for (int64_t i = 0; i < ndof; i++) {
                const double dv = ind_of[i] >= 0 ? 0.0 + 1.0 * x[ind_of[i]] : 0.0;
                u_work[i] += dv;
                last_du[i] = dv;
            }
            for (int64_t p : pres) u_work[p] = (1.0 - lam_new) * 0.0 + lam_new * d.u00[p];
            have_du = true;
            if (norm_seq(x) < 1e-14) {
                inc_ok = true;
                break;
            }
        }
        if (inc_ok) {
            u = u_work;
            lam = lam_new;
            f_now = f_inc;
            done_inc++;
            retries = 0;
            const double remain = 1.0 - lam;
            if (remain > 1e-14) dt = std::min(std::min(dt * 1.5, 1.0 / ninc), remain);
        } else {
            retries++;
            dt *= 0.5;
            if (dt < 1e-6 || retries > 16) {
                char buf[96];
                std::snprintf(buf, sizeof buf, "NLGEOM Inkrement bei λ=%.4f fehlgeschlagen (", lam);
                throw Error{AXG_E_NOT_CONVERGED,
                            std::string(buf) + (inc_err.empty() ? "unbekannt" : inc_err) + ")."};
            }
        }
    }
    std::copy(u.begin(), u.end(), u_full);
    if (rf)
        for (int64_t i = 0; i < ndof; i++) rf[i] = last_fint[i] - f_now[i];
    if (stress || vm || strain || stress_gp) {
        const int64_t nn6 = std::max<int64_t>(d.nnode, 1) * 6;
        DevBuf<double> dst(nn6), dvm(std::max<int64_t>(d.nnode, 1)), dsn(nn6), dgp(6 * ne);
        upload(u);
        nl_nodal_device(mesh, du.p, cau.p, glb.p, dst.p, dvm.p, dsn.p, dgp.p);
        if (stress) AXG_HIP(hipMemcpyAsync(stress, dst.p, 8 * d.nnode * 6, hipMemcpyDeviceToHost, s));
        if (vm) AXG_HIP(hipMemcpyAsync(vm, dvm.p, 8 * d.nnode, hipMemcpyDeviceToHost, s));
        if (strain) AXG_HIP(hipMemcpyAsync(strain, dsn.p, 8 * d.nnode * 6, hipMemcpyDeviceToHost, s));
        if (stress_gp)
            AXG_HIP(hipMemcpyAsync(stress_gp, dgp.p, 8 * d.nelem * 6, hipMemcpyDeviceToHost, s));
        AXG_HIP(hipStreamSynchronize(s));
    }
    if (ninc_done) *ninc_done = done_inc;
    if (newton_iters) *newton_iters = iters;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iters = pcg_iters;
        info->outer_iters = iters;
        info->nrhs = 1;
        info->residual = residual;
        info->rel_residual = residual / std::max(norm_seq(f_now), 1e-300);
        info->solve_ms = pcg_ms;
        std::snprintf(info->name, sizeof(info->name), "Newton (%.44s, %d inc)",
                      solver_name.empty() ? "equilibrium" : solver_name.c_str(), done_inc);
    }
    API_END
}

// The *BUCKLE branch of solve_linear (analysis.rs:1176-1260) for a solid
// deck: the static solution of the step's loads, Kg_ff from its element mean
// stresses, subspace_ab on (K_ff, -Kg_ff), mode shapes over the full DOF space
// (unit Euclidean length in the free DOFs, as subspace_ab leaves them).
int axg_deck_buckle(axg_ctx *ctx, axg_deck *deck, const axg_solve_opts *opts, double *factors,
                    double *modes_full, double *u_static, int32_t *nfound, axg_solve_info *info) {
    API_BEGIN
    AXG_REQUIRE(ctx && deck && factors && nfound, AXG_E_INVALID,
                "axg_deck_buckle: invalid argument");
    axg_deck &d = *deck;
    AXG_REQUIRE(d.nbuckle > 0, AXG_E_INVALID, "axg_deck_buckle: the deck has no *BUCKLE step");
    DeckMesh g;
    deck_mesh(ctx, d, g);
    AXG_REQUIRE(d.n_free > 0, AXG_E_INVALID, "Eigenproblem ohne freie DOF.");
    std::vector<double> u(d.ndof);
    axg_solve_info si;
    static_solve(d, g, opts, u.data(), &si);
    if (u_static) std::copy(u.begin(), u.end(), u_static);
    const int32_t nb = d.nbuckle;
    std::vector<double> vals(nb), vecs(modes_full ? (size_t)nb * d.n_free : 0);
    int32_t k = 0;
    check_status(axg_buckle(g.m, u.data(), nb, opts, vals.data(),
                            modes_full ? vecs.data() : nullptr, &k, info));
    // DofMap::reconstruct adds u0 on the prescribed DOFs (zero for a pinned support)
    for (int32_t j = 0; j < k; j++) {
        factors[j] = vals[j];
        if (modes_full) reconstruct(d, vecs.data() + (size_t)j * d.n_free, modes_full + (size_t)j * d.ndof);
    }
    *nfound = k;
    API_END
}

// ------------------------------------------------------------- writers
// Key of a writer call: the writer, a count, the input addresses and their
// values (FNV-1a over the 64-bit words).
static uint64_t writer_key(int kind, int64_t extra,
                           std::initializer_list<std::pair<const double *, int64_t>> in) {
    uint64_t h = 14695981039346656037ull;
    auto mix = [&h](uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    mix((uint64_t)kind);
    mix((uint64_t)extra);
    for (auto &a : in) {
        mix((uint64_t)(uintptr_t)a.first);
        if (!a.first) continue;
        for (int64_t i = 0; i < a.second; i++) {
            uint64_t w;
            std::memcpy(&w, a.first + i, 8);
            mix(w);
        }
    }
    return h;
}

// The size query (buf == NULL) keeps the formatted text; the fill call with
// the same key takes it instead of formatting again.
static bool serve_cached(axg_deck &d, uint64_t key, char *buf, int64_t cap, int64_t *len) {
    if (!d.wcache.valid || d.wcache.key != key) return false;
    *len = (int64_t)d.wcache.text.size();
    if (!buf) return true;
    AXG_REQUIRE(cap >= *len, AXG_E_INVALID, "output buffer too small");
    std::memcpy(buf, d.wcache.text.data(), d.wcache.text.size());
    d.wcache.valid = false;
    std::string().swap(d.wcache.text);
    return true;
}

static void copy_out(axg_deck &d, uint64_t key, std::string &o, char *buf, int64_t cap,
                     int64_t *len) {
    *len = (int64_t)o.size();
    if (!buf) {
        d.wcache.key = key;
        d.wcache.valid = true;
        d.wcache.text = std::move(o);
        return;
    }
    AXG_REQUIRE(cap >= (int64_t)o.size(), AXG_E_INVALID, "output buffer too small");
    std::memcpy(buf, o.data(), o.size());
}

// dat.rs:3-94 for a static or frequency result: displacements of every node,
// forces where |fx| + |fy| + |fz| > 1e-14, element stresses (stress_gp, one
// point per element), then the nodal stresses and the frequencies when given.
int axg_deck_write_dat(axg_deck *deck, const double *u_full, const double *rf,
                       const double *stress_gp, const double *nodal_stress, const double *freqs,
                       int32_t nfreq, char *buf, int64_t cap, int64_t *len) {
    API_BEGIN
    AXG_REQUIRE(deck && u_full && rf && len, AXG_E_INVALID, "axg_deck_write_dat: NULL argument");
    const axg_deck &d = *deck;
    const uint64_t key = writer_key(1, nfreq, {{u_full, d.ndof}, {rf, d.ndof},
                                               {stress_gp, 6 * d.nelem}, {nodal_stress, 6 * d.nnode},
                                               {freqs, nfreq}});
    if (serve_cached(*deck, key, buf, cap, len)) return AXG_OK;
    std::string o;
    auto time0 = [&](std::string &s) { rust_exp(s, 0.0, 7, 14); };
    auto e14 = [](std::string &s, double v) {
        s += ' ';
        rust_exp(s, v, 6, 14);
    };
    o += "\n displacements (vx,vy,vz) for set NALL and time  ";
    time0(o);
    o += "\n\n";
    format_rows(o, d.nnode, [&](int64_t k, std::string &s) {
        rust_int(s, d.node_ids[k], 10);
        for (int c = 0; c < 3; c++) e14(s, u_full[3 * k + c]);
        s += '\n';
    });
    o += "\n forces (fx,fy,fz) for set NALL and time  ";
    time0(o);
    o += "\n\n";
    format_rows(o, d.nnode, [&](int64_t k, std::string &s) {
        const double *r = rf + 3 * k;
        if (std::fabs(r[0]) + std::fabs(r[1]) + std::fabs(r[2]) > 1e-14) {
            rust_int(s, d.node_ids[k], 10);
            for (int c = 0; c < 3; c++) e14(s, r[c]);
            s += '\n';
        }
    });
    o += "\n stresses (elem, integ.pnt.,sxx,syy,szz,sxy,syz,szx) for set EALL and time  ";
    time0(o);
    o += "\n\n";
    if (stress_gp)
        format_rows(o, d.nelem, [&](int64_t e, std::string &s) {
            rust_int(s, d.elements[e].id, 10);
            s += ' ';
            rust_int(s, 1, 4);
            for (int c = 0; c < 6; c++) e14(s, stress_gp[6 * e + c]);
            s += '\n';
        });
    o += '\n';
    if (nodal_stress) {  // append_nodal_stress, dat.rs:70-83
        o += "\n stresses (node,sxx,syy,szz,sxy,syz,szx) for set NALL\n\n";
        format_rows(o, d.nnode, [&](int64_t k, std::string &s) {
            rust_int(s, d.node_ids[k], 10);
            for (int c = 0; c < 6; c++) e14(s, nodal_stress[6 * k + c]);
            s += '\n';
        });
        o += '\n';
    }
    if (freqs && nfreq > 0) {  // append_frequencies, dat.rs:85-94
        o += "\n EIGENVALUE OUTPUT\n\n MODE  FREQUENCY (CYCLES/TIME)\n\n";
        for (int32_t i = 0; i < nfreq; i++) {
            rust_int(o, i + 1, 10);
            e14(o, freqs[i]);
            o += '\n';
        }
        o += '\n';
    }
    copy_out(*deck, key, o, buf, cap, len);
    API_END
}

// frd.rs:42-92: an integer right-aligned in [start, start + width) (the last
// width characters when longer)
static void put_int(std::string &line, size_t start, size_t width, int64_t v) {
    std::string t;
    rust_int(t, v, (int)width);
    const size_t n = std::min(t.size(), width);
    line.replace(start + width - n, n, t, t.size() - n, n);
}

// frd.rs:173-452 for one static frame (time 1, increment 1): the heading and
// fixed header records, the nodes used by elements (2C) and the elements (3C,
// C3D20 midsides in FRD order), then DISP, STRESS, FORC and TOSTRAIN blocks
// as the deck's *NODE FILE / *EL FILE select them, each with its 1PSTEP and
// 100CL records. stress, rf, strain may be NULL (zeros); without a strain the
// linear path's rule applies: strain from the nodal stress with the first
// material's E, nu (analysis.rs:1163-1174).
int axg_deck_write_frd(axg_deck *deck, const double *u_full, const double *stress,
                       const double *rf, const double *strain, char *buf, int64_t cap,
                       int64_t *len) {
    API_BEGIN
    AXG_REQUIRE(deck && u_full && len, AXG_E_INVALID, "axg_deck_write_frd: NULL argument");
    const axg_deck &d = *deck;
    const uint64_t key = writer_key(2, 0, {{u_full, d.ndof}, {stress, 6 * d.nnode}, {rf, d.ndof},
                                          {strain, 6 * d.nnode}});
    if (serve_cached(*deck, key, buf, cap, len)) return AXG_OK;
    std::string o;
    o.reserve(1 << 16);
    // the heading's first 66 characters (UTF-8 code points)
    std::string h = d.heading.empty() ? std::string("Axia") : d.heading;
    {
        size_t cut = 0, count = 0;
        while (cut < h.size() && count < 66) {
            const unsigned char c = (unsigned char)h[cut];
            cut += c < 0x80 ? 1 : c < 0xE0 ? 2 : c < 0xF0 ? 3 : 4;
            count++;
        }
        h = h.substr(0, std::min(cut, h.size()));
    }
    o += "    1C" + h + "\n";
    o += "    1UDATE              24.September.2026\n";
    o += "    1UTIME              21:00:00\n";
    o += "    1UHOST              axia\n";
    o += "    1UPGM               Axia FEM\n";
    o += "    1UVERSION           1.31.10\n";
    o += "    1UCODE              CalculiX-compatible Axia FEM\n";
    // frd_output_nodes (frd.rs:134-170): the nodes elements use, in node order
    std::vector<int64_t> out;
    {
        std::vector<uint8_t> used(d.nnode, 0);
        for (int32_t c : d.conn) used[c] = 1;
        for (int64_t i = 0; i < d.nnode; i++)
            if (used[i]) out.push_back(i);
        if (out.empty())
            for (int64_t i = 0; i < d.nnode; i++) out.push_back(i);
    }
    const int64_t nout = (int64_t)out.size();
    auto block_header = [&](char kind, int64_t n) {  // frd.rs:52-55
        o += "    ";
        o += kind;
        o += 'C';
        o.append(18, ' ');
        rust_int(o, n, 12);
        rust_int(o, 1, 38);
        o += '\n';
    };
    auto e12 = [](std::string &s, double v) {
        char t[12];
        frd_e12(v, t);
        s.append(t, 12);
    };
    block_header('2', nout);
    format_rows(o, nout, [&](int64_t k, std::string &s) {
        const int64_t i = out[k];
        s += " -1";
        rust_int(s, d.node_ids[i], 10);
        for (int q = 0; q < 3; q++) e12(s, d.coords[3 * i + q]);
        s += '\n';
    });
    o += " -3\n";
    block_header('3', d.nelem);
    const int ty = d.nn == 8 ? 1 : 4;  // ElemKind::frd_type (model.rs:364-369)
    format_rows(o, d.nelem, [&](int64_t e, std::string &s) {
        const DeckElem &el = d.elements[e];
        s += " -1";
        rust_int(s, el.id, 10);
        rust_int(s, ty, 5);
        rust_int(s, 0, 5);
        rust_int(s, 1, 5);
        s += '\n';
        // frd_nodes (frd.rs:110-132): C3D20 INP 1-12, 17-20, 13-16
        static const int k20[20] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 16, 17, 18, 19, 12, 13, 14, 15};
        for (int a0 = 0; a0 < d.nn; a0 += 10) {  // write_minus2, 10 per line
            s += " -2";
            for (int a = a0; a < std::min(d.nn, a0 + 10); a++)
                rust_int(s, el.nodes[d.nn == 20 ? k20[a] : a], 10);
            s += '\n';
        }
    });
    o += " -3\n";
    int counter = 0;
    const double time = 1.0;
    auto result_block = [&](const char *name, int ncomp, const std::vector<std::array<int, 3>> &meta,
                            const std::vector<const char *> &cnames, bool last_all,
                            const std::function<double(int64_t, int)> &val) {
        counter++;
        std::string p(70, ' ');  // 1PSTEP (frd.rs:57-64)
        p.replace(0, 10, "    1PSTEP");
        put_int(p, 24, 12, counter);
        put_int(p, 36, 12, 1);
        put_int(p, 48, 12, 1);
        o += p + "\n";
        std::string c(75, ' ');  // 100CL (frd.rs:81-92)
        c.replace(0, 7, "  100CL");
        put_int(c, 7, 5, 101);
        char t[12];
        frd_e12(time, t);
        c.replace(12, 12, t, 12);
        put_int(c, 24, 12, nout);
        c[57] = '0';
        put_int(c, 58, 5, 1);
        c[74] = '1';
        o += c + "\n";
        std::string nm(name);
        o += " -4  " + nm + std::string(nm.size() < 8 ? 8 - nm.size() : 0, ' ');
        rust_int(o, ncomp, 5);
        o += "    1\n";
        for (size_t i = 0; i < meta.size(); i++) {
            std::string cn(cnames[i]);
            o += " -5  " + cn + std::string(cn.size() < 8 ? 8 - cn.size() : 0, ' ') + "    1";
            rust_int(o, meta[i][0], 5);
            rust_int(o, meta[i][1], 5);
            rust_int(o, meta[i][2], 5);
            if (last_all && i + 1 == meta.size()) o += "    1ALL";
            o += '\n';
        }
        const int nvals = last_all ? ncomp - 1 : ncomp;
        format_rows(o, nout, [&](int64_t k, std::string &s) {
            const int64_t i = out[k];
            s += " -1";
            rust_int(s, d.node_ids[i], 10);
            for (int q = 0; q < nvals; q++) e12(s, val(i, q));
            s += '\n';
        });
        o += " -3\n";
    };
    if (d.output_u)
        result_block("DISP", 4, {{2, 1, 0}, {2, 2, 0}, {2, 3, 0}, {2, 0, 0}},
                     {"D1", "D2", "D3", "ALL"}, true,
                     [&](int64_t i, int q) { return u_full[3 * i + q]; });
    if (d.output_s)
        result_block("STRESS", 6, {{4, 1, 1}, {4, 2, 2}, {4, 3, 3}, {4, 1, 2}, {4, 2, 3}, {4, 3, 1}},
                     {"SXX", "SYY", "SZZ", "SXY", "SYZ", "SZX"}, false,
                     [&](int64_t i, int q) { return stress ? stress[6 * i + q] : 0.0; });
    if (d.output_rf)
        result_block("FORC", 4, {{2, 1, 0}, {2, 2, 0}, {2, 3, 0}, {2, 0, 0}},
                     {"F1", "F2", "F3", "ALL"}, true,
                     [&](int64_t i, int q) { return rf ? rf[3 * i + q] : 0.0; });
    if (d.output_e) {
        std::vector<double> e_lin;
        const double *ev = strain;
        if (!ev && stress) {
            AXG_REQUIRE(!d.materials.empty(), AXG_E_INVALID, "Kein Material");
            const double e = d.first_material().first, nu = d.first_material().second;
            e_lin.resize((size_t)d.nnode * 6);
            for (int64_t i = 0; i < d.nnode; i++) {
                const double *sv = stress + 6 * i;
                const double tr = sv[0] + sv[1] + sv[2];
                const double g2 = e / (1.0 + nu);
                for (int c = 0; c < 3; c++) e_lin[6 * i + c] = ((1.0 + nu) * sv[c] - nu * tr) / e;
                for (int c = 3; c < 6; c++) e_lin[6 * i + c] = sv[c] / g2;
            }
            ev = e_lin.data();
        }
        result_block("TOSTRAIN", 6, {{4, 1, 1}, {4, 2, 2}, {4, 3, 3}, {4, 1, 2}, {4, 2, 3}, {4, 3, 1}},
                     {"EXX", "EYY", "EZZ", "EXY", "EYZ", "EZX"}, false,
                     [&](int64_t i, int q) { return ev ? ev[6 * i + q] : 0.0; });
    }
    o += " 9999\n";
    copy_out(*deck, key, o, buf, cap, len);
    API_END
}

}  // extern "C"
