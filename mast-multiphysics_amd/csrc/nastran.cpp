// Nastran bulk-data reader: the mesh, constraints, loads and properties of a
// .bdf model as a mastg_problem_desc.
//
// MAST::NastranIO::read (src/mesh/nastran_io.cpp:217-255) reads a BDF through
// pyNastran (src/mesh/pynastran_io.pyx) into a libMesh mesh: GRID points as
// nodes whose libMesh id is the Nastran id (read_nodes, :120-143), elements
// by type with libMesh id = EID and subdomains by (PID, type) (read_elements,
// :146-206), and every SPC set as a node boundary (read_node_boundaries,
// :209-223).  libMesh then numbers the DOFs over the elements in id order.
// This reader parses the cards itself (pyNastran is a Python dependency that
// is not part of the reference) and returns the same mesh in that order:
// nodes sorted by Nastran id, elements sorted by EID, DOF bases handed out
// first-touch over the elements with a node's 6 structural variables
// contiguous.  It also turns the SPC / SPC1 components into constrained DOFs
// (homogeneous only), FORCE cards into nodal loads, MAT1 into E and nu, and
// PSHELL thicknesses into the per-element parameter of a plate.
//
// Cards: GRID (basic coordinates), CHEXA (8 nodes), CQUAD4, SPC1 (with THRU),
// SPC, FORCE (basic frame), MAT1, PSHELL, PSOLID; small-field, large-field
// (GRID*) and free-field (comma) formats with continuations; Nastran reals
// such as 1.-3 or 7.2+10.  Other cards are skipped.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "mastg/mastg.h"

struct mastg_mesh_file {
    mastg_problem_desc d{};
    std::vector<double> xyz, point_loads, params;
    std::vector<int32_t> conn;
    std::vector<int64_t> base, constrained, node_ids, elem_ids;
    std::vector<int32_t> elem_pid;
    int n_spc_sets = 0;
};

namespace {

thread_local std::string g_nastran_error;

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) a++;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) b--;
    return s.substr(a, b - a);
}

std::string upper(std::string s) {
    for (auto& ch : s) ch = char(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

// a Nastran real: 1.5, -2.E4, 1.-3 (= 1e-3), 7.2+10, 3 (an integer field read as real)
bool parse_real(const std::string& f, double* out) {
    std::string s = upper(trim(f));
    if (s.empty()) return false;
    // insert the exponent letter of the implicit-exponent form (a sign after
    // the mantissa that does not follow an E or D)
    for (size_t i = 1; i < s.size(); i++) {
        if ((s[i] == '+' || s[i] == '-') && s[i - 1] != 'E' && s[i - 1] != 'D') {
            s.insert(i, "E");
            break;
        }
    }
    for (auto& ch : s)
        if (ch == 'D') ch = 'E';
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (!end || *end != '\0') return false;
    *out = v;
    return true;
}

bool parse_int(const std::string& f, int64_t* out) {
    const std::string s = trim(f);
    if (s.empty()) return false;
    char* end = nullptr;
    const long long v = std::strtoll(s.c_str(), &end, 10);
    if (!end || *end != '\0') return false;
    *out = v;
    return true;
}

struct Card {
    std::string name;
    std::vector<std::string> f;   // data fields (field 2 onwards), continuations appended
    int line = 0;
    const std::string& at(size_t i) const {
        static const std::string empty;
        return i < f.size() ? f[i] : empty;
    }
};

// splits the bulk data into cards
bool read_cards(std::istream& in, std::vector<Card>& cards, std::string& err) {
    std::string raw;
    std::vector<std::string> lines;
    std::vector<int> line_no;
    int ln = 0;
    while (std::getline(in, raw)) {
        ln++;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        const size_t c = raw.find('$');
        if (c != std::string::npos) raw = raw.substr(0, c);
        const std::string t = upper(trim(raw));
        if (t.empty()) continue;
        if (t.rfind("BEGIN", 0) == 0 && t.find("BULK") != std::string::npos) {
            // executive and case control before BEGIN BULK are not read
            lines.clear();
            line_no.clear();
            continue;
        }
        if (t.rfind("ENDDATA", 0) == 0) break;
        lines.push_back(raw);
        line_no.push_back(ln);
    }
    for (size_t k = 0; k < lines.size(); k++) {
        const std::string& s = lines[k];
        std::string name;
        std::vector<std::string> fields;
        bool large = false;
        if (s.find(',') != std::string::npos) {
            std::stringstream ss(s);
            std::string tok;
            bool first = true;
            while (std::getline(ss, tok, ',')) {
                if (first) name = trim(tok);
                else fields.push_back(trim(tok));
                first = false;
            }
            if (!s.empty() && s.back() == ',') fields.push_back("");
        } else {
            name = trim(s.substr(0, std::min<size_t>(8, s.size())));
            // large field: a card name ending in '*', or its continuation lines,
            // which start with '*'
            large = !name.empty() && (name.back() == '*' || name[0] == '*');
            const size_t w = large ? 16 : 8, nf = large ? 4 : 8;
            for (size_t i = 0; i < nf; i++) {
                const size_t p = 8 + i * w;
                fields.push_back(p < s.size() ? trim(s.substr(p, std::min(w, s.size() - p))) : std::string());
            }
        }
        const bool cont = name.empty() || name[0] == '+' || name[0] == '*';
        if (cont) {
            if (cards.empty()) {
                err = "continuation line " + std::to_string(line_no[k]) + " without a card";
                return false;
            }
            auto& f = cards.back().f;
            f.insert(f.end(), fields.begin(), fields.end());
        } else {
            Card c;
            c.name = upper(large && name.back() == '*' ? name.substr(0, name.size() - 1) : name);
            c.f = std::move(fields);
            c.line = line_no[k];
            cards.push_back(std::move(c));
        }
    }
    return true;
}

int fail(const std::string& m) {
    g_nastran_error = m;
    return MASTG_ERR_INVALID;
}

}  // namespace

extern "C" {

const char* mastg_nastran_last_error(void) { return g_nastran_error.c_str(); }

int mastg_nastran_read(const char* path, mastg_mesh_file** out) {
    if (!path || !out) return MASTG_ERR_INVALID;
    std::ifstream in(path);
    if (!in) return fail(std::string("cannot open ") + path);
    std::vector<Card> cards;
    std::string err;
    if (!read_cards(in, cards, err)) return fail(err);

    struct Grid { double x[3]; };
    std::map<int64_t, Grid> grids;
    struct Elem { int64_t pid; std::vector<int64_t> g; };
    std::map<int64_t, Elem> hexa, quad;
    // sid -> (grid, components, from a THRU range: a grid of such a range
    // need not exist, a listed one must)
    struct Spc { int64_t g; std::string comp; bool thru; };
    std::map<int64_t, std::vector<Spc>> spcs;
    struct Force { int64_t g; double f[3]; };
    std::vector<Force> forces;
    std::map<int64_t, double> pshell_t;
    double E = 0, nu = -2;
    bool have_mat = false;
    auto bad = [&](const Card& c, const std::string& what) {
        return fail(c.name + " card at line " + std::to_string(c.line) + ": " + what);
    };
    for (const Card& c : cards) {
        if (c.name == "GRID") {
            int64_t id, cp = 0;
            if (!parse_int(c.at(0), &id)) return bad(c, "grid id");
            if (!c.at(1).empty() && (!parse_int(c.at(1), &cp) || cp != 0))
                return bad(c, "only the basic coordinate system (CP blank or 0) is read");
            Grid g{};
            for (int i = 0; i < 3; i++)
                if (!c.at(2 + i).empty() && !parse_real(c.at(2 + i), &g.x[i])) return bad(c, "coordinate");
            if (grids.count(id)) return bad(c, "duplicate grid id");
            grids[id] = g;
        } else if (c.name == "CHEXA" || c.name == "CQUAD4") {
            const int nn = c.name == "CHEXA" ? 8 : 4;
            int64_t eid, pid;
            if (!parse_int(c.at(0), &eid) || !parse_int(c.at(1), &pid)) return bad(c, "element or property id");
            Elem e{pid, {}};
            for (int i = 0; i < nn; i++) {
                int64_t g;
                if (!parse_int(c.at(2 + i), &g)) return bad(c, "node id");
                e.g.push_back(g);
            }
            if (nn == 8) {
                for (size_t i = 10; i < c.f.size(); i++)
                    if (!trim(c.f[i]).empty()) return bad(c, "only 8-node CHEXA elements are read");
            }
            auto& m = nn == 8 ? hexa : quad;
            if (m.count(eid)) return bad(c, "duplicate element id");
            m[eid] = e;
        } else if (c.name == "SPC1") {
            int64_t sid;
            if (!parse_int(c.at(0), &sid)) return bad(c, "set id");
            const std::string comp = trim(c.at(1));
            std::vector<Spc>& set = spcs[sid];
            for (size_t i = 2; i < c.f.size(); i++) {
                const std::string t = upper(trim(c.f[i]));
                if (t.empty()) continue;
                if (t == "THRU") {
                    int64_t hi;
                    if (set.empty() || i + 1 >= c.f.size() || !parse_int(c.f[i + 1], &hi) || hi < set.back().g)
                        return bad(c, "THRU range");
                    set.back().thru = true;
                    for (int64_t v = set.back().g + 1; v <= hi; v++) set.push_back(Spc{v, comp, true});
                    i++;
                    continue;
                }
                int64_t v;
                if (!parse_int(t, &v)) return bad(c, "grid id");
                set.push_back(Spc{v, comp, false});
            }
        } else if (c.name == "SPC") {
            int64_t sid;
            if (!parse_int(c.at(0), &sid)) return bad(c, "set id");
            for (size_t k = 1; k + 1 < c.f.size(); k += 3) {
                int64_t g;
                if (trim(c.f[k]).empty()) continue;
                if (!parse_int(c.f[k], &g)) return bad(c, "grid id");
                double dv = 0;
                if (k + 2 < c.f.size() && !trim(c.f[k + 2]).empty() && (!parse_real(c.f[k + 2], &dv) || dv != 0))
                    return bad(c, "only homogeneous constraints (D = 0) are read");
                spcs[sid].push_back(Spc{g, trim(c.f[k + 1]), false});
            }
        } else if (c.name == "FORCE") {
            int64_t g, cid = 0;
            double F, n[3] = {0, 0, 0};
            if (!parse_int(c.at(1), &g) || !parse_real(c.at(3), &F)) return bad(c, "grid or magnitude");
            if (!c.at(2).empty() && (!parse_int(c.at(2), &cid) || cid != 0))
                return bad(c, "only the basic frame (CID blank or 0) is read");
            for (int i = 0; i < 3; i++)
                if (!c.at(4 + i).empty() && !parse_real(c.at(4 + i), &n[i])) return bad(c, "direction");
            Force fr{g, {F * n[0], F * n[1], F * n[2]}};
            forces.push_back(fr);
        } else if (c.name == "MAT1") {
            double G = 0;
            const bool hE = parse_real(c.at(1), &E), hG = parse_real(c.at(2), &G), hn = parse_real(c.at(3), &nu);
            if (!hE) return bad(c, "Young's modulus");
            if (!hn) {
                if (!hG) return bad(c, "needs NU or G");
                nu = E / (2 * G) - 1;
            }
            if (have_mat) return bad(c, "one material is read");
            have_mat = true;
        } else if (c.name == "PSHELL") {
            int64_t pid;
            double t;
            if (!parse_int(c.at(0), &pid) || !parse_real(c.at(2), &t)) return bad(c, "property id or thickness");
            pshell_t[pid] = t;
        }
    }
    if (grids.empty()) return fail("no GRID cards");
    if (hexa.empty() == quad.empty()) return fail("the model must hold CHEXA or CQUAD4 elements, not both or neither");
    const bool plate = !quad.empty();
    const auto& elems = plate ? quad : hexa;
    const int npe = plate ? 4 : 8;

    auto* m = new mastg_mesh_file;
    // nodes in Nastran id order
    std::map<int64_t, int32_t> node_index;
    for (auto& kv : grids) {
        node_index[kv.first] = int32_t(m->node_ids.size());
        m->node_ids.push_back(kv.first);
        for (int i = 0; i < 3; i++) m->xyz.push_back(kv.second.x[i]);
    }
    // elements in EID order
    for (auto& kv : elems) {
        m->elem_ids.push_back(kv.first);
        m->elem_pid.push_back(int32_t(kv.second.pid));
        for (int64_t g : kv.second.g) {
            auto it = node_index.find(g);
            if (it == node_index.end()) {
                delete m;
                return fail("element " + std::to_string(kv.first) + " references missing grid " + std::to_string(g));
            }
            m->conn.push_back(it->second);
        }
        if (plate) {
            auto t = pshell_t.find(kv.second.pid);
            if (t == pshell_t.end()) {
                delete m;
                return fail("element " + std::to_string(kv.first) + ": no PSHELL for property " +
                            std::to_string(kv.second.pid));
            }
            m->params.push_back(t->second);
        } else {
            m->params.push_back(1.0);
        }
    }
    const int64_t n_nodes = int64_t(m->node_ids.size()), n_elems = int64_t(m->elem_ids.size());
    // first-touch DOF bases, 6 structural variables per node
    const int nv = 6;
    m->base.assign(n_nodes, -1);
    int64_t next = 0;
    for (int64_t e = 0; e < n_elems; e++)
        for (int a = 0; a < npe; a++) {
            const int32_t nd = m->conn[e * npe + a];
            if (m->base[nd] < 0) {
                m->base[nd] = next;
                next += nv;
            }
        }
    // nodes no element uses still get DOFs (after the others, in id order)
    for (int64_t nd = 0; nd < n_nodes; nd++)
        if (m->base[nd] < 0) {
            m->base[nd] = next;
            next += nv;
        }
    const int64_t n_dofs = next;
    std::set<int64_t> con;
    for (auto& set : spcs) {
        m->n_spc_sets++;
        for (auto& gc : set.second) {
            auto it = node_index.find(gc.g);
            if (it == node_index.end()) {
                if (gc.thru) continue;   // grids of a THRU range need not exist
                delete m;
                return fail("SPC set " + std::to_string(set.first) + " references missing grid " +
                            std::to_string(gc.g));
            }
            for (char ch : gc.comp) {
                if (ch < '1' || ch > '6') {
                    delete m;
                    return fail("SPC component " + gc.comp + " is not a subset of 123456");
                }
                con.insert(m->base[it->second] + (ch - '1'));
            }
        }
    }
    m->constrained.assign(con.begin(), con.end());
    m->point_loads.assign(n_dofs, 0.0);
    for (auto& f : forces) {
        auto it = node_index.find(f.g);
        if (it == node_index.end()) {
            delete m;
            return fail("FORCE references missing grid " + std::to_string(f.g));
        }
        for (int i = 0; i < 3; i++) m->point_loads[m->base[it->second] + i] += f.f[i];
    }
    mastg_problem_desc& d = m->d;
    d.physics = plate ? MASTG_QUAD4_MINDLIN : MASTG_HEX8_ELASTIC;
    d.n_nodes = n_nodes;
    d.n_elems = n_elems;
    d.xyz = m->xyz.data();
    d.conn = m->conn.data();
    d.n_vars = nv;
    d.node_dof_base = m->base.data();
    d.n_constrained = int64_t(m->constrained.size());
    d.constrained_dofs = m->constrained.data();
    d.point_loads = m->point_loads.data();
    d.E = have_mat ? E : 0.0;
    d.nu = have_mat ? nu : 0.0;
    d.penalty = 3.0;
    d.kappa = 5.0 / 6.0;
    d.offset = 0.0;
    *out = m;
    return MASTG_OK;
}

const mastg_problem_desc* mastg_nastran_desc(const mastg_mesh_file* m) { return m ? &m->d : nullptr; }
const double* mastg_nastran_params(const mastg_mesh_file* m) { return m ? m->params.data() : nullptr; }

int mastg_nastran_ids(const mastg_mesh_file* m, int64_t* node_ids, int64_t* elem_ids, int32_t* elem_pids) {
    if (!m) return MASTG_ERR_INVALID;
    if (node_ids) std::copy(m->node_ids.begin(), m->node_ids.end(), node_ids);
    if (elem_ids) std::copy(m->elem_ids.begin(), m->elem_ids.end(), elem_ids);
    if (elem_pids) std::copy(m->elem_pid.begin(), m->elem_pid.end(), elem_pids);
    return MASTG_OK;
}

int mastg_nastran_spc_sets(const mastg_mesh_file* m) { return m ? m->n_spc_sets : 0; }

int mastg_nastran_destroy(mastg_mesh_file* m) {
    delete m;
    return MASTG_OK;
}

}  // extern "C"
