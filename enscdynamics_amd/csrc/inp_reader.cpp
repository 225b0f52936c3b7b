// Abaqus .inp reader (include/ensc_deck.h).
//
// Restates the semantics of the reference's DataIn
// (Explicit3D/source/dataIn.cpp:300-1788) over a list of lines: a line that
// starts with '*' (and not "**") is a keyword line, upper-cased, dispatched
// first on the structural keywords (*HEADING/*PART/*ASSEMBLY/*MATERIAL/
// *STEP, 343-398) and then on the first handler key that is a substring of
// the line, in the byte order of the reference's std::map (401-409); other
// keywords fall to the per-state handling (412-519).  Data lines are consumed
// by the handlers; stray data lines are skipped.  Set and instance
// resolution follows transferToExDyna (1550-1788): only assembly-level sets
// reach the solver, instance-local ids are offset by the instance's first
// node, and sets of several instances are concatenated in instance-name order.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ensc_deck.h"
#include "guard.h"

namespace ensc {
extern thread_local std::string g_last_error;
}

namespace {

using std::map;
using std::string;
using std::vector;

struct Point3 {
    double v[3];
};

struct Part {
    string name;
    vector<Point3> nodes;
    vector<vector<uint32_t>> elems;
};

struct Instance {
    string name, part;
    double tr[3] = {0, 0, 0};
    size_t node0 = 0, elem0 = 0;
};

struct Assembly {
    string name;
    vector<Instance> inst;
    map<string, map<string, vector<size_t>>> nsets, esets;
};

struct SpcLine {
    string set;
    size_t d0, d1, cond;
};
struct VelLine {
    string set;
    size_t d0, d1;
    double value;
    string amp;
};
struct DsloadLine {
    string surf, amp;
    double value;
};
struct StepRec {
    string name;
    double period = 0.0;
    vector<SpcLine> spc;
    vector<VelLine> vel;
    vector<DsloadLine> dsl;
    bool reset_spc = false, reset_vel = false, reset_dload = false, reset_dsload = false;
    bool grav = false;
    string grav_amp;
    double grav_v = 0, gdir[3] = {0, 0, 0};
    size_t n_dsload = 0;
};

struct IniVel {
    string ref;
    size_t dof;
    double value;
};

enum State { GLOBAL, PART, ASSEMBLY, STEP, MATERIAL, PREDEFINED };

struct Parser {
    vector<string> L;
    size_t pos = 0;
    string str;
    State state = GLOBAL;
    vector<Part> parts;
    vector<Assembly> assemblies;
    Part* cur_part = nullptr;
    Assembly* cur_asm = nullptr;
    vector<StepRec> steps;
    vector<IniVel> ini;
    map<string, vector<std::pair<double, double>>> amps;
    // *Surface, type=ELEMENT: name -> (element set, face 0..5 for S1..S6)
    map<string, vector<std::pair<string, int>>> surfaces;
    double E = 2.06e11, nu = 0.3, rho = 7850.0;  // Material_elastic.h defaults
    double cvisl = 0.06, cvisq = 1.2, total_time = 0.0;
    bool plastic = false;
    double yield0 = 0.0, hard = 0.0;
    string err;

    bool more() const { return pos < L.size(); }
    int peek() const { return more() ? (L[pos].empty() ? '\n' : (unsigned char)L[pos][0]) : -1; }
    bool getline() {
        if (!more()) return false;
        str = L[pos++];
        return true;
    }

    static void upper(string& s) {
        for (char& c : s) c = (char)std::toupper((unsigned char)c);
    }
    static string nospace(string s) {
        s.erase(std::remove_if(s.begin(), s.end(), [](unsigned char c) { return std::isspace(c); }), s.end());
        return s;
    }
    // extractNameFromKeyword (dataIn.cpp:100-115)
    static string name_of(const string& line, const string& prefix = "NAME=") {
        size_t p = line.find(prefix);
        if (p == string::npos) return "";
        p += prefix.size();
        size_t e = line.find(',', p);
        if (e == string::npos) e = line.size();
        return nospace(line.substr(p, e - p));
    }
    // std::getline(iss, token, ',') over a line: successive fields
    struct Fields {
        const string& s;
        size_t p = 0;
        bool done = false;
        explicit Fields(const string& x) : s(x) {}
        bool next(string& out) {
            out.clear();
            if (done || p > s.size() || (p == s.size())) {
                done = true;
                return false;
            }
            size_t e = s.find(',', p);
            if (e == string::npos) {
                out = s.substr(p);
                p = s.size() + 1;
            } else {
                out = s.substr(p, e - p);
                p = e + 1;
            }
            return true;
        }
    };
    static bool to_double(const string& t, double& v) {
        const char* b = t.c_str();
        char* e = nullptr;
        v = std::strtod(b, &e);
        return e != b;
    }
    static bool to_long(const string& t, long& v) {
        const char* b = t.c_str();
        char* e = nullptr;
        v = std::strtol(b, &e, 10);
        return e != b;
    }
    static bool to_ulong(const string& t, unsigned long long& v) {
        const char* b = t.c_str();
        char* e = nullptr;
        v = std::strtoull(b, &e, 10);
        return e != b;
    }

    // whitespace tokens with one trailing comma dropped, leading integer
    // parsed (dataIn.cpp:686-696)
    static void id_tokens(const string& line, vector<size_t>& out) {
        std::istringstream iss(line);
        string tok;
        while (iss >> tok) {
            if (!tok.empty() && tok.back() == ',') tok.pop_back();
            unsigned long long v;
            if (to_ulong(tok, v)) out.push_back((size_t)v);
        }
    }

    void NODE() {
        if (!cur_part) return;
        while (peek() == '*') getline();
        while (more() && peek() != '*') {
            getline();
            if (str.empty()) continue;
            Fields f(str);
            string t;
            Point3 x{{0, 0, 0}};
            f.next(t);
            for (int d = 0; d < 3; ++d) {
                f.next(t);
                to_double(t, x.v[d]);
            }
            cur_part->nodes.push_back(x);
        }
    }

    void ELEMENT() {
        if (!cur_part) return;
        while (peek() == '*') getline();
        while (more() && peek() != '*') {
            getline();
            if (str.empty()) continue;
            Fields f(str);
            string t;
            f.next(t);
            vector<uint32_t> v;
            while (f.next(t)) {
                unsigned long long n;
                if (to_ulong(t, n)) v.push_back((uint32_t)(n - 1));
            }
            cur_part->elems.push_back(v);
        }
    }

    void SET(bool node) {
        string set = name_of(str, "NSET=");
        if (set.empty()) set = name_of(str, "ELSET=");
        const string inst = name_of(str, "INSTANCE=");
        const bool gen = str.find("GENERATE") != string::npos;
        vector<size_t> ids;
        while (more() && peek() != '*') {
            getline();
            if (str.empty()) continue;
            vector<size_t> lv;
            id_tokens(str, lv);
            if (!node && lv.size() == 3) {
                for (size_t i = lv[0]; i <= lv[1]; i += lv[2]) ids.push_back(i);
            } else {
                ids.insert(ids.end(), lv.begin(), lv.end());
            }
        }
        if (node && gen && ids.size() >= 2) {
            const size_t a = ids[0], b = ids[1], s = ids.size() >= 3 ? ids[2] : 1;
            ids.clear();
            for (size_t i = a; i <= b && s > 0; i += s) ids.push_back(i);
        }
        for (auto& v : ids) v -= 1;
        if (cur_asm && state == ASSEMBLY) {
            auto& dst = node ? cur_asm->nsets : cur_asm->esets;
            if (!inst.empty()) {
                dst[set][inst] = ids;
            } else {
                map<string, vector<size_t>> g;
                g["__GLOBAL__"] = ids;
                dst[set] = g;
            }
        }
        // part-level sets never reach the solver (dataIn.cpp:704-706, 1653)
    }

    StepRec* bc_target() {
        if (state == STEP && !steps.empty()) return &steps.back();
        return steps.empty() ? nullptr : &steps.front();
    }

    void BOUNDARY() {
        const bool op_new = str.find("OP=NEW") != string::npos;
        const bool vel = str.find("TYPE=VELOCITY") != string::npos;
        const string amp = name_of(str, "AMPLITUDE=");
        if (op_new) {
            StepRec* s = (state == STEP && !steps.empty()) ? &steps.back() : (steps.empty() ? nullptr : &steps.front());
            if (s) (vel ? s->reset_vel : s->reset_spc) = true;
        }
        if (!more() || L[pos].empty() || L[pos][0] == '*') return;  // empty command
        if (vel)
            VELOCITY_NODE(amp);
        else
            SPC_NODE();
    }

    void SPC_NODE() {
        while (peek() == '*') getline();
        if (peek() == '*') return;
        while (more() && peek() != '*') {
            getline();
            upper(str);
            SpcLine b{"", 0, 2, 1};
            Fields f(str);
            string t;
            f.next(t);
            b.set = (str.find("SET-") != string::npos) ? t : "#" + t;
            if (str.find("PINNED") != string::npos || str.find("ENCASTRE") != string::npos) {
                b.d0 = 0, b.d1 = 2, b.cond = 1;
            } else if (str.find("XSYMM") != string::npos) {
                b.d0 = 0, b.d1 = 0, b.cond = 1;
            } else if (str.find("YSYMM") != string::npos) {
                b.d0 = 1, b.d1 = 1, b.cond = 1;
            } else if (str.find("ZSYMM") != string::npos) {
                b.d0 = 2, b.d1 = 2, b.cond = 1;
            } else {
                f.next(t);
                if (t.empty()) {
                    b.d0 = 0, b.d1 = 2, b.cond = 1;
                } else {
                    long v = 0;
                    to_long(t, v);
                    b.d0 = (size_t)(v - 1);
                    f.next(t);
                    if (t.empty()) {
                        b.d1 = b.d0;
                    } else {
                        to_long(t, v);
                        b.d1 = (size_t)(v - 1);
                    }
                    f.next(t);
                    if (t.empty()) {
                        b.cond = 1;
                    } else {
                        to_long(t, v);
                        b.cond = (size_t)v;
                    }
                }
            }
            if (StepRec* s = bc_target()) s->spc.push_back(b);
        }
    }

    void VELOCITY_NODE(const string& amp) {
        while (peek() == '*') getline();
        if (peek() == '*') return;
        while (more() && peek() != '*') {
            getline();
            upper(str);
            VelLine b{"", 0, 0, 0.0, amp};
            Fields f(str);
            string t;
            f.next(t);
            b.set = (str.find("SET-") != string::npos) ? t : "#" + t;
            f.next(t);
            if (t.empty()) continue;  // missing dof: the reference skips the line
            long v = 0;
            to_long(t, v);
            b.d0 = (size_t)(v - 1);
            f.next(t);
            if (t.empty()) {
                b.d1 = b.d0;
            } else {
                to_long(t, v);
                b.d1 = (size_t)(v - 1);
            }
            if (f.next(t)) to_double(t, b.value);
            if (StepRec* s = bc_target()) s->vel.push_back(b);
        }
    }

    void INITIAL_VELOCITY() {
        while (peek() == '*') getline();
        while (more() && peek() != '*') {
            getline();
            if (str.empty()) continue;
            upper(str);
            Fields f(str);
            string t;
            f.next(t);
            string ref = nospace(t);
            const bool is_set = ref.find("SET-") != string::npos || ref.find("NSET-") != string::npos;
            string inst;
            const size_t dot = ref.find('.');
            if (dot != string::npos) {
                inst = ref.substr(0, dot);
                ref = ref.substr(dot + 1);
            }
            IniVel iv;
            iv.ref = is_set ? (inst.empty() ? ref : inst + "." + ref) : "#" + ref;
            long dof = 1;
            if (f.next(t)) {
                if (!to_long(t, dof)) continue;  // the reference drops the entry
            }
            iv.dof = (size_t)(dof - 1);
            iv.value = 0.0;
            if (f.next(t)) {
                if (!to_double(t, iv.value)) continue;
            }
            ini.push_back(iv);
        }
    }

    void DLOAD() {
        const bool op_new = str.find("OP=NEW") != string::npos;
        if (op_new && state == STEP && !steps.empty()) steps.back().reset_dload = true;
        const string amp = name_of(str, "AMPLITUDE=");
        while (more() && peek() != '*') {
            getline();
            if (str.find("GRAV") == string::npos) continue;
            Fields f(str);
            string t;
            f.next(t);
            f.next(t);
            double vals[4] = {0, 0, 0, 0};
            for (int i = 0; i < 4; ++i) {
                f.next(t);
                to_double(t, vals[i]);
            }
            if (state == STEP && !steps.empty()) {
                StepRec& s = steps.back();
                s.grav = true;
                s.grav_amp = amp;
                s.grav_v = vals[0];
                s.gdir[0] = vals[1];
                s.gdir[1] = vals[2];
                s.gdir[2] = vals[3];
            }
        }
    }

    void DSLOAD() {
        const bool op_new = str.find("OP=NEW") != string::npos;
        if (op_new && state == STEP && !steps.empty()) steps.back().reset_dsload = true;
        const string amp = name_of(str, "AMPLITUDE=");
        while (more() && peek() != '*') {
            getline();
            upper(str);
            // surface, load type, value (dataIn.cpp:1441-1456); pressure (P) only
            Fields f(str);
            string surf, type, t;
            f.next(surf);
            f.next(type);
            f.next(t);
            surf = nospace(surf);
            type = nospace(type);
            double v = 0.0;
            to_double(t, v);
            if (state == STEP && !steps.empty()) {
                steps.back().n_dsload++;
                if (type == "P") steps.back().dsl.push_back({surf, amp, v});
            }
        }
    }

    void TIME() {
        while (peek() == '*') getline();
        while (more() && peek() != '*') {
            getline();
            Fields f(str);
            string t;
            f.next(t);
            f.next(t);
            double p = 0.0;
            to_double(t, p);
            if (state == STEP && !steps.empty()) steps.back().period = p;
            total_time += p;
        }
    }

    void AMP() {
        const string name = name_of(str);
        while (more() && peek() != '*') {
            getline();
            vector<string> fl;
            Fields f(str);
            string t;
            while (f.next(t)) fl.push_back(t);
            vector<std::pair<double, double>> tab(fl.size() / 2);
            for (size_t i = 0; i < tab.size(); ++i) {
                to_double(fl[2 * i], tab[i].first);
                to_double(fl[2 * i + 1], tab[i].second);
            }
            amps[name] = tab;  // each data line replaces the table (dataIn.cpp:1267)
        }
    }

    void BULK_VISCOSITY() {
        if (!getline()) return;
        Fields f(str);
        string t;
        f.next(t);
        to_double(t, cvisl);
        f.next(t);
        to_double(t, cvisq);
    }

    // *Surface, type=ELEMENT, name=S: lines "element set, Sn" (extension:
    // every line is kept; the reference keeps the last, dataIn.cpp:1339-1414)
    void SURFACE() {
        const string name = name_of(str, "NAME=");
        while (more() && peek() != '*') {
            getline();
            upper(str);
            Fields f(str);
            string set, face;
            f.next(set);
            f.next(face);
            set = nospace(set);
            face = nospace(face);
            if (name.empty() || set.empty() || face.size() != 2 || face[0] != 'S') continue;
            const int k = face[1] - '1';
            if (k >= 0 && k < 6) surfaces[name].push_back({set, k});
        }
    }

    void OUTPUT() {
        while (more() && peek() != '*') getline();
    }

    void INSTANCE() {
        if (!cur_asm) return;
        Instance in;
        in.name = name_of(str, "NAME=");
        in.part = name_of(str, "PART=");
        if (more() && peek() != '*') {
            std::istringstream iss(L[pos]);
            double x, y, z;
            bool ok = false;
            if (iss >> x) {
                if (iss.peek() == ',') iss.ignore(1);
                if (iss >> y) {
                    if (iss.peek() == ',') iss.ignore(1);
                    if (iss >> z) ok = true;
                }
            }
            if (ok) {
                in.tr[0] = x;
                in.tr[1] = y;
                in.tr[2] = z;
                ++pos;
            }
        }
        cur_asm->inst.push_back(in);
    }

    bool dispatch() {
        // handler keys in std::map (byte) order (dataIn.cpp:69-92, 402-409)
        static const char* keys[] = {"*AMPLITUDE", "*BOUNDARY", "*BULK VISCOSITY", "*DLOAD", "*DSLOAD",
                                     "*DYNAMIC", "*ELEMENT", "*ELSET", "*END ASSEMBLY", "*END INSTANCE",
                                     "*END PART", "*END STEP", "*INITIAL CONDITION", "*INSTANCE", "*MATERIAL",
                                     "*NODE", "*NSET", "*OUTPUT", "*OUTPUT INTERVAL", "*SURFACE"};
        for (const char* k : keys) {
            if (str.find(k) == string::npos) continue;
            const string key(k);
            if (key == "*AMPLITUDE") AMP();
            else if (key == "*BOUNDARY") BOUNDARY();
            else if (key == "*BULK VISCOSITY") BULK_VISCOSITY();
            else if (key == "*DLOAD") DLOAD();
            else if (key == "*DSLOAD") DSLOAD();
            else if (key == "*DYNAMIC") TIME();
            else if (key == "*ELEMENT") ELEMENT();
            else if (key == "*ELSET") SET(false);
            else if (key == "*END ASSEMBLY") {
                cur_asm = nullptr;
                state = GLOBAL;
            } else if (key == "*END INSTANCE") {
            } else if (key == "*END PART") {
                cur_part = nullptr;
                state = GLOBAL;
            } else if (key == "*END STEP") {
                state = GLOBAL;
            } else if (key == "*INITIAL CONDITION") {
                if (str.find("TYPE") != string::npos && str.find("VELOCITY") != string::npos) INITIAL_VELOCITY();
            } else if (key == "*INSTANCE") INSTANCE();
            else if (key == "*MATERIAL") {
                // never reached: *MATERIAL is a structural keyword
            } else if (key == "*NODE") NODE();
            else if (key == "*NSET") SET(true);
            else if (key == "*OUTPUT") OUTPUT();
            else if (key == "*SURFACE") SURFACE();
            return true;
        }
        return false;
    }

    void by_state() {
        switch (state) {
            case PART:
                if (str.find("*SOLID SECTION") != string::npos)
                    if (more() && peek() != '*') getline();
                break;
            case MATERIAL:
                if (str.find("*DENSITY") != string::npos) {
                    if (more() && peek() != '*') {
                        getline();
                        std::istringstream iss(str);
                        double d;
                        if (iss >> d) rho = d;
                    }
                } else if (str.find("*ELASTIC") != string::npos) {
                    if (more() && peek() != '*') {
                        getline();
                        Fields f(str);
                        string t;
                        f.next(t);
                        to_double(t, E);
                        f.next(t);
                        to_double(t, nu);
                    }
                } else if (str.find("*PLASTIC") != string::npos) {
                    // extension: J2 yield curve (yield stress, plastic strain) pairs
                    vector<std::pair<double, double>> pts;
                    while (more() && peek() != '*') {
                        getline();
                        Fields f(str);
                        string a, b;
                        if (!f.next(a)) continue;
                        f.next(b);
                        double s = 0, e = 0;
                        if (!to_double(a, s)) continue;
                        to_double(b, e);
                        pts.push_back({s, e});
                    }
                    if (!pts.empty()) {
                        plastic = true;
                        yield0 = pts[0].first;
                        hard = (pts.size() >= 2 && pts[1].second != pts[0].second)
                                   ? (pts[1].first - pts[0].first) / (pts[1].second - pts[0].second)
                                   : 0.0;
                    }
                }
                break;
            case STEP:
                if (str.find("*RESTART") != string::npos) {
                    while (more() && peek() != '*') getline();
                }
                break;
            default:
                break;
        }
    }

    void parse() {
        StepRec init;
        init.name = "Initial";
        steps.push_back(init);
        while (more()) {
            if (peek() != '*') {
                getline();
                continue;
            }
            getline();
            if (str.size() <= 1 || str[1] == '*') continue;
            upper(str);
            if (str.find("*HEADING") != string::npos) {
                while (more() && peek() != '*') getline();
                continue;
            }
            if (str.find("*PART") != string::npos) {
                parts.emplace_back();
                cur_part = &parts.back();
                cur_part->name = name_of(str);
                state = PART;
                continue;
            }
            if (str.find("*ASSEMBLY") != string::npos) {
                assemblies.emplace_back();
                cur_asm = &assemblies.back();
                cur_asm->name = name_of(str);
                if (cur_asm->name.empty()) cur_asm->name = "MainAssembly";
                state = ASSEMBLY;
                continue;
            }
            if (str.find("*MATERIAL") != string::npos) {
                state = MATERIAL;
                continue;
            }
            if (str.find("*STEP") != string::npos) {
                StepRec s;
                s.name = name_of(str);
                if (s.name.empty()) s.name = "Step-" + std::to_string(steps.size() + 1);
                steps.push_back(s);
                state = STEP;
                continue;
            }
            if (str.find("*INITIAL CONDITIONS") != string::npos) state = PREDEFINED;
            if (!dispatch()) by_state();
        }
    }
};

}  // namespace

struct ensc_deck {
    // flattened model
    std::vector<double> x0, v_init;
    std::vector<uint32_t> conn;
    std::map<std::string, std::vector<uint32_t>> nsets, esets;
    // surface pressures (ENSC_DECK_DSLOAD): per step, per line
    std::vector<std::vector<ensc_dsload_entry>> dsl;
    std::vector<std::vector<std::vector<uint32_t>>> dsl_elems;
    std::vector<std::vector<std::vector<int32_t>>> dsl_faces;
    std::vector<std::string> amp_names;
    std::vector<std::vector<double>> amp_t, amp_a;
    std::vector<ensc_amplitude> amps;
    std::vector<std::vector<ensc_bc_entry>> spc, vel;
    std::vector<std::vector<std::vector<uint32_t>>> spc_nodes, vel_nodes;
    std::vector<ensc_step_def> steps;
    ensc_model_desc desc{};
    std::string json;
};

namespace {

std::string jnum(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}
std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

bool build_deck(Parser& P, const ensc_project* proj, uint32_t flags, ensc_deck& D, std::string& err) {
    // transferToExDyna (dataIn.cpp:1550-1788)
    std::vector<Point3> verts;
    std::vector<std::vector<uint32_t>> elems;
    if (!P.assemblies.empty()) {
        for (auto& A : P.assemblies) {
            size_t gn = 0, ge = 0;
            for (auto& I : A.inst) {
                auto it = std::find_if(P.parts.begin(), P.parts.end(), [&](const Part& p) { return p.name == I.part; });
                if (it == P.parts.end()) continue;
                I.node0 = gn;
                I.elem0 = ge;
                const bool tr = I.tr[0] != 0.0 || I.tr[1] != 0.0 || I.tr[2] != 0.0;
                for (const auto& n : it->nodes) {
                    Point3 q = n;
                    if (tr)
                        for (int d = 0; d < 3; ++d) q.v[d] += I.tr[d];
                    verts.push_back(q);
                    ++gn;
                }
                for (const auto& e : it->elems) {
                    std::vector<uint32_t> g;
                    for (uint32_t v : e) g.push_back((uint32_t)(v + I.node0));
                    elems.push_back(g);
                    ++ge;
                }
            }
            for (const auto& ns : A.nsets) {
                std::vector<uint32_t> glob;
                for (const auto& ip : ns.second) {
                    if (ip.first == "__GLOBAL__") {
                        for (size_t v : ip.second) glob.push_back((uint32_t)v);
                    } else {
                        auto ii = std::find_if(A.inst.begin(), A.inst.end(),
                                               [&](const Instance& x) { return x.name == ip.first; });
                        if (ii != A.inst.end())
                            for (size_t v : ip.second) glob.push_back((uint32_t)(v + ii->node0));
                    }
                }
                D.nsets[ns.first] = glob;
            }
            for (const auto& es : A.esets) {
                std::vector<uint32_t> glob;
                for (const auto& ip : es.second) {
                    if (ip.first == "__GLOBAL__") {
                        for (size_t v : ip.second) glob.push_back((uint32_t)v);
                    } else {
                        auto ii = std::find_if(A.inst.begin(), A.inst.end(),
                                               [&](const Instance& x) { return x.name == ip.first; });
                        if (ii != A.inst.end())
                            for (size_t v : ip.second) glob.push_back((uint32_t)(v + ii->elem0));
                    }
                }
                D.esets[es.first] = glob;
            }
        }
    } else {
        std::map<std::string, size_t> off;
        size_t gn = 0;
        for (const auto& p : P.parts) {
            off[p.name] = gn;
            for (const auto& n : p.nodes) {
                verts.push_back(n);
                ++gn;
            }
        }
        for (const auto& p : P.parts)
            for (const auto& e : p.elems) {
                std::vector<uint32_t> g;
                for (uint32_t v : e) g.push_back((uint32_t)(v + off[p.name]));
                elems.push_back(g);
            }
    }
    const size_t Nn = verts.size(), Ne = elems.size();
    D.x0.resize(3 * Nn);
    for (size_t i = 0; i < Nn; ++i)
        for (int d = 0; d < 3; ++d) D.x0[3 * i + d] = verts[i].v[d];
    D.conn.resize(8 * Ne);
    for (size_t e = 0; e < Ne; ++e) {
        if (elems[e].size() != 8) {
            err = "element " + std::to_string(e + 1) + " does not have 8 nodes (only C3D8R is supported)";
            return false;
        }
        for (int k = 0; k < 8; ++k) D.conn[8 * e + k] = elems[e][k];
    }
    auto resolve = [&](const std::string& name, std::vector<uint32_t>& out) {
        out.clear();
        if (!name.empty() && name[0] == '#') {
            char* endp = nullptr;
            const long v = std::strtol(name.c_str() + 1, &endp, 10);
            if (endp != name.c_str() + 1) out.push_back((uint32_t)(v - 1));
            return;
        }
        auto it = D.nsets.find(name);
        if (it != D.nsets.end()) out = it->second;  // unknown set: skipped, as the reference warns
    };
    // initial conditions (exDyna3D.cpp:338-377)
    D.v_init.assign(3 * Nn, 0.0);
    for (const auto& iv : P.ini) {
        std::vector<uint32_t> ids;
        resolve(iv.ref, ids);
        for (uint32_t n : ids)
            if (n < Nn && iv.dof < 3) D.v_init[3 * (size_t)n + iv.dof] = iv.value;
    }
    // amplitudes in name order
    for (const auto& a : P.amps) {
        D.amp_names.push_back(a.first);
        std::vector<double> t, v;
        for (const auto& p : a.second) {
            t.push_back(p.first);
            v.push_back(p.second);
        }
        D.amp_t.push_back(t);
        D.amp_a.push_back(v);
    }
    auto amp_index = [&](const std::string& n) -> int32_t {
        if (n.empty()) return -1;
        for (size_t i = 0; i < D.amp_names.size(); ++i)
            if (D.amp_names[i] == n) return (int32_t)i;
        return -1;
    };
    const size_t S = P.steps.size();
    D.spc.resize(S);
    D.vel.resize(S);
    D.spc_nodes.resize(S);
    D.vel_nodes.resize(S);
    D.steps.resize(S);
    for (size_t s = 0; s < S; ++s) {
        const StepRec& R = P.steps[s];
        for (const auto& b : R.spc) {
            std::vector<uint32_t> ids;
            resolve(b.set, ids);
            D.spc_nodes[s].push_back(ids);
        }
        for (const auto& b : R.vel) {
            std::vector<uint32_t> ids;
            resolve(b.set, ids);
            D.vel_nodes[s].push_back(ids);
        }
    }
    for (size_t s = 0; s < S; ++s) {
        const StepRec& R = P.steps[s];
        for (size_t i = 0; i < R.spc.size(); ++i) {
            ensc_bc_entry e{};
            e.nodes = D.spc_nodes[s][i].data();
            e.n_nodes = (uint32_t)D.spc_nodes[s][i].size();
            e.d0 = (int32_t)R.spc[i].d0;
            e.d1 = (int32_t)R.spc[i].d1;
            e.cond = (int32_t)R.spc[i].cond;
            e.amp = -1;
            D.spc[s].push_back(e);
        }
        for (size_t i = 0; i < R.vel.size(); ++i) {
            ensc_bc_entry e{};
            e.nodes = D.vel_nodes[s][i].data();
            e.n_nodes = (uint32_t)D.vel_nodes[s][i].size();
            e.d0 = (int32_t)R.vel[i].d0;
            e.d1 = (int32_t)R.vel[i].d1;
            e.cond = 1;
            e.value = R.vel[i].value;
            // the reference's velocity BCs are constant (BoundaryCondition.h:18,
            // dataIn.cpp:1114-1195 drops AMPLITUDE=): kept only on request
            e.amp = (flags & ENSC_DECK_VEL_AMPLITUDE) ? amp_index(R.vel[i].amp) : -1;
            D.vel[s].push_back(e);
        }
        ensc_step_def& sd = D.steps[s];
        std::memset(&sd, 0, sizeof(sd));
        sd.period = R.period;
        sd.reset_spc = R.reset_spc;
        sd.reset_vel = R.reset_vel;
        sd.reset_dload = R.reset_dload;
        sd.grav_on = R.grav;
        sd.grav_value = R.grav_v;
        for (int d = 0; d < 3; ++d) sd.grav_dir[d] = R.gdir[d];
        sd.grav_amp = amp_index(R.grav_amp);
        sd.n_spc = (uint32_t)D.spc[s].size();
        sd.spc = D.spc[s].data();
        sd.n_vel = (uint32_t)D.vel[s].size();
        sd.vel = D.vel[s].data();
        sd.reset_dsload = R.reset_dsload;
    }
    // surface pressures (extension, on request): each *Dsload line's surface
    // resolved to (element, face) pairs through its element sets
    D.dsl.assign(S, {});
    D.dsl_elems.assign(S, {});
    D.dsl_faces.assign(S, {});
    if (flags & ENSC_DECK_DSLOAD) {
        for (size_t s = 0; s < S; ++s) {
            const StepRec& R = P.steps[s];
            for (const auto& ln : R.dsl) {
                std::vector<uint32_t> el;
                std::vector<int32_t> fc;
                auto sf = P.surfaces.find(ln.surf);
                if (sf != P.surfaces.end())
                    for (const auto& part : sf->second) {
                        auto es = D.esets.find(part.first);
                        if (es == D.esets.end()) continue;
                        for (uint32_t e : es->second)
                            if (e < Ne) {
                                el.push_back(e);
                                fc.push_back(part.second);
                            }
                    }
                D.dsl_elems[s].push_back(std::move(el));
                D.dsl_faces[s].push_back(std::move(fc));
                ensc_dsload_entry en{};
                en.value = ln.value;
                en.amp = amp_index(ln.amp);
                D.dsl[s].push_back(en);
            }
            for (size_t i = 0; i < D.dsl[s].size(); ++i) {
                D.dsl[s][i].elems = D.dsl_elems[s][i].data();
                D.dsl[s][i].faces = D.dsl_faces[s][i].data();
                D.dsl[s][i].n_faces = (uint32_t)D.dsl_elems[s][i].size();
            }
            D.steps[s].n_dsload = (uint32_t)D.dsl[s].size();
            D.steps[s].dsload = D.dsl[s].empty() ? nullptr : D.dsl[s].data();
        }
    }
    for (size_t i = 0; i < D.amp_t.size(); ++i) {
        ensc_amplitude a{};
        a.t = D.amp_t[i].data();
        a.a = D.amp_a[i].data();
        a.n = (uint32_t)D.amp_t[i].size();
        D.amps.push_back(a);
    }
    ensc_model_desc& m = D.desc;
    std::memset(&m, 0, sizeof(m));
    m.n_nodes = (uint32_t)Nn;
    m.n_elems = (uint32_t)Ne;
    m.x0 = D.x0.data();
    m.conn = D.conn.data();
    m.E = P.E;
    m.nu = P.nu;
    m.rho = P.rho;
    // *Plastic has no effect in the reference (dataIn.cpp:69-92 has no such
    // keyword): J2 only on request
    const bool pl = (flags & ENSC_DECK_PLASTIC) && P.plastic;
    m.plastic = pl ? 1 : 0;
    m.yield0 = pl ? P.yield0 : 0.0;
    m.hard = pl ? P.hard : 0.0;
    m.c_dr = proj ? proj->c_dr : 0.013;
    m.c_cr = proj ? proj->c_cr : 0.05;
    m.factor = proj ? proj->factor : 0.71;
    m.cvisl = P.cvisl;
    m.cvisq = P.cvisq;
    m.total_time = P.total_time;
    m.v_init = D.v_init.data();
    m.n_steps = (uint32_t)S;
    m.n_amps = (uint32_t)D.amps.size();
    m.steps = D.steps.data();
    m.amps = D.amps.empty() ? nullptr : D.amps.data();
    // JSON of what was parsed, in the reference harness's layout
    std::string j = "{\"E\": " + jnum(P.E) + ", \"nu\": " + jnum(P.nu) + ", \"rho\": " + jnum(P.rho) +
                    ", \"totalTime\": " + jnum(P.total_time) + ", \"Cvisl\": " + jnum(P.cvisl) +
                    ", \"Cvisq\": " + jnum(P.cvisq) + ", \"c_dr\": " + jnum(m.c_dr) + ", \"c_cr\": " + jnum(m.c_cr) +
                    ", \"factor\": " + jnum(m.factor) + ", \"n_nodes\": " + std::to_string(Nn) +
                    ", \"n_elems\": " + std::to_string(Ne) + ", \"plastic\": " + (P.plastic ? "1" : "0") +
                    ", \"yield0\": " + jnum(P.yield0) + ", \"hard\": " + jnum(P.hard) + ", \"ini_vel\": [";
    for (size_t i = 0; i < P.ini.size(); ++i)
        j += (i ? ", [" : "[") + jstr(P.ini[i].ref) + ", " + std::to_string(P.ini[i].dof) + ", " +
             jnum(P.ini[i].value) + "]";
    j += "], \"node_sets\": {";
    bool first = true;
    for (const auto& kv : D.nsets) {
        j += (first ? "" : ", ") + jstr(kv.first) + ": [";
        for (size_t i = 0; i < kv.second.size(); ++i) j += (i ? ", " : "") + std::to_string(kv.second[i]);
        j += "]";
        first = false;
    }
    j += "}, \"amplitudes\": {";
    for (size_t i = 0; i < D.amp_names.size(); ++i) {
        j += (i ? ", " : "") + jstr(D.amp_names[i]) + ": [";
        for (size_t q = 0; q < D.amp_t[i].size(); ++q)
            j += (q ? ", [" : "[") + jnum(D.amp_t[i][q]) + ", " + jnum(D.amp_a[i][q]) + "]";
        j += "]";
    }
    j += "}, \"steps\": [";
    for (size_t s = 0; s < S; ++s) {
        const StepRec& R = P.steps[s];
        j += (s ? ", " : "") + std::string("{\"name\": ") + jstr(R.name) + ", \"period\": " + jnum(R.period) +
             ", \"reset_spc\": " + std::to_string(R.reset_spc) + ", \"reset_vel\": " + std::to_string(R.reset_vel) +
             ", \"reset_dload\": " + std::to_string(R.reset_dload) +
             ", \"reset_dsload\": " + std::to_string(R.reset_dsload) + ", \"spc\": [";
        for (size_t i = 0; i < R.spc.size(); ++i)
            j += (i ? ", [" : "[") + jstr(R.spc[i].set) + ", " + std::to_string(R.spc[i].d0) + ", " +
                 std::to_string(R.spc[i].d1) + ", " + std::to_string(R.spc[i].cond) + "]";
        j += "], \"vel\": [";
        for (size_t i = 0; i < R.vel.size(); ++i)
            j += (i ? ", [" : "[") + jstr(R.vel[i].set) + ", " + std::to_string(R.vel[i].d0) + ", " +
                 std::to_string(R.vel[i].d1) + ", " + jnum(R.vel[i].value) + "]";
        j += "], \"gravity\": [" + std::string(R.grav ? "1" : "0") + ", " + jstr(R.grav_amp) + ", " +
             jnum(R.grav_v) + ", " + jnum(R.gdir[0]) + ", " + jnum(R.gdir[1]) + ", " + jnum(R.gdir[2]) +
             "], \"n_dsload\": " + std::to_string(R.n_dsload) + ", \"dsload\": [";
        for (size_t i = 0; i < R.dsl.size(); ++i)
            j += (i ? ", [" : "[") + jstr(R.dsl[i].surf) + ", " + jnum(R.dsl[i].value) + ", " + jstr(R.dsl[i].amp) + "]";
        j += "]}";
    }
    j += "]}";
    D.json = j;
    return true;
}

ensc_status parse_lines(std::vector<std::string>&& lines, const ensc_project* proj, uint32_t flags, ensc_deck** out) {
    if (!out) {
        ensc::g_last_error = "out is null";
        return ENSC_ERR_ARG;
    }
    *out = nullptr;
    Parser P;
    P.L = std::move(lines);
    P.parse();
    ensc_deck* D = new ensc_deck();
    std::string err;
    if (!build_deck(P, proj, flags, *D, err)) {
        delete D;
        ensc::g_last_error = err;
        return ENSC_ERR_ARG;
    }
    *out = D;
    return ENSC_OK;
}

std::vector<std::string> split_lines(const char* text, size_t len) {
    std::vector<std::string> L;
    size_t a = 0;
    for (size_t i = 0; i <= len; ++i) {
        if (i == len || text[i] == '\n') {
            if (i == len && a == len) break;
            std::string s(text + a, i - a);
            if (!s.empty() && s.back() == '\r') s.pop_back();
            L.push_back(std::move(s));
            a = i + 1;
        }
    }
    return L;
}

}  // namespace

extern "C" {

ensc_status ensc_project_read(const char* path, ensc_project* out) {
    if (!path || !out) {
        ensc::g_last_error = "null argument";
        return ENSC_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    out->c_dr = 0.013;
    out->c_cr = 0.05;
    out->factor = 0.71;
    std::ifstream f(path);
    if (!f.is_open()) {
        ensc::g_last_error = std::string("cannot open ") + path;
        return ENSC_ERR_ARG;
    }
    std::vector<std::string> L;
    std::string s;
    while (std::getline(f, s)) {
        if (!s.empty() && s.back() == '\r') s.pop_back();
        L.push_back(s);
    }
    // main.cpp:31-49: the deck is the first non-empty, non-comment line
    for (const auto& l : L)
        if (!l.empty() && l[0] != '#') {
            std::snprintf(out->deck, sizeof(out->deck), "%s", l.c_str());
            break;
        }
    // read_project_txt (exDyna3D.cpp:714-769): positional lines 5, 8, 11
    bool more = false;
    for (size_t i = 2; i < L.size(); ++i)
        if (!L[i].empty() && L[i][0] != '#') {
            more = true;
            break;
        }
    if (more) {
        auto num = [&](size_t line, double& v) {
            if (line - 1 < L.size()) {
                const char* b = L[line - 1].c_str();
                char* e = nullptr;
                const double x = std::strtod(b, &e);
                if (e != b) v = x;
            }
        };
        num(5, out->c_dr);
        num(8, out->c_cr);
        num(11, out->factor);
    }
    return ENSC_OK;
}

ensc_status ensc_deck_read_ex(const char* path, const ensc_project* project, uint32_t flags, ensc_deck** out) {
    if (!path) {
        ensc::g_last_error = "null path";
        return ENSC_ERR_ARG;
    }
    return ensc::guarded([&]() -> ensc_status {
        std::ifstream f(path, std::ios::binary);
        if (!f.is_open()) {
            ensc::g_last_error = std::string("cannot open deck ") + path;
            return ENSC_ERR_ARG;
        }
        std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        return parse_lines(split_lines(text.data(), text.size()), project, flags, out);
    });
}

ensc_status ensc_deck_parse_ex(const char* text, size_t len, const ensc_project* project, uint32_t flags,
                               ensc_deck** out) {
    if (!text) {
        ensc::g_last_error = "null text";
        return ENSC_ERR_ARG;
    }
    return ensc::guarded([&] { return parse_lines(split_lines(text, len), project, flags, out); });
}

ensc_status ensc_deck_read(const char* path, const ensc_project* project, ensc_deck** out) {
    return ensc_deck_read_ex(path, project, 0u, out);
}

ensc_status ensc_deck_parse(const char* text, size_t len, const ensc_project* project, ensc_deck** out) {
    return ensc_deck_parse_ex(text, len, project, 0u, out);
}

ensc_status ensc_deck_model(const ensc_deck* deck, ensc_model_desc* out) {
    if (!deck || !out) {
        ensc::g_last_error = "null argument";
        return ENSC_ERR_ARG;
    }
    *out = deck->desc;
    return ENSC_OK;
}

int64_t ensc_deck_json(const ensc_deck* deck, char* buf, int64_t cap) {
    if (!deck) return -1;
    const int64_t n = (int64_t)deck->json.size();
    if (buf && cap > 0) {
        const int64_t w = std::min<int64_t>(n, cap - 1);
        std::memcpy(buf, deck->json.data(), (size_t)w);
        buf[w] = '\0';
    }
    return n;
}

ensc_status ensc_deck_free(ensc_deck* deck) {
    delete deck;
    return ENSC_OK;
}

}  // extern "C"
