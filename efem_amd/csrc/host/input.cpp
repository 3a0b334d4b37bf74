// Reference input formats (see input.hpp).
#include "input.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace efem {
namespace {

// ---- a small JSON reader (objects, arrays, numbers, strings, literals) ----
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* key) const
    {
        if (kind != Obj) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class JsonReader {
public:
    explicit JsonReader(const std::string& t) : s_(t) {}
    bool parse(JVal& v, std::string& err)
    {
        if (!value(v)) {
            err = "JSON: " + err_ + " at offset " + std::to_string(p_);
            return false;
        }
        ws();
        if (p_ != s_.size()) {
            err = "JSON: trailing characters at offset " + std::to_string(p_);
            return false;
        }
        return true;
    }

private:
    const std::string& s_;
    size_t p_ = 0;
    std::string err_;
    void ws()
    {
        while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
    }
    bool fail(const char* m)
    {
        err_ = m;
        return false;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if (s_.compare(p_, n, w) != 0) return false;
        p_ += n;
        return true;
    }
    bool string(std::string& out)
    {
        if (s_[p_] != '"') return fail("expected string");
        ++p_;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\' && p_ < s_.size()) {
                const char e = s_[p_++];
                switch (e) {
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'u': {
                        // keep ASCII escapes; others become '?'
                        if (p_ + 4 > s_.size()) return fail("bad escape");
                        const long cp = std::strtol(s_.substr(p_, 4).c_str(), nullptr, 16);
                        p_ += 4;
                        c = cp < 128 ? static_cast<char>(cp) : '?';
                        break;
                    }
                    default: c = e;
                }
            }
            out.push_back(c);
        }
        if (p_ >= s_.size()) return fail("unterminated string");
        ++p_;
        return true;
    }
    bool value(JVal& v)
    {
        ws();
        if (p_ >= s_.size()) return fail("unexpected end");
        const char c = s_[p_];
        if (c == '{') {
            v.kind = JVal::Obj;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return true;
            }
            for (;;) {
                ws();
                std::string key;
                if (p_ >= s_.size() || !string(key)) return fail("expected key");
                ws();
                if (p_ >= s_.size() || s_[p_] != ':') return fail("expected ':'");
                ++p_;
                JVal item;
                if (!value(item)) return false;
                v.obj.emplace_back(std::move(key), std::move(item));
                ws();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == '}') {
                    ++p_;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JVal::Arr;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return true;
            }
            for (;;) {
                JVal item;
                if (!value(item)) return false;
                v.arr.push_back(std::move(item));
                ws();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == ']') {
                    ++p_;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return string(v.str);
        }
        if (lit("true")) {
            v.kind = JVal::Bool;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JVal::Bool;
            return true;
        }
        if (lit("null")) return true;
        char* end = nullptr;
        const double d = std::strtod(s_.c_str() + p_, &end);
        if (end == s_.c_str() + p_) return fail("unexpected character");
        p_ = static_cast<size_t>(end - s_.c_str());
        v.kind = JVal::Num;
        v.num = d;
        return true;
    }
};

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// integers of a line split on commas and blanks (repeated separators count once)
bool splitInts(const std::string& line, std::vector<long long>& out)
{
    out.clear();
    std::string tok;
    auto flush = [&]() {
        if (tok.empty()) return true;
        char* end = nullptr;
        const long long v = std::strtoll(tok.c_str(), &end, 10);
        if (end == tok.c_str()) return false;
        out.push_back(v);
        tok.clear();
        return true;
    };
    for (char c : line) {
        if (c == ',' || std::isspace(static_cast<unsigned char>(c))) {
            if (!flush()) return false;
        } else {
            tok.push_back(c);
        }
    }
    return flush();
}

bool readText(const std::string& path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

std::vector<std::string> lines(const std::string& text)
{
    std::vector<std::string> v;
    std::istringstream ss(text);
    std::string l;
    while (std::getline(ss, l)) v.push_back(l);
    return v;
}

// IO::readFile: every trimmed line with no '#' and no '!' and not empty
std::vector<std::string> dataLines(const std::string& text)
{
    std::vector<std::string> v;
    for (const std::string& l0 : lines(text)) {
        const std::string l = trim(l0);
        if (l.empty() || l.find('#') != std::string::npos || l.find('!') != std::string::npos) continue;
        v.push_back(l);
    }
    return v;
}

struct Material {
    int law = -1;
    double p[4] = {0, 0, 0, 0};
    double density = 0.0;  // NeoHookean.cpp:27-28 (MooneyRivlin.cpp, Yeoh.cpp alike), default 0
};

double num(const JVal* v, double d) { return v && v->kind == JVal::Num ? v->num : d; }

}  // namespace

bool loadInput(const std::string& jsonPath, InputDeck& d, std::string& msg)
{
    d = InputDeck();
    std::string text;
    if (!readText(jsonPath, text)) {
        msg = "config file " + jsonPath + " not found";
        return false;
    }
    JVal root;
    if (!JsonReader(text).parse(root, msg)) return false;
    const size_t slash = jsonPath.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : jsonPath.substr(0, slash + 1);
    // analysis (Config.cpp:23-81; Integrator.cpp:13-14 sets eps = eps_rel = residual_criterion)
    if (const JVal* an = root.get("analysis")) {
        if (const JVal* v = an->get("nl_solver"); v && v->kind == JVal::Str) {
            if (v->str == "lbfgs_ah") d.usePrecond = 1;
            else if (v->str == "lbfgs") d.usePrecond = 0;
        }
        d.maxSolverIters = static_cast<int>(num(an->get("max_solver_iters"), d.maxSolverIters));
        d.maxLineSearch = static_cast<int>(num(an->get("max_line_search"), d.maxLineSearch));
        d.dxCriterion = num(an->get("dx_criterion"), d.dxCriterion);
        d.residualCriterion = num(an->get("residual_criterion"), d.residualCriterion);
        // gravity strings of Config.cpp:45-61 (other strings: unchanged)
        if (const JVal* g = an->get("gravity"); g && g->kind == JVal::Str) {
            static const char* names[7] = {"non", "x", "-x", "y", "-y", "z", "-z"};
            for (int k = 0; k < 7; ++k)
                if (g->str == names[k]) d.gravity = k;
        }
        if (const JVal* ci = an->get("consistent_inertia"); ci && ci->kind == JVal::Bool)
            d.consistentInertia = ci->b ? 1 : 0;
    }
    // model: one body
    const JVal* models = root.get("model");
    if (!models || models->kind != JVal::Arr || models->arr.empty()) {
        msg = "config: no model";
        return false;
    }
    if (models->arr.size() > 1) {
        msg = "config: more than one body (model[]) is not supported";
        return false;
    }
    const JVal* mf = models->arr[0].get("mesh_file");
    if (!mf || mf->kind != JVal::Str) {
        msg = "config: model[0].mesh_file missing";
        return false;
    }
    std::string mesh;
    if (!readText(dir + mf->str, mesh)) {
        msg = "mesh file " + dir + mf->str + " not found";
        return false;
    }
    // MeshParser::parse: "#node" rows "x y z", "#element" rows of 1-based
    // node ids (+ material id); "##" / "!!" lines and blank lines skipped;
    // a section ends at the next line starting with '#'
    std::vector<std::vector<long long>> rawElems;
    {
        int section = 0;
        std::vector<long long> ints;
        for (const std::string& l0 : lines(mesh)) {
            const std::string l = trim(l0);
            if (l.empty() || l.compare(0, 2, "##") == 0 || l.compare(0, 2, "!!") == 0) continue;
            if (l[0] == '#') {
                section = l.compare(0, 5, "#node") == 0 ? 1 : (l.compare(0, 8, "#element") == 0 ? 2 : 0);
                continue;
            }
            if (section == 1) {
                std::istringstream is(l);
                double x[3];
                for (double& v : x) {
                    if (!(is >> v)) {
                        msg = "mesh: bad node row: " + l;
                        return false;
                    }
                    if (is.peek() == ',') is.get();
                }
                d.X.insert(d.X.end(), x, x + 3);
            } else if (section == 2) {
                if (!splitInts(l, ints)) {
                    msg = "mesh: bad element row: " + l;
                    return false;
                }
                rawElems.push_back(ints);
            }
        }
    }
    const int N = static_cast<int>(d.X.size() / 3);
    if (N == 0 || rawElems.empty()) {
        msg = "mesh: no nodes or no elements";
        return false;
    }
    // materials (loadMaterials)
    std::map<int, Material> mats;
    if (const JVal* ma = root.get("material"); ma && ma->kind == JVal::Arr) {
        for (const JVal& m : ma->arr) {
            const JVal* t = m.get("type");
            if (!t || t->kind != JVal::Str) continue;
            Material mm;
            const int id = static_cast<int>(num(m.get("id"), 0));
            if (t->str == "neo_hookean") {
                mm.law = EFEM_NEOHOOKEAN;
                mm.p[0] = num(m.get("C10"), 0);
            } else if (t->str == "mo_riv") {
                mm.law = EFEM_MOONEY_RIVLIN;
                mm.p[0] = num(m.get("C10"), 0);
                mm.p[1] = num(m.get("C01"), 0);
            } else if (t->str == "yeoh") {
                mm.law = EFEM_YEOH;
                mm.p[0] = num(m.get("C10"), 0);
                mm.p[1] = num(m.get("C20"), 0);
                mm.p[2] = num(m.get("C30"), 0);
            } else {
                continue;  // unsupported type: ignored as the reference does
            }
            mm.p[3] = num(m.get("D1"), 0);
            mm.density = num(m.get("density"), 0);
            mats[id] = mm;
        }
    }
    // elements (Body::buildElement): the count selects the kind, one extra
    // value is the material id (default -1); elements whose material is not
    // defined get no element constraint
    std::vector<int> elemLaw;
    std::vector<std::array<double, 4>> elemP;
    std::vector<double> elemRho;
    for (const auto& row : rawElems) {
        const int n = static_cast<int>(row.size());
        const int nen = n <= 5 ? 4 : (n <= 9 ? 8 : (n <= 11 ? 10 : (n <= 21 ? 20 : 0)));
        if (nen == 0 || n < nen) {
            msg = "mesh: element row with " + std::to_string(n) + " values";
            return false;
        }
        if (d.nen == 0) d.nen = nen;
        if (nen != d.nen) {
            msg = "mesh: mixed element kinds are not supported";
            return false;
        }
        const int mid = n > nen ? static_cast<int>(row[nen]) : -1;
        auto it = mats.find(mid);
        if (it == mats.end()) {
            ++d.skipped;
            continue;
        }
        for (int a = 0; a < nen; ++a) {
            const long long id = row[a] - 1;
            if (id < 0 || id >= N) {
                msg = "mesh: node id out of range in an element row";
                return false;
            }
            d.conn.push_back(static_cast<int>(id));
        }
        elemLaw.push_back(it->second.law);
        elemP.push_back({it->second.p[0], it->second.p[1], it->second.p[2], it->second.p[3]});
        elemRho.push_back(it->second.density);
    }
    const int E = static_cast<int>(elemLaw.size());
    if (E == 0) {
        msg = "mesh: no element has a defined material";
        return false;
    }
    // one law for the problem: Neo-Hookean is Mooney-Rivlin with C01 = 0 and
    // Yeoh with C20 = C30 = 0 (the same stress and coefficient bit for bit)
    bool hasNH = false, hasMR = false, hasY = false;
    for (int l : elemLaw) {
        hasNH |= l == EFEM_NEOHOOKEAN;
        hasMR |= l == EFEM_MOONEY_RIVLIN;
        hasY |= l == EFEM_YEOH;
    }
    if (hasMR && hasY) {
        msg = "materials: Mooney-Rivlin and Yeoh elements in one mesh are not supported";
        return false;
    }
    d.law = hasMR ? EFEM_MOONEY_RIVLIN : (hasY ? EFEM_YEOH : EFEM_NEOHOOKEAN);
    (void)hasNH;
    // parameters and densities share one indexing: shared if every element
    // has the same ones, else per element
    bool uniform = true;
    for (int e = 1; e < E; ++e) {
        for (int k = 0; k < 4; ++k) uniform &= elemP[e][k] == elemP[0][k];
        uniform &= elemRho[e] == elemRho[0];
    }
    d.paramsPerElem = uniform ? 0 : 1;
    if (uniform) {
        d.params.assign(elemP[0].begin(), elemP[0].end());
        d.density.assign(1, elemRho[0]);
    } else {
        for (const auto& p : elemP) d.params.insert(d.params.end(), p.begin(), p.end());
        d.density = elemRho;
    }
    // constraints
    d.fixed.assign(N, 0);
    const JVal* cons = root.get("constraint");
    if (cons && cons->kind == JVal::Arr) {
        for (const JVal& c : cons->arr) {
            const JVal* t = c.get("type");
            if (!t || t->kind != JVal::Str) continue;
            std::vector<std::string> rows;
            if (const JVal* cf = c.get("constraint_file"); cf && cf->kind == JVal::Str && !cf->str.empty()) {
                std::string ct;
                if (!readText(dir + cf->str, ct)) {
                    msg = "constraint file " + dir + cf->str + " not found";
                    return false;
                }
                rows = dataLines(ct);
            }
            if (const JVal* mid = c.get("model_id"); mid && !(mid->kind == JVal::Num && mid->num == -1)) {
                if (!(mid->kind == JVal::Arr && mid->arr.size() == 1 && mid->arr[0].num == -1)) {
                    msg = "constraint model_id: only the single body (-1) is supported";
                    return false;
                }
            }
            std::vector<long long> ints;
            auto node = [&](long long v, int& out) {
                if (v < 1 || v > N) {
                    msg = t->str + ": node id " + std::to_string(v) + " out of range";
                    return false;
                }
                out = static_cast<int>(v - 1);
                return true;
            };
            if (t->str == "fix") {
                for (const std::string& r : rows) {
                    if (!splitInts(r, ints)) {
                        msg = "fix: bad row: " + r;
                        return false;
                    }
                    for (long long v : ints) {
                        int k;
                        if (!node(v, k)) return false;
                        d.fixed[k] = 1;
                    }
                }
            } else if (t->str == "force") {
                const JVal* v = c.get("value");
                if (v && v->kind == JVal::Arr && v->arr.size() >= 3)
                    for (int k = 0; k < 3; ++k) d.fvec[k] = v->arr[k].num;
                if (!d.fnodes.empty()) {
                    msg = "force: only one force constraint is supported";
                    return false;
                }
                for (const std::string& r : rows) {
                    if (!splitInts(r, ints)) {
                        msg = "force: bad row: " + r;
                        return false;
                    }
                    for (long long vv : ints) {
                        int k;
                        if (!node(vv, k)) return false;
                        d.fnodes.push_back(k);
                    }
                }
            } else if (t->str == "pressure") {
                // value > 0: follower pressure; [x, y, z]: directed traction;
                // otherwise no load (ExternalConstraint.cpp:184-199)
                const JVal* v = c.get("value");
                double pv = 0.0, pd[3] = {0, 0, 0};
                if (v && v->kind == JVal::Num) pv = v->num;
                if (v && v->kind == JVal::Arr && v->arr.size() >= 3)
                    for (int k = 0; k < 3; ++k) pd[k] = v->arr[k].num;
                if (pv <= 0.0 && std::sqrt(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]) <= 0.0) continue;
                if (!d.faces.empty()) {
                    msg = "pressure: only one pressure constraint is supported";
                    return false;
                }
                d.directed = pv > 0.0 ? 0 : 1;
                d.pressure = pv > 0.0 ? pv : 0.0;
                for (int k = 0; k < 3; ++k) d.pdir[k] = pv > 0.0 ? 0.0 : pd[k];
                for (const std::string& r : rows) {
                    if (!splitInts(r, ints)) {
                        msg = "pressure: bad row: " + r;
                        return false;
                    }
                    std::vector<int> face;
                    for (long long vv : ints) {
                        int k;
                        if (!node(vv, k)) return false;
                        if (std::find(face.begin(), face.end(), k) == face.end()) face.push_back(k);
                    }
                    if (d.fnn == 0) d.fnn = static_cast<int>(face.size());
                    if (static_cast<int>(face.size()) != d.fnn) {
                        msg = "pressure: faces with different node counts";
                        return false;
                    }
                    d.faces.insert(d.faces.end(), face.begin(), face.end());
                }
            }
        }
    }
    return true;
}

}  // namespace efem

// ---- C ABI -----------------------------------------------------------------
struct efem_input {
    efem::InputDeck d;
};

namespace {
thread_local std::string g_input_error;
}

extern "C" efem_input* efem_input_load(const char* json_path)
{
    g_input_error.clear();
    if (!json_path) {
        g_input_error = "null path";
        return nullptr;
    }
    try {
        auto h = std::make_unique<efem_input>();
        if (!efem::loadInput(json_path, h->d, g_input_error)) return nullptr;
        return h.release();
    } catch (const std::exception& e) {
        g_input_error = e.what();
        return nullptr;
    }
}

extern "C" const char* efem_input_error(void) { return g_input_error.c_str(); }

extern "C" void efem_input_free(efem_input* h) { delete h; }

extern "C" int efem_input_sizes(const efem_input* h, long long* s)
{
    if (!h || !s) return EFEM_ERR_ARG;
    const efem::InputDeck& d = h->d;
    s[0] = static_cast<long long>(d.X.size() / 3);
    s[1] = d.nen ? static_cast<long long>(d.conn.size() / d.nen) : 0;
    s[2] = d.nen;
    s[3] = d.law;
    s[4] = d.paramsPerElem;
    s[5] = d.fnn ? static_cast<long long>(d.faces.size() / d.fnn) : 0;
    s[6] = d.fnn;
    s[7] = std::count(d.fixed.begin(), d.fixed.end(), 1);
    s[8] = static_cast<long long>(d.fnodes.size());
    s[9] = d.skipped;
    return EFEM_OK;
}

extern "C" int efem_input_problem(const efem_input* h, efem_problem* pb, int precision)
{
    if (!h || !pb) return EFEM_ERR_ARG;
    const efem::InputDeck& d = h->d;
    std::memset(pb, 0, sizeof(*pb));
    pb->n_nodes = static_cast<int>(d.X.size() / 3);
    pb->X = d.X.data();
    pb->n_elems = static_cast<int>(d.conn.size() / d.nen);
    pb->nen = d.nen;
    pb->conn = d.conn.data();
    pb->reduced = 0;
    pb->law = d.law;
    pb->params = d.params.data();
    pb->params_per_elem = d.paramsPerElem;
    pb->fixed = d.fixed.data();
    pb->n_faces = d.fnn ? static_cast<int>(d.faces.size() / d.fnn) : 0;
    pb->fnn = d.fnn;
    pb->faces = d.faces.empty() ? nullptr : d.faces.data();
    pb->pressure = d.pressure;
    pb->pdir = d.directed ? d.pdir : nullptr;
    pb->n_fnodes = static_cast<int>(d.fnodes.size());
    pb->fnodes = d.fnodes.empty() ? nullptr : d.fnodes.data();
    pb->fvec = d.fnodes.empty() ? nullptr : d.fvec;
    pb->volumes = nullptr;
    pb->precision = precision;
    pb->gravity = d.gravity;
    pb->consistent_inertia = d.consistentInertia;
    pb->density = d.density.empty() ? nullptr : d.density.data();
    return EFEM_OK;
}

extern "C" int efem_input_lbfgs_params(const efem_input* h, efem_lbfgs_params* prm)
{
    if (!h || !prm) return EFEM_ERR_ARG;
    const efem::InputDeck& d = h->d;
    std::memset(prm, 0, sizeof(*prm));
    prm->epsilon = d.residualCriterion;
    prm->epsilon_rel = d.residualCriterion;
    prm->max_iterations = d.maxSolverIters;
    prm->max_linesearch = d.maxLineSearch;
    prm->m = 6;
    prm->lns_fea = 1.0;
    prm->use_precond = d.usePrecond;
    prm->line_search = EFEM_LS_PROJECTION;
    return EFEM_OK;
}
