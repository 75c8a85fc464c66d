// Input and output formats of the reference driver (include/dsim_io.h):
// sim.cfg subset parser, gradient and mesh readers, the PGSE state machine,
// the Stejskal-Tanner tensor fit and the appended MATLAB .m writer.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dsim_io.h"

namespace {

thread_local std::string g_ioerr;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_ioerr = buf;
    return code;
}

bool read_file(const char *path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool parse_number(const std::string &tok, double &v) {
    if (tok.empty()) return false;
    const char *s = tok.c_str();
    char *end = nullptr;
    errno = 0;
    v = std::strtod(s, &end);
    return end && *end == '\0' && end != s;
}

void copy_str(char *dst, const std::string &s) {
    size_t n = std::min(s.size(), (size_t)DSIM_PATH_MAX - 1);
    memcpy(dst, s.data(), n);
    dst[n] = '\0';
}

// ---------------------------------------------------------------------------
// libconfig subset: groups `name : { ... };`, settings `name = value;`,
// scalars (numbers, true/false, "strings"), comments //, #, /* */.
// ---------------------------------------------------------------------------
struct CfgLexer {
    const std::string &s;
    size_t i = 0;
    explicit CfgLexer(const std::string &src) : s(src) {}
    void skip() {
        for (;;) {
            while (i < s.size() && isspace((unsigned char)s[i])) ++i;
            if (i < s.size() && s[i] == '#') {
                while (i < s.size() && s[i] != '\n') ++i;
            } else if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') ++i;
            } else if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                size_t e = s.find("*/", i + 2);
                i = (e == std::string::npos) ? s.size() : e + 2;
            } else {
                return;
            }
        }
    }
    // kinds: 'n' name, 's' string, 'v' scalar word, punctuation char, 0 end
    char next(std::string &tok) {
        skip();
        tok.clear();
        if (i >= s.size()) return 0;
        char c = s[i];
        if (c == '"') {
            ++i;
            for (;;) {
                while (i < s.size() && s[i] != '"') {
                    if (s[i] == '\\' && i + 1 < s.size()) ++i;
                    tok += s[i++];
                }
                ++i;
                // adjacent string literals concatenate
                size_t save = i;
                skip();
                if (i < s.size() && s[i] == '"') { ++i; continue; }
                i = save;
                break;
            }
            return 's';
        }
        if (strchr("{}[]();:=,", c)) {
            ++i;
            tok = std::string(1, c);
            return c;
        }
        while (i < s.size() && !isspace((unsigned char)s[i]) && !strchr("{}[]();:=,\"#", s[i])) {
            if (s[i] == '/' && i + 1 < s.size() && (s[i + 1] == '/' || s[i + 1] == '*')) break;
            tok += s[i++];
        }
        return (isalpha((unsigned char)tok[0]) || tok[0] == '*') ? 'n' : 'v';
    }
};

bool parse_cfg(const std::string &src, std::map<std::string, std::string> &out, std::string &err) {
    CfgLexer lx(src);
    std::vector<std::string> path;
    std::string tok;
    for (;;) {
        char k = lx.next(tok);
        if (k == 0) break;
        if (k == '}') {
            if (path.empty()) { err = "unbalanced }"; return false; }
            path.pop_back();
            continue;
        }
        if (k == ';' || k == ',') continue;
        if (k != 'n') { err = "expected a setting name near '" + tok + "'"; return false; }
        std::string name = tok;
        char sep = lx.next(tok);
        if (sep != '=' && sep != ':') { err = "expected = or : after " + name; return false; }
        char v = lx.next(tok);
        if (v == '{') {
            path.push_back(name);
            continue;
        }
        std::string full;
        for (auto &p : path) full += p + ".";
        full += name;
        if (v == '[' || v == '(') {  // arrays/lists: kept as their raw items, unused by dsim
            char close = v == '[' ? ']' : ')';
            int depth = 1;
            std::string acc;
            while (depth > 0) {
                char q = lx.next(tok);
                if (q == 0) { err = "unterminated list"; return false; }
                if (q == v) depth++;
                if (q == close) depth--;
                if (depth > 0) acc += tok + " ";
            }
            out[full] = acc;
            continue;
        }
        if (v == 's' || v == 'v' || v == 'n') {
            out[full] = tok;
            continue;
        }
        err = "bad value for " + full;
        return false;
    }
    if (!path.empty()) { err = "unterminated group " + path.back(); return false; }
    return true;
}

bool as_bool(const std::string &v, int32_t &o) {
    std::string l = v;
    for (auto &c : l) c = (char)tolower((unsigned char)c);
    if (l == "true") { o = 1; return true; }
    if (l == "false") { o = 0; return true; }
    double d;
    if (parse_number(v, d)) { o = d != 0.0; return true; }
    return false;
}

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
void tokenize(const std::string &s, std::vector<std::string> &toks) {
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && isspace((unsigned char)s[i])) ++i;
        size_t b = i;
        while (i < s.size() && !isspace((unsigned char)s[i])) ++i;
        if (i > b) toks.emplace_back(s.substr(b, i - b));
    }
}

// A membrane section opener: "A", "M" or "G", optionally followed by digits.
bool section_token(const std::string &t, char &kind, bool &has_num, double &num) {
    if (t.empty() || (t[0] != 'A' && t[0] != 'M' && t[0] != 'G')) return false;
    if (t.size() == 1) { kind = t[0]; has_num = false; return true; }
    double d;
    if (!parse_number(t.substr(1), d)) return false;
    kind = t[0];
    has_num = true;
    num = d;
    return true;
}

}  // namespace

struct dsim_mesh_file {
    std::vector<float> vertices;
    std::vector<uint32_t> triangles, tri_info;
    dsim_mesh_info info{};
};

extern "C" int dsim_config_defaults(dsim_config *c) {
    if (!c) return fail(DSIM_ERR_ARG, "NULL config");
    memset(c, 0, sizeof *c);
    // diffusion.cpp:89-107
    c->timeStep = 0.01f;
    c->stepsPerUpdate = 1;
    c->useGpu = 1;
    c->useDisplay = 1;
    c->searchMethod = 0;
    c->writeOut = 1;
    c->numSpins = 30000;
    c->gyroMagneticRatio = 42576.0f;
    c->extraAdc = 2.1f;
    c->intraAdc = 2.1f;
    c->myelinAdc = 0.1f;
    c->spaceScale = 100.0f;
    c->permeability = 0.000002f;
    c->extraT2 = 79.0f;
    c->intraT2 = 79.0f;
    c->myelinT2 = 7.6f;
    c->startBox = 0.6f;
    copy_str(c->configFile, "sim.cfg");
    return DSIM_OK;
}

extern "C" int dsim_config_read(const char *path, dsim_config *c) {
    if (!path || !c) return fail(DSIM_ERR_ARG, "NULL argument");
    std::string src, err;
    if (!read_file(path, src)) return fail(DSIM_ERR_IO, "cannot open %s", path);
    copy_str(c->configFile, path);
    std::map<std::string, std::string> kv;
    if (!parse_cfg(src, kv, err)) return fail(DSIM_ERR_FORMAT, "%s: %s", path, err.c_str());
    auto get = [&](std::initializer_list<const char *> keys, std::string &v) {
        for (const char *k : keys) {
            auto it = kv.find(k);
            if (it != kv.end()) { v = it->second; return true; }
        }
        return false;
    };
    std::string v;
    double d;
    int32_t b;
    struct F { std::initializer_list<const char *> k; float *p; };
    F floats[] = {{{"sim.gyroMagneticRatio", "sim.gyromagneticRatio"}, &c->gyroMagneticRatio},
                  {{"sim.timeStep", "sim.timestep"}, &c->timeStep},
                  {{"sim.extraAdc"}, &c->extraAdc}, {{"sim.intraAdc"}, &c->intraAdc},
                  {{"sim.myelinAdc"}, &c->myelinAdc}, {{"sim.extraT2"}, &c->extraT2},
                  {{"sim.intraT2"}, &c->intraT2}, {{"sim.myelinT2"}, &c->myelinT2},
                  {{"sim.spaceScale"}, &c->spaceScale}, {{"sim.startBox"}, &c->startBox},
                  {{"sim.permeability"}, &c->permeability}};
    for (auto &f : floats)
        if (get(f.k, v) && parse_number(v, d)) *f.p = (float)d;
    if (get({"app.stepsPerUpdate"}, v) && parse_number(v, d)) c->stepsPerUpdate = (int32_t)d;
    if (get({"sim.numSpins"}, v) && parse_number(v, d)) c->numSpins = (uint32_t)d;
    if (get({"sim.searchMethod"}, v) && parse_number(v, d)) c->searchMethod = (uint32_t)d;
    if (get({"app.useGpu"}, v) && as_bool(v, b)) c->useGpu = b;
    if (get({"app.useDisplay"}, v) && as_bool(v, b)) c->useDisplay = b;
    if (get({"app.writeOut"}, v) && as_bool(v, b)) c->writeOut = b;
    if (get({"sim.outFile"}, v)) copy_str(c->outFile, v);
    if (get({"sim.gradsFile"}, v)) copy_str(c->gradsFile, v);
    if (get({"fibers.fiberFile", "fibers.file"}, v)) copy_str(c->fiberFile, v);
    if (get({"sim.seed"}, v) && parse_number(v, d)) c->seed = (uint64_t)d;
    return DSIM_OK;
}

extern "C" int dsim_config_apply_args(dsim_config *c, int argc, const char *const *argv) {
    if (!c) return fail(DSIM_ERR_ARG, "NULL config");
    for (int a = 1; a < argc; ++a) {
        if (!argv[a]) continue;
        std::string s = argv[a];
        if (s.empty() || s[0] != '-') continue;
        s = s.substr(s.size() > 1 && s[1] == '-' ? 2 : 1);
        std::string key = s, val;
        size_t eq = s.find('=');
        bool has_val = eq != std::string::npos;
        if (has_val) { key = s.substr(0, eq); val = s.substr(eq + 1); }
        double d = 0;
        bool num = has_val && parse_number(val, d);
        if (key == "disp") c->useDisplay = 1;
        else if (key == "nodisp") c->useDisplay = 0;
        else if (key == "gpu") c->useGpu = 1;
        else if (key == "cpu") c->useGpu = 0;
        else if (key == "w") c->writeOut = 1;
        else if (key == "nw") c->writeOut = 0;
        else if (key == "stepsPerUpdate" && num) c->stepsPerUpdate = (int32_t)d;
        else if (key == "numSpins" && num) c->numSpins = (uint32_t)d;
        else if (key == "searchMethod" && num) c->searchMethod = (uint32_t)d;
        else if (key == "perm" && num) c->permeability = (float)d;
        else if (key == "startBox" && num) c->startBox = (float)d;
        else if (key == "seed" && num) c->seed = (uint64_t)d;
        else if (key == "fiberFile" && has_val) copy_str(c->fiberFile, val);
        else if (key == "config" && has_val) copy_str(c->configFile, val);
        else if (key == "outFile" && has_val) copy_str(c->outFile, val);
        else if (key == "gradsFile" && has_val) copy_str(c->gradsFile, val);
    }
    return DSIM_OK;
}

extern "C" int dsim_grads_read(const char *path, float *rows, uint32_t cap, uint32_t *n) {
    if (!path || !n) return fail(DSIM_ERR_ARG, "NULL argument");
    std::string src;
    if (!read_file(path, src)) return fail(DSIM_ERR_IO, "cannot open %s", path);
    std::vector<std::string> toks;
    tokenize(src, toks);
    std::vector<double> vals;
    for (size_t i = 0; i < toks.size(); ++i) {
        double d;
        if (toks[i][0] == '#' || toks[i][0] == '%') {  // comment to end of line
            continue;
        }
        if (!parse_number(toks[i], d)) break;
        vals.push_back(d);
    }
    uint32_t k = (uint32_t)(vals.size() / 6);
    for (uint32_t j = 0; j < k && j < cap; ++j)
        for (int q = 0; q < 6; ++q) rows[j * 6 + q] = (float)vals[j * 6 + q];
    *n = k;
    return DSIM_OK;
}

extern "C" int dsim_mesh_read(const char *path, float space_scale, dsim_mesh_file **out) {
    if (!path || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (!(space_scale > 0.0f)) return fail(DSIM_ERR_ARG, "space_scale must be > 0");
    std::string src;
    if (!read_file(path, src)) return fail(DSIM_ERR_IO, "cannot open %s", path);
    // comments (%, #) run to the end of the line
    std::string clean;
    clean.reserve(src.size());
    bool in_c = false;
    for (char ch : src) {
        if (ch == '%' || ch == '#') in_c = true;
        if (ch == '\n') in_c = false;
        clean += in_c ? ' ' : ch;
    }
    std::vector<std::string> t;
    tokenize(clean, t);
    auto *M = new dsim_mesh_file();
    struct Sec { char kind; double fiber; std::vector<uint32_t> tris; };
    std::vector<Sec> secs;
    std::vector<double> verts, tris;
    size_t i = 0;
    double d;
    while (i < t.size()) {
        const std::string &w = t[i];
        char kind;
        bool has_num;
        double num;
        if (w == "V" || w == "T") {
            std::vector<double> &dst = (w == "V") ? verts : tris;
            ++i;
            std::vector<double> buf;
            while (i < t.size() && parse_number(t[i], d)) { buf.push_back(d); ++i; }
            size_t whole = buf.size() - buf.size() % 3;
            dst.insert(dst.end(), buf.begin(), buf.begin() + whole);
        } else if (section_token(w, kind, has_num, num)) {
            ++i;
            if (!has_num) {
                if (i < t.size() && parse_number(t[i], d)) { num = d; ++i; }
                else num = 0.0;
            }
            Sec s{kind, num, {}};
            while (i < t.size() && parse_number(t[i], d)) { s.tris.push_back((uint32_t)d); ++i; }
            secs.push_back(std::move(s));
        } else if (w.rfind("Number_of_", 0) == 0 || w == "Max_triangles_on_membrane") {
            ++i;
            if (i < t.size() && parse_number(t[i], d)) ++i;  // header value
        } else {
            ++i;  // unknown word
        }
    }
    dsim_mesh_info &I = M->info;
    I.n_vertices = (uint32_t)(verts.size() / 3);
    I.n_triangles = (uint32_t)(tris.size() / 3);
    for (auto &s : secs) {
        if (s.kind == 'A') I.n_axon++;
        else if (s.kind == 'M') I.n_myelin++;
        else I.n_glia++;
        I.max_triangles_on_surface = std::max(I.max_triangles_on_surface, (uint32_t)s.tris.size());
    }
    I.n_fibers = I.n_axon + I.n_glia;
    if (I.n_axon != I.n_myelin) {
        delete M;
        return fail(DSIM_ERR_FORMAT,
                    "%s: %u axon and %u myelin membranes; they must pair up", path, I.n_axon,
                    I.n_myelin);
    }
    M->vertices.resize(verts.size());
    I.xmax = I.ymax = I.zmax = 1.0f;
    for (size_t k = 0; k < verts.size(); ++k) {
        float x = (float)verts[k];
        M->vertices[k] = x / space_scale;
        float a = std::fabs(x) / space_scale;
        float &m = (k % 3 == 0) ? I.xmax : (k % 3 == 1) ? I.ymax : I.zmax;
        m = std::fmax(a, m);
    }
    M->triangles.resize(tris.size());
    M->tri_info.assign(tris.size(), 0);
    for (size_t k = 0; k < tris.size(); ++k) {
        uint32_t v = (uint32_t)(float)tris[k];
        if (v >= I.n_vertices) {
            delete M;
            return fail(DSIM_ERR_FORMAT, "%s: triangle %zu uses vertex %u of %u", path, k / 3, v,
                        I.n_vertices);
        }
        M->triangles[k] = v;
        if (k % 3 == 0) M->tri_info[k + 2] = v;
    }
    for (auto &s : secs) {
        uint32_t mem = s.kind == 'A' ? 0u : s.kind == 'M' ? 1u : 2u;
        uint32_t f = (uint32_t)(float)s.fiber;
        if (f >= I.n_fibers || f >= 65535u) {
            delete M;
            return fail(DSIM_ERR_FORMAT, "%s: membrane %c%u beyond %u fibres", path, s.kind, f,
                        I.n_fibers);
        }
        for (uint32_t tri : s.tris) {
            if (tri >= I.n_triangles) {
                delete M;
                return fail(DSIM_ERR_FORMAT, "%s: membrane %c%u lists triangle %u of %u", path,
                            s.kind, f, tri, I.n_triangles);
            }
            M->tri_info[tri * 3] = f;
            M->tri_info[tri * 3 + 1] = mem;
        }
    }
    *out = M;
    return DSIM_OK;
}

extern "C" int dsim_mesh_view(const dsim_mesh_file *m, dsim_mesh *view, dsim_mesh_info *info) {
    if (!m) return fail(DSIM_ERR_ARG, "NULL mesh");
    if (view) {
        view->n_vertices = m->info.n_vertices;
        view->vertices = m->vertices.data();
        view->n_triangles = m->info.n_triangles;
        view->triangles = m->triangles.data();
        view->tri_info = m->tri_info.data();
    }
    if (info) *info = m->info;
    return DSIM_OK;
}

extern "C" int dsim_mesh_free(dsim_mesh_file *m) {
    delete m;
    return DSIM_OK;
}

extern "C" int dsim_mesh_write(const char *path, const dsim_mesh *mesh, float space_scale) {
    if (!path || !mesh) return fail(DSIM_ERR_ARG, "NULL argument");
    FILE *f = fopen(path, "w");
    if (!f) return fail(DSIM_ERR_IO, "cannot write %s", path);
    fprintf(f, "V\n");
    for (uint32_t v = 0; v < mesh->n_vertices; ++v)
        fprintf(f, "%.9g %.9g %.9g\n", mesh->vertices[3 * v] * space_scale,
                mesh->vertices[3 * v + 1] * space_scale, mesh->vertices[3 * v + 2] * space_scale);
    fprintf(f, "T\n");
    for (uint32_t t = 0; t < mesh->n_triangles; ++t)
        fprintf(f, "%u %u %u\n", mesh->triangles[3 * t], mesh->triangles[3 * t + 1],
                mesh->triangles[3 * t + 2]);
    // membrane sections grouped by (membrane, fibre)
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> groups;
    for (uint32_t t = 0; t < mesh->n_triangles; ++t)
        groups[{mesh->tri_info[3 * t + 1], mesh->tri_info[3 * t]}].push_back(t);
    const char tag[3] = {'A', 'M', 'G'};
    for (auto &g : groups) {
        fprintf(f, "%c%u\n", tag[std::min(2u, g.first.first)], g.first.second);
        for (uint32_t t : g.second) fprintf(f, "%u\n", t);
    }
    int bad = ferror(f);
    fclose(f);
    return bad ? fail(DSIM_ERR_IO, "write error on %s", path) : DSIM_OK;
}

extern "C" int dsim_pgse_schedule(float ld, float bd, float ro, float dt, int32_t S, int32_t *sign,
                                  uint32_t cap, int32_t *Lo, int32_t *Bo, int32_t *Ro,
                                  int32_t *readout) {
    if (S <= 0 || !(dt > 0.0f)) return fail(DSIM_ERR_ARG, "bad time step");
    // diffusion.cpp:605-608: (int)(x/(timeStep*(float)stepsPerUpdate)+0.5)
    int32_t L = (int32_t)(ld / (dt * (float)S) + 0.5);
    int32_t B = (int32_t)(bd / (dt * (float)S) + 0.5);
    int32_t R = (int32_t)(ro / (dt * (float)S) + 0.5);
    int32_t n = B + L + R, g = 0, rd = 0;
    for (int32_t k = 0; k <= n; ++k) {
        if (k == 0) g = 1;
        else if (k == L && k < B) g = 0;
        else if (k == B) g = -1;
        else if (k == B + L) g = 0;
        else if (k == B + L + R) rd = 1;
        if (sign && k < n && (uint32_t)k < cap) sign[k] = g;
    }
    if (Lo) *Lo = L;
    if (Bo) *Bo = B;
    if (Ro) *Ro = R;
    if (readout) *readout = rd;
    return DSIM_OK;
}

// ---------------------------------------------------------------------------
// Tensor fit
// ---------------------------------------------------------------------------
namespace {

double round6(double x) { return std::floor(x * 1000000.0 + 0.5) / 1000000.0; }

// One-sided Jacobi SVD (Hestenes) of A (m x n, row-major): A = U S V^T with
// singular values descending; U is m x n, V is n x n.
void svd(std::vector<double> A, int m, int n, std::vector<double> &U, std::vector<double> &S,
         std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double a = 0, b = 0, c = 0;
                for (int i = 0; i < m; ++i) {
                    double x = A[i * n + p], y = A[i * n + q];
                    a += x * x; b += y * y; c += x * y;
                }
                if (c == 0.0 || std::fabs(c) <= 1e-300) continue;
                off = std::max(off, std::fabs(c) / std::sqrt(a * b + 1e-300));
                double zeta = (b - a) / (2 * c);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                for (int i = 0; i < m; ++i) {
                    double x = A[i * n + p], y = A[i * n + q];
                    A[i * n + p] = cs * x - sn * y;
                    A[i * n + q] = sn * x + cs * y;
                }
                for (int i = 0; i < n; ++i) {
                    double x = V[i * n + p], y = V[i * n + q];
                    V[i * n + p] = cs * x - sn * y;
                    V[i * n + q] = sn * x + cs * y;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<int> order(n);
    std::vector<double> s(n);
    for (int j = 0; j < n; ++j) {
        double z = 0;
        for (int i = 0; i < m; ++i) z += A[i * n + j] * A[i * n + j];
        s[j] = std::sqrt(z);
        order[j] = j;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s[a] > s[b]; });
    U.assign((size_t)m * n, 0.0);
    S.assign(n, 0.0);
    std::vector<double> V2((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        int j = order[k];
        S[k] = s[j];
        for (int i = 0; i < m; ++i) U[i * n + k] = s[j] > 0 ? A[i * n + j] / s[j] : 0.0;
        for (int i = 0; i < n; ++i) V2[i * n + k] = V[i * n + j];
    }
    V = V2;
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi; values ascending.
void eig3(const double D[3][3], double val[3], double vec[3][3]) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    memcpy(a, D, sizeof a);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    double x = a[k][p], y = a[k][q];
                    a[k][p] = c * x - s * y;
                    a[k][q] = s * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    double x = a[p][k], y = a[q][k];
                    a[p][k] = c * x - s * y;
                    a[q][k] = s * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    double x = v[k][p], y = v[k][q];
                    v[k][p] = c * x - s * y;
                    v[k][q] = s * x + c * y;
                }
            }
    }
    int o[3] = {0, 1, 2};
    std::sort(o, o + 3, [&](int x, int y) { return a[x][x] < a[y][y]; });
    for (int k = 0; k < 3; ++k) {
        val[k] = a[o[k]][o[k]];
        // sign convention: largest-magnitude component positive
        int big = 0;
        for (int r = 1; r < 3; ++r)
            if (std::fabs(v[r][o[k]]) > std::fabs(v[big][o[k]])) big = r;
        double sg = v[big][o[k]] < 0 ? -1.0 : 1.0;
        for (int r = 0; r < 3; ++r) vec[r][k] = sg * v[r][o[k]];
    }
}

}  // namespace

extern "C" int dsim_tensor_fit_compute(uint32_t ng, const float *rows, const float *signal,
                                       dsim_tensor_fit *out) {
    if (!rows || !signal || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    memset(out, 0, sizeof *out);
    const double offset = 0.000001;
    std::vector<double> dir(3 * (size_t)ng), bval(ng), logdw;
    std::vector<uint32_t> nz;
    double logb0 = 0;
    uint32_t n0 = 0;
    for (uint32_t j = 0; j < ng; ++j) {
        const float *r = rows + 6 * (size_t)j;
        double norm = (double)std::sqrt(r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
        if (norm == 0) {
            logb0 += std::log((double)signal[j] + offset);
            n0++;
        } else {
            for (int k = 0; k < 3; ++k) dir[3 * j + k] = (double)r[3 + k] / norm;
            logdw.push_back(std::log((double)signal[j] + offset));
            nz.push_back(j);
        }
        // diffusion.cpp:317
        bval[j] = std::pow((2.0 * 3.1415926535 * 42576.0) * norm / 1000000000 * r[0], 2) *
                  (r[1] - r[0] / 3.0);
    }
    if (n0 == 0 || nz.size() < 1) return fail(DSIM_ERR_ARG, "need b=0 and diffusion-weighted lines");
    logb0 /= n0;
    const int K = (int)nz.size();
    std::vector<double> adc(K), M((size_t)K * 6);
    for (int k = 0; k < K; ++k) {
        adc[k] = -1 / bval[nz[k]] * (logdw[k] - logb0);
        const double *b = &dir[3 * nz[k]];
        double row[6] = {b[0] * b[0], b[1] * b[1], b[2] * b[2], 2 * b[0] * b[1], 2 * b[0] * b[2],
                         2 * b[1] * b[2]};
        for (int q = 0; q < 6; ++q) M[k * 6 + q] = round6(row[q]);
    }
    std::vector<double> U, S, V;
    svd(M, K, 6, U, S, V);
    for (auto &x : S) x = round6(x);
    for (auto &x : U) x = round6(x);
    for (auto &x : V) x = round6(x);
    // pinv = V S^-1 U^T, rounded (diffusion.cpp:379-420)
    std::vector<double> P((size_t)6 * K, 0.0);
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < K; ++k) {
            double acc = 0;
            for (int q = 0; q < 6; ++q) {
                double si = S[q] == 0 ? 0.0 : 1.0 / S[q];
                acc += V[i * 6 + q] * si * U[k * 6 + q];
            }
            P[i * K + k] = round6(acc);
        }
    double coef[6];
    for (int i = 0; i < 6; ++i) {
        double acc = 0;
        for (int k = 0; k < K; ++k) acc += P[i * K + k] * adc[k];
        coef[i] = acc;
    }
    double Dm[3][3] = {{coef[0], coef[3], coef[4]}, {coef[3], coef[1], coef[5]},
                       {coef[4], coef[5], coef[2]}};
    double val[3], vec[3][3];
    eig3(Dm, val, vec);
    memcpy(out->tensor, Dm, sizeof Dm);
    memcpy(out->eigvec, vec, sizeof vec);
    for (int k = 0; k < 3; ++k) out->eigval[k] = (float)val[k];
    float *l = out->eigval;
    out->md = (l[0] + l[1] + l[2]) / 3.0f;
    out->fa = (float)(std::sqrt(3.0 / 2.0) *
                      std::sqrt(std::pow(l[0] - out->md, 2) + std::pow(l[1] - out->md, 2) +
                                std::pow(l[2] - out->md, 2)) /
                      std::sqrt(std::pow(l[0], 2) + std::pow(l[1], 2) + std::pow(l[2], 2)));
    out->rd = (l[0] + l[1]) / 2.0f;
    return DSIM_OK;
}

extern "C" int dsim_write_m(const char *path, const dsim_config *c, uint32_t ng, const float *rows,
                            const float *sig, const dsim_tensor_fit *fits) {
    if (!path || !c || !rows || !sig || !fits) return fail(DSIM_ERR_ARG, "NULL argument");
    FILE *f = fopen(path, "at");
    if (!f) return fail(DSIM_ERR_IO, "cannot open %s", path);
    time_t now = time(nullptr);
    // diffusion.cpp:668-728
    fprintf(f, "\n%% * * * %s", ctime(&now));
    fprintf(f, "if (exist('m','var')), m=m+1; else, m=1; end\n");
    fprintf(f, "s=0;\n");
    const char *bar = "%%%%%%%%%%%%%%%%%";
    fprintf(f, "\n%s \n%% Input parameters \n%s\n", bar, bar);
    fprintf(f, "numSpins(m) = %u;\n", c->numSpins);
    fprintf(f, "useGpu(m) = %d;\n", c->useGpu);
    fprintf(f, "fibersFile{m} = '%s';\n", c->fiberFile);
    fprintf(f, "useDisplay(m) = %d;\n", c->useDisplay);
    fprintf(f, "stepsPerUpdate(m) = %i;\n", c->stepsPerUpdate);
    fprintf(f, "gyroMagneticRatio(m) = %g;\n", c->gyroMagneticRatio);
    fprintf(f, "timeStep(m) = %g;\n", c->timeStep);
    fprintf(f, "extraAdc(m) = %g;\n", c->extraAdc);
    fprintf(f, "intraAdc(m) = %g;\n", c->intraAdc);
    fprintf(f, "myelinAdc(m) = %g;\n", c->myelinAdc);
    fprintf(f, "spaceScale(m) = %g;\n", c->spaceScale);
    fprintf(f, "permeability(m) = %g;\n", c->permeability);
    fprintf(f, "\n\n%s \n%% Simulation data \n%s\n\n", bar, bar);
    fprintf(f, "%%The following variables give the diffusion tensor, eigenvector matrix, eigenvalues\n");
    fprintf(f, "%%(eigenvalue no. i corresponds to column no. i in the eigenvector matrix), mean diffusivity,\n");
    fprintf(f, "%%fractional anisotropy and radial diffusivity for the whole volume (index 1) and each\n");
    fprintf(f, "%%compartment (index 2 for compartment 0, index 3 for compartment 1, etc).\n");
    // the reference writes indices 1..4 (total and compartments 0-2)
    for (unsigned k = 0; k < 4; ++k) {
        const dsim_tensor_fit &t = fits[k];
        fprintf(f, "tensor{m}{%u} = [%g, %g, %g; %g, %g, %g; %g, %g, %g];\n", k + 1, t.tensor[0][0],
                t.tensor[0][1], t.tensor[0][2], t.tensor[1][0], t.tensor[1][1], t.tensor[1][2],
                t.tensor[2][0], t.tensor[2][1], t.tensor[2][2]);
        fprintf(f, "eigenvecs{m}{%u} = [%g, %g, %g; %g, %g, %g; %g, %g, %g];\n", k + 1,
                t.eigvec[0][0], t.eigvec[0][1], t.eigvec[0][2], t.eigvec[1][0], t.eigvec[1][1],
                t.eigvec[1][2], t.eigvec[2][0], t.eigvec[2][1], t.eigvec[2][2]);
        fprintf(f, "eigenvals{m}{%u} = [%g,%g,%g];\n", k + 1, t.eigval[0], t.eigval[1], t.eigval[2]);
        fprintf(f, "md{m}(%u) = %g;\n", k + 1, t.md);
        fprintf(f, "fa{m}(%u) = %g;\n", k + 1, t.fa);
        fprintf(f, "rd{m}(%u) = %g;\n", k + 1, t.rd);
    }
    fprintf(f, "%%The following variables give results for each gradient run.\n");
    fprintf(f, "%%mrSigTotal{m}(c,s) is the total signal after gradient run s during diffusion simulation no. m\n");
    fprintf(f, "%%mrSigCompartments{m}(c,s) is the contribution of compartment no. (c-2) to the \n");
    fprintf(f, "%%total signal after gradient run s during diffusion simulation no. m. c=1 represents all compartments. \n");
    for (uint32_t j = 0; j < ng; ++j) {
        const float *r = rows + 6 * (size_t)j;
        const float *s = sig + 5 * (size_t)j;
        fprintf(f, "\ns=s+1;\n");
        fprintf(f, "delta{m}(s) = %g;\n", r[0]);
        fprintf(f, "Delta{m}(s) = %g;\n", r[1]);
        fprintf(f, "readOut{m}(s) = %g;\n", r[2]);
        fprintf(f, "dwGrads{m}(:,s) = [%g,%g,%g];\n\n", r[3], r[4], r[5]);
        fprintf(f, "mrSigTotal{m}(s) = %g;\n\n", s[0]);
        for (unsigned k = 0; k < 4; ++k) fprintf(f, "mrSigCompartments{m}(%u,s) = %g;\n", k + 1, s[k + 1]);
        fflush(f);
    }
    int bad = ferror(f);
    fclose(f);
    return bad ? fail(DSIM_ERR_IO, "write error on %s", path) : DSIM_OK;
}

extern "C" const char *dsim_io_last_error(void) { return g_ioerr.c_str(); }
