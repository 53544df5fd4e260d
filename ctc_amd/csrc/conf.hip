// Configuration: the subset of the libconfig 1.7 format the reference's
// parameter files use (groups, arrays, lists, int / float / bool / string
// scalars, comments, @include), read into the four layers of the
// reference's Config (src/conf/imp/type.h): program setters, argument
// string, optional file, default file, looked up in that order
// (src/conf/imp/get.h:1-20).  Lookups are typed as in libconfig: "2" is an
// int and "2.0" a float, and asking for the other type is an error.
//
// Host code only; no GPU is touched.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "udx_internal.h"

namespace udx {
namespace {

enum Kind { K_GROUP, K_ARRAY, K_LIST, K_INT, K_FLOAT, K_BOOL, K_STRING };

struct Node {
    Kind kind = K_GROUP;
    std::string name;
    long long i = 0;
    double f = 0;
    bool b = false;
    std::string s;
    std::vector<std::unique_ptr<Node>> kids;
    Node *child(const std::string &n) const {
        for (auto &k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
};

struct Parser {
    std::string src, file, incdir;
    size_t p = 0;
    int line = 1, depth = 0;
    std::string err;

    bool fail(const std::string &m) {
        if (err.empty()) {
            std::ostringstream o;
            o << (file.empty() ? "<string>" : file) << ":" << line << ": " << m;
            err = o.str();
        }
        return false;
    }
    int peek() const { return p < src.size() ? (unsigned char)src[p] : -1; }
    void skip() {
        for (;;) {
            while (p < src.size() && std::isspace((unsigned char)src[p])) {
                if (src[p] == '\n') line++;
                p++;
            }
            if (p < src.size() && (src[p] == '#' || (src[p] == '/' && p + 1 < src.size() && src[p + 1] == '/'))) {
                while (p < src.size() && src[p] != '\n') p++;
                continue;
            }
            if (p + 1 < src.size() && src[p] == '/' && src[p + 1] == '*') {
                p += 2;
                while (p + 1 < src.size() && !(src[p] == '*' && src[p + 1] == '/')) {
                    if (src[p] == '\n') line++;
                    p++;
                }
                p += 2;
                continue;
            }
            return;
        }
    }
    bool name(std::string &out) {
        skip();
        int c = peek();
        if (!(std::isalpha(c) || c == '*')) return fail("expected a setting name");
        size_t b = p;
        while (p < src.size() &&
               (std::isalnum((unsigned char)src[p]) || src[p] == '_' || src[p] == '-' || src[p] == '*'))
            p++;
        out = src.substr(b, p - b);
        return true;
    }
    bool string_lit(std::string &out) {
        out.clear();
        for (;;) {
            skip();
            if (peek() != '"') return true;
            p++;
            while (p < src.size() && src[p] != '"') {
                char c = src[p++];
                if (c == '\\' && p < src.size()) {
                    char e = src[p++];
                    if (e == 'n') c = '\n';
                    else if (e == 't') c = '\t';
                    else if (e == 'r') c = '\r';
                    else if (e == 'f') c = '\f';
                    else if (e == 'x' && p + 1 < src.size()) {
                        c = (char)std::strtol(src.substr(p, 2).c_str(), nullptr, 16);
                        p += 2;
                    } else c = e;
                }
                if (c == '\n') line++;
                out.push_back(c);
            }
            if (peek() != '"') return fail("unterminated string");
            p++;
        }
    }
    bool scalar(Node &n) {
        skip();
        int c = peek();
        if (c == '"') {
            n.kind = K_STRING;
            return string_lit(n.s);
        }
        size_t b = p;
        while (p < src.size() &&
               (std::isalnum((unsigned char)src[p]) || src[p] == '.' || src[p] == '-' || src[p] == '+'))
            p++;
        std::string t = src.substr(b, p - b);
        if (t.empty()) return fail("expected a value");
        std::string lo = t;
        for (auto &ch : lo) ch = (char)std::tolower((unsigned char)ch);
        if (lo == "true" || lo == "false") {
            n.kind = K_BOOL;
            n.b = lo == "true";
            return true;
        }
        const bool hex = lo.size() > 2 && (lo.compare(0, 2, "0x") == 0 || lo.compare(0, 3, "-0x") == 0);
        const bool flt = !hex && (lo.find('.') != std::string::npos || lo.find('e') != std::string::npos);
        char *end = nullptr;
        errno = 0;
        if (flt) {
            n.kind = K_FLOAT;
            n.f = std::strtod(t.c_str(), &end);
        } else {
            n.kind = K_INT;
            std::string u = t;
            if (!u.empty() && (u.back() == 'L' || u.back() == 'l')) u.pop_back();
            if (!u.empty() && (u.back() == 'L' || u.back() == 'l')) u.pop_back();
            n.i = std::strtoll(u.c_str(), &end, hex ? 16 : 10);
            if (end && *end == '\0') return errno == 0 || fail("integer out of range: " + t);
            return fail("bad number: " + t);
        }
        if (!end || *end != '\0') return fail("bad number: " + t);
        return true;
    }
    bool value(Node &n) {
        skip();
        int c = peek();
        if (c == '{') {
            p++;
            n.kind = K_GROUP;
            if (!settings(n, '}')) return false;
            p++;
            return true;
        }
        if (c == '[' || c == '(') {
            const char close = c == '[' ? ']' : ')';
            n.kind = c == '[' ? K_ARRAY : K_LIST;
            p++;
            skip();
            if (peek() == close) {
                p++;
                return true;
            }
            for (;;) {
                std::unique_ptr<Node> e(new Node);
                if (!value(*e)) return false;
                if (n.kind == K_ARRAY) {
                    if (e->kind == K_GROUP || e->kind == K_ARRAY || e->kind == K_LIST)
                        return fail("array elements must be scalars");
                    if (!n.kids.empty() && n.kids[0]->kind != e->kind)
                        return fail("array elements must share one type");
                }
                n.kids.push_back(std::move(e));
                skip();
                if (peek() == ',') {
                    p++;
                    continue;
                }
                if (peek() == close) {
                    p++;
                    return true;
                }
                return fail(std::string("expected ',' or '") + close + "'");
            }
        }
        return scalar(n);
    }
    bool include(Node &g) {
        std::string f;
        if (!string_lit(f)) return false;
        if (depth > 10) return fail("@include nested too deeply");
        std::string path = (!f.empty() && f[0] == '/') || incdir.empty() ? f : incdir + "/" + f;
        std::ifstream in(path);
        if (!in) return fail("cannot open included file '" + path + "'");
        std::stringstream ss;
        ss << in.rdbuf();
        Parser sub;
        sub.src = ss.str();
        sub.file = path;
        sub.incdir = incdir;
        sub.depth = depth + 1;
        if (!sub.settings(g, -1)) {
            err = sub.err;
            return false;
        }
        return true;
    }
    // settings until `close` (or the end of the input for -1)
    bool settings(Node &g, int close) {
        for (;;) {
            skip();
            int c = peek();
            if (c == close || (close == -1 && c == -1)) return true;
            if (c == -1) return fail("unexpected end of input");
            if (c == '@') {
                size_t b = ++p;
                while (p < src.size() && std::isalpha((unsigned char)src[p])) p++;
                if (src.compare(b, p - b, "include") != 0) return fail("unknown directive");
                if (!include(g)) return false;
                continue;
            }
            std::unique_ptr<Node> n(new Node);
            if (!name(n->name)) return false;
            skip();
            if (peek() != '=' && peek() != ':') return fail("expected '=' or ':' after " + n->name);
            p++;
            if (!value(*n)) return false;
            skip();
            if (peek() == ';' || peek() == ',') p++;
            if (g.child(n->name)) return fail("duplicate setting name '" + n->name + "'");
            g.kids.push_back(std::move(n));
        }
    }
};

const char *kind_name(Kind k) {
    switch (k) {
        case K_INT: return "int";
        case K_FLOAT: return "float";
        case K_BOOL: return "bool";
        case K_STRING: return "string";
        case K_ARRAY: return "array";
        case K_LIST: return "list";
        default: return "group";
    }
}

}  // namespace
}  // namespace udx

using namespace udx;

// layers in lookup order (src/conf/imp/type.h:2-8)
enum { L_EXE, L_ARG, L_OPT, L_DEF, NLAYER };

struct udx_conf {
    Node layer[NLAYER];
    std::string incdir;
    std::vector<std::string> strings;   // storage of returned strings
};

static const Node *find(const udx_conf *c, const char *desc) {
    for (int l = 0; l < NLAYER; l++) {
        const Node *n = &c->layer[l];
        std::string d(desc);
        size_t b = 0;
        while (n) {
            size_t e = d.find('.', b);
            std::string part = d.substr(b, e == std::string::npos ? std::string::npos : e - b);
            if (n->kind != K_GROUP) {
                n = nullptr;
                break;
            }
            n = n->child(part);
            if (e == std::string::npos) break;
            b = e + 1;
        }
        if (n) return n;
    }
    return nullptr;
}

static int parse_into(Node &g, const std::string &text, const std::string &file, const std::string &incdir) {
    Parser P;
    P.src = text;
    P.file = file;
    P.incdir = incdir;
    if (!P.settings(g, -1)) {
        set_error("%s", P.err.c_str());
        return UDX_EINVAL;
    }
    return UDX_OK;
}

static int read_file(Node &g, const char *path, const std::string &incdir) {
    std::ifstream in(path);
    UDX_CHECK(in, UDX_EINVAL, "conf: cannot open '%s'", path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse_into(g, ss.str(), path, incdir);
}

static Node *make_path(Node &root, const char *desc) {
    std::string d(desc);
    Node *n = &root;
    size_t b = 0;
    for (;;) {
        size_t e = d.find('.', b);
        std::string part = d.substr(b, e == std::string::npos ? std::string::npos : e - b);
        Node *k = n->child(part);
        if (!k) {
            n->kids.emplace_back(new Node);
            k = n->kids.back().get();
            k->name = part;
        }
        if (e == std::string::npos) return k;
        if (k->kind != K_GROUP) {
            k->kind = K_GROUP;
            k->kids.clear();
        }
        n = k;
        b = e + 1;
    }
}

#define CONF_FIND(kind_, what)                                                                           \
    const Node *n = find(c, desc);                                                                       \
    UDX_CHECK(n, UDX_EINVAL, "conf: could not find <%s>", desc);                                         \
    UDX_CHECK(n->kind == (kind_), UDX_EINVAL, "conf: <%s> has wrong type: expected %s, found %s", desc, what, \
              kind_name(n->kind))

static int vlookup(const udx_conf *c, const char *desc, Kind ek, const char *what, int maxn, int *n_out,
                   const Node **out) {
    CONF_FIND(K_ARRAY, what);
    const int len = (int)n->kids.size();
    UDX_CHECK(len <= maxn, UDX_EINVAL, "conf: %s: too many components: %d / %d", desc, len, maxn);
    for (int j = 0; j < len; j++)
        UDX_CHECK(n->kids[j]->kind == ek, UDX_EINVAL, "conf: <%s> has wrong type: expected %s", desc, what);
    *n_out = len;
    *out = n;
    return UDX_OK;
}

extern "C" {

int udx_conf_ini(udx_conf **c) {
    UDX_CHECK(c, UDX_EINVAL, "udx_conf_ini: null output");
    *c = new udx_conf();
    return UDX_OK;
}

int udx_conf_fin(udx_conf *c) {
    delete c;
    return UDX_OK;
}

int udx_conf_set_include_dir(udx_conf *c, const char *dir) {
    UDX_CHECK(c && dir, UDX_EINVAL, "udx_conf_set_include_dir: null argument");
    c->incdir = dir;
    return UDX_OK;
}

int udx_conf_read(udx_conf *c, const char *default_file, const char *opt_file, const char *args) {
    UDX_CHECK(c, UDX_EINVAL, "udx_conf_read: null handle");
    if (default_file) UDX_TRY(read_file(c->layer[L_DEF], default_file, c->incdir));
    if (opt_file) UDX_TRY(read_file(c->layer[L_OPT], opt_file, c->incdir));
    if (args) UDX_TRY(parse_into(c->layer[L_ARG], args, "", c->incdir));
    return UDX_OK;
}

int udx_conf_lookup_int(const udx_conf *c, const char *desc, int *a) {
    UDX_CHECK(c && desc && a, UDX_EINVAL, "udx_conf_lookup_int: null argument");
    CONF_FIND(K_INT, "int");
    *a = (int)n->i;
    return UDX_OK;
}

int udx_conf_lookup_float(const udx_conf *c, const char *desc, float *a) {
    UDX_CHECK(c && desc && a, UDX_EINVAL, "udx_conf_lookup_float: null argument");
    CONF_FIND(K_FLOAT, "float");
    *a = (float)n->f;
    return UDX_OK;
}

int udx_conf_lookup_bool(const udx_conf *c, const char *desc, int *a) {
    UDX_CHECK(c && desc && a, UDX_EINVAL, "udx_conf_lookup_bool: null argument");
    CONF_FIND(K_BOOL, "bool");
    *a = n->b ? 1 : 0;
    return UDX_OK;
}

int udx_conf_lookup_string(const udx_conf *c, const char *desc, const char **a) {
    UDX_CHECK(c && desc && a, UDX_EINVAL, "udx_conf_lookup_string: null argument");
    CONF_FIND(K_STRING, "string");
    *a = n->s.c_str();
    return UDX_OK;
}

int udx_conf_lookup_vint(const udx_conf *c, const char *desc, int maxn, int *n, int *a) {
    UDX_CHECK(c && desc && n && a, UDX_EINVAL, "udx_conf_lookup_vint: null argument");
    const Node *arr;
    UDX_TRY(vlookup(c, desc, K_INT, "array of int", maxn, n, &arr));
    for (int j = 0; j < *n; j++) a[j] = (int)arr->kids[j]->i;
    return UDX_OK;
}

int udx_conf_lookup_vfloat(const udx_conf *c, const char *desc, int maxn, int *n, float *a) {
    UDX_CHECK(c && desc && n && a, UDX_EINVAL, "udx_conf_lookup_vfloat: null argument");
    const Node *arr;
    UDX_TRY(vlookup(c, desc, K_FLOAT, "array of float", maxn, n, &arr));
    for (int j = 0; j < *n; j++) a[j] = (float)arr->kids[j]->f;
    return UDX_OK;
}

int udx_conf_lookup_vstring(const udx_conf *c, const char *desc, int maxn, int *n, const char **a) {
    UDX_CHECK(c && desc && n && a, UDX_EINVAL, "udx_conf_lookup_vstring: null argument");
    const Node *arr;
    UDX_TRY(vlookup(c, desc, K_STRING, "array of string", maxn, n, &arr));
    for (int j = 0; j < *n; j++) a[j] = arr->kids[j]->s.c_str();
    return UDX_OK;
}

int udx_conf_lookup_int3(const udx_conf *c, const char *desc, int a[3]) {
    int n = 0;
    UDX_TRY(udx_conf_lookup_vint(c, desc, 3, &n, a));
    UDX_CHECK(n == 3, UDX_EINVAL, "conf: fail to read `%s`: int3 must have 3 components, found %d", desc, n);
    return UDX_OK;
}

int udx_conf_lookup_float3(const udx_conf *c, const char *desc, float a[3]) {
    int n = 0;
    UDX_TRY(udx_conf_lookup_vfloat(c, desc, 3, &n, a));
    UDX_CHECK(n == 3, UDX_EINVAL, "conf: fail to read `%s`: float3 must have 3 components, found %d", desc, n);
    return UDX_OK;
}

int udx_conf_has(const udx_conf *c, const char *desc) { return c && desc && find(c, desc) ? 1 : 0; }

int udx_conf_set_int(udx_conf *c, const char *desc, int a) {
    UDX_CHECK(c && desc, UDX_EINVAL, "udx_conf_set_int: null argument");
    Node *n = make_path(c->layer[L_EXE], desc);
    n->kids.clear();
    n->kind = K_INT;
    n->i = a;
    return UDX_OK;
}

int udx_conf_set_float(udx_conf *c, const char *desc, float a) {
    UDX_CHECK(c && desc, UDX_EINVAL, "udx_conf_set_float: null argument");
    Node *n = make_path(c->layer[L_EXE], desc);
    n->kids.clear();
    n->kind = K_FLOAT;
    n->f = a;
    return UDX_OK;
}

int udx_conf_set_bool(udx_conf *c, const char *desc, int a) {
    UDX_CHECK(c && desc, UDX_EINVAL, "udx_conf_set_bool: null argument");
    Node *n = make_path(c->layer[L_EXE], desc);
    n->kids.clear();
    n->kind = K_BOOL;
    n->b = a != 0;
    return UDX_OK;
}

int udx_conf_set_string(udx_conf *c, const char *desc, const char *a) {
    UDX_CHECK(c && desc && a, UDX_EINVAL, "udx_conf_set_string: null argument");
    Node *n = make_path(c->layer[L_EXE], desc);
    n->kids.clear();
    n->kind = K_STRING;
    n->s = a;
    return UDX_OK;
}

// ---- parameter groups ---------------------------------------------------
// <ns>.{dpd, a, g, spow}: the grey (single colour) DPD view, sigma from kBT
// and dt (src/pair/conf.cpp:17-57, src/pair/imp/main.h:54-62)
int udx_pair_set_conf(const udx_conf *c, const char *ns, float kBT, float dt, udx_pair_dpd *p) {
    UDX_CHECK(c && ns && p, UDX_EINVAL, "udx_pair_set_conf: null argument");
    std::string b(ns);
    int dpd = 0, na = 0, ng = 0;
    UDX_TRY(udx_conf_lookup_bool(c, (b + ".dpd").c_str(), &dpd));
    UDX_CHECK(dpd, UDX_EINVAL, "conf: %s.dpd is false; only DPD pairs are supported", ns);
    float a[16], g[16], spow;
    UDX_TRY(udx_conf_lookup_vfloat(c, (b + ".a").c_str(), 16, &na, a));
    UDX_TRY(udx_conf_lookup_vfloat(c, (b + ".g").c_str(), 16, &ng, g));
    UDX_TRY(udx_conf_lookup_float(c, (b + ".spow").c_str(), &spow));
    UDX_CHECK(na == ng && na > 0, UDX_EINVAL, "conf: %s.a and %s.g must have the same length: %d / %d", ns, ns, na,
              ng);
    p->a = a[0];
    p->g = g[0];
    p->spow = spow;
    udx_pair_compute_dpd_sigma(kBT, dt, p);
    return UDX_OK;
}

// <ns>.{gammaC, ..., totVolume} (src/rbc/params/conf.cpp:11-34)
int udx_rbc_params_set_conf(const udx_conf *c, const char *ns, udx_rbc_params *p) {
    UDX_CHECK(c && ns && p, UDX_EINVAL, "udx_rbc_params_set_conf: null argument");
    const char *keys[] = {"gammaC", "gammaT", "kBT", "kb", "phi", "ks", "x0", "mpow", "ka", "kd", "kv",
                          "totArea", "totVolume"};
    float *dst[] = {&p->gammaC, &p->gammaT, &p->kBT, &p->kb, &p->phi, &p->ks, &p->x0, &p->mpow,
                    &p->ka, &p->kd, &p->kv, &p->totArea, &p->totVolume};
    for (int k = 0; k < 13; k++) UDX_TRY(udx_conf_lookup_float(c, (std::string(ns) + "." + keys[k]).c_str(), dst[k]));
    return UDX_OK;
}

// wvel.type and its keys: "constant" u; "shear" gdot, vdir, gdir, half;
// "shear sin" the same and w; "hele shaw" u, h (src/wall/wvel/conf.cpp:10-47)
int udx_wall_vel_set_conf(const udx_conf *c, udx_wall_vel *v) {
    UDX_CHECK(c && v, UDX_EINVAL, "udx_wall_vel_set_conf: null argument");
    const char *type;
    UDX_TRY(udx_conf_lookup_string(c, "wvel.type", &type));
    udx_wall_vel w{};
    const std::string t(type);
    if (t == "constant") {
        w.type = UDX_WVEL_CSTE;
        UDX_TRY(udx_conf_lookup_float3(c, "wvel.u", w.u));
    } else if (t == "shear" || t == "shear sin") {
        w.type = t == "shear" ? UDX_WVEL_SHEAR : UDX_WVEL_SHEAR_SIN;
        UDX_TRY(udx_conf_lookup_float(c, "wvel.gdot", &w.gdot));
        UDX_TRY(udx_conf_lookup_int(c, "wvel.vdir", &w.vdir));
        UDX_TRY(udx_conf_lookup_int(c, "wvel.gdir", &w.gdir));
        UDX_TRY(udx_conf_lookup_int(c, "wvel.half", &w.half));
        if (w.type == UDX_WVEL_SHEAR_SIN) UDX_TRY(udx_conf_lookup_float(c, "wvel.w", &w.w));
    } else if (t == "hele shaw") {
        w.type = UDX_WVEL_HS;
        UDX_TRY(udx_conf_lookup_float(c, "wvel.u", &w.hs_u));
        UDX_TRY(udx_conf_lookup_float(c, "wvel.h", &w.hs_h));
    } else {
        UDX_CHECK(false, UDX_EINVAL, "unknown type <%s>", type);
    }
    *v = w;
    return UDX_OK;
}

int udx_wall_repulse_set_conf(const udx_conf *c, const char *ns, float *lambda) {
    UDX_CHECK(c && ns && lambda, UDX_EINVAL, "udx_wall_repulse_set_conf: null argument");
    const char *grp;
    UDX_TRY(udx_conf_lookup_string(c, (std::string(ns) + ".repulsion").c_str(), &grp));
    if (std::string(grp) == "none") {
        *lambda = 0.f;
        return UDX_OK;
    }
    float l;
    UDX_TRY(udx_conf_lookup_float(c, (std::string(grp) + ".lambda").c_str(), &l));
    *lambda = l;
    return UDX_OK;
}

}  // extern "C"
