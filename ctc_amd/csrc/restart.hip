// Restart files (src/io/restart/imp/main.h:1-141) in the bop particle format
// (https://github.com/amlucas/bop, un-vendored in the reference; the header
// layout below is the one of the reference's own golden files
// src/test_data/bop.ref.noflu.txt and src/test_data/wall.ref.freeze.wall.t0.txt):
//
//   BASE/CODE/ID.bop     text header
//       <n>
//       DATA_FILE: <ID>.values
//       DATA_FORMAT: float | int        (the reader also takes ascii, iascii)
//       VARIABLES: x y z vx vy vz | i
//       NRANK: 1                        (optional on read)
//       <n>                             (one count per rank)
//   BASE/CODE/ID.values  n * nvars little-endian 4-byte values
//
// ID is "templ" (UDX_RESTART_TEMPL), "final" (UDX_RESTART_FINAL) or %05d.
// Host code only: the buffers are host memory, as in the reference; the
// solvent and wall helpers copy between the device and a host staging buffer.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "udx_internal.h"

namespace {

std::string id_code(int id) {
    if (id == UDX_RESTART_TEMPL) return "templ";
    if (id == UDX_RESTART_FINAL) return "final";
    char b[32];
    snprintf(b, sizeof b, "%05d", id);
    return b;
}

std::string stem(const char *base, const char *code, int id) {
    return std::string(base) + "/" + code + "/" + id_code(id);
}

// BopFLOAT with 6 variables, or BopINT with 1 (src/io/restart/imp/main.h:104-118)
int write_bop(const char *base, const char *code, int id, long n, const void *data, bool is_int) {
    UDX_CHECK(base && code && n >= 0 && (n == 0 || data), UDX_EINVAL, "restart: bad arguments");
    UDX_CHECK(id >= 0 || id == UDX_RESTART_TEMPL || id == UDX_RESTART_FINAL, UDX_EINVAL, "restart: id %d", id);
    const std::string s = stem(base, code, id);
    const int nvars = is_int ? 1 : 6;
    FILE *f = fopen((s + ".values").c_str(), "wb");
    UDX_CHECK(f, UDX_EINVAL, "restart: could not open <%s.values>: %s", s.c_str(), strerror(errno));
    const size_t cnt = (size_t)n * nvars;
    const bool ok = cnt == 0 || fwrite(data, 4, cnt, f) == cnt;
    fclose(f);
    UDX_CHECK(ok, UDX_EINVAL, "restart: write failed <%s.values>", s.c_str());
    f = fopen((s + ".bop").c_str(), "w");
    UDX_CHECK(f, UDX_EINVAL, "restart: could not open <%s.bop>: %s", s.c_str(), strerror(errno));
    fprintf(f, "%ld\nDATA_FILE: %s.values\nDATA_FORMAT: %s\nVARIABLES: %s\nNRANK: 1\n%ld\n", n,
            id_code(id).c_str(), is_int ? "int" : "float", is_int ? "i" : "x y z vx vy vz", n);
    const bool closed = fclose(f) == 0;
    UDX_CHECK(closed, UDX_EINVAL, "restart: write failed <%s.bop>", s.c_str());
    return UDX_OK;
}

int count_words(const std::string &v) {
    int k = 0;
    bool in = false;
    for (char c : v) {
        const bool sp = c == ' ' || c == '\t' || c == '\r' || c == '\n';
        if (!sp && !in) k++;
        in = !sp;
    }
    return k;
}

int read_bop(const char *base, const char *code, int id, int maxn, int *n_out, void *data, bool is_int) {
    UDX_CHECK(base && code && n_out && maxn >= 0 && (maxn == 0 || data), UDX_EINVAL, "restart: bad arguments");
    const std::string s = stem(base, code, id);
    const std::string hname = s + ".bop";
    FILE *f = fopen(hname.c_str(), "r");
    UDX_CHECK(f, UDX_EINVAL, "restart: could not open <%s>: %s", hname.c_str(), strerror(errno));
    std::vector<std::string> lines;
    char buf[4096];
    while (fgets(buf, sizeof buf, f)) {
        std::string l(buf);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        lines.push_back(l);
    }
    fclose(f);
    UDX_CHECK(!lines.empty(), UDX_EINVAL, "restart: empty header <%s>", hname.c_str());
    char *end = nullptr;
    const long n = strtol(lines[0].c_str(), &end, 10);
    UDX_CHECK(end != lines[0].c_str() && n >= 0, UDX_EINVAL, "restart: bad count in <%s>", hname.c_str());
    std::string dfile, fmt, vars;
    for (size_t k = 1; k < lines.size(); k++) {
        const std::string &l = lines[k];
        auto val = [&](const char *key) -> const char * {
            const size_t m = strlen(key);
            if (l.compare(0, m, key) != 0) return nullptr;
            const char *p = l.c_str() + m;
            while (*p == ' ') p++;
            return p;
        };
        if (const char *v = val("DATA_FILE:")) dfile = v;
        else if (const char *v = val("DATA_FORMAT:")) fmt = v;
        else if (const char *v = val("VARIABLES:")) vars = v;
    }
    UDX_CHECK(!dfile.empty() && !fmt.empty(), UDX_EINVAL, "restart: <%s> lacks DATA_FILE or DATA_FORMAT",
              hname.c_str());
    const int nvars = count_words(vars);
    const int want = is_int ? 1 : 6;
    UDX_CHECK(nvars == want, UDX_EINVAL, "restart: <%s> has %d variables, expected %d", hname.c_str(), nvars, want);
    UDX_CHECK(n <= maxn, UDX_ECAP, "restart: <%s> holds %ld entries, capacity %d", hname.c_str(), n, maxn);
    // the values file sits beside the header
    const size_t slash = hname.rfind('/');
    const std::string dname = (slash == std::string::npos ? std::string() : hname.substr(0, slash + 1)) + dfile;
    const bool ascii = fmt == "ascii" || fmt == "iascii";
    const bool fint = fmt == "int" || fmt == "iascii";
    UDX_CHECK(fmt == "float" || fmt == "int" || ascii, UDX_EINVAL, "restart: <%s>: format <%s> not supported",
              hname.c_str(), fmt.c_str());
    UDX_CHECK(fint == is_int, UDX_EINVAL, "restart: <%s>: format <%s> for %s data", hname.c_str(), fmt.c_str(),
              is_int ? "integer" : "particle");
    f = fopen(dname.c_str(), ascii ? "r" : "rb");
    UDX_CHECK(f, UDX_EINVAL, "restart: could not open <%s>: %s", dname.c_str(), strerror(errno));
    const size_t cnt = (size_t)n * nvars;
    bool ok = true;
    if (!ascii) {
        ok = cnt == 0 || fread(data, 4, cnt, f) == cnt;
    } else {
        for (size_t k = 0; k < cnt && ok; k++) {
            if (is_int) ok = fscanf(f, "%d", (int *)data + k) == 1;
            else ok = fscanf(f, "%f", (float *)data + k) == 1;
        }
    }
    fclose(f);
    UDX_CHECK(ok, UDX_EINVAL, "restart: <%s> holds fewer than %zu values", dname.c_str(), cnt);
    *n_out = (int)n;
    return UDX_OK;
}

}  // namespace

extern "C" {

int udx_restart_write_pp(const char *base, const char *code, int id, long n, const float *pp) {
    return write_bop(base, code, id, n, pp, false);
}

int udx_restart_read_pp(const char *base, const char *code, int id, int maxn, int *n, float *pp) {
    return read_bop(base, code, id, maxn, n, pp, false);
}

int udx_restart_write_ii(const char *base, const char *code, int id, long n, const int *ii) {
    return write_bop(base, code, id, n, ii, true);
}

int udx_restart_read_ii(const char *base, const char *code, int id, int maxn, int *n, int *ii) {
    return read_bop(base, code, id, maxn, n, ii, true);
}

}  // extern "C"
