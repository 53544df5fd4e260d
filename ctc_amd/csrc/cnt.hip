// Membrane-membrane contact (src/cnt) for one object type on one periodic
// domain: DPD plus the capped repulsive Lennard-Jones term between membrane
// vertices, and between the periodic images of the vertices outside the
// domain (the object halo of src/exch/obj) and the vertices.
//
// Passes: the vertex cell list (cells clamped into the domain: vertices
// outside it are projected onto the boundary cells; stable within a cell), the
// image list (flags per (receive fragment, vertex), one exclusive scan: the
// order receive fragment 0..25, vertices ascending), the image cell list, and
// one thread per vertex summing in one accumulator its bulk pairs, the images
// acting on it and its own images' pairs, added to its force once.  The
// reference adds with float atomics; this order is fixed.
#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {
namespace {

constexpr int CNT_THREADS = 128;

struct DpdLj {
    float a, g, s, spow, ljs, lje;
};

// DPD + LJ force on a from b (src/pair/dev.h:34-83); r and 1/r correctly
// rounded, the DPD part as the solvent's
__device__ __forceinline__ bool cnt_pair(const DpdLj &p, const float *a, const float *b, float rnd, float f[3]) {
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
    f[0] = f[1] = f[2] = 0.0f;
    if (r2 >= 1.0f) return false;
    float ex, ey, ez, r, invr;
    if (r2 < 1e-6f) {
        ex = ey = ez = r = invr = 0.0f;
    } else {
        r = __builtin_sqrtf(r2);
        invr = 1.0f / r;
        ex = dx * invr;
        ey = dy * invr;
        ez = dz * invr;
    }
    const float dvx = a[3] - b[3], dvy = a[4] - b[4], dvz = a[5] - b[5];
    const float ev = __builtin_fmaf(ez, dvz, __builtin_fmaf(ex, dvx, ey * dvy));
    float rm = 1.0f - r;
    rm = rm > 0.0f ? rm : 0.0f;
    const float wr = p.spow == 0.25f ? __builtin_sqrtf(__builtin_sqrtf(rm))
                     : (p.spow == 0.5f ? __builtin_sqrtf(rm) : (p.spow == 1.0f ? rm : __builtin_powf(rm, p.spow)));
    float f0 = __builtin_fmaf(-p.g * wr, ev, p.s * rnd) * wr;
    f0 = __builtin_fmaf(p.a, rm, f0);
    const float t2 = p.ljs * p.ljs * invr * invr;
    const float t4 = t2 * t2, t6 = t4 * t2;
    float flj = p.lje * 24 * invr * t6 * (2 * t6 - 1);
    flj = flj > 1e4f ? 1e4f : (flj < 0.0f ? 0.0f : flj);
    f0 = f0 + flj;
    f[0] = f0 * ex;
    f[1] = f0 * ey;
    f[2] = f0 * ez;
    return true;
}

// floor(x) + L/2, clamped into the domain (vertices) or into [-2, L + 1]
// (image centres: a block that far out reaches no cell of the domain)
__device__ __forceinline__ int ccell(float x, int L, bool clampit) {
    const float f = floorf(x);
    const int lo = clampit ? 0 : -2, hi = clampit ? L - 1 : L + 1;
    if (!(f >= (float)(lo - L / 2))) return lo;
    if (!(f <= (float)(hi - L / 2))) return hi;
    return L / 2 + (int)f;
}

__device__ __forceinline__ void load6(const float *pp, int i, float o[6]) {
    for (int c = 0; c < 6; c++) o[c] = pp[6 * (size_t)i + c];
}

// clamped cell of each point; per-cell counts (n_dev: a device-side count
// below nmax, or null)
__global__ __launch_bounds__(CNT_THREADS) void k_cnt_cells(int nmax, const int *n_dev, int Lx, int Ly, int Lz,
                                                           const float *pp, int *cid, int *counts) {
    const int i = blockIdx.x * CNT_THREADS + threadIdx.x;
    if (i >= nmax || (n_dev && i >= *n_dev)) return;
    const float *p = pp + 6 * (size_t)i;
    const int c = ccell(p[0], Lx, true) + Lx * (ccell(p[1], Ly, true) + Ly * ccell(p[2], Lz, true));
    cid[i] = c;
    atomicAdd(counts + c, 1);
}

__global__ __launch_bounds__(CNT_THREADS) void k_cnt_scatter(int nmax, const int *n_dev, const int *cid,
                                                             const int *starts, int *fill, int *order) {
    const int i = blockIdx.x * CNT_THREADS + threadIdx.x;
    if (i >= nmax || (n_dev && i >= *n_dev)) return;
    const int c = cid[i];
    order[starts[c] + atomicAdd(fill + c, 1)] = i;
}

// ascending index within each cell; clears the counters for the next call
__global__ __launch_bounds__(CNT_THREADS) void k_cnt_sortcells(int ncell, const int *starts, int *counts, int *fill,
                                                               int *order) {
    const int c = blockIdx.x * CNT_THREADS + threadIdx.x;
    if (c >= ncell) return;
    const int s = starts[c], e = starts[c + 1];
    counts[c] = 0;
    fill[c] = 0;
    for (int a = s + 1; a < e; a++) {
        const int x = order[a];
        int b = a - 1;
        while (b >= s && order[b] > x) {
            order[b + 1] = order[b];
            b--;
        }
        order[b + 1] = x;
    }
}

__device__ __forceinline__ void frag_dir(int i, int d[3]) {
    const int t[3] = {i % 3, (i / 3) % 3, i / 9};
    for (int c = 0; c < 3; c++) d[c] = t[c] == 0 ? 1 : (t[c] == 1 ? -1 : 0);
}

// does vertex r (outside the domain) go to receive fragment i: its code on
// every axis where d_i is nonzero is -d_i (src/exch/map/dev.h:8-83)
__device__ __forceinline__ bool sends(int Lx, int Ly, int Lz, const float *r, int i) {
    const int L[3] = {Lx, Ly, Lz};
    int code[3], d[3];
    frag_dir(i, d);
    bool out = false;
    for (int c = 0; c < 3; c++) {
        code[c] = -1 + (r[c] >= (float)(-L[c] / 2)) + (r[c] >= (float)(L[c] / 2));
        out |= code[c] != 0;
    }
    if (!out) return false;
    for (int c = 0; c < 3; c++)
        if (d[c] != 0 && code[c] != -d[c]) return false;
    return true;
}

__global__ __launch_bounds__(CNT_THREADS) void k_cnt_flags(int n, int Lx, int Ly, int Lz, const float *pp,
                                                           int *flags) {
    const int g = blockIdx.x * CNT_THREADS + threadIdx.x;
    if (g >= 26 * n) return;
    const int i = g / n, v = g - i * n;
    flags[g] = sends(Lx, Ly, Lz, pp + 6 * (size_t)v, i) ? 1 : 0;
}

__global__ __launch_bounds__(CNT_THREADS) void k_cnt_images(int n, int Lx, int Ly, int Lz, const float *pp,
                                                            const int *flags, const int *pos, float *img,
                                                            int *img_v, int *n_img) {
    const int g = blockIdx.x * CNT_THREADS + threadIdx.x;
    if (g == 0) *n_img = pos[26 * n];
    if (g >= 26 * n || !flags[g]) return;
    const int i = g / n, v = g - i * n;
    int d[3];
    frag_dir(i, d);
    const int L[3] = {Lx, Ly, Lz};
    const int h = pos[g];
    for (int c = 0; c < 6; c++) img[6 * (size_t)h + c] = pp[6 * (size_t)v + c];
    for (int c = 0; c < 3; c++) img[6 * (size_t)h + c] = pp[6 * (size_t)v + c] + (float)(L[c] * d[c]);
    img_v[h] = v;
}

struct CntK {
    int n, Lx, Ly, Lz;
    DpdLj p;
    float seed_bulk, seed_halo;
    const float *pp, *img;
    const int *starts, *order, *hstarts, *horder, *flags, *pos;
};

__global__ __launch_bounds__(CNT_THREADS) void k_cnt_force(CntK K, float *ff, unsigned long long *pairs) {
    const int v = blockIdx.x * CNT_THREADS + threadIdx.x;
    if (v >= K.n) return;
    const int L[3] = {K.Lx, K.Ly, K.Lz};
    float rv[6];
    load6(K.pp, v, rv);
    int cv[3];
    for (int c = 0; c < 3; c++) cv[c] = ccell(rv[c], L[c], true);
    float acc[3] = {0, 0, 0};
    unsigned np = 0;
    // bulk pairs
    for (int z = cv[2] - 1; z <= cv[2] + 1; z++)
        for (int y = cv[1] - 1; y <= cv[1] + 1; y++)
            for (int x = cv[0] - 1; x <= cv[0] + 1; x++) {
                if (x < 0 || x >= L[0] || y < 0 || y >= L[1] || z < 0 || z >= L[2]) continue;
                const int cid = x + L[0] * (y + L[1] * z);
                for (int s = K.starts[cid]; s < K.starts[cid + 1]; s++) {
                    const int u = K.order[s];
                    if (u == v) continue;
                    float ru[6], f[3];
                    load6(K.pp, u, ru);
                    const int hi = u > v ? u : v, lo = u > v ? v : u;
                    const float rnd = mean0var1ii(K.seed_bulk, hi, lo);
                    if (!(v == hi ? cnt_pair(K.p, rv, ru, rnd, f) : cnt_pair(K.p, ru, rv, rnd, f))) continue;
                    if (v == hi) {
                        acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
                        np++;
                    } else {
                        acc[0] -= f[0]; acc[1] -= f[1]; acc[2] -= f[2];
                    }
                }
            }
    // images acting on v
    for (int z = cv[2] - 1; z <= cv[2] + 1; z++)
        for (int y = cv[1] - 1; y <= cv[1] + 1; y++)
            for (int x = cv[0] - 1; x <= cv[0] + 1; x++) {
                if (x < 0 || x >= L[0] || y < 0 || y >= L[1] || z < 0 || z >= L[2]) continue;
                const int cid = x + L[0] * (y + L[1] * z);
                for (int s = K.hstarts[cid]; s < K.hstarts[cid + 1]; s++) {
                    const int h = K.horder[s];
                    float ri[6], f[3];
                    load6(K.img, h, ri);
                    bool ok = true;
                    for (int c = 0; c < 3; c++) {
                        const int ch = ccell(ri[c], L[c], false);
                        if (cv[c] < ch - 1 || cv[c] > ch + 1) ok = false;
                    }
                    if (!ok) continue;
                    if (!cnt_pair(K.p, ri, rv, mean0var1ii(K.seed_halo, h, v), f)) continue;
                    acc[0] -= f[0]; acc[1] -= f[1]; acc[2] -= f[2];
                }
            }
    // v's own images against the vertices of their blocks
    for (int i = 0; i < 26; i++) {
        if (!K.flags[i * K.n + v]) continue;
        const int h = K.pos[i * K.n + v];
        float ri[6];
        load6(K.img, h, ri);
        int ch[3];
        for (int c = 0; c < 3; c++) ch[c] = ccell(ri[c], L[c], false);
        for (int z = ch[2] - 1; z <= ch[2] + 1; z++)
            for (int y = ch[1] - 1; y <= ch[1] + 1; y++)
                for (int x = ch[0] - 1; x <= ch[0] + 1; x++) {
                    if (x < 0 || x >= L[0] || y < 0 || y >= L[1] || z < 0 || z >= L[2]) continue;
                    const int cid = x + L[0] * (y + L[1] * z);
                    for (int s = K.starts[cid]; s < K.starts[cid + 1]; s++) {
                        const int b = K.order[s];
                        float rb[6], f[3];
                        load6(K.pp, b, rb);
                        if (!cnt_pair(K.p, ri, rb, mean0var1ii(K.seed_halo, h, b), f)) continue;
                        acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
                        np++;
                    }
                }
    }
    float *o = ff + 3 * (size_t)v;
    o[0] += acc[0];
    o[1] += acc[1];
    o[2] += acc[2];
    if (np) atomicAdd(pairs, (unsigned long long)np);
}

}  // namespace
}  // namespace udx

using namespace udx;

struct udx_cnt {
    udx_ctx *ctx;
    int L[3], max_n, ncell;
    udx_rnd trunk;   // (7119, 187, 18278, 15674): src/cnt/imp/main.h:22 on rank 0
    int *cid, *counts, *starts, *fill, *order, *scan_tmp;           // vertex cell list
    int *flags, *pos, *img_v, *n_img;                               // image list
    float *img;
    int *hcid, *hcounts, *hstarts, *hfill, *horder;                 // image cell list
    unsigned long long *pairs;
};

static void cnt_free(udx_cnt *c) {
    dev_free(c->cid);
    dev_free(c->counts);
    dev_free(c->starts);
    dev_free(c->fill);
    dev_free(c->order);
    dev_free(c->scan_tmp);
    dev_free(c->flags);
    dev_free(c->pos);
    dev_free(c->img_v);
    dev_free(c->n_img);
    dev_free(c->img);
    dev_free(c->hcid);
    dev_free(c->hcounts);
    dev_free(c->hstarts);
    dev_free(c->hfill);
    dev_free(c->horder);
    dev_free(c->pairs);
}

namespace udx {

udx_ctx *cnt_ctx(const udx_cnt *c) { return c->ctx; }

// contact forces of n vertices (AoS, device) added to ff (3 floats each)
int cnt_forces(hipStream_t s, udx_cnt *c, const udx_pair_dpd_lj &pr, int n, const float *pp, float seed_bulk,
               float seed_halo, float *ff) {
    UDX_CHECK(n <= c->max_n, UDX_ECAP, "contact: %d vertices above the capacity %d", n, c->max_n);
    if (n <= 0) return UDX_OK;
    const int *L = c->L;
    // vertex cell list (counts and fills are zero here: cleared at allocation
    // and after every sort)
    hipLaunchKernelGGL(k_cnt_cells, dim3(ceil_div(n, CNT_THREADS)), dim3(CNT_THREADS), 0, s, n, nullptr, L[0], L[1],
                       L[2], pp, c->cid, c->counts);
    UDX_TRY(scan_exclusive(s, c->counts, c->ncell, c->starts, c->scan_tmp));
    hipLaunchKernelGGL(k_cnt_scatter, dim3(ceil_div(n, CNT_THREADS)), dim3(CNT_THREADS), 0, s, n, nullptr, c->cid,
                       c->starts, c->fill, c->order);
    hipLaunchKernelGGL(k_cnt_sortcells, dim3(ceil_div(c->ncell, CNT_THREADS)), dim3(CNT_THREADS), 0, s, c->ncell,
                       c->starts, c->counts, c->fill, c->order);
    // images of the vertices outside the domain
    hipLaunchKernelGGL(k_cnt_flags, dim3(ceil_div(26 * n, CNT_THREADS)), dim3(CNT_THREADS), 0, s, n, L[0], L[1],
                       L[2], pp, c->flags);
    UDX_TRY(scan_exclusive(s, c->flags, 26 * n, c->pos, c->scan_tmp));
    hipLaunchKernelGGL(k_cnt_images, dim3(ceil_div(26 * n, CNT_THREADS)), dim3(CNT_THREADS), 0, s, n, L[0], L[1],
                       L[2], pp, c->flags, c->pos, c->img, c->img_v, c->n_img);
    const int maxh = 7 * n;   // a vertex has at most 7 images (3 faces, 3 edges, 1 corner)
    hipLaunchKernelGGL(k_cnt_cells, dim3(ceil_div(maxh, CNT_THREADS)), dim3(CNT_THREADS), 0, s, maxh, c->n_img, L[0],
                       L[1], L[2], c->img, c->hcid, c->hcounts);
    UDX_TRY(scan_exclusive(s, c->hcounts, c->ncell, c->hstarts, c->scan_tmp));
    hipLaunchKernelGGL(k_cnt_scatter, dim3(ceil_div(maxh, CNT_THREADS)), dim3(CNT_THREADS), 0, s, maxh, c->n_img,
                       c->hcid, c->hstarts, c->hfill, c->horder);
    hipLaunchKernelGGL(k_cnt_sortcells, dim3(ceil_div(c->ncell, CNT_THREADS)), dim3(CNT_THREADS), 0, s, c->ncell,
                       c->hstarts, c->hcounts, c->hfill, c->horder);
    CntK K{n, L[0], L[1], L[2], {pr.a, pr.g, pr.s, pr.spow, pr.ljs, pr.lje}, seed_bulk, seed_halo, pp, c->img,
           c->starts, c->order, c->hstarts, c->horder, c->flags, c->pos};
    hipLaunchKernelGGL(k_cnt_force, dim3(ceil_div(n, CNT_THREADS)), dim3(CNT_THREADS), 0, s, K, ff, c->pairs);
    UDX_LAUNCHED();
    return UDX_OK;
}

float cnt_draw(udx_cnt *c) { return udx_rnd_get(&c->trunk); }

}  // namespace udx

extern "C" {

int udx_cnt_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int max_n, udx_cnt **out) {
    UDX_CHECK(ctx && out && Lx > 0 && Ly > 0 && Lz > 0 && max_n > 0, UDX_EINVAL, "udx_cnt_ini: bad arguments");
    UDX_CHECK((long)26 * max_n < (1L << 31), UDX_ECAP, "udx_cnt_ini: %d vertices exceed the index range", max_n);
    udx_cnt *c = new udx_cnt();
    c->ctx = ctx;
    c->L[0] = Lx;
    c->L[1] = Ly;
    c->L[2] = Lz;
    c->max_n = max_n;
    c->ncell = Lx * Ly * Lz;
    udx_rnd_ini(7119, 187, 18278, 15674, &c->trunk);
    const size_t nc1 = (size_t)c->ncell + 1, n = (size_t)max_n, h = 7 * (size_t)max_n;
    const int scan_n = 26 * max_n > c->ncell ? 26 * max_n : c->ncell;
    int st = dev_alloc(&c->cid, n);
    if (!st) st = dev_alloc(&c->counts, nc1);
    if (!st) st = dev_alloc(&c->starts, nc1);
    if (!st) st = dev_alloc(&c->fill, nc1);
    if (!st) st = dev_alloc(&c->order, n);
    if (!st) st = dev_alloc(&c->scan_tmp, (size_t)scan_scratch_size(scan_n));
    if (!st) st = dev_alloc(&c->flags, 26 * n);
    if (!st) st = dev_alloc(&c->pos, 26 * n + 1);
    if (!st) st = dev_alloc(&c->img_v, h);
    if (!st) st = dev_alloc(&c->n_img, 1);
    if (!st) st = dev_alloc(&c->img, 6 * h);
    if (!st) st = dev_alloc(&c->hcid, h);
    if (!st) st = dev_alloc(&c->hcounts, nc1);
    if (!st) st = dev_alloc(&c->hstarts, nc1);
    if (!st) st = dev_alloc(&c->hfill, nc1);
    if (!st) st = dev_alloc(&c->horder, h);
    if (!st) st = dev_alloc(&c->pairs, 1);
    hipStream_t s = ctx->stream;
    if (!st && (hipMemsetAsync(c->counts, 0, sizeof(int) * nc1, s) != hipSuccess ||
                hipMemsetAsync(c->fill, 0, sizeof(int) * nc1, s) != hipSuccess ||
                hipMemsetAsync(c->hcounts, 0, sizeof(int) * nc1, s) != hipSuccess ||
                hipMemsetAsync(c->hfill, 0, sizeof(int) * nc1, s) != hipSuccess ||
                hipMemsetAsync(c->pairs, 0, sizeof(unsigned long long), s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        st = UDX_EHIP;
    if (st) {
        cnt_free(c);
        delete c;
        set_error("udx_cnt_ini: allocation failed");
        return st;
    }
    *out = c;
    return UDX_OK;
}

int udx_cnt_fin(udx_cnt *c) {
    if (!c) return UDX_OK;
    UDX_HIP(hipStreamSynchronize(c->ctx->stream));
    cnt_free(c);
    delete c;
    return UDX_OK;
}

int udx_cnt_apply_seed(udx_cnt *c, const udx_pair_dpd_lj *p, int n, const float *pp, float seed_bulk,
                       float seed_halo, float *ff) {
    UDX_CHECK(c && p && n >= 0 && (n == 0 || (pp && ff)), UDX_EINVAL, "udx_cnt_apply: bad arguments");
    UDX_TRY(cnt_forces(c->ctx->stream, c, *p, n, pp, seed_bulk, seed_halo, ff));
    UDX_HIP(hipStreamSynchronize(c->ctx->stream));
    return UDX_OK;
}

int udx_cnt_apply(udx_cnt *c, const udx_pair_dpd_lj *p, int n, const float *pp, float *ff) {
    UDX_CHECK(c, UDX_EINVAL, "udx_cnt_apply: null handle");
    const float sb = udx_rnd_get(&c->trunk), sh = udx_rnd_get(&c->trunk);
    return udx_cnt_apply_seed(c, p, n, pp, sb, sh, ff);
}

int udx_cnt_pairs(udx_cnt *c, unsigned long long *npairs, int reset) {
    UDX_CHECK(c && npairs, UDX_EINVAL, "udx_cnt_pairs: null argument");
    UDX_HIP(hipMemcpyAsync(npairs, c->pairs, sizeof *npairs, hipMemcpyDeviceToHost, c->ctx->stream));
    UDX_HIP(hipStreamSynchronize(c->ctx->stream));
    if (reset) UDX_HIP(hipMemsetAsync(c->pairs, 0, sizeof *npairs, c->ctx->stream));
    return UDX_OK;
}

}  // extern "C"
