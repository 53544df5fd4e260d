// Juelicher bending forces (src/rbc/force/bending/juelicher) for nc cells of
// one template, added to the membrane force.
//
// The terms are the reference's (kb = 1, H0 = -1/2 hard-coded there; the
// dihedral-gradient formula is minus the gradient of its angle; each
// triangle's area gradient carries the coefficient of the vertex it moves).
// Geometry in double, every term rounded to float once.  The reference
// accumulates with float atomics; here every sum has a fixed order:
//   k_jul_edges      one thread per (cell, edge): theta and l theta
//   k_jul_tris       one thread per (cell, triangle): area / 3
//   k_jul_vertices   one thread per (cell, vertex): A_v and L_v over its
//                    triangles / edges in ascending order
//   k_jul_cells      one thread per cell: the totals (ascending vertices) and
//                    the area-difference coefficient C
//   k_jul_force      one thread per (cell, vertex): edge-length terms, then
//                    dihedral terms (ascending edges), then area terms
//                    (ascending triangles); ff += f + fad
#include <unordered_map>
#include <vector>

#include "udx_internal.h"

namespace udx {
namespace {

constexpr int JUL_THREADS = 128;

__device__ __forceinline__ double vabs3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ __forceinline__ void vsub(const double a[3], const double b[3], double c[3]) {
    c[0] = a[0] - b[0];
    c[1] = a[1] - b[1];
    c[2] = a[2] - b[2];
}
__device__ __forceinline__ double vdot(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void vcross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - b[1] * a[2];
    c[1] = b[0] * a[2] - a[0] * b[2];
    c[2] = a[0] * b[1] - b[0] * a[1];
}

// Kahan's triangle area from the sorted side lengths (src/math/tri/dev.h:38-50)
__device__ double kahan_area(const double r0[3], const double r1[3], const double r2[3]) {
    double u[3], a, b, c, t;
    vsub(r0, r1, u);
    a = vabs3(u);
    vsub(r1, r2, u);
    b = vabs3(u);
    vsub(r2, r0, u);
    c = vabs3(u);
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    return sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c))) / 4;
}

__device__ double orient3d(const double a[3], const double b[3], const double c[3], const double d[3]) {
    const double ax = a[0] - d[0], ay = a[1] - d[1], az = a[2] - d[2];
    const double bx = b[0] - d[0], by = b[1] - d[1], bz = b[2] - d[2];
    const double cx = c[0] - d[0], cy = c[1] - d[1], cz = c[2] - d[2];
    return ax * (by * cz - bz * cy) + bx * (cy * az - cz * ay) + cx * (ay * bz - az * by);
}

// (a - c) x (b - c)
__device__ void ac_bc_cross(const double a[3], const double b[3], const double c[3], double r[3]) {
    r[0] = (a[1] - c[1]) * (b[2] - c[2]) - (a[2] - c[2]) * (b[1] - c[1]);
    r[1] = (a[2] - c[2]) * (b[0] - c[0]) - (a[0] - c[0]) * (b[2] - c[2]);
    r[2] = (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0]);
}

// the dihedral angle at edge b-c between triangles (a, b, c) and (c, b, d)
// (src/math/tri/dev.h:105-121, juelicher/dev/main.h:10-14)
__device__ double dih_theta(const double a[3], const double b[3], const double c[3], const double d[3]) {
    double k[3], l[3];
    ac_bc_cross(b, c, a, k);
    ac_bc_cross(c, b, d, l);
    const double bcx = b[0] - c[0], bcy = b[1] - c[1], bcz = b[2] - c[2];
    const double bc = sqrt(bcx * bcx + bcy * bcy + bcz * bcz);
    const double y = -orient3d(a, b, c, d) * bc;
    const double x = k[0] * l[0] + k[1] * l[1] + k[2] * l[2];
    return -atan2(y, x);
}

__device__ void vnorm(const double a[3], double b[3]) {
    const double s = vabs3(a);
    if (s > 1e-12 || s < -1e-12) {
        for (int x = 0; x < 3; x++) b[x] = (1 / s) * a[x];
    } else {
        for (int x = 0; x < 3; x++) b[x] = a[x];
    }
}

__device__ void tri_normal(const double a[3], const double b[3], const double c[3], double e[3]) {
    double u[3], v[3], n[3];
    vsub(b, a, u);
    vsub(c, a, v);
    vcross(u, v, n);
    vnorm(n, e);
}

__device__ double tri_cot(const double a[3], const double b[3], const double c[3]) {
    double u[3], v[3];
    const double y = 2 * kahan_area(a, b, c);
    vsub(a, b, u);
    vsub(c, b, v);
    return vdot(u, v) / y;
}

// the reference's dihedral-gradient formula at role q (0: a, 1: b, 2: c, 3: d)
__device__ void ddih(const double a[3], const double b[3], const double c[3], const double d[3], int q,
                     double g[3]) {
    double n[3], k[3], u[3];
    tri_normal(a, b, c, n);
    tri_normal(c, b, d, k);
    vsub(b, c, u);
    const double e = vabs3(u);
    if (q == 0) {
        const double s = e / (2 * kahan_area(a, b, c));
        for (int x = 0; x < 3; x++) g[x] = s * n[x];
    } else if (q == 3) {
        const double s = e / (2 * kahan_area(c, b, d));
        for (int x = 0; x < 3; x++) g[x] = s * k[x];
    } else if (q == 2) {
        const double bn = tri_cot(a, b, c), bk = tri_cot(c, b, d);
        for (int x = 0; x < 3; x++) g[x] = (-bn / e) * n[x] + (-bk / e) * k[x];
    } else {
        const double cn = tri_cot(b, c, a), ck = tri_cot(d, c, b);
        for (int x = 0; x < 3; x++) g[x] = (-cn / e) * n[x] + (-ck / e) * k[x];
    }
}

__device__ __forceinline__ void getv(const float *pp, int v, double r[3]) {
    r[0] = pp[6 * (size_t)v];
    r[1] = pp[6 * (size_t)v + 1];
    r[2] = pp[6 * (size_t)v + 2];
}

struct JulK {
    int nv, nt, ne, nc;
    const int *tt, *dih;                 // 3 per triangle, 4 per edge (i, j, k, l)
    const int *vt_start, *vt;            // vertex -> (triangle << 2 | corner), ascending
    const int *vjk_start, *vjk;          // vertex -> (edge << 2 | role 1 or 2), ascending
    const int *ve_start, *ve;            // vertex -> (edge << 2 | role 0..3), ascending
    float *theta, *ltheta, *tarea;       // per (cell, edge) and (cell, triangle)
    float *area, *lt;                    // per (cell, vertex)
    float *C;                            // per cell
};

__global__ __launch_bounds__(JUL_THREADS) void k_jul_edges(JulK K, const float *pp) {
    const int g = blockIdx.x * JUL_THREADS + threadIdx.x;
    if (g >= K.nc * K.ne) return;
    const int c = g / K.ne, e = g - c * K.ne;
    const float *p = pp + 6 * (size_t)c * K.nv;
    const int *d = K.dih + 4 * e;
    double a[3], b[3], cc[3], dd[3], u[3];
    getv(p, d[0], a);
    getv(p, d[1], b);
    getv(p, d[2], cc);
    getv(p, d[3], dd);
    const double t0 = dih_theta(a, b, cc, dd);
    vsub(cc, b, u);
    K.theta[g] = (float)t0;
    K.ltheta[g] = (float)(vabs3(u) * t0);
}

__global__ __launch_bounds__(JUL_THREADS) void k_jul_tris(JulK K, const float *pp) {
    const int g = blockIdx.x * JUL_THREADS + threadIdx.x;
    if (g >= K.nc * K.nt) return;
    const int c = g / K.nt, t = g - c * K.nt;
    const float *p = pp + 6 * (size_t)c * K.nv;
    double a[3], b[3], cc[3];
    getv(p, K.tt[3 * t], a);
    getv(p, K.tt[3 * t + 1], b);
    getv(p, K.tt[3 * t + 2], cc);
    K.tarea[g] = (float)(kahan_area(a, b, cc) / 3);
}

__global__ __launch_bounds__(JUL_THREADS) void k_jul_vertices(JulK K) {
    const int g = blockIdx.x * JUL_THREADS + threadIdx.x;
    if (g >= K.nc * K.nv) return;
    const int c = g / K.nv, v = g - c * K.nv;
    float A = 0, L = 0;
    for (int s = K.vt_start[v]; s < K.vt_start[v + 1]; s++) A += K.tarea[(size_t)c * K.nt + (K.vt[s] >> 2)];
    for (int s = K.vjk_start[v]; s < K.vjk_start[v + 1]; s++) L += K.ltheta[(size_t)c * K.ne + (K.vjk[s] >> 2)];
    K.area[g] = A;
    K.lt[g] = L;
}

__global__ __launch_bounds__(JUL_THREADS) void k_jul_cells(JulK K) {
    const int c = blockIdx.x * JUL_THREADS + threadIdx.x;
    if (c >= K.nc) return;
    const float H0 = -1.0f / 2.0f, kb = 1.0f, pi = (float)3.141592653589793, kad = 2 * kb / pi;
    float A = 0, Lt = 0;
    for (int v = 0; v < K.nv; v++) {
        A += K.area[(size_t)c * K.nv + v];
        Lt += K.lt[(size_t)c * K.nv + v];
    }
    K.C[c] = (Lt / 4 - H0 * A) * (4 * kad * pi / A);
}

__global__ __launch_bounds__(JUL_THREADS) void k_jul_force(JulK K, const float *pp, float *ff) {
    const int g = blockIdx.x * JUL_THREADS + threadIdx.x;
    if (g >= K.nc * K.nv) return;
    const int c = g / K.nv, v = g - c * K.nv;
    const float H0 = -1.0f / 2.0f;
    const float *p = pp + 6 * (size_t)c * K.nv;
    const float *area = K.area + (size_t)c * K.nv, *lt = K.lt + (size_t)c * K.nv;
    const float C = K.C[c];
    float f[3] = {0, 0, 0}, fad[3] = {0, 0, 0};
    // edge-length terms (v is j or k)
    for (int s = K.vjk_start[v]; s < K.vjk_start[v + 1]; s++) {
        const int e = K.vjk[s] >> 2, role = K.vjk[s] & 3;
        const int *d = K.dih + 4 * e;
        const int j = d[1], k = d[2];
        double b[3], cc[3], u[3], w[3], gr[3];
        getv(p, j, b);
        getv(p, k, cc);
        vsub(cc, b, u);
        vnorm(u, w);
        for (int x = 0; x < 3; x++) gr[x] = role == 2 ? w[x] : -w[x];
        const float th = K.theta[(size_t)c * K.ne + e];
        double coef = -((lt[j] / area[j] / 4 - H0) + (lt[k] / area[k] / 4 - H0)) * th;
        for (int x = 0; x < 3; x++) f[x] += (float)(coef * gr[x]);
        coef = -C / 4 * th;
        for (int x = 0; x < 3; x++) fad[x] += (float)(coef * gr[x]);
    }
    // dihedral terms (v is i, j, k or l)
    for (int s = K.ve_start[v]; s < K.ve_start[v + 1]; s++) {
        const int e = K.ve[s] >> 2, role = K.ve[s] & 3;
        const int *d = K.dih + 4 * e;
        double r[4][3], gr[3], u[3];
        for (int q = 0; q < 4; q++) getv(p, d[q], r[q]);
        ddih(r[0], r[1], r[2], r[3], role, gr);
        vsub(r[2], r[1], u);
        const double len0 = vabs3(u);
        const int j = d[1], k = d[2];
        double coef = -((lt[j] / area[j] / 4 - H0) + (lt[k] / area[k] / 4 - H0)) * len0;
        for (int x = 0; x < 3; x++) f[x] += (float)(coef * gr[x]);
        coef = -C / 2.0 * len0;
        for (int x = 0; x < 3; x++) fad[x] += (float)(coef * gr[x]);
    }
    // area terms, the coefficient of v on the gradient of each triangle at v
    const double coefa = (1.0 / 3.0) * (lt[v] * lt[v] / 8.0 / area[v] / area[v] - 2.0 * H0 * H0);
    for (int s = K.vt_start[v]; s < K.vt_start[v + 1]; s++) {
        const int t = K.vt[s] >> 2, q = K.vt[s] & 3;
        double r[3][3], n[3], n2[3], e[3], gr[3];
        for (int m = 0; m < 3; m++) getv(p, K.tt[3 * t + m], r[m]);
        tri_normal(r[0], r[1], r[2], n);
        for (int x = 0; x < 3; x++) n2[x] = 0.5 * n[x];
        // opposite edge: b->c for a, c->a for b, a->b for c
        vsub(r[(q + 2) % 3], r[(q + 1) % 3], e);
        vcross(n2, e, gr);
        for (int x = 0; x < 3; x++) f[x] += (float)(coefa * gr[x]);
    }
    float *o = ff + 3 * (size_t)g;
    for (int x = 0; x < 3; x++) o[x] = o[x] + (f[x] + fad[x]);
}

}  // namespace

struct JulTemplate {
    int nv = 0, nt = 0, ne = 0, max_cells = 0;
    int *tt = nullptr, *dih = nullptr;
    int *vt_start = nullptr, *vt = nullptr, *vjk_start = nullptr, *vjk = nullptr, *ve_start = nullptr,
        *ve = nullptr;
    float *theta = nullptr, *ltheta = nullptr, *tarea = nullptr, *area = nullptr, *lt = nullptr, *C = nullptr;
};

void jul_free(JulTemplate *j) {
    if (!j) return;
    dev_free(j->tt);
    dev_free(j->dih);
    dev_free(j->vt_start);
    dev_free(j->vt);
    dev_free(j->vjk_start);
    dev_free(j->vjk);
    dev_free(j->ve_start);
    dev_free(j->ve);
    dev_free(j->theta);
    dev_free(j->ltheta);
    dev_free(j->tarea);
    dev_free(j->area);
    dev_free(j->lt);
    dev_free(j->C);
    delete j;
}

// CSR of (index << 2 | tag) lists per vertex, in the order of the items
static void csr(int nv, const std::vector<std::pair<int, int>> &items, std::vector<int> &start,
                std::vector<int> &list) {
    start.assign(nv + 1, 0);
    for (auto &it : items) start[it.first + 1]++;
    for (int v = 0; v < nv; v++) start[v + 1] += start[v];
    std::vector<int> fill(start.begin(), start.end() - 1);
    list.resize(items.size());
    for (auto &it : items) list[fill[it.first]++] = it.second;
}

int jul_build(hipStream_t s, int nv, int nt, const int *tt_host, int max_cells, JulTemplate **out) {
    // edges by first appearance, half-edges x->y, y->z, z->x of triangles in order
    std::unordered_map<long long, int> half;   // (a, b) -> triangle * 3 + corner of a
    half.reserve(3 * (size_t)nt);
    for (int t = 0; t < nt; t++)
        for (int q = 0; q < 3; q++) half[(long long)tt_host[3 * t + q] * nv + tt_host[3 * t + (q + 1) % 3]] = 3 * t + q;
    std::vector<char> seen(3 * (size_t)nt, 0);
    std::vector<int> dih;
    for (int t = 0; t < nt; t++)
        for (int q = 0; q < 3; q++) {
            if (seen[3 * t + q]) continue;
            const int j = tt_host[3 * t + q], k = tt_host[3 * t + (q + 1) % 3], i = tt_host[3 * t + (q + 2) % 3];
            auto it = half.find((long long)k * nv + j);
            UDX_CHECK(it != half.end(), UDX_EINVAL, "juelicher: the mesh is not closed (edge %d-%d)", j, k);
            const int u = it->second / 3, w = it->second % 3;
            seen[it->second] = 1;
            seen[3 * t + q] = 1;
            const int l = tt_host[3 * u + (w + 2) % 3];
            dih.insert(dih.end(), {i, j, k, l});
        }
    const int ne = (int)dih.size() / 4;
    std::vector<std::pair<int, int>> it_t, it_jk, it_e;
    for (int t = 0; t < nt; t++)
        for (int q = 0; q < 3; q++) it_t.push_back({tt_host[3 * t + q], t << 2 | q});
    for (int e = 0; e < ne; e++) {
        for (int q = 0; q < 4; q++) it_e.push_back({dih[4 * e + q], e << 2 | q});
        it_jk.push_back({dih[4 * e + 1], e << 2 | 1});
        it_jk.push_back({dih[4 * e + 2], e << 2 | 2});
    }
    std::vector<int> vt_start, vt, vjk_start, vjk, ve_start, ve;
    csr(nv, it_t, vt_start, vt);
    csr(nv, it_jk, vjk_start, vjk);
    csr(nv, it_e, ve_start, ve);
    JulTemplate *j = new JulTemplate();
    j->nv = nv;
    j->nt = nt;
    j->ne = ne;
    j->max_cells = max_cells;
    int st = dev_alloc(&j->tt, 3 * (size_t)nt);
    if (!st) st = dev_alloc(&j->dih, 4 * (size_t)ne);
    if (!st) st = dev_alloc(&j->vt_start, (size_t)nv + 1);
    if (!st) st = dev_alloc(&j->vt, vt.size());
    if (!st) st = dev_alloc(&j->vjk_start, (size_t)nv + 1);
    if (!st) st = dev_alloc(&j->vjk, vjk.size());
    if (!st) st = dev_alloc(&j->ve_start, (size_t)nv + 1);
    if (!st) st = dev_alloc(&j->ve, ve.size());
    if (!st) st = dev_alloc(&j->theta, (size_t)max_cells * ne);
    if (!st) st = dev_alloc(&j->ltheta, (size_t)max_cells * ne);
    if (!st) st = dev_alloc(&j->tarea, (size_t)max_cells * nt);
    if (!st) st = dev_alloc(&j->area, (size_t)max_cells * nv);
    if (!st) st = dev_alloc(&j->lt, (size_t)max_cells * nv);
    if (!st) st = dev_alloc(&j->C, (size_t)max_cells);
    auto up = [&](int *d, const std::vector<int> &h) {
        return hipMemcpyAsync(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s) == hipSuccess;
    };
    if (!st && !(hipMemcpyAsync(j->tt, tt_host, 3 * sizeof(int) * (size_t)nt, hipMemcpyHostToDevice, s) ==
                     hipSuccess &&
                 up(j->dih, dih) && up(j->vt_start, vt_start) && up(j->vt, vt) && up(j->vjk_start, vjk_start) &&
                 up(j->vjk, vjk) && up(j->ve_start, ve_start) && up(j->ve, ve) &&
                 hipStreamSynchronize(s) == hipSuccess))
        st = UDX_EHIP;
    if (st) {
        jul_free(j);
        set_error("juelicher: allocation failed");
        return st;
    }
    *out = j;
    return UDX_OK;
}

int jul_apply(hipStream_t s, const JulTemplate *j, int nc, const float *pp, float *ff) {
    if (nc <= 0) return UDX_OK;
    JulK K{j->nv, j->nt, j->ne, nc, j->tt, j->dih, j->vt_start, j->vt, j->vjk_start, j->vjk, j->ve_start, j->ve,
           j->theta, j->ltheta, j->tarea, j->area, j->lt, j->C};
    hipLaunchKernelGGL(k_jul_edges, dim3(ceil_div(nc * j->ne, JUL_THREADS)), dim3(JUL_THREADS), 0, s, K, pp);
    hipLaunchKernelGGL(k_jul_tris, dim3(ceil_div(nc * j->nt, JUL_THREADS)), dim3(JUL_THREADS), 0, s, K, pp);
    hipLaunchKernelGGL(k_jul_vertices, dim3(ceil_div(nc * j->nv, JUL_THREADS)), dim3(JUL_THREADS), 0, s, K);
    hipLaunchKernelGGL(k_jul_cells, dim3(ceil_div(nc, JUL_THREADS)), dim3(JUL_THREADS), 0, s, K);
    hipLaunchKernelGGL(k_jul_force, dim3(ceil_div(nc * j->nv, JUL_THREADS)), dim3(JUL_THREADS), 0, s, K, pp, ff);
    UDX_LAUNCHED();
    return UDX_OK;
}

}  // namespace udx
