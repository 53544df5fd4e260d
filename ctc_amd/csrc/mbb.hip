// Bounce-back of the solvent on moving membrane meshes (src/mesh_bounce) on
// one periodic domain.
//
// Per call: the periodic images of the meshes that leave the domain cropped
// by 1 (the mesh halo of src/exch/mesh), the earliest crossing of every
// solvent particle with a triangle instance (one thread per triangle
// instance over the solvent cells its start-of-step bounding box + 0.1
// covers; 64-bit atomicMin of (time bits, instance triangle) per particle),
// the bounce of each crossing particle (one thread per particle; in the
// solvent step this pass also applies the periodic wrap the move left out),
// the momentum of each triangle summed in ascending particle order (counting
// sort of the crossings by triangle), and the velocity change of the
// vertices (one thread per vertex over its triangles in ascending order).
//
// Arithmetic follows the reference's mixed precision: float particle and
// vertex data, double plane/root computations, float results; -ffp-contract
// is off, so every product and sum rounds as written.
#include <cstdint>
#include <vector>

#include "udx_internal.h"

namespace udx {
namespace {

constexpr int MBB_THREADS = 256;
constexpr int MBB_MAX_IMAGES = 8;   // a mesh and at most 7 images (3 faces, 3 edges, 1 corner)
constexpr unsigned long long MBB_EMPTY = ~0ull;
constexpr float MBB_MAX_SPAN = 8.0f;     // the widest triangle searched (a start-of-step bounding box side)
constexpr int MBB_WQ = 256;              // per-wave LDS list of candidate pairs (4 per lane)
constexpr int MBB_CAND_BLOCKS = 1024;    // grid of the dense crossing-test pass
enum { BB_SUCCESS = 0, BB_NOCROSS = 1, BB_WTRIANGLE = 2, BB_HFAIL = 3 };

struct MPart {
    float r[3], v[3];
};

// ---- crossing time ---------------------------------------------------------
// The crossing test of src/mesh_bounce/dev/intersection.h:30-127, derived
// here.  Over the step every vertex moves linearly, X(t) = X0 + t Xv, and so
// does the particle, p(t) = p0 + t pv.  With the edges e1(t) = B(t) - A(t)
// and e2(t) = C(t) - A(t) the plane normal N(t) = e1 x e2 is the quadratic
//   N(t) = N0 + t N1 + t^2 N2,
//   N0 = e1 x e2,  N1 = e1 x e2v + e1v x e2,  N2 = e1v x e2v,
// and the particle's offset from A is g(t) = g0 + t gv.  The signed volume
// phi(t) = g . N is the cubic k0 + k1 t + k2 t^2 + k3 t^3 with
//   k0 = g0.N0,  k1 = gv.N0 + g0.N1,  k2 = gv.N1 + g0.N2,  k3 = gv.N2,
// and the particle passes through the plane at the first root of phi in
// [0, dt] where phi changes sign.  [0, dt] is cut at the turning points of
// phi (roots of phi' by the cancellation-free quadratic formula); phi is
// monotone on each piece, the first piece whose ends differ in sign holds
// the root, and a Newton iteration kept inside that bracket (bisection when
// a step leaves it) converges to it.  The coefficients are used as they are
// (the reference drops a leading coefficient below 1e-8 and solves the
// lower degree); a root where phi touches zero without a sign change is not
// a crossing.
__device__ __forceinline__ double cub_eval(const double k[4], double t) {
    return ((k[3] * t + k[2]) * t + k[1]) * t + k[0];
}
__device__ __forceinline__ double cub_slope(const double k[4], double t) {
    return (3 * k[3] * t + 2 * k[2]) * t + k[1];
}

// turning points of the cubic strictly inside (0, T), ascending
__device__ int cub_turning(const double k[4], double T, double out[2]) {
    const double qa = 3 * k[3], qb = 2 * k[2], qc = k[1];
    double r[2];
    int m = 0, n = 0;
    if (qa == 0) {
        if (qb != 0) r[m++] = -qc / qb;
    } else {
        const double disc = qb * qb - 4 * qa * qc;
        if (disc > 0) {
            const double w = -0.5 * (qb + (qb >= 0 ? sqrt(disc) : -sqrt(disc)));
            r[m++] = w / qa;
            if (w != 0) r[m++] = qc / w;
        }
    }
    if (m == 2 && r[1] < r[0]) {
        const double t = r[0];
        r[0] = r[1];
        r[1] = t;
    }
    for (int i = 0; i < m; i++)
        if (r[i] > 0 && r[i] < T) out[n++] = r[i];
    return n;
}

// the root in (lo, hi) of a cubic monotone there; flo = value at lo (the
// value at hi has the other sign)
__device__ double cub_bracketed(const double k[4], double lo, double hi, double flo) {
    double t = 0.5 * (lo + hi);
    for (int it = 0; it < 200; it++) {
        const double f = cub_eval(k, t);
        if (f == 0) return t;
        if ((f < 0) == (flo < 0)) {
            lo = t;
            flo = f;
        } else {
            hi = t;
        }
        const double d = cub_slope(k, t);
        double tn = d != 0 ? t - f / d : lo;
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == t || !(hi - lo > 4e-16 * hi)) return tn;
        t = tn;
    }
    return t;
}

// first sign-changing root of the cubic in [0, T]
__device__ bool cub_first_root(const double k[4], double T, double *h) {
    double s[4];
    int m = 0;
    s[m++] = 0;
    m += cub_turning(k, T, s + 1);
    s[m++] = T;
    double fprev = cub_eval(k, s[0]);
    if (fprev == 0) {
        *h = s[0];
        return true;
    }
    for (int i = 1; i < m; i++) {
        const double f = cub_eval(k, s[i]);
        if (f == 0) {
            *h = s[i];
            return true;
        }
        if ((f < 0) != (fprev < 0)) {
            *h = cub_bracketed(k, s[i - 1], s[i], fprev);
            return true;
        }
        fprev = f;
    }
    return false;
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Triangle A B C and a particle P over [0, dt]: BB_* and (h, u, v, s) =
// crossing time, barycentric coordinates of the crossing point, the side of
// the plane the particle starts on (+1: phi(0) > 0)
__device__ int intersect(float dt_f, const MPart &A, const MPart &B, const MPart &C, const MPart &P, float out[4]) {
    const double T = dt_f;
    double e1[3], e2[3], e1v[3], e2v[3], g0[3], gv[3];
    for (int c = 0; c < 3; c++) {
        e1[c] = (double)(B.r[c] - A.r[c]);
        e2[c] = (double)(C.r[c] - A.r[c]);
        e1v[c] = (double)(B.v[c] - A.v[c]);
        e2v[c] = (double)(C.v[c] - A.v[c]);
        g0[c] = (double)(P.r[c] - A.r[c]);
        gv[c] = (double)(P.v[c] - A.v[c]);
    }
    double N0[3], N1[3], N2[3], xa[3], xb[3];
    cross3(e1, e2, N0);
    cross3(e1, e2v, xa);
    cross3(e1v, e2, xb);
    for (int c = 0; c < 3; c++) N1[c] = xa[c] + xb[c];
    cross3(e1v, e2v, N2);
    // sides of the plane at the start and at the end of the step
    double gT[3], NT[3];
    for (int c = 0; c < 3; c++) {
        gT[c] = g0[c] + T * gv[c];
        NT[c] = N0[c] + T * (N1[c] + T * N2[c]);
    }
    const double phi0 = dot3(g0, N0), phiT = dot3(gT, NT);
    out[3] = phi0 > 0 ? 1.f : -1.f;
    if (phi0 * phiT > 0) return BB_NOCROSS;
    const double k[4] = {phi0, dot3(gv, N0) + dot3(g0, N1), dot3(gv, N1) + dot3(g0, N2), dot3(gv, N2)};
    double h;
    if (!cub_first_root(k, T, &h)) return BB_HFAIL;
    // the crossing point on the triangle at time h: g = u e1 + v e2 in the
    // least-squares sense (normal equations of the two edge directions)
    double g[3], f1[3], f2[3];
    for (int c = 0; c < 3; c++) {
        g[c] = g0[c] + h * gv[c];
        f1[c] = e1[c] + h * e1v[c];
        f2[c] = e2[c] + h * e2v[c];
    }
    const double m11 = dot3(f1, f1), m12 = dot3(f1, f2), m22 = dot3(f2, f2);
    const double r1 = dot3(g, f1), r2 = dot3(g, f2);
    const double inv = 1.0 / (m11 * m22 - m12 * m12);
    const double u = (r1 * m22 - r2 * m12) * inv;
    const double v = (r2 * m11 - r1 * m12) * inv;
    if (u < 0 || v < 0 || u + v > 1) return BB_WTRIANGLE;
    out[0] = (float)h;
    out[1] = (float)u;
    out[2] = (float)v;
    return BB_SUCCESS;
}

// ---- particle and mesh access --------------------------------------------
// solvent as AoS particles (the reference's Particle) or as the solvent
// step's float4 position/velocity streams
struct PartsAoS {
    float *pp;
    const float *pp0;
    __device__ void load(int i, float r[3], float v[3]) const {
        for (int c = 0; c < 3; c++) {
            r[c] = pp[6 * (size_t)i + c];
            v[c] = pp[6 * (size_t)i + 3 + c];
        }
    }
    __device__ void load_r0(int i, float r[3]) const {
        for (int c = 0; c < 3; c++) r[c] = pp0[6 * (size_t)i + c];
    }
    __device__ float r1(int i, int c) const { return pp[6 * (size_t)i + c]; }
    __device__ void store(int i, const float r[3], const float v[3]) const {
        for (int c = 0; c < 3; c++) {
            pp[6 * (size_t)i + c] = r[c];
            pp[6 * (size_t)i + 3 + c] = v[c];
        }
    }
};

struct PartsF4 {
    float4 *pos, *vel;
    const float4 *pos0;
    __device__ void load(int i, float r[3], float v[3]) const {
        const float4 p = pos[i], u = vel[i];
        r[0] = p.x; r[1] = p.y; r[2] = p.z;
        v[0] = u.x; v[1] = u.y; v[2] = u.z;
    }
    __device__ void load_r0(int i, float r[3]) const {
        const float4 p = pos0[i];
        r[0] = p.x; r[1] = p.y; r[2] = p.z;
    }
    __device__ float r1(int i, int c) const {
        const float4 p = pos[i];
        return c == 0 ? p.x : (c == 1 ? p.y : p.z);
    }
    __device__ void store(int i, const float r[3], const float v[3]) const {
        const float4 p = pos[i], u = vel[i];
        pos[i] = make_float4(r[0], r[1], r[2], p.w);
        vel[i] = make_float4(v[0], v[1], v[2], u.w);
    }
};

// the particle's start and its straight-line velocity (dev/utils.h:43-48)
template <class P>
__device__ void particle_path(const P &parts, float dt, int i, MPart *p0) {
    float r1[3], v1[3];
    parts.load(i, r1, v1);
    parts.load_r0(i, p0->r);
    for (int c = 0; c < 3; c++) p0->v[c] = (float)((double)(r1[c] - p0->r[c]) / (double)dt);
}

struct MeshView {
    int nt, nv;
    const int *tt;             // 3 per triangle
    const float *rp, *rc;      // vertices (6 floats) at the start of the step / now
    const int *inst_mesh;
    const float *inst_shift;   // 3 per instance
};

// vertex j of instance k: r = rc + shift, v = (r - (rp + shift)) / dt
__device__ void fetch_vertex(const MeshView &m, float dt, const float *sh, int j, MPart *o) {
    for (int c = 0; c < 3; c++) {
        const float pc = m.rc[6 * (size_t)j + c] + sh[c];
        const float pp = m.rp[6 * (size_t)j + c] + sh[c];
        o->r[c] = pc;
        o->v[c] = (pc - pp) / dt;
    }
}

__device__ void fetch_tri(const MeshView &m, float dt, int gid, MPart *A, MPart *B, MPart *C) {
    const int k = gid / m.nt, t = gid - k * m.nt;
    const int base = m.inst_mesh[k] * m.nv;
    const float *sh = m.inst_shift + 3 * k;
    fetch_vertex(m, dt, sh, base + m.tt[3 * t], A);
    fetch_vertex(m, dt, sh, base + m.tt[3 * t + 1], B);
    fetch_vertex(m, dt, sh, base + m.tt[3 * t + 2], C);
}

__device__ __forceinline__ void revert(float h, MPart *p) {
    for (int c = 0; c < 3; c++) p->r[c] = p->r[c] - p->v[c] * h;
}

__device__ __forceinline__ int cell_of(int L, float x) {
    const int c = (int)floor((double)x + L / 2);
    return c < 0 ? 0 : (c > L - 1 ? L - 1 : c);
}

__device__ __forceinline__ unsigned long long make_key(float h, int gid) {
    unsigned b = __float_as_uint(h);
    if (h == 0) b = 0;
    return ((unsigned long long)b << 32) | (unsigned)gid;
}

// ---- instances: meshes and their periodic images --------------------------
__device__ __forceinline__ int frag_d2i(int x, int y, int z) {
    return ((x + 2) % 3) + 3 * ((y + 2) % 3) + 9 * ((z + 2) % 3);
}

// per mesh: the box code of its current bounding box against the domain
// cropped by 1 (src/exch/map/dev.h:21-30)
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_codes(int nv, int Lx, int Ly, int Lz, const float *rc,
                                                           int *codes) {
    const int m = blockIdx.x;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int v = threadIdx.x; v < nv; v += MBB_THREADS)
        for (int c = 0; c < 3; c++) {
            const float x = rc[6 * ((size_t)m * nv + v) + c];
            lo[c] = fminf(lo[c], x);
            hi[c] = fmaxf(hi[c], x);
        }
    __shared__ float slo[3][MBB_THREADS], shi[3][MBB_THREADS];
    for (int c = 0; c < 3; c++) {
        slo[c][threadIdx.x] = lo[c];
        shi[c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int s = MBB_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; c++) {
                slo[c][threadIdx.x] = fminf(slo[c][threadIdx.x], slo[c][threadIdx.x + s]);
                shi[c][threadIdx.x] = fmaxf(shi[c][threadIdx.x], shi[c][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int L[3] = {Lx, Ly, Lz};
        for (int c = 0; c < 3; c++) {
            const int h = (L[c] - 2) / 2;
            codes[3 * m + c] = -1 + (slo[c][0] >= -h) + (shi[c][0] >= h);
        }
    }
}

// does the box code send to fragment fid (a face, edge or corner it reaches)
__device__ bool code_sends(const int *code, const int d[3]) {
    bool any = false;
    for (int c = 0; c < 3; c++) {
        if (d[c] != 0 && d[c] != code[c]) return false;
        any |= d[c] != 0;
    }
    return any;
}

// one block: meshes unshifted, then images by fragment id, meshes ascending.
// Pass 1 counts the meshes of every fragment, pass 2 places them: fragment
// f's images start after the meshes and the images of fragments below f.
__device__ __forceinline__ unsigned code_mask(const int *code) {
    unsigned mask = 0;
    for (int fid = 0; fid < 26; fid++) {
        const int t3[3] = {fid % 3, (fid / 3) % 3, fid / 9};
        int d[3];
        for (int c = 0; c < 3; c++) d[c] = t3[c] == 0 ? 1 : (t3[c] == 1 ? -1 : 0);
        if (code_sends(code, d)) mask |= 1u << fid;
    }
    return mask;
}

__global__ __launch_bounds__(1024) void k_mbb_instances(int nm, int Lx, int Ly, int Lz, const int *codes,
                                                        int *inst_mesh, float *inst_shift, int *n_inst) {
    __shared__ int fcount[26], fbase[26];
    __shared__ int wcount[16][26];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid < 26) fcount[tid] = 0;
    for (int m = tid; m < nm; m += 1024) {
        inst_mesh[m] = m;
        inst_shift[3 * m] = inst_shift[3 * m + 1] = inst_shift[3 * m + 2] = 0.f;
    }
    __syncthreads();
    for (int m = tid; m < nm; m += 1024) {
        const unsigned mask = code_mask(codes + 3 * m);
        for (int fid = 0; fid < 26; fid++)
            if (mask >> fid & 1u) atomicAdd(&fcount[fid], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int b = nm;
        for (int fid = 0; fid < 26; fid++) {
            fbase[fid] = b;
            b += fcount[fid];
        }
        *n_inst = b;
    }
    __syncthreads();
    const int L[3] = {Lx, Ly, Lz};
    for (int m0 = 0; m0 < nm; m0 += 1024) {
        const int m = m0 + tid;
        const unsigned mask = m < nm ? code_mask(codes + 3 * m) : 0u;
        unsigned long long bal[26];
        for (int fid = 0; fid < 26; fid++) {
            bal[fid] = __ballot((mask >> fid) & 1u);
            if (lane == 0) wcount[wid][fid] = __popcll(bal[fid]);
        }
        __syncthreads();
        for (int fid = 0; fid < 26; fid++) {
            if (!((mask >> fid) & 1u)) continue;
            int off = fbase[fid] + __popcll(bal[fid] & ((1ull << lane) - 1ull));
            for (int w = 0; w < wid; w++) off += wcount[w][fid];
            const int t3[3] = {fid % 3, (fid / 3) % 3, fid / 9};
            inst_mesh[off] = m;
            for (int c = 0; c < 3; c++) {
                const int d = t3[c] == 0 ? 1 : (t3[c] == 1 ? -1 : 0);
                inst_shift[3 * off + c] = (float)(-d * L[c]);
            }
        }
        __syncthreads();
        if (tid < 26) {
            int tot = 0;
            for (int w = 0; w < 16; w++) tot += wcount[w][tid];
            fbase[tid] += tot;
        }
        __syncthreads();
    }
}

// ---- search --------------------------------------------------------------
// the straight-line velocity of every particle over the step, once per call
// (the search tests each particle against several triangles)
template <class P>
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_paths(int nmax, const int *n_dev, float dt, P parts,
                                                           float4 *v0) {
    const int i = blockIdx.x * MBB_THREADS + threadIdx.x;
    if (i >= nmax || i >= *n_dev) return;
    MPart p0;
    particle_path(parts, dt, i, &p0);
    v0[i] = make_float4(p0.v[0], p0.v[1], p0.v[2], 0.f);
}

// The plane test (the particle on both sides of the triangle's plane over
// the step) rejects almost every candidate; the few that pass go through the
// full crossing test in a second, dense pass, so that no wave serialises on
// one lane's cubic.  Each wave gathers its (triangle, particle) pairs in LDS
// and appends them to a global list with one atomic per MBB_WQ pairs.
template <class P>
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_find(MeshView m, int max_gid, const int *n_inst, float dt,
                                                          int Lx, int Ly, int Lz, const int *starts, P parts,
                                                          const float4 *v0, unsigned long long *keys, int cap,
                                                          unsigned *cand_n, uint2 *cand) {
    __shared__ uint2 wq_all[MBB_THREADS / 64][MBB_WQ];
    const int lane = threadIdx.x & 63;
    uint2 *wq = wq_all[threadIdx.x >> 6];
    const unsigned long long below = (1ull << lane) - 1ull;
    const int gid = blockIdx.x * MBB_THREADS + threadIdx.x;
    bool active = gid < max_gid && gid < *n_inst * m.nt;
    if (!wave_any(active)) return;   // whole waves only: the queue below is wave-wide
    MPart A, B, C;
    const double ddt = dt;
    double N0[3] = {0, 0, 0}, NT[3] = {0, 0, 0};
    int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    if (active) {
        fetch_tri(m, dt, gid, &A, &B, &C);
        revert(dt, &A);
        revert(dt, &B);
        revert(dt, &C);
        const float tol = 0.1f;
        const int L[3] = {Lx, Ly, Lz};
        for (int c = 0; c < 3; c++) {
            const float mn = fminf(fminf(A.r[c], B.r[c]), C.r[c]);
            const float mx = fmaxf(fmaxf(A.r[c], B.r[c]), C.r[c]);
            // a triangle wider than MBB_MAX_SPAN (or not finite) is no
            // membrane triangle: skipped rather than swept over the domain
            if (!(mx - mn <= MBB_MAX_SPAN)) active = false;
            lo[c] = cell_of(L[c], mn - tol);
            hi[c] = cell_of(L[c], mx + tol);
        }
        // the plane normal at the start and at the end of the step, with the
        // operations of intersect(), so the same decisions
        double e1[3], e2[3], e1v[3], e2v[3], N1[3], N2[3], xa[3], xb[3];
        for (int c = 0; c < 3; c++) {
            e1[c] = (double)(B.r[c] - A.r[c]);
            e2[c] = (double)(C.r[c] - A.r[c]);
            e1v[c] = (double)(B.v[c] - A.v[c]);
            e2v[c] = (double)(C.v[c] - A.v[c]);
        }
        cross3(e1, e2, N0);
        cross3(e1, e2v, xa);
        cross3(e1v, e2, xb);
        for (int c = 0; c < 3; c++) N1[c] = xa[c] + xb[c];
        cross3(e1v, e2v, N2);
        for (int c = 0; c < 3; c++) NT[c] = N0[c] + ddt * (N1[c] + ddt * N2[c]);
    }
    const int ny = active ? hi[1] - lo[1] + 1 : 0, nrows = active ? ny * (hi[2] - lo[2] + 1) : 0;
    int fill = 0;   // the same in every lane of the wave
    auto flush = [&]() {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(cand_n, (unsigned)fill);
        base = __shfl(base, 0);
        __builtin_amdgcn_wave_barrier();
        for (int k = lane; k < fill; k += 64)
            if (base + k < (unsigned)cap) cand[base + k] = wq[k];   // beyond: k_mbb_find_all redoes the search
        __builtin_amdgcn_wave_barrier();
        fill = 0;
    };
    // rows of cells: the particles of cells lo.x..hi.x are contiguous
    for (int r = 0; wave_any(r < nrows); r++) {
        int b = 0, len = 0;
        if (r < nrows) {
            const int iy = lo[1] + r % ny, iz = lo[2] + r / ny;
            const int row = Lx * (iy + Ly * iz);
            b = starts[row + lo[0]];
            len = starts[row + hi[0] + 1] - b;
        }
        // four candidates per iteration: their loads in flight together
        for (int c0 = 0; wave_any(c0 < len); c0 += 4) {
            unsigned pass = 0;
            MPart p[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = b + (c0 + u < len ? c0 + u : (len > 0 ? len - 1 : 0));
                if (len > 0) {
                    parts.load_r0(i, p[u].r);
                    const float4 w = v0[i];
                    p[u].v[0] = w.x;
                    p[u].v[1] = w.y;
                    p[u].v[2] = w.z;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (c0 + u >= len) continue;
                double g0[3], gT[3];
                for (int c = 0; c < 3; c++) {
                    g0[c] = (double)(p[u].r[c] - A.r[c]);
                    gT[c] = g0[c] + ddt * (double)(p[u].v[c] - A.v[c]);
                }
                if (!(dot3(g0, N0) * dot3(gT, NT) > 0)) pass |= 1u << u;   // not BB_NOCROSS
            }
            unsigned long long bu[4];
            int tot = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                bu[u] = __ballot((pass >> u) & 1u);
                tot += __popcll(bu[u]);
            }
            if (tot == 0) continue;
            if (fill + tot > MBB_WQ) flush();
            int off = fill;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if ((pass >> u) & 1u) wq[off + __popcll(bu[u] & below)] = make_uint2((unsigned)gid, (unsigned)(b + c0 + u));
                off += __popcll(bu[u]);
            }
            fill += tot;
        }
    }
    if (fill > 0) flush();
}

// the whole search with the full test in place, run only when the pair list
// overflowed (the list pass keeps its results: atomicMin is idempotent)
template <class P>
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_find_all(MeshView m, int max_gid, const int *n_inst, float dt,
                                                              int Lx, int Ly, int Lz, const int *starts, P parts,
                                                              const float4 *v0, unsigned long long *keys, int cap,
                                                              const unsigned *cand_n) {
    if (*cand_n <= (unsigned)cap) return;
    const int ngid = *n_inst * m.nt < max_gid ? *n_inst * m.nt : max_gid;
    for (int gid = blockIdx.x * MBB_THREADS + threadIdx.x; gid < ngid; gid += gridDim.x * MBB_THREADS) {
        MPart A, B, C;
        fetch_tri(m, dt, gid, &A, &B, &C);
        revert(dt, &A);
        revert(dt, &B);
        revert(dt, &C);
        const float tol = 0.1f;
        const int L[3] = {Lx, Ly, Lz};
        int lo[3], hi[3];
        bool ok = true;
        for (int c = 0; c < 3; c++) {
            const float mn = fminf(fminf(A.r[c], B.r[c]), C.r[c]);
            const float mx = fmaxf(fmaxf(A.r[c], B.r[c]), C.r[c]);
            if (!(mx - mn <= MBB_MAX_SPAN)) ok = false;
            lo[c] = cell_of(L[c], mn - tol);
            hi[c] = cell_of(L[c], mx + tol);
        }
        if (!ok) continue;
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++) {
                const int row = Lx * (iy + Ly * iz);
                const int e = starts[row + hi[0] + 1];
                for (int i = starts[row + lo[0]]; i < e; i++) {
                    MPart p0;
                    parts.load_r0(i, p0.r);
                    const float4 w = v0[i];
                    p0.v[0] = w.x;
                    p0.v[1] = w.y;
                    p0.v[2] = w.z;
                    float o[4];
                    if (intersect(dt, A, B, C, p0, o) == BB_SUCCESS) atomicMin(keys + i, make_key(o[0], gid));
                }
            }
    }
}

// the full crossing test of the listed (triangle, particle) pairs
template <class P>
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_cand(MeshView m, float dt, P parts, const float4 *v0, int cap,
                                                          const unsigned *cand_n, const uint2 *cand,
                                                          unsigned long long *keys) {
    const unsigned n = *cand_n < (unsigned)cap ? *cand_n : (unsigned)cap;
    for (unsigned k = blockIdx.x * MBB_THREADS + threadIdx.x; k < n; k += gridDim.x * MBB_THREADS) {
        const uint2 e = cand[k];
        MPart a, b, c, p0;
        fetch_tri(m, dt, (int)e.x, &a, &b, &c);
        revert(dt, &a);
        revert(dt, &b);
        revert(dt, &c);
        parts.load_r0((int)e.y, p0.r);
        const float4 w = v0[e.y];
        p0.v[0] = w.x;
        p0.v[1] = w.y;
        p0.v[2] = w.z;
        float o[4];
        if (intersect(dt, a, b, c, p0, o) == BB_SUCCESS) atomicMin(keys + e.y, make_key(o[0], (int)e.x));
    }
}

// ---- bounce ----------------------------------------------------------------
__device__ __forceinline__ float wrap1(float x, float L) {
    const float h = 0.5f * L;
    if (x < -h) x = x + L;
    if (x >= h) x = x - L;
    return x;
}

// bounce of particle i off instance triangle gid (src/mesh_bounce/dev/main.h:
// 150-210); returns its momentum change (P, L) and writes the particle
template <class P>
__device__ void perform_one(const MeshView &m, const P &parts, float dt, float mass, int i, int gid, float pos[3],
                            float vel[3], float dm[6]) {
    MPart p1, p0, A, B, C;
    parts.load(i, p1.r, p1.v);
    particle_path(parts, dt, i, &p0);
    fetch_tri(m, dt, gid, &A, &B, &C);
    float d[4];
    {
        MPart a = A, b = B, c = C;
        revert(dt, &a);
        revert(dt, &b);
        revert(dt, &c);
        (void)intersect(dt, a, b, c, p0, d);
    }
    float rw[3], vw[3];
    {
        MPart a = A, b = B, c = C;
        const float h = (float)((double)dt - (double)d[0]);
        revert(h, &a);
        revert(h, &b);
        revert(h, &c);
        const float u = d[1], v = d[2], w = 1 - u - v;
        for (int k = 0; k < 3; k++) {
            rw[k] = w * a.r[k] + u * b.r[k] + v * c.r[k];
            vw[k] = w * a.v[k] + u * b.v[k] + v * c.v[k];
        }
    }
    MPart pn;
    for (int k = 0; k < 3; k++) {
        pn.v[k] = 2 * vw[k] - p0.v[k];
        pn.r[k] = (float)((double)rw[k] + ((double)dt - (double)d[0]) * (double)pn.v[k]);
    }
    {
        float a[3], b[3], n[3];
        for (int k = 0; k < 3; k++) {
            a[k] = B.r[k] - A.r[k];
            b[k] = C.r[k] - A.r[k];
        }
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        float s = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        s = 1.0f / sqrtf(s);
        const float l = d[3] * 1e-2f;
        for (int k = 0; k < 3; k++) pn.r[k] = pn.r[k] + l * (n[k] * s);
    }
    float rcen[3], r0[3], r1[3];
    for (int k = 0; k < 3; k++) {
        rcen[k] = (float)(0.333333 * (double)(A.r[k] + B.r[k] + C.r[k]));
        r0[k] = p1.r[k] - rcen[k];
        r1[k] = pn.r[k] - rcen[k];
        dm[k] = -mass * (pn.v[k] - p1.v[k]);
        pos[k] = pn.r[k];
        vel[k] = pn.v[k];
    }
    const float *v0 = p1.v, *v1 = pn.v;
    dm[3] = -mass * (r1[1] * v1[2] - r1[2] * v1[1] - r0[1] * v0[2] + r0[2] - v0[1]);
    dm[4] = -mass * (r1[2] * v1[0] - r1[0] * v1[2] - r0[2] * v0[0] + r0[0] - v0[2]);
    dm[5] = -mass * (r1[0] * v1[1] - r1[1] * v1[0] - r0[0] * v0[1] + r0[1] - v0[0]);
}

// one thread per particle: the bounce of the crossing particles, the reset
// of the search keys and, when wrap is set, the periodic wrap of every
// particle
template <class P>
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_perform(MeshView m, int nmax, const int *n_dev, float dt,
                                                             float mass, int wrap, float Lx, float Ly, float Lz,
                                                             P parts, unsigned long long *keys, int *rec_tri,
                                                             float *rec_m, int *tri_cnt,
                                                             unsigned long long *ncol) {
    const int i = blockIdx.x * MBB_THREADS + threadIdx.x;
    if (i >= nmax || i >= *n_dev) return;
    const unsigned long long key = keys[i];
    int tri = -1;
    float r[3], v[3];
    if (key != MBB_EMPTY) {
        keys[i] = MBB_EMPTY;
        const int gid = (int)(key & 0xffffffffu);
        float dm[6];
        perform_one(m, parts, dt, mass, i, gid, r, v, dm);
        const int k = gid / m.nt;
        tri = m.inst_mesh[k] * m.nt + (gid - k * m.nt);
        for (int c = 0; c < 6; c++) rec_m[6 * (size_t)i + c] = dm[c];
        atomicAdd(tri_cnt + tri, 1);
        atomicAdd(ncol, 1ull);
    } else if (wrap) {
        float u[3];
        parts.load(i, r, u);
        for (int c = 0; c < 3; c++) v[c] = u[c];
    }
    rec_tri[i] = tri;
    if (tri >= 0 || wrap) {
        if (wrap) {
            r[0] = wrap1(r[0], Lx);
            r[1] = wrap1(r[1], Ly);
            r[2] = wrap1(r[2], Lz);
        }
        parts.store(i, r, v);
    }
}

// slot of each crossing in its triangle's segment (order fixed afterwards)
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_scatter(int nmax, const int *n_dev, const int *rec_tri,
                                                             const int *tri_start, int *tri_fill, int *tri_ids) {
    const int i = blockIdx.x * MBB_THREADS + threadIdx.x;
    if (i >= nmax || i >= *n_dev) return;
    const int t = rec_tri[i];
    if (t < 0) return;
    const int slot = atomicAdd(tri_fill + t, 1);
    tri_ids[tri_start[t] + slot] = i;
}

// ---- momentum to the vertices (src/mesh_bounce/dev/collect.h:84-103) ------
enum { XX, XY, XZ, YY, YZ, ZZ };

__device__ void inertia(float mass, const float a[3], const float b[3], const float c[3], float I[6]) {
    const float inv_12 = 1.f / 12.f, inv_24 = 1.f / 24.f;
    float M[3], d[3], e[3];
    for (int k = 0; k < 3; k++)
        M[k] = inv_12 * (a[k] * a[k] - a[k] * b[k] + b[k] * b[k] - c[k] * (a[k] + b[k]) + c[k] * c[k]);
    d[0] = 2 * a[0] - b[0] - c[0];
    d[1] = 2 * b[0] - a[0] - c[0];
    d[2] = 2 * c[0] - a[0] - b[0];
    e[0] = 2 * a[1] - b[1] - c[1];
    e[1] = 2 * b[1] - a[1] - c[1];
    e[2] = 2 * c[1] - a[1] - b[1];
    I[XX] = mass * (M[1] + M[2]);
    I[YY] = mass * (M[0] + M[2]);
    I[ZZ] = mass * (M[0] + M[1]);
    I[XY] = -inv_24 * mass * (d[0] * a[1] + d[1] * b[1] + d[2] * c[1]);
    I[XZ] = -inv_24 * mass * (d[0] * a[2] + d[1] * b[2] + d[2] * c[2]);
    I[YZ] = -inv_24 * mass * (e[0] * a[2] + e[1] * b[2] + e[2] * c[2]);
}

__device__ void inverse(const float A[6], float I[6]) {
    const float mx = A[YY] * A[ZZ] - A[YZ] * A[YZ];
    const float my = A[XY] * A[ZZ] - A[XZ] * A[YZ];
    const float mz = A[XY] * A[YZ] - A[XZ] * A[YY];
    float idet = mx * A[XX] - my * A[XY] + mz * A[XZ];
    idet = 1.f / idet;
    I[XX] = idet * mx;
    I[XY] = -idet * my;
    I[XZ] = idet * mz;
    I[YY] = idet * (A[XX] * A[ZZ] - A[XZ] * A[XZ]);
    I[YZ] = idet * (A[XY] * A[XZ] - A[XX] * A[YZ]);
    I[ZZ] = idet * (A[XX] * A[YY] - A[XY] * A[XY]);
}

__device__ __forceinline__ bool nz(float a) { return fabsf(a) > 1e-6f; }

// one thread per triangle: its momentum (crossings in ascending particle
// order), and for a nonzero momentum the velocity change of its vertices
// (rbc_M2v, src/mesh_bounce/dev/utils.h:172-213)
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_triangles(int ntri, int nt, int nv, const int *tt,
                                                               const float *rc, float rbc_mass, const int *tri_start,
                                                               int *tri_cnt, int *tri_fill, int *tri_ids,
                                                               const float *rec_m, float *mm, float *tri_dv,
                                                               int *tri_active) {
    const int t = blockIdx.x * MBB_THREADS + threadIdx.x;
    if (t >= ntri) return;
    const int s = tri_start[t], e = s + tri_cnt[t];
    tri_cnt[t] = 0;
    tri_fill[t] = 0;
    for (int a = s + 1; a < e; a++) {   // insertion sort: segments hold a few crossings
        const int x = tri_ids[a];
        int b = a - 1;
        while (b >= s && tri_ids[b] > x) {
            tri_ids[b + 1] = tri_ids[b];
            b--;
        }
        tri_ids[b + 1] = x;
    }
    float m[6] = {0, 0, 0, 0, 0, 0};
    for (int a = s; a < e; a++) {
        const int i = tri_ids[a];
        for (int c = 0; c < 6; c++) m[c] += rec_m[6 * (size_t)i + c];
    }
    if (mm)
        for (int c = 0; c < 6; c++) mm[6 * (size_t)t + c] = m[c];
    const bool act = nz(m[0]) && nz(m[1]) && nz(m[2]) && nz(m[3]) && nz(m[4]) && nz(m[5]);
    tri_active[t] = act ? 1 : 0;
    if (!act) return;
    const int cell = t / nt, tl = t - cell * nt;
    float x[3][3];
    for (int q = 0; q < 3; q++) {
        const int j = cell * nv + tt[3 * tl + q];
        for (int c = 0; c < 3; c++) x[q][c] = rc[6 * (size_t)j + c];
    }
    float I[6], Ii[6], om[3], v[3], com[3];
    const float one_third = 1.f / 3.f;
    inertia(rbc_mass, x[0], x[1], x[2], I);
    inverse(I, Ii);
    const float *L = m + 3;
    om[0] = one_third * (Ii[XX] * L[0] + Ii[XY] * L[1] + Ii[XZ] * L[2]);
    om[1] = one_third * (Ii[XY] * L[0] + Ii[YY] * L[1] + Ii[YZ] * L[2]);
    om[2] = one_third * (Ii[XZ] * L[0] + Ii[YZ] * L[1] + Ii[ZZ] * L[2]);
    const float inv_mass = (float)(1.0 / (double)rbc_mass);
    for (int k = 0; k < 3; k++) {
        v[k] = one_third * m[k] * inv_mass;
        com[k] = one_third * (x[0][k] + x[1][k] + x[2][k]);
    }
    for (int q = 0; q < 3; q++) {
        float r[3];
        for (int k = 0; k < 3; k++) r[k] = x[q][k] - com[k];
        float *o = tri_dv + 9 * (size_t)t + 3 * q;
        o[0] = v[0] + r[1] * om[2] - r[2] * om[1];
        o[1] = v[1] + r[2] * om[0] - r[0] * om[2];
        o[2] = v[2] + r[0] * om[1] - r[1] * om[0];
    }
}

// one thread per vertex: += the changes of its triangles in ascending order
__global__ __launch_bounds__(MBB_THREADS) void k_mbb_vertices(int nvtx, int nt, int nv, const int *vstart,
                                                              const int *vadj, const int *tri_active,
                                                              const float *tri_dv, float *rc) {
    const int j = blockIdx.x * MBB_THREADS + threadIdx.x;
    if (j >= nvtx) return;
    const int cell = j / nv, jl = j - cell * nv;
    float v[3];
    bool any = false;
    for (int a = vstart[jl]; a < vstart[jl + 1]; a++) {
        const int t = cell * nt + (vadj[a] >> 2), q = vadj[a] & 3;
        if (!tri_active[t]) continue;
        if (!any) {
            for (int c = 0; c < 3; c++) v[c] = rc[6 * (size_t)j + 3 + c];
            any = true;
        }
        for (int c = 0; c < 3; c++) v[c] += tri_dv[9 * (size_t)t + 3 * q + c];
    }
    if (any)
        for (int c = 0; c < 3; c++) rc[6 * (size_t)j + 3 + c] = v[c];
}

}  // namespace
}  // namespace udx

using namespace udx;

struct udx_mbb {
    udx_ctx *ctx;
    int nt, nv, max_meshes, maxp;
    int *tt;                       // 3 per triangle
    int *vstart, *vadj;            // vertex -> (triangle << 2 | corner), ascending triangle
    int *codes, *inst_mesh, *n_inst;
    float *inst_shift;
    unsigned long long *keys, *ncol;
    int *rec_tri, *tri_cnt, *tri_start, *tri_fill, *tri_ids, *tri_active, *scan_tmp;
    float *rec_m, *tri_dv;
    float4 *v0;                    // the particles' straight-line velocities
    unsigned *cand_n;              // (triangle instance, particle) pairs past the plane test
    uint2 *cand;
    int cand_cap;
    int *n_dev;                    // the particle count for the AoS entry point
};

namespace udx {

static void mbb_free(udx_mbb *b) {
    dev_free(b->tt);
    dev_free(b->vstart);
    dev_free(b->vadj);
    dev_free(b->codes);
    dev_free(b->inst_mesh);
    dev_free(b->n_inst);
    dev_free(b->inst_shift);
    dev_free(b->keys);
    dev_free(b->ncol);
    dev_free(b->rec_tri);
    dev_free(b->tri_cnt);
    dev_free(b->tri_start);
    dev_free(b->tri_fill);
    dev_free(b->tri_ids);
    dev_free(b->tri_active);
    dev_free(b->scan_tmp);
    dev_free(b->rec_m);
    dev_free(b->tri_dv);
    dev_free(b->n_dev);
    dev_free(b->v0);
    dev_free(b->cand_n);
    dev_free(b->cand);
}

udx_ctx *mbb_ctx(const udx_mbb *b) { return b->ctx; }
int mbb_nv(const udx_mbb *b) { return b->nv; }

// the bounce on either particle layout; wrap: apply the periodic wrap to
// every particle (the solvent step defers it to here)
template <class P>
static int mbb_run(hipStream_t s, udx_mbb *b, float dt, float flu_mass, float rbc_mass, const int L[3], int nm,
                   const float *rp, float *rc, const int *starts, int nmax, const int *n_dev, P parts, int wrap,
                   float *mm) {
    MeshView m{b->nt, b->nv, b->tt, rp, rc, b->inst_mesh, b->inst_shift};
    const int ntri = nm * b->nt;
    UDX_HIP(hipMemsetAsync(b->ncol, 0, sizeof(unsigned long long), s));
    if (nm > 0) {
        hipLaunchKernelGGL(k_mbb_codes, dim3(nm), dim3(MBB_THREADS), 0, s, b->nv, L[0], L[1], L[2], rc, b->codes);
        hipLaunchKernelGGL(k_mbb_instances, dim3(1), dim3(1024), 0, s, nm, L[0], L[1], L[2], b->codes, b->inst_mesh,
                           b->inst_shift, b->n_inst);
        const int max_gid = MBB_MAX_IMAGES * ntri;
        if (nmax > 0)
            hipLaunchKernelGGL((k_mbb_paths<P>), dim3(ceil_div(nmax, MBB_THREADS)), dim3(MBB_THREADS), 0, s, nmax,
                               n_dev, dt, parts, b->v0);
        UDX_HIP(hipMemsetAsync(b->cand_n, 0, sizeof(unsigned), s));
        hipLaunchKernelGGL((k_mbb_find<P>), dim3(ceil_div(max_gid, MBB_THREADS)), dim3(MBB_THREADS), 0, s, m,
                           max_gid, b->n_inst, dt, L[0], L[1], L[2], starts, parts, b->v0, b->keys, b->cand_cap,
                           b->cand_n, b->cand);
        hipLaunchKernelGGL((k_mbb_cand<P>), dim3(MBB_CAND_BLOCKS), dim3(MBB_THREADS), 0, s, m, dt, parts, b->v0,
                           b->cand_cap, b->cand_n, b->cand, b->keys);
        hipLaunchKernelGGL((k_mbb_find_all<P>), dim3(MBB_CAND_BLOCKS), dim3(MBB_THREADS), 0, s, m, max_gid,
                           b->n_inst, dt, L[0], L[1], L[2], starts, parts, b->v0, b->keys, b->cand_cap, b->cand_n);
    }
    if (nmax > 0)
        hipLaunchKernelGGL((k_mbb_perform<P>), dim3(ceil_div(nmax, MBB_THREADS)), dim3(MBB_THREADS), 0, s, m, nmax,
                           n_dev, dt, flu_mass, wrap, (float)L[0], (float)L[1], (float)L[2], parts, b->keys,
                           b->rec_tri, b->rec_m, b->tri_cnt, b->ncol);
    if (nm > 0) {
        UDX_TRY(scan_exclusive(s, b->tri_cnt, ntri, b->tri_start, b->scan_tmp));
        if (nmax > 0)
            hipLaunchKernelGGL(k_mbb_scatter, dim3(ceil_div(nmax, MBB_THREADS)), dim3(MBB_THREADS), 0, s, nmax,
                               n_dev, b->rec_tri, b->tri_start, b->tri_fill, b->tri_ids);
        hipLaunchKernelGGL(k_mbb_triangles, dim3(ceil_div(ntri, MBB_THREADS)), dim3(MBB_THREADS), 0, s, ntri, b->nt,
                           b->nv, b->tt, rc, rbc_mass, b->tri_start, b->tri_cnt, b->tri_fill, b->tri_ids, b->rec_m,
                           mm, b->tri_dv, b->tri_active);
        hipLaunchKernelGGL(k_mbb_vertices, dim3(ceil_div(nm * b->nv, MBB_THREADS)), dim3(MBB_THREADS), 0, s,
                           nm * b->nv, b->nt, b->nv, b->vstart, b->vadj, b->tri_active, b->tri_dv, rc);
    }
    UDX_LAUNCHED();
    return UDX_OK;
}

int mbb_step(hipStream_t s, udx_mbb *b, float dt, float flu_mass, float rbc_mass, const Box &box, int nm,
             const float *rp, float *rc, const int *starts, int nmax, const int *n_dev, float4 *pos, float4 *vel,
             const float4 *pos0) {
    UDX_CHECK(nm <= b->max_meshes && nmax <= b->maxp, UDX_ECAP, "membrane bounce: %d meshes / %d particles above "
              "the capacity %d / %d", nm, nmax, b->max_meshes, b->maxp);
    PartsF4 parts{pos, vel, pos0};
    return mbb_run(s, b, dt, flu_mass, rbc_mass, box.L, nm, rp, rc, starts, nmax, n_dev, parts, 1, nullptr);
}

}  // namespace udx

extern "C" {

int udx_mbb_ini(udx_ctx *ctx, int nv, int nt, const int *tt_host, int max_meshes, int maxp, udx_mbb **out) {
    UDX_CHECK(ctx && tt_host && out && nv > 0 && nt > 0 && max_meshes >= 0 && maxp >= 0, UDX_EINVAL,
              "udx_mbb_ini: bad arguments");
    UDX_CHECK((long)MBB_MAX_IMAGES * max_meshes * nt < (1L << 31), UDX_ECAP, "udx_mbb_ini: %d meshes of %d "
              "triangles exceed the triangle index range", max_meshes, nt);
    for (int k = 0; k < 3 * nt; k++)
        UDX_CHECK(tt_host[k] >= 0 && tt_host[k] < nv, UDX_EINVAL, "udx_mbb_ini: vertex index %d out of range",
                  tt_host[k]);
    // vertex -> (triangle, corner) in ascending triangle order
    std::vector<int> cnt(nv + 1, 0), adj(3 * (size_t)nt);
    for (int k = 0; k < 3 * nt; k++) cnt[tt_host[k] + 1]++;
    for (int v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int t = 0; t < nt; t++)
        for (int q = 0; q < 3; q++) adj[fill[tt_host[3 * t + q]]++] = t << 2 | q;
    udx_mbb *b = new udx_mbb();
    b->ctx = ctx;
    b->nt = nt;
    b->nv = nv;
    b->max_meshes = max_meshes;
    b->maxp = maxp;
    const int mm1 = max_meshes > 0 ? max_meshes : 1, np1 = maxp > 0 ? maxp : 1;
    const int ntri = mm1 * nt;
    int st = dev_alloc(&b->tt, 3 * (size_t)nt);
    if (!st) st = dev_alloc(&b->vstart, (size_t)nv + 1);
    if (!st) st = dev_alloc(&b->vadj, 3 * (size_t)nt);
    if (!st) st = dev_alloc(&b->codes, 3 * (size_t)mm1);
    if (!st) st = dev_alloc(&b->inst_mesh, (size_t)MBB_MAX_IMAGES * mm1);
    if (!st) st = dev_alloc(&b->inst_shift, 3 * (size_t)MBB_MAX_IMAGES * mm1);
    if (!st) st = dev_alloc(&b->n_inst, 1);
    if (!st) st = dev_alloc(&b->keys, (size_t)np1);
    if (!st) st = dev_alloc(&b->ncol, 1);
    if (!st) st = dev_alloc(&b->rec_tri, (size_t)np1);
    if (!st) st = dev_alloc(&b->rec_m, 6 * (size_t)np1);
    if (!st) st = dev_alloc(&b->tri_cnt, (size_t)ntri + 1);
    if (!st) st = dev_alloc(&b->tri_start, (size_t)ntri + 1);
    if (!st) st = dev_alloc(&b->tri_fill, (size_t)ntri);
    if (!st) st = dev_alloc(&b->tri_ids, (size_t)np1);
    if (!st) st = dev_alloc(&b->tri_active, (size_t)ntri);
    if (!st) st = dev_alloc(&b->tri_dv, 9 * (size_t)ntri);
    if (!st) st = dev_alloc(&b->scan_tmp, (size_t)scan_scratch_size(ntri));
    if (!st) st = dev_alloc(&b->n_dev, 1);
    if (!st) st = dev_alloc(&b->v0, (size_t)np1);
    // the pairs past the plane test: a few per crossing particle; beyond the
    // list the search tests pairs in place
    b->cand_cap = 2 * np1 > (1 << 20) ? 2 * np1 : (1 << 20);
    if (!st) st = dev_alloc(&b->cand_n, 1);
    if (!st) st = dev_alloc(&b->cand, (size_t)b->cand_cap);
    hipStream_t s = ctx->stream;
    if (!st && (hipMemcpyAsync(b->tt, tt_host, 3 * sizeof(int) * (size_t)nt, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(b->vstart, cnt.data(), sizeof(int) * ((size_t)nv + 1), hipMemcpyHostToDevice, s) !=
                    hipSuccess ||
                hipMemcpyAsync(b->vadj, adj.data(), 3 * sizeof(int) * (size_t)nt, hipMemcpyHostToDevice, s) !=
                    hipSuccess ||
                hipMemsetAsync(b->keys, 0xff, sizeof(unsigned long long) * (size_t)np1, s) != hipSuccess ||
                hipMemsetAsync(b->tri_cnt, 0, sizeof(int) * ((size_t)ntri + 1), s) != hipSuccess ||
                hipMemsetAsync(b->tri_fill, 0, sizeof(int) * (size_t)ntri, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        st = UDX_EHIP;
    if (st) {
        mbb_free(b);
        delete b;
        set_error("udx_mbb_ini: allocation failed");
        return st;
    }
    *out = b;
    return UDX_OK;
}

int udx_mbb_fin(udx_mbb *b) {
    if (!b) return UDX_OK;
    UDX_HIP(hipStreamSynchronize(b->ctx->stream));
    mbb_free(b);
    delete b;
    return UDX_OK;
}

int udx_mbb_bounce(udx_mbb *b, float dt, float flu_mass, float rbc_mass, int Lx, int Ly, int Lz, int nm,
                   const float *rp, float *rc, const int *starts, int n, const float *pp0, float *pp, float *mm) {
    UDX_CHECK(b && dt > 0 && flu_mass > 0 && rbc_mass > 0 && Lx > 0 && Ly > 0 && Lz > 0 && nm >= 0 && n >= 0,
              UDX_EINVAL, "udx_mbb_bounce: bad arguments");
    UDX_CHECK(nm == 0 || (rp && rc), UDX_EINVAL, "udx_mbb_bounce: null vertices");
    UDX_CHECK(n == 0 || (pp && pp0 && starts), UDX_EINVAL, "udx_mbb_bounce: null solvent");
    UDX_CHECK(nm <= b->max_meshes && n <= b->maxp, UDX_ECAP, "udx_mbb_bounce: %d meshes / %d particles above the "
              "capacity %d / %d", nm, n, b->max_meshes, b->maxp);
    hipStream_t s = b->ctx->stream;
    UDX_HIP(hipMemcpyAsync(b->n_dev, &n, sizeof n, hipMemcpyHostToDevice, s));
    const int L[3] = {Lx, Ly, Lz};
    PartsAoS parts{pp, pp0};
    UDX_TRY(mbb_run(s, b, dt, flu_mass, rbc_mass, L, nm, rp, rc, starts, n, b->n_dev, parts, 0, mm));
    UDX_HIP(hipStreamSynchronize(s));
    return UDX_OK;
}

int udx_mbb_count(udx_mbb *b, unsigned long long *ncol) {
    UDX_CHECK(b && ncol, UDX_EINVAL, "udx_mbb_count: null argument");
    UDX_HIP(hipMemcpyAsync(ncol, b->ncol, sizeof *ncol, hipMemcpyDeviceToHost, b->ctx->stream));
    UDX_HIP(hipStreamSynchronize(b->ctx->stream));
    return UDX_OK;
}

const int *udx_mbb_collisions(const udx_mbb *b) { return b ? b->rec_tri : nullptr; }

int udx_mbb_set_pair_capacity(udx_mbb *b, int cap) {
    UDX_CHECK(b && cap >= 1, UDX_EINVAL, "udx_mbb_set_pair_capacity: bad arguments");
    UDX_HIP(hipStreamSynchronize(b->ctx->stream));
    dev_free(b->cand);
    b->cand_cap = 0;
    UDX_TRY(dev_alloc(&b->cand, (size_t)cap));
    b->cand_cap = cap;
    return UDX_OK;
}

}  // extern "C"
