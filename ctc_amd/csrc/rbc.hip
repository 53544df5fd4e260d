// RBC membrane on MI355X: per-cell area and volume (FP64), the membrane
// force (global and local area, volume, WLC + POW springs, viscosity) and
// Kantor bending, and constant stretching loads.
//
// Reference: one thread per (vertex, ring slot) evaluates its triangle and
// dihedrals in double and adds the result to the vertex by a float atomic
// (src/rbc/force/dev/main.h:48-77, src/rbc/force/bending/kantor/dev/main.h);
// area and volume are FP32 warp sums with atomics per cell
// (src/rbc/force/area_volume/dev/main.h).  Here one thread per vertex walks
// its ring slots in order and sums in double, so forces are deterministic and
// rounded to float once; area and volume are FP64 block reductions.  The
// double arithmetic runs at the FP32 rate on CDNA4's vector units.
#include <cmath>
#include <vector>

#include "udx_internal.h"

namespace udx {

namespace {

constexpr int RBC_THREADS = 128;

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 mk(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ D3 rpos(const float *pp, int i) {
    const float *p = pp + 6 * (long)i;
    return mk(p[0], p[1], p[2]);
}
__device__ __forceinline__ D3 rvel(const float *pp, int i) {
    const float *p = pp + 6 * (long)i;
    return mk(p[3], p[4], p[5]);
}

struct RbcK {
    int md, nv, nc;
    const int *adj0, *adj1;
    const float *ll, *aa;   // per slot (stress free) or null
    float l0, a0;           // stress-full references
    double ka, kd, kv, ks, x0, mpow, gammaC, gammaT, totArea, totVolume;
    double kb, sphi, cphi;  // bending; sin/cos of the spontaneous angle
    int bending;
};

// (vertex, slot) -> i1..i4 of src/rbc/adj/dev.h:2-42 (template indices)
__device__ __forceinline__ bool adj_map(const RbcK &K, int v, int j, int *i1, int *i2, int *i3, int *i4) {
    const int md = K.md, k = (v % K.nv) * md;
    *i1 = K.adj0[k + j];
    if (*i1 == -1) return false;
    int b = K.adj0[k + (j + 1) % md], c;
    if (b == -1) {
        b = K.adj0[k];
        c = K.adj0[k + 1];
    } else {
        c = K.adj0[k + (j + 2) % md];
        if (c == -1) c = K.adj0[k];
    }
    *i2 = b;
    *i3 = c;
    *i4 = K.adj1[k + j];
    return true;
}

__device__ __forceinline__ double wlc(double lmax, double ks, double r) {
    const double x = r / lmax;
    return ks / lmax * ((4 * x * x - 9 * x + 6) / (4 * (x - 1) * (x - 1)));
}

// force on r1 of triangle (r1, r2, r3): src/rbc/force/dev/common.h:1-84
__device__ D3 ftri(const RbcK &K, D3 r1, D3 r2, D3 r3, double l0, double a0, double A, double V) {
    const D3 x21 = r2 - r1, x32 = r3 - r2, x31 = r3 - r1;
    const D3 nn = cross(x21, x31);
    D3 f = cross(nn, x32);
    const double a = 0.5 * sqrt(dot(nn, nn));
    const double fA = -K.ka * (A - K.totArea) / (4 * K.totArea * a);
    const double fa = -K.kd * (a - a0) / (4 * a0 * a);
    f = (fA + fa) * f;
    f = f + (K.kv * (V - K.totVolume) / (6 * K.totVolume)) * cross(r3, r2);
    // spring length in float as the reference's sqrtf
    double r = (double)__builtin_sqrtf((float)dot(x21, x21));
    const double lmax = l0 / K.x0;
    if (r >= lmax) r = 0.99 * lmax;
    const double fwlc = wlc(lmax, K.ks, r);
    const float m1 = (float)(K.mpow + 1);
    const double fpow = -wlc(lmax, K.ks, l0) * powf((float)l0, m1) / powf((float)r, m1);
    return f + (fwlc + fpow) * x21;
}

// src/rbc/force/dev/common.h:86-98
__device__ __forceinline__ D3 fvisc(const RbcK &K, D3 r1, D3 r2, D3 u1, D3 u2) {
    const D3 du = u2 - u1, dr = r1 - r2;
    const double fac = dot(du, dr) / dot(dr, dr);
    return K.gammaT * du + (K.gammaC * fac) * dr;
}

// one Kantor dihedral (a, b, c | b, c, d): force on a (which == 1) or on b
// (which == 2) (src/mesh/force/kantor0/dev.h:28-120)
template <int WHICH>
__device__ D3 dih(const RbcK &K, D3 a, D3 b, D3 c, D3 d) {
    const D3 ab = a - b, ac = a - c, cd = c - d, bd = b - d, bc = b - c, da = d - a;
    const D3 k = cross(ab, ac), n = cross(bd, bc);
    const double over_k = 1.0 / sqrt(dot(k, k)), over_n = 1.0 / sqrt(dot(n, n));
    const double cost = dot(k, n) * over_k * over_n;
    const double s2 = 1.0 - cost * cost;
    double over_sin = 1.0 / sqrt(s2 > 1e-6 ? s2 : 1e-6);
    over_sin = copysign(over_sin, dot(k - n, da));
    const double beta = K.cphi * K.kb - cost * (K.sphi * K.kb) * over_sin;
    const double b11 = -beta * cost * over_k * over_k, b12 = beta * over_k * over_n;
    if (WHICH == 1) {
        const D3 r32 = c - b;
        return b11 * cross(k, r32) + b12 * cross(n, r32);
    }
    const double b22 = -beta * cost * over_n * over_n;
    D3 f = b11 * cross(k, ac);
    f = f + b12 * (cross(k, cd) + cross(n, ac));
    return f + b22 * cross(n, cd);
}

// one thread per vertex: membrane terms over the ring slots, then bending,
// summed in double and added to the float force once
__global__ __launch_bounds__(RBC_THREADS) void k_rbc_force(RbcK K, const float *pp, const double *av, float *ff) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= K.nc * K.nv) return;
    const int cell = v / K.nv, off = cell * K.nv;
    const double A = av[2 * cell], V = av[2 * cell + 1];
    const D3 r0 = rpos(pp, v), u0 = rvel(pp, v);
    D3 fm = mk(0, 0, 0), fb = mk(0, 0, 0);
    for (int j = 0; j < K.md; j++) {
        int i1, i2, i3, i4;
        if (!adj_map(K, v, j, &i1, &i2, &i3, &i4)) continue;
        const int s = (v % K.nv) * K.md + j;
        const double l0 = K.ll ? (double)K.ll[s] : (double)K.l0;
        const double a0 = K.aa ? (double)K.aa[s] : (double)K.a0;
        const D3 r1 = rpos(pp, i1 + off), r2 = rpos(pp, i2 + off);
        fm = fm + ftri(K, r0, r1, r2, l0, a0, A, V);
        fm = fm + fvisc(K, r0, r1, u0, rvel(pp, i1 + off));
        if (K.bending) {
            const D3 r3 = rpos(pp, i3 + off), r4 = rpos(pp, i4 + off);
            fb = fb + dih<1>(K, r0, r2, r1, r4);
            fb = fb + dih<2>(K, r1, r0, r2, r3);
        }
    }
    const D3 f = fm + fb;
    float *o = ff + 3 * (long)v;
    o[0] += (float)f.x;
    o[1] += (float)f.y;
    o[2] += (float)f.z;
}

// per cell [area, volume] in FP64: one block per cell, fixed-order tree
// reduction (src/rbc/force/area_volume/dev/main.h:1-46)
constexpr int AV_THREADS = 256;
__global__ __launch_bounds__(AV_THREADS) void k_rbc_av(int nt, int nv, const int *tt, const float *pp, double *av) {
    __shared__ double sa[AV_THREADS / 64], sv[AV_THREADS / 64];
    const int cell = blockIdx.x, off = cell * nv;
    double a = 0, vol = 0;
    for (int t = threadIdx.x; t < nt; t += AV_THREADS) {
        const D3 r0 = rpos(pp, tt[3 * t] + off), r1 = rpos(pp, tt[3 * t + 1] + off), r2 = rpos(pp, tt[3 * t + 2] + off);
        const D3 n = cross(r1 - r0, r2 - r0);
        a += 0.5 * sqrt(dot(n, n));
        vol += dot(r0, cross(r1, r2)) / 6.0;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        vol += __shfl_xor(vol, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        sv[threadIdx.x >> 6] = vol;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double A = 0, V = 0;
        for (int w = 0; w < AV_THREADS / 64; w++) {
            A += sa[w];
            V += sv[w];
        }
        av[2 * cell] = A;
        av[2 * cell + 1] = V;
    }
}

// src/rbc/stretch/dev/main.h:1-17
__global__ void k_rbc_stretch(int n, int nv, const float *f, float *ff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = i % nv;
    float *o = ff + 3 * (long)i;
    o[0] += f[3 * k];
    o[1] += f[3 * k + 1];
    o[2] += f[3 * k + 2];
}

// ---------------------------------------------------------------- adjoint
// F = -grad E with E = ka E_A + kd E_a + kv E_V + ks E_s + kb E_b, so dF/dx
// is symmetric and the vector-Jacobian product with a cotangent g is the
// derivative of F along g.  One thread per vertex re-evaluates its ring's
// triangle and dihedral terms (nothing is stored from the forward pass) in
// forward-mode dual numbers with tangent g, with the cell's area and volume
// tangents from k_rbc_av_dual; the unit-coefficient term values give
// dL/dk = g . F_k.  The viscous term is excluded (zero relative velocities)
// and the spring length is the double root.
struct Dd { double v, d; };
__device__ __forceinline__ Dd dc(double v) { return Dd{v, 0.0}; }
__device__ __forceinline__ Dd operator+(Dd a, Dd b) { return Dd{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dd operator-(Dd a, Dd b) { return Dd{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dd operator*(Dd a, Dd b) { return Dd{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dd operator/(Dd a, Dd b) { return Dd{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Dd dsqrt(Dd a) {
    const double s = sqrt(a.v);
    return Dd{s, a.d / (2 * s)};
}
__device__ __forceinline__ Dd dpowc(Dd a, double c) {
    const double p = pow(a.v, c);
    return Dd{p, c * p / a.v * a.d};
}
struct V3 { Dd x, y, z; };
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(Dd s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ Dd vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vzero() { return V3{dc(0), dc(0), dc(0)}; }
__device__ __forceinline__ V3 dpos(const float *pp, const float *g, int i) {
    const float *p = pp + 6 * (long)i, *t = g + 3 * (long)i;
    return V3{Dd{p[0], t[0]}, Dd{p[1], t[1]}, Dd{p[2], t[2]}};
}

// per cell [A, dA, V, dV]
__global__ __launch_bounds__(AV_THREADS) void k_rbc_av_dual(int nt, int nv, const int *tt, const float *pp,
                                                            const float *g, double *avd) {
    __shared__ double s[4][AV_THREADS / 64];
    const int cell = blockIdx.x, off = cell * nv;
    Dd a = dc(0), vol = dc(0);
    for (int t = threadIdx.x; t < nt; t += AV_THREADS) {
        const V3 r0 = dpos(pp, g, tt[3 * t] + off), r1 = dpos(pp, g, tt[3 * t + 1] + off),
                 r2 = dpos(pp, g, tt[3 * t + 2] + off);
        const V3 n = vcross(r1 - r0, r2 - r0);
        a = a + dc(0.5) * dsqrt(vdot(n, n));
        vol = vol + vdot(r0, vcross(r1, r2)) / dc(6.0);
    }
    double q[4] = {a.v, a.d, vol.v, vol.d};
    for (int k = 0; k < 4; k++)
        for (int d = 32; d >= 1; d >>= 1) q[k] += __shfl_xor(q[k], d, 64);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; k++) s[k][threadIdx.x >> 6] = q[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0;
        for (int w = 0; w < AV_THREADS / 64; w++) t += s[threadIdx.x][w];
        avd[4 * cell + threadIdx.x] = t;
    }
}

struct AdjK {
    int md, nv, nc;
    const int *adj0, *adj1;
    const float *ll, *aa;
    float l0, a0;
    double coef[5];           // ka, kd, kv, ks, kb
    double totArea, totVolume, x0, mpow, sphi, cphi;
};

__device__ __forceinline__ int adj_map_a(const AdjK &K, int v, int j, int *i1, int *i2, int *i3, int *i4) {
    const int md = K.md, k = (v % K.nv) * md;
    *i1 = K.adj0[k + j];
    if (*i1 == -1) return 0;
    int b = K.adj0[k + (j + 1) % md], c;
    if (b == -1) {
        b = K.adj0[k];
        c = K.adj0[k + 1];
    } else {
        c = K.adj0[k + (j + 2) % md];
        if (c == -1) c = K.adj0[k];
    }
    *i2 = b;
    *i3 = c;
    *i4 = K.adj1[k + j];
    return 1;
}

__device__ void tri_terms(const AdjK &K, V3 r1, V3 r2, V3 r3, double l0, double a0, Dd A, Dd V, V3 out[4]) {
    const V3 x21 = r2 - r1, x32 = r3 - r2, x31 = r3 - r1;
    const V3 nn = vcross(x21, x31), f = vcross(nn, x32);
    const Dd a = dc(0.5) * dsqrt(vdot(nn, nn));
    const Dd fA = (dc(-1.0) * (A - dc(K.totArea))) / (dc(4 * K.totArea) * a);
    const Dd fa = (dc(-1.0) * (a - dc(a0))) / (dc(4 * a0) * a);
    out[0] = fA * f;
    out[1] = fa * f;
    out[2] = ((V - dc(K.totVolume)) / dc(6 * K.totVolume)) * vcross(r3, r2);
    const double lmax = l0 / K.x0, m = K.mpow;
    Dd r = dsqrt(vdot(x21, x21));
    if (r.v >= lmax) r = dc(0.99 * lmax);
    const Dd xr = r / dc(lmax);
    const Dd w = ((dc(4) * (xr * xr) - dc(9) * xr) + dc(6)) / (dc(4) * ((xr - dc(1)) * (xr - dc(1))));
    const Dd fwlc = dc(1.0 / lmax) * w;
    const double xl = l0 / lmax, wl = (4 * xl * xl - 9 * xl + 6) / (4 * (xl - 1) * (xl - 1)) / lmax;
    const Dd fpow = dc(-wl * pow(l0, m + 1)) / dpowc(r, m + 1);
    out[3] = (fwlc + fpow) * x21;
}

template <int WHICH>
__device__ V3 dih_dual(const AdjK &K, V3 a, V3 b, V3 c, V3 d) {
    const V3 ab = a - b, ac = a - c, cd = c - d, bd = b - d, bc = b - c, da = d - a;
    const V3 k = vcross(ab, ac), n = vcross(bd, bc);
    const Dd over_k = dc(1) / dsqrt(vdot(k, k)), over_n = dc(1) / dsqrt(vdot(n, n));
    const Dd cost = (vdot(k, n) * over_k) * over_n;
    const Dd s2 = dc(1) - cost * cost;
    Dd over_sin = dc(1) / dsqrt(s2.v > 1e-6 ? s2 : dc(1e-6));
    if (vdot(k - n, da).v < 0) over_sin = Dd{-over_sin.v, -over_sin.d};
    const Dd beta = dc(K.cphi) - (cost * dc(K.sphi)) * over_sin;
    const Dd b11 = ((dc(-1) * beta) * cost) * (over_k * over_k), b12 = beta * (over_k * over_n);
    if (WHICH == 1) {
        const V3 r32 = c - b;
        return b11 * vcross(k, r32) + b12 * vcross(n, r32);
    }
    const Dd b22 = ((dc(-1) * beta) * cost) * (over_n * over_n);
    V3 f = b11 * vcross(k, ac);
    f = f + b12 * (vcross(k, cd) + vcross(n, ac));
    return f + b22 * vcross(n, cd);
}

// one vertex: gx and q[k] = g . F_k
__device__ __forceinline__ void adjoint_vertex(const AdjK &K, int v, const float *pp, const float *g,
                                               const double *avd, float *gx, double q[5]) {
    const int cell = v / K.nv, off = cell * K.nv;
    const Dd A{avd[4 * cell], avd[4 * cell + 1]}, Vol{avd[4 * cell + 2], avd[4 * cell + 3]};
    const V3 r0 = dpos(pp, g, v);
    V3 t[5] = {vzero(), vzero(), vzero(), vzero(), vzero()};
    for (int j = 0; j < K.md; j++) {
        int i1, i2, i3, i4;
        if (!adj_map_a(K, v, j, &i1, &i2, &i3, &i4)) continue;
        const int s = (v % K.nv) * K.md + j;
        const double l0 = K.ll ? (double)K.ll[s] : (double)K.l0, a0 = K.aa ? (double)K.aa[s] : (double)K.a0;
        const V3 r1 = dpos(pp, g, i1 + off), r2 = dpos(pp, g, i2 + off);
        V3 f4[4];
        tri_terms(K, r0, r1, r2, l0, a0, A, Vol, f4);
        for (int k = 0; k < 4; k++) t[k] = t[k] + f4[k];
        const V3 r3 = dpos(pp, g, i3 + off), r4 = dpos(pp, g, i4 + off);
        t[4] = t[4] + dih_dual<1>(K, r0, r2, r1, r4);
        t[4] = t[4] + dih_dual<2>(K, r1, r0, r2, r3);
    }
    const float *gv = g + 3 * (long)v;
    double dx = 0, dy = 0, dz = 0;
    for (int k = 0; k < 5; k++) {
        dx += K.coef[k] * t[k].x.d;
        dy += K.coef[k] * t[k].y.d;
        dz += K.coef[k] * t[k].z.d;
        q[k] = ((double)gv[0] * t[k].x.v + (double)gv[1] * t[k].y.v) + (double)gv[2] * t[k].z.v;
    }
    float *o = gx + 3 * (long)v;
    o[0] = (float)dx;
    o[1] = (float)dy;
    o[2] = (float)dz;
}

// gpb[5 b + k]: block b's sum of g . F_k over its vertices (wave shuffles,
// then the block's waves in order)
__global__ __launch_bounds__(RBC_THREADS) void k_rbc_adjoint(AdjK K, const float *pp, const float *g,
                                                             const double *avd, float *gx, double *gpb) {
    __shared__ double sg[5][RBC_THREADS / 64];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double q[5] = {0, 0, 0, 0, 0};
    if (v < K.nc * K.nv) adjoint_vertex(K, v, pp, g, avd, gx, q);
    for (int k = 0; k < 5; k++)
        for (int d = 32; d >= 1; d >>= 1) q[k] += __shfl_xor(q[k], d, 64);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; k++) sg[k][threadIdx.x >> 6] = q[k];
    __syncthreads();
    if (threadIdx.x < 5) {
        double t = 0;
        for (int w = 0; w < RBC_THREADS / 64; w++) t += sg[threadIdx.x][w];
        gpb[5 * (long)blockIdx.x + threadIdx.x] = t;
    }
}
// gp[k] = sum over blocks of gpb[5 b + k]: one block, fixed order
__global__ __launch_bounds__(256) void k_rbc_gp(int nb, const double *gpb, double *gp) {
    __shared__ double s[5][4];
    double q[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < 5; k++) q[k] += gpb[5 * (long)b + k];
    for (int k = 0; k < 5; k++)
        for (int d = 32; d >= 1; d >>= 1) q[k] += __shfl_xor(q[k], d, 64);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; k++) s[k][threadIdx.x >> 6] = q[k];
    __syncthreads();
    if (threadIdx.x < 5) gp[threadIdx.x] = s[threadIdx.x][0] + s[threadIdx.x][1] + s[threadIdx.x][2] + s[threadIdx.x][3];
}

// ---------------------------------------------------------------- host
// ring adjacency of every vertex (src/rbc/adj/imp/ini.h:1-45): the
// neighbours in face order starting from the smallest, and the vertex across
// each ring edge; -1 pads vertices of degree < md
int build_adjacency(int md, int nt, int nv, const int *tt, std::vector<int> &adj0, std::vector<int> &adj1) {
    const int n = md * nv;
    std::vector<int> key(n, -1), val(n, -1);
    adj0.assign(n, -1);
    adj1.assign(n, -1);
    auto put = [&](int f, int x, int y) -> bool {
        for (int j = f * md; j < (f + 1) * md; j++)
            if (key[j] == -1) {
                key[j] = x;
                val[j] = y;
                return true;
            }
        return false;
    };
    auto get = [&](int f, int x) -> int {
        for (int j = f * md; j < (f + 1) * md; j++)
            if (key[j] == x) return val[j];
        return -1;
    };
    for (int t = 0; t < nt; t++) {
        const int a = tt[3 * t], b = tt[3 * t + 1], c = tt[3 * t + 2];
        UDX_CHECK(a >= 0 && a < nv && b >= 0 && b < nv && c >= 0 && c < nv, UDX_EINVAL,
                  "rbc: triangle %d = [%d %d %d] out of range (nv = %d)", t, a, b, c, nv);
        UDX_CHECK(put(a, b, c) && put(b, c, a) && put(c, a, b), UDX_EINVAL, "rbc: vertex degree above %d", md);
    }
    for (int v = 0; v < nv; v++) {
        UDX_CHECK(key[v * md] != -1, UDX_EINVAL, "rbc: isolated vertex %d", v);
        int mi = key[v * md];
        for (int j = v * md + 1; j < (v + 1) * md && key[j] != -1; j++) mi = key[j] < mi ? key[j] : mi;
        int c = mi, k = v * md;
        do {
            UDX_CHECK(k < (v + 1) * md, UDX_EINVAL, "rbc: ring of vertex %d does not close", v);
            const int c0 = c;
            c = get(v, c0);
            UDX_CHECK(c != -1, UDX_EINVAL, "rbc: open mesh at vertex %d", v);
            adj0[k] = c0;
            adj1[k] = get(c, c0);
            UDX_CHECK(adj1[k] != -1, UDX_EINVAL, "rbc: edge %d-%d has one face", c, c0);
            k++;
        } while (c != mi);
    }
    return UDX_OK;
}

// ---------------------------------------------------------------- redistribution
// Membranes move as whole cells: a cell whose bounding-box centre
// 0.5 (min + max) has left [-L/2, L/2) on an axis goes to the fragment of
// that direction (src/distr/rbc/dev.h:3-20, src/distr/map/dev.h:17-26) and,
// on one periodic domain, comes back shifted by -L d (src/distr/common/dev.h:
// 20-32, src/distr/rbc/imp/unpack.h:30-51).  The reference appends the moved
// cells after the kept ones, kept cells first, then fragments 0..25; within
// each group the order here is ascending (the reference's comes from
// atomics).
constexpr int RD_THREADS = 256;
__global__ __launch_bounds__(RD_THREADS) void k_rbc_cell_fid(int nv, int Lx, int Ly, int Lz, const float *pp,
                                                            int *cfid) {
    __shared__ float s[6][RD_THREADS / 64];
    const int cell = blockIdx.x;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int k = threadIdx.x; k < nv; k += RD_THREADS) {
        const float *p = pp + 6 * ((long)cell * nv + k);
        for (int c = 0; c < 3; c++) {
            lo[c] = fminf(lo[c], p[c]);
            hi[c] = fmaxf(hi[c], p[c]);
        }
    }
    for (int c = 0; c < 3; c++)
        for (int d = 32; d >= 1; d >>= 1) {
            lo[c] = fminf(lo[c], __shfl_xor(lo[c], d, 64));
            hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], d, 64));
        }
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; c++) {
            s[c][threadIdx.x >> 6] = lo[c];
            s[3 + c][threadIdx.x >> 6] = hi[c];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int L[3] = {Lx, Ly, Lz};
        int d[3];
        for (int c = 0; c < 3; c++) {
            float a = s[c][0], b = s[3 + c][0];
            for (int w = 1; w < RD_THREADS / 64; w++) {
                a = fminf(a, s[c][w]);
                b = fmaxf(b, s[3 + c][w]);
            }
            const float r = 0.5f * (a + b);
            d[c] = -1 + (r >= (float)(-L[c] / 2)) + (r >= (float)(L[c] / 2));
        }
        // fragment id of direction d (src/frag/dev.h:19-36)
        cfid[cell] = (d[0] + 2) % 3 + 3 * ((d[1] + 2) % 3) + 9 * ((d[2] + 2) % 3);
    }
}

// new position of every cell: kept cells (fragment 26) first, then
// fragments 0..25, ascending within each: one block; counts per fragment,
// then ranks chunk by chunk from wave ballots
constexpr int CO_THREADS = 1024;
__global__ __launch_bounds__(CO_THREADS) void k_rbc_cell_order(int nc, const int *cfid, int *cnew) {
    __shared__ int start[27];
    __shared__ int wcnt[CO_THREADS / 64][27];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid < 27) start[tid] = 0;
    __syncthreads();
    for (int c = tid; c < nc; c += CO_THREADS) atomicAdd(&start[cfid[c]], 1);
    __syncthreads();
    if (tid == 0) {
        int acc = start[26];
        start[26] = 0;
        for (int f = 0; f < 26; f++) {
            const int k = start[f];
            start[f] = acc;
            acc += k;
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int base = 0; base < nc; base += CO_THREADS) {
        const int c = base + tid;
        const int f = c < nc ? cfid[c] : -1;
        int rank = 0;
        for (int k = 0; k < 27; k++) {
            const unsigned long long m = __ballot(f == k);
            if (lane == 0) wcnt[w][k] = __popcll(m);
            if (f == k) rank = __popcll(m & below);
        }
        __syncthreads();
        if (f >= 0) {
            int before = 0;
            for (int ww = 0; ww < w; ww++) before += wcnt[ww][f];
            cnew[c] = start[f] + before + rank;
        }
        __syncthreads();
        if (tid < 27) {
            int tot = 0;
            for (int ww = 0; ww < CO_THREADS / 64; ww++) tot += wcnt[ww][tid];
            start[tid] += tot;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_rbc_permute(int n, int nv, int Lx, int Ly, int Lz, const int *cfid,
                                                     const int *cnew, const float *src, float *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = i / nv, k = i - c * nv, f = cfid[c];
    const int dx = (f + 2) % 3 - 1, dy = (f / 3 + 2) % 3 - 1, dz = (f / 9 + 2) % 3 - 1;
    const float *p = src + 6 * (long)i;
    float *q = dst + 6 * ((long)cnew[c] * nv + k);
    q[0] = p[0] + (float)(-dx * Lx);
    q[1] = p[1] + (float)(-dy * Ly);
    q[2] = p[2] + (float)(-dz * Lz);
    q[3] = p[3];
    q[4] = p[4];
    q[5] = p[5];
}

}  // namespace

int rbc_redistribute(hipStream_t s, const Box &b, int nc, int nv, float *pp, float *tmp, int *cfid, int *cnew) {
    if (nc == 0) return UDX_OK;
    hipLaunchKernelGGL(k_rbc_cell_fid, dim3(nc), dim3(RD_THREADS), 0, s, nv, b.L[0], b.L[1], b.L[2], pp, cfid);
    UDX_LAUNCHED();
    hipLaunchKernelGGL(k_rbc_cell_order, dim3(1), dim3(CO_THREADS), 0, s, nc, cfid, cnew);
    UDX_LAUNCHED();
    const int n = nc * nv;
    hipLaunchKernelGGL(k_rbc_permute, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, nv, b.L[0], b.L[1], b.L[2], cfid,
                       cnew, pp, tmp);
    UDX_LAUNCHED();
    UDX_HIP(hipMemcpyAsync(pp, tmp, 6 * sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, s));
    return UDX_OK;
}

}  // namespace udx

using namespace udx;

struct udx_rbc {
    udx_ctx *ctx;
    int nv, nt, md, max_cells;
    int *tt, *adj0, *adj1;       // device
    std::vector<int> adj0_h, adj1_h;
    float *ll, *aa;              // device, stress free (or null)
    float l0, a0;
    int stress_free, stressful_set, bending;
    int juelicher;               // Juelicher bending instead of Kantor's
    JulTemplate *jt;             // its edges and scratch, built on first use
    double *av;                  // device, 2 * max_cells
    double *avd, *gpb, *gp;      // adjoint scratch: 4 * max_cells, 5 per vertex block, 5
    float *rd_tmp;               // redistribution scratch: 6 floats per vertex, 2 ints per cell
    int *rd_fid, *rd_new;
};

namespace udx {
udx_ctx *rbc_ctx(const udx_rbc *r) { return r->ctx; }
int rbc_nv(const udx_rbc *r) { return r->nv; }
int rbc_max_cells(const udx_rbc *r) { return r->max_cells; }
}  // namespace udx

extern "C" {

int udx_rbc_fin(udx_rbc *r) {
    if (!r) return UDX_OK;
    dev_free(r->tt);
    dev_free(r->adj0);
    dev_free(r->adj1);
    dev_free(r->ll);
    dev_free(r->aa);
    dev_free(r->av);
    dev_free(r->avd);
    dev_free(r->gpb);
    dev_free(r->gp);
    dev_free(r->rd_tmp);
    dev_free(r->rd_fid);
    dev_free(r->rd_new);
    jul_free(r->jt);
    delete r;
    return UDX_OK;
}

int udx_rbc_ini(udx_ctx *ctx, int nv, int nt, const int *tt, int max_cells, udx_rbc **out) {
    UDX_CHECK(ctx && tt && out && nv > 0 && nt > 0 && max_cells > 0, UDX_EINVAL, "udx_rbc_ini: bad arguments");
    std::vector<int> deg(nv, 0);
    for (int t = 0; t < 3 * nt; t++) {
        UDX_CHECK(tt[t] >= 0 && tt[t] < nv, UDX_EINVAL, "udx_rbc_ini: vertex index %d out of range", tt[t]);
        deg[tt[t]]++;
    }
    int md = 0;
    for (int v = 0; v < nv; v++) md = deg[v] > md ? deg[v] : md;
    udx_rbc *r = new udx_rbc();
    r->ctx = ctx;
    r->nv = nv;
    r->nt = nt;
    r->md = md;
    r->max_cells = max_cells;
    r->bending = 1;
    int s = build_adjacency(md, nt, nv, tt, r->adj0_h, r->adj1_h);
    if (!s) s = dev_alloc(&r->tt, 3 * (size_t)nt);
    if (!s) s = dev_alloc(&r->adj0, (size_t)md * nv);
    if (!s) s = dev_alloc(&r->adj1, (size_t)md * nv);
    if (!s) s = dev_alloc(&r->av, 2 * (size_t)max_cells);
    if (!s && (hipMemcpy(r->tt, tt, sizeof(int) * 3 * (size_t)nt, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(r->adj0, r->adj0_h.data(), sizeof(int) * (size_t)md * nv, hipMemcpyHostToDevice) !=
                   hipSuccess ||
               hipMemcpy(r->adj1, r->adj1_h.data(), sizeof(int) * (size_t)md * nv, hipMemcpyHostToDevice) !=
                   hipSuccess)) {
        set_error("udx_rbc_ini: copy failed");
        s = UDX_EHIP;
    }
    if (s) {
        udx_rbc_fin(r);
        return s;
    }
    *out = r;
    return UDX_OK;
}

int udx_rbc_md(const udx_rbc *r, int *md) {
    UDX_CHECK(r && md, UDX_EINVAL, "udx_rbc_md: null argument");
    *md = r->md;
    return UDX_OK;
}

int udx_rbc_adjacency(const udx_rbc *r, int *adj0, int *adj1) {
    UDX_CHECK(r && adj0 && adj1, UDX_EINVAL, "udx_rbc_adjacency: null argument");
    for (size_t k = 0; k < r->adj0_h.size(); k++) {
        adj0[k] = r->adj0_h[k];
        adj1[k] = r->adj1_h[k];
    }
    return UDX_OK;
}

int udx_rbc_set_stressful(udx_rbc *r, float totArea) {
    UDX_CHECK(r && totArea > 0, UDX_EINVAL, "udx_rbc_set_stressful: bad arguments");
    const float a0 = totArea / r->nt;
    r->a0 = a0;
    r->l0 = (float)std::sqrt(a0 * 4.0 / std::sqrt(3.0));
    r->stress_free = 0;
    r->stressful_set = 1;
    return UDX_OK;
}

int udx_rbc_set_stressfree(udx_rbc *r, const float *rr) {
    UDX_CHECK(r && rr, UDX_EINVAL, "udx_rbc_set_stressfree: null argument");
    const int n = r->md * r->nv;
    std::vector<float> ll(n, 0.f), aa(n, 0.f);
    for (int v = 0; v < r->nv; v++)
        for (int j = 0; j < r->md; j++) {
            const int k = v * r->md + j;
            const int i1 = r->adj0_h[k];
            if (i1 == -1) continue;
            int i2 = r->adj0_h[v * r->md + (j + 1) % r->md];
            if (i2 == -1) i2 = r->adj0_h[v * r->md];
            // edge length: float differences, float sum of squares, double
            // root (src/rbc/shape/imp/main.h:1-25)
            const float *p0 = rr + 3 * v, *p1 = rr + 3 * i1, *p2 = rr + 3 * i2;
            const float d0 = p0[0] - p1[0], d1 = p0[1] - p1[1], d2 = p0[2] - p1[2];
            const float q = d0 * d0 + d1 * d1 + d2 * d2;
            ll[k] = (float)std::sqrt((double)q);
            // Kahan's triangle area in double (src/math/tri/dev.h:38-50)
            auto len = [](const float *a, const float *b) {
                const double x = (double)a[0] - b[0], y = (double)a[1] - b[1], z = (double)a[2] - b[2];
                return std::sqrt(x * x + y * y + z * z);
            };
            double a = len(p0, p1), b = len(p1, p2), c = len(p2, p0), t;
            if (a < b) { t = a; a = b; b = t; }
            if (b < c) { t = b; b = c; c = t; }
            if (a < b) { t = a; a = b; b = t; }
            aa[k] = (float)(std::sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c))) / 4);
        }
    if (!r->ll) UDX_TRY(dev_alloc(&r->ll, (size_t)n));
    if (!r->aa) UDX_TRY(dev_alloc(&r->aa, (size_t)n));
    UDX_HIP(hipMemcpy(r->ll, ll.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    UDX_HIP(hipMemcpy(r->aa, aa.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    r->stress_free = 1;
    return UDX_OK;
}

int udx_rbc_set_bending(udx_rbc *r, int kantor) {
    UDX_CHECK(r, UDX_EINVAL, "udx_rbc_set_bending: null handle");
    r->bending = kantor ? 1 : 0;
    r->juelicher = 0;
    return UDX_OK;
}

int udx_rbc_set_bending_type(udx_rbc *r, int type) {
    UDX_CHECK(r, UDX_EINVAL, "udx_rbc_set_bending_type: null handle");
    UDX_CHECK(type == UDX_BENDING_NONE || type == UDX_BENDING_KANTOR || type == UDX_BENDING_JUELICHER, UDX_EINVAL,
              "udx_rbc_set_bending_type: unknown bending type %d", type);
    if (type == UDX_BENDING_JUELICHER && !r->jt) {
        std::vector<int> tt_h(3 * (size_t)r->nt);
        UDX_HIP(hipMemcpy(tt_h.data(), r->tt, sizeof(int) * tt_h.size(), hipMemcpyDeviceToHost));
        UDX_TRY(jul_build(r->ctx->stream, r->nv, r->nt, tt_h.data(), r->max_cells, &r->jt));
    }
    r->bending = type == UDX_BENDING_KANTOR ? 1 : 0;
    r->juelicher = type == UDX_BENDING_JUELICHER ? 1 : 0;
    return UDX_OK;
}

int udx_rbc_area_volume(udx_rbc *r, int nc, const float *pp, double *av) {
    UDX_CHECK(r && pp, UDX_EINVAL, "udx_rbc_area_volume: null argument");
    UDX_CHECK(nc >= 0 && nc <= r->max_cells, UDX_ECAP, "udx_rbc_area_volume: nc = %d above max_cells = %d", nc,
              r->max_cells);
    if (nc == 0) return UDX_OK;
    hipLaunchKernelGGL(k_rbc_av, dim3(nc), dim3(AV_THREADS), 0, r->ctx->stream, r->nt, r->nv, r->tt, pp,
                       av ? av : r->av);
    UDX_LAUNCHED();
    return UDX_OK;
}

const double *udx_rbc_av(const udx_rbc *r) { return r ? r->av : nullptr; }

int udx_rbc_force_apply(udx_rbc *r, const udx_rbc_params *p, float dt, int nc, const float *pp, float *ff) {
    (void)dt;  // only the membrane's thermal noise uses dt; it is not part of this path
    UDX_CHECK(r && p && pp && ff, UDX_EINVAL, "udx_rbc_force_apply: null argument");
    UDX_CHECK(r->stress_free || r->stressful_set, UDX_ESTATE,
              "udx_rbc_force_apply: call udx_rbc_set_stressful or udx_rbc_set_stressfree first");
    if (nc <= 0) return UDX_OK;
    UDX_TRY(udx_rbc_area_volume(r, nc, pp, r->av));
    RbcK K;
    K.md = r->md;
    K.nv = r->nv;
    K.nc = nc;
    K.adj0 = r->adj0;
    K.adj1 = r->adj1;
    K.ll = r->stress_free ? r->ll : nullptr;
    K.aa = r->stress_free ? r->aa : nullptr;
    K.l0 = r->l0;
    K.a0 = r->a0;
    K.ka = p->ka;
    K.kd = p->kd;
    K.kv = p->kv;
    K.ks = p->ks;
    K.x0 = p->x0;
    K.mpow = p->mpow;
    K.gammaC = p->gammaC;
    K.gammaT = p->gammaT;
    K.totArea = p->totArea;
    K.totVolume = p->totVolume;
    K.kb = p->kb;
    const double phi = (double)p->phi / 180.0 * 3.14159265358979323846;
    K.sphi = std::sin(phi);
    K.cphi = std::cos(phi);
    K.bending = r->bending;
    const int n = nc * r->nv;
    hipLaunchKernelGGL(k_rbc_force, dim3(ceil_div(n, RBC_THREADS)), dim3(RBC_THREADS), 0, r->ctx->stream, K, pp,
                       r->av, ff);
    UDX_LAUNCHED();
    // bending_apply (src/rbc/force/imp/forces.h:54) for the Juelicher model
    if (r->juelicher) UDX_TRY(jul_apply(r->ctx->stream, r->jt, nc, pp, ff));
    return UDX_OK;
}

int udx_rbc_force_adjoint(udx_rbc *r, const udx_rbc_params *p, int nc, const float *pp, const float *g,
                          float *gx, double gp[5]) {
    UDX_CHECK(r && p && pp && g && gx && gp, UDX_EINVAL, "udx_rbc_force_adjoint: null argument");
    UDX_CHECK(r->stress_free || r->stressful_set, UDX_ESTATE,
              "udx_rbc_force_adjoint: call udx_rbc_set_stressful or udx_rbc_set_stressfree first");
    UDX_CHECK(nc >= 0 && nc <= r->max_cells, UDX_ECAP, "udx_rbc_force_adjoint: nc = %d above max_cells = %d", nc,
              r->max_cells);
    UDX_CHECK(!r->juelicher, UDX_EINVAL, "udx_rbc_force_adjoint: the adjoint covers Kantor bending only");
    for (int k = 0; k < 5; k++) gp[k] = 0;
    if (nc == 0) return UDX_OK;
    hipStream_t s = r->ctx->stream;
    if (!r->avd) {
        UDX_TRY(dev_alloc(&r->avd, 4 * (size_t)r->max_cells));
        UDX_TRY(dev_alloc(&r->gpb, 5 * (size_t)ceil_div(r->max_cells * r->nv, RBC_THREADS)));
        UDX_TRY(dev_alloc(&r->gp, 5));
    }
    hipLaunchKernelGGL(k_rbc_av_dual, dim3(nc), dim3(AV_THREADS), 0, s, r->nt, r->nv, r->tt, pp, g, r->avd);
    UDX_LAUNCHED();
    AdjK K;
    K.md = r->md;
    K.nv = r->nv;
    K.nc = nc;
    K.adj0 = r->adj0;
    K.adj1 = r->adj1;
    K.ll = r->stress_free ? r->ll : nullptr;
    K.aa = r->stress_free ? r->aa : nullptr;
    K.l0 = r->l0;
    K.a0 = r->a0;
    K.coef[0] = p->ka;
    K.coef[1] = p->kd;
    K.coef[2] = p->kv;
    K.coef[3] = p->ks;
    K.coef[4] = r->bending ? p->kb : 0.0;
    K.totArea = p->totArea;
    K.totVolume = p->totVolume;
    K.x0 = p->x0;
    K.mpow = p->mpow;
    const double phi = (double)p->phi / 180.0 * 3.14159265358979323846;
    K.sphi = std::sin(phi);
    K.cphi = std::cos(phi);
    const int nb = ceil_div(nc * r->nv, RBC_THREADS);
    hipLaunchKernelGGL(k_rbc_adjoint, dim3(nb), dim3(RBC_THREADS), 0, s, K, pp, g, r->avd, gx, r->gpb);
    UDX_LAUNCHED();
    hipLaunchKernelGGL(k_rbc_gp, dim3(1), dim3(256), 0, s, nb, r->gpb, r->gp);
    UDX_LAUNCHED();
    UDX_HIP(hipMemcpyAsync(gp, r->gp, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    UDX_HIP(hipStreamSynchronize(s));
    return UDX_OK;
}

int udx_rbc_distribute(udx_rbc *r, int Lx, int Ly, int Lz, int nc, float *pp) {
    UDX_CHECK(r && (nc == 0 || pp), UDX_EINVAL, "udx_rbc_distribute: null argument");
    UDX_CHECK(Lx > 0 && Ly > 0 && Lz > 0, UDX_EINVAL, "udx_rbc_distribute: bad box");
    UDX_CHECK(nc >= 0 && nc <= r->max_cells, UDX_ECAP, "udx_rbc_distribute: nc = %d above max_cells = %d", nc,
              r->max_cells);
    if (nc == 0) return UDX_OK;
    if (!r->rd_tmp) {
        UDX_TRY(dev_alloc(&r->rd_tmp, 6 * (size_t)r->max_cells * r->nv));
        UDX_TRY(dev_alloc(&r->rd_fid, (size_t)r->max_cells));
        UDX_TRY(dev_alloc(&r->rd_new, (size_t)r->max_cells));
    }
    Box b{{Lx, Ly, Lz}};
    return rbc_redistribute(r->ctx->stream, b, nc, r->nv, pp, r->rd_tmp, r->rd_fid, r->rd_new);
}

int udx_rbc_stretch_apply(udx_ctx *ctx, int nc, int nv, const float *f, float *ff) {
    UDX_CHECK(ctx && f && ff && nv > 0, UDX_EINVAL, "udx_rbc_stretch_apply: bad arguments");
    const int n = nc * nv;
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_rbc_stretch, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n, nv, f, ff);
    UDX_LAUNCHED();
    return UDX_OK;
}

}  // extern "C"
