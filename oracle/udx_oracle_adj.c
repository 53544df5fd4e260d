/*
 * CPU oracle (FP64) of the membrane adjoint.
 *
 * TEST INFRASTRUCTURE ONLY (see udx_oracle.h).
 *
 * The conservative membrane force is F = -grad E with
 * E = ka E_A + kd E_a + kv E_V + ks E_s + kb E_b, so F is linear in each
 * coefficient and its Jacobian J = dF/dx = -Hess E is symmetric: the
 * vector-Jacobian product J^T g equals the Jacobian-vector product J g, the
 * derivative of F along g.  Here J g is computed by forward-mode dual
 * numbers through the same formulas as orc_rbc_forces / orc_rbc_bending
 * (the global area and volume carry their own tangents), together with the
 * unit-coefficient term forces F_k whose dot products with g are
 * dL/dk = g . F_k.  The viscous term is not part of it (zero relative
 * velocities); the spring length is the double root (the production force
 * rounds it to float).  Positions are double (3 per vertex), so with g = 0
 * the term forces are also an FP64 evaluation of the membrane force.
 */
#include "udx_oracle.h"

#include <math.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

typedef struct { double v, d; } dd;
static dd D(double v, double d) { dd r = {v, d}; return r; }
static dd C(double v) { return D(v, 0); }
static dd dadd(dd a, dd b) { return D(a.v + b.v, a.d + b.d); }
static dd dsub(dd a, dd b) { return D(a.v - b.v, a.d - b.d); }
static dd dmul(dd a, dd b) { return D(a.v * b.v, a.d * b.v + a.v * b.d); }
static dd ddiv(dd a, dd b) { return D(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
static dd dsqrt(dd a) { double s = sqrt(a.v); return D(s, a.d / (2 * s)); }
static dd dpowc(dd a, double c) { double p = pow(a.v, c); return D(p, c * p / a.v * a.d); }

typedef struct { dd x, y, z; } v3;
static v3 V(dd x, dd y, dd z) { v3 r = {x, y, z}; return r; }
static v3 vsub(v3 a, v3 b) { return V(dsub(a.x, b.x), dsub(a.y, b.y), dsub(a.z, b.z)); }
static v3 vadd(v3 a, v3 b) { return V(dadd(a.x, b.x), dadd(a.y, b.y), dadd(a.z, b.z)); }
static v3 vscal(dd s, v3 a) { return V(dmul(s, a.x), dmul(s, a.y), dmul(s, a.z)); }
static dd vdot(v3 a, v3 b) { return dadd(dadd(dmul(a.x, b.x), dmul(a.y, b.y)), dmul(a.z, b.z)); }
static v3 vcross(v3 a, v3 b) {
    return V(dsub(dmul(a.y, b.z), dmul(a.z, b.y)), dsub(dmul(a.z, b.x), dmul(a.x, b.z)),
             dsub(dmul(a.x, b.y), dmul(a.y, b.x)));
}
static v3 vzero(void) { return V(C(0), C(0), C(0)); }

static v3 dpos(const double *x, const double *g, int i) {
    return V(D(x[3 * i], g[3 * i]), D(x[3 * i + 1], g[3 * i + 1]), D(x[3 * i + 2], g[3 * i + 2]));
}

/* per-cell area and volume with their tangents */
static void av_dual(int nt, int nv, int c, const int *tt, const double *pp, const double *g, dd *A, dd *Vol) {
    dd a = C(0), vol = C(0);
    for (int t = 0; t < nt; t++) {
        v3 r0 = dpos(pp, g, tt[3 * t] + c * nv), r1 = dpos(pp, g, tt[3 * t + 1] + c * nv),
           r2 = dpos(pp, g, tt[3 * t + 2] + c * nv);
        v3 n = vcross(vsub(r1, r0), vsub(r2, r0));
        a = dadd(a, dmul(C(0.5), dsqrt(vdot(n, n))));
        vol = dadd(vol, ddiv(vdot(r0, vcross(r1, r2)), C(6.0)));
    }
    *A = a;
    *Vol = vol;
}

/* the five unit-coefficient terms of one triangle slot: 0 global area,
 * 1 local area, 2 volume, 3 spring */
static void tri_terms(const orc_rbc_par *p, v3 r1, v3 r2, v3 r3, double l0, double a0, dd A, dd Vol, v3 out[4]) {
    v3 x21 = vsub(r2, r1), x32 = vsub(r3, r2), x31 = vsub(r3, r1);
    v3 nn = vcross(x21, x31), f = vcross(nn, x32);
    dd a = dmul(C(0.5), dsqrt(vdot(nn, nn)));
    dd fA = ddiv(dmul(C(-1.0), dsub(A, C(p->totArea))), dmul(C(4 * (double)p->totArea), a));
    dd fa = ddiv(dmul(C(-1.0), dsub(a, C(a0))), dmul(C(4 * a0), a));
    out[0] = vscal(fA, f);
    out[1] = vscal(fa, f);
    out[2] = vscal(ddiv(dsub(Vol, C(p->totVolume)), C(6 * (double)p->totVolume)), vcross(r3, r2));
    /* spring with ks = 1 */
    double x0 = p->x0, m = p->mpow, lmax = l0 / x0;
    dd r = dsqrt(vdot(x21, x21));
    if (r.v >= lmax) r = C(0.99 * lmax);
    dd xr = ddiv(r, C(lmax));
    dd w = ddiv(dadd(dsub(dmul(C(4), dmul(xr, xr)), dmul(C(9), xr)), C(6)),
                dmul(C(4), dmul(dsub(xr, C(1)), dsub(xr, C(1)))));
    dd fwlc = dmul(C(1.0 / lmax), w);
    double xl = l0 / lmax, wl = (4 * xl * xl - 9 * xl + 6) / (4 * (xl - 1) * (xl - 1)) / lmax;
    dd fpow = ddiv(C(-wl * pow(l0, m + 1)), dpowc(r, m + 1));
    out[3] = vscal(dadd(fwlc, fpow), x21);
}

/* one Kantor dihedral with kb = 1: force on a (which 1) or b (which 2) */
static v3 dih_dual(int which, double sphi, double cphi, v3 a, v3 b, v3 c, v3 d) {
    v3 ab = vsub(a, b), ac = vsub(a, c), cd = vsub(c, d), bd = vsub(b, d), bc = vsub(b, c), da = vsub(d, a);
    v3 k = vcross(ab, ac), n = vcross(bd, bc);
    dd over_k = ddiv(C(1), dsqrt(vdot(k, k))), over_n = ddiv(C(1), dsqrt(vdot(n, n)));
    dd cost = dmul(dmul(vdot(k, n), over_k), over_n);
    dd s2 = dsub(C(1), dmul(cost, cost));
    dd over_sin = ddiv(C(1), dsqrt(s2.v > 1e-6 ? s2 : C(1e-6)));
    if (vdot(vsub(k, n), da).v < 0) over_sin = D(-over_sin.v, -over_sin.d);
    dd beta = dsub(C(cphi), dmul(dmul(cost, C(sphi)), over_sin));
    dd b11 = dmul(dmul(dmul(C(-1), beta), cost), dmul(over_k, over_k));
    dd b12 = dmul(beta, dmul(over_k, over_n));
    if (which == 1) {
        v3 r32 = vsub(c, b);
        return vadd(vscal(b11, vcross(k, r32)), vscal(b12, vcross(n, r32)));
    }
    dd b22 = dmul(dmul(dmul(C(-1), beta), cost), dmul(over_n, over_n));
    v3 f = vscal(b11, vcross(k, ac));
    f = vadd(f, vscal(b12, vadd(vcross(k, cd), vcross(n, ac))));
    return vadd(f, vscal(b22, vcross(n, cd)));
}

void orc_rbc_adjoint(const orc_rbc_par *p, int md, int nv, int nc, int nt, const int *tt, const int *adj0,
                     const int *adj1, const double *pp, const float *ll, const float *aa, float l0, float a0,
                     const double *g, double *gx, double *terms, double gp[5]) {
    int m[6];
    const double coef[5] = {p->ka, p->kd, p->kv, p->ks, p->kb};
    const double phi = p->phi / 180.0 * ORC_PI, sphi = sin(phi), cphi = cos(phi);
    for (int k = 0; k < 5; k++) gp[k] = 0;
    for (int c = 0; c < nc; c++) {
        dd A, Vol;
        av_dual(nt, nv, c, tt, pp, g, &A, &Vol);
        for (int v = c * nv; v < (c + 1) * nv; v++) {
            v3 t[5] = {vzero(), vzero(), vzero(), vzero(), vzero()};
            for (int j = 0; j < md; j++) {
                int i = v * md + j;
                if (!orc_adj_map(md, nv, adj0, adj1, i, m)) continue;
                int s = i % (md * nv);
                double L0 = ll ? ll[s] : l0, A0 = aa ? aa[s] : a0;
                v3 r0 = dpos(pp, g, m[0]), r1 = dpos(pp, g, m[1]), r2 = dpos(pp, g, m[2]);
                v3 f4[4];
                tri_terms(p, r0, r1, r2, L0, A0, A, Vol, f4);
                for (int k = 0; k < 4; k++) t[k] = vadd(t[k], f4[k]);
                v3 r3 = dpos(pp, g, m[3]), r4 = dpos(pp, g, m[4]);
                t[4] = vadd(t[4], dih_dual(1, sphi, cphi, r0, r2, r1, r4));
                t[4] = vadd(t[4], dih_dual(2, sphi, cphi, r1, r0, r2, r3));
            }
            double dx = 0, dy = 0, dz = 0;
            for (int k = 0; k < 5; k++) {
                dx += coef[k] * t[k].x.d;
                dy += coef[k] * t[k].y.d;
                dz += coef[k] * t[k].z.d;
                gp[k] += g[3 * v] * t[k].x.v + g[3 * v + 1] * t[k].y.v + g[3 * v + 2] * t[k].z.v;
                if (terms) {
                    double *o = terms + ((size_t)k * nc * nv + v) * 3;
                    o[0] = t[k].x.v; o[1] = t[k].y.v; o[2] = t[k].z.v;
                }
            }
            gx[3 * v] = dx;
            gx[3 * v + 1] = dy;
            gx[3 * v + 2] = dz;
        }
    }
}
