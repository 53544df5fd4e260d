/*
 * CPU oracle of the RBC membrane path: mesh adjacency, per-cell area and
 * volume, membrane forces (area, volume, WLC + POW springs, viscosity) and
 * Kantor bending, plus the membrane energy whose gradient the forces are.
 *
 * TEST INFRASTRUCTURE ONLY (see udx_oracle.h).  The reference evaluates the
 * membrane terms in double from float positions and adds each (vertex, ring
 * slot) contribution to the float force by an atomic; here each vertex sums
 * its slots in double, in slot order, and the caller rounds once.
 */
#include "udx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* ---- small double vector helpers ---------------------------------------- */
typedef struct { double x, y, z; } d3;
static d3 mk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 sub3(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross3(d3 a, d3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static d3 scal3(double s, d3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static d3 add3(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 pos(const float *pp, int i) { return mk(pp[6 * i], pp[6 * i + 1], pp[6 * i + 2]); }
static d3 vel(const float *pp, int i) { return mk(pp[6 * i + 3], pp[6 * i + 4], pp[6 * i + 5]); }

/* ======================================================================== */
/* mesh: max degree (src/io/mesh_read/imp/main.h:109-129) and the ring      */
/* adjacency (src/rbc/adj/imp/ini.h:1-45, src/algo/edg/imp/main.h)          */
/* ======================================================================== */
int orc_mesh_md(int nv, int nt, const int *tt) {
    int *d = (int *)calloc((size_t)nv, sizeof(int)), m = 0;
    for (int t = 0; t < 3 * nt; t++) d[tt[t]]++;
    for (int i = 0; i < nv; i++) if (d[i] > m) m = d[i];
    free(d);
    return m;
}

/* slot table of vertex f: key x -> value y (first free slot) */
static void edg_set(int md, int f, int x, int y, int *hx, int *hy) {
    int j = f * md;
    while (hx[j] != -1) j++;
    hx[j] = x;
    hy[j] = y;
}

static int edg_get(int md, int f, int x, const int *hx, const int *hy) {
    int j = f * md;
    while (hx[j] != x) j++;
    return hy[j];
}

int orc_adj_ini(int md, int nt, int nv, const int *tt, int *adj0, int *adj1) {
    int n = md * nv;
    int *hx = (int *)malloc(sizeof(int) * (size_t)n), *hy = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) adj0[i] = adj1[i] = hx[i] = -1;
    /* directed edge f0 -> f1 of a face has third vertex f2 */
    for (int t = 0; t < nt; t++) {
        int f0 = tt[3 * t], f1 = tt[3 * t + 1], f2 = tt[3 * t + 2];
        edg_set(md, f0, f1, f2, hx, hy);
        edg_set(md, f1, f2, f0, hx, hy);
        edg_set(md, f2, f0, f1, hx, hy);
    }
    for (int v = 0; v < nv; v++) {
        /* the ring of v, starting from its smallest neighbour; adj1 holds
         * the vertex across the edge (ring[k], ring[k+1]) from v */
        int lo = v * md, mi = hx[lo];
        for (int k = lo + 1; k < lo + md && hx[k] != -1; k++)
            if (hx[k] < mi) mi = hx[k];
        int c = mi, k = lo;
        do {
            int c0 = c;
            c = edg_get(md, v, c0, hx, hy);
            adj0[k] = c0;
            adj1[k] = edg_get(md, c, c0, hx, hy);
            k++;
        } while (c != mi);
    }
    free(hx);
    free(hy);
    return 0;
}

/* (vertex, slot) i -> i0..i4 and the cell (src/rbc/adj/dev.h:2-42) */
int orc_adj_map(int md, int nv, const int *adj0, const int *adj1, int i, int m[6]) {
    int i0 = i / md, j = i % md, k = (i0 % nv) * md;
    int i1 = adj0[k + j], i2, i3, i4;
    if (i1 == -1) return 0;
    i2 = adj0[k + (j + 1) % md];
    if (i2 == -1) {
        i2 = adj0[k];
        i3 = adj0[k + 1];
    } else {
        i3 = adj0[k + (j + 2) % md];
        if (i3 == -1) i3 = adj0[k];
    }
    i4 = adj1[k + j];
    int cell = i0 / nv, off = cell * nv;
    m[0] = i0; m[1] = i1 + off; m[2] = i2 + off; m[3] = i3 + off; m[4] = i4 + off; m[5] = cell;
    return 1;
}

/* ======================================================================== */
/* area and volume per cell, FP64 (src/rbc/force/area_volume/dev/main.h)    */
/* ======================================================================== */
void orc_rbc_area_volume(int nt, int nv, int nc, const int *tt, const float *pp, double *av) {
    for (int c = 0; c < nc; c++) {
        double A = 0, V = 0;
        for (int t = 0; t < nt; t++) {
            d3 r0 = pos(pp, tt[3 * t] + c * nv), r1 = pos(pp, tt[3 * t + 1] + c * nv),
               r2 = pos(pp, tt[3 * t + 2] + c * nv);
            d3 n = cross3(sub3(r1, r0), sub3(r2, r0));
            A += 0.5 * sqrt(dot3(n, n));
            V += dot3(r0, cross3(r1, r2)) / 6.0;
        }
        av[2 * c] = A;
        av[2 * c + 1] = V;
    }
}

/* ======================================================================== */
/* stress-free references per (vertex, slot): edge length and triangle      */
/* area of the reference shape (src/rbc/shape/imp/main.h:1-60,               */
/* kahan_area of src/math/tri/dev.h:38-50)                                  */
/* ======================================================================== */
static double vabs3(const float *a, const float *b) {
    float d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return sqrt((double)(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
}

static double dabs3(const double *a, const double *b) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

static double kahan_area(const float *f0, const float *f1, const float *f2) {
    double r0[3] = {f0[0], f0[1], f0[2]}, r1[3] = {f1[0], f1[1], f1[2]}, r2[3] = {f2[0], f2[1], f2[2]};
    double a = dabs3(r0, r1), b = dabs3(r1, r2), c = dabs3(r2, r0), t;
    /* a >= b >= c */
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    return sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c))) / 4;
}

void orc_rbc_shape(int md, int nv, const int *adj0, const int *adj1, const float *rr, float *ll, float *aa) {
    int m[6];
    for (int i = 0; i < md * nv; i++) {
        ll[i] = aa[i] = 0;
        if (!orc_adj_map(md, nv, adj0, adj1, i, m)) continue;
        ll[i] = (float)vabs3(rr + 3 * m[0], rr + 3 * m[1]);
        aa[i] = (float)kahan_area(rr + 3 * m[0], rr + 3 * m[1], rr + 3 * m[2]);
    }
}

/* equilibrium spring length and triangle area of the stress-full model
 * (src/rbc/force/imp/main.h:50-59) */
void orc_rbc_stressful(int nt, float totArea, float *l0, float *a0) {
    float a = totArea / nt;
    *a0 = a;
    *l0 = (float)sqrt(a * 4.0 / sqrt(3.0));
}

/* ======================================================================== */
/* membrane force terms (src/rbc/force/dev/common.h:1-108)                  */
/* ======================================================================== */
static double wlc0(double r) { return (4 * r * r - 9 * r + 6) / (4 * (r - 1) * (r - 1)); }
static double wlc(double lmax, double ks, double r) { return ks / lmax * wlc0(r / lmax); }

/* force on r1 of the triangle (r1, r2, r3): local and global area, volume
 * and the spring r1 - r2 */
static d3 ftri(const orc_rbc_par *p, d3 r1, d3 r2, d3 r3, double l0, double a0, double A, double V) {
    d3 x21 = sub3(r2, r1), x32 = sub3(r3, r2), x31 = sub3(r3, r1);
    /* area */
    d3 nn = cross3(x21, x31), f = cross3(nn, x32);
    double a = 0.5 * sqrt(dot3(nn, nn));
    double fA = -(double)p->ka * (A - p->totArea) / (4 * (double)p->totArea * a);
    double fa = -(double)p->kd * (a - a0) / (4 * a0 * a);
    f = scal3(fA + fa, f);
    /* volume */
    d3 fv = scal3((double)p->kv * (V - p->totVolume) / (6 * (double)p->totVolume), cross3(r3, r2));
    f = add3(f, fv);
    /* WLC + POW spring; r in float as the reference's sqrtf */
    double ks = p->ks, m = p->mpow, x0 = p->x0;
    double r = sqrtf((float)dot3(x21, x21));
    double lmax = l0 / x0;
    if (r >= lmax) r = 0.99 * lmax;
    double fwlc = wlc(lmax, ks, r);
    double fpow = -wlc(lmax, ks, l0) * powf((float)l0, (float)(m + 1)) / powf((float)r, (float)(m + 1));
    return add3(f, scal3(fwlc + fpow, x21));
}

static d3 fvisc(const orc_rbc_par *p, d3 r1, d3 r2, d3 u1, d3 u2) {
    d3 du = sub3(u2, u1), dr = sub3(r1, r2);
    double fac = dot3(du, dr) / dot3(dr, dr);
    return add3(scal3(p->gammaT, du), scal3(p->gammaC * fac, dr));
}

void orc_rbc_forces(const orc_rbc_par *p, int md, int nv, int nc, const int *adj0, const int *adj1,
                    const float *pp, const double *av, const float *ll, const float *aa, float l0, float a0,
                    double *ff) {
    int m[6];
    for (int v = 0; v < nc * nv; v++) {
        d3 f = mk(0, 0, 0);
        for (int j = 0; j < md; j++) {
            int i = v * md + j;
            if (!orc_adj_map(md, nv, adj0, adj1, i, m)) continue;
            int s = i % (md * nv);
            double L0 = ll ? ll[s] : l0, A0 = aa ? aa[s] : a0;
            d3 r0 = pos(pp, m[0]), r1 = pos(pp, m[1]), r2 = pos(pp, m[2]);
            f = add3(f, ftri(p, r0, r1, r2, L0, A0, av[2 * m[5]], av[2 * m[5] + 1]));
            f = add3(f, fvisc(p, r0, r1, vel(pp, m[0]), vel(pp, m[1])));
        }
        ff[3 * v] += f.x;
        ff[3 * v + 1] += f.y;
        ff[3 * v + 2] += f.z;
    }
}

/* ======================================================================== */
/* Kantor bending (src/rbc/force/bending/kantor/dev/main.h:23-66,           */
/* src/mesh/force/kantor0/dev.h:28-120)                                     */
/* ======================================================================== */
/* gradient of kb (1 - cos(theta - phi)) of the dihedral (a, b, c | b, c, d)
 * with respect to a (which == 1) or to b (which == 2), negated */
static d3 dih(int which, double sphi, double cphi, double kb, d3 a, d3 b, d3 c, d3 d) {
    d3 ab = sub3(a, b), ac = sub3(a, c), cd = sub3(c, d), bd = sub3(b, d), bc = sub3(b, c), da = sub3(d, a);
    d3 k = cross3(ab, ac), n = cross3(bd, bc);
    double kk = dot3(k, k), nn = dot3(n, n), kn = dot3(k, n);
    double over_k = 1.0 / sqrt(kk), over_n = 1.0 / sqrt(nn);
    double cost = kn * over_k * over_n;
    double s2 = 1.0 - cost * cost;
    double over_sin = 1.0 / sqrt(s2 > 1e-6 ? s2 : 1e-6);
    over_sin = copysign(over_sin, dot3(sub3(k, n), da));
    double beta = cphi * kb - cost * (sphi * kb) * over_sin;
    double b11 = -beta * cost * over_k * over_k, b12 = beta * over_k * over_n;
    if (which == 1) {
        d3 r32 = sub3(c, b);
        return add3(scal3(b11, cross3(k, r32)), scal3(b12, cross3(n, r32)));
    }
    double b22 = -beta * cost * over_n * over_n;
    d3 f = scal3(b11, cross3(k, ac));
    f = add3(f, scal3(b12, add3(cross3(k, cd), cross3(n, ac))));
    return add3(f, scal3(b22, cross3(n, cd)));
}

void orc_rbc_bending(const orc_rbc_par *p, int md, int nv, int nc, const int *adj0, const int *adj1,
                     const float *pp, double *ff) {
    int m[6];
    double phi = p->phi / 180.0 * ORC_PI, kb = p->kb;
    double sphi = sin(phi), cphi = cos(phi);
    for (int v = 0; v < nc * nv; v++) {
        d3 f = mk(0, 0, 0);
        for (int j = 0; j < md; j++) {
            int i = v * md + j;
            if (!orc_adj_map(md, nv, adj0, adj1, i, m)) continue;
            d3 r0 = pos(pp, m[0]), r1 = pos(pp, m[1]), r2 = pos(pp, m[2]), r3 = pos(pp, m[3]), r4 = pos(pp, m[4]);
            f = add3(f, dih(1, sphi, cphi, kb, r0, r2, r1, r4));
            f = add3(f, dih(2, sphi, cphi, kb, r1, r0, r2, r3));
        }
        ff[3 * v] += f.x;
        ff[3 * v + 1] += f.y;
        ff[3 * v + 2] += f.z;
    }
}

/* ======================================================================== */
/* energy whose negative gradient the forces above are (positions in        */
/* double; springs in closed form, wlc: ks lmax/4 (3x^2 - 2x^3)/(1 - x),    */
/* pow: C r^(1-m)/(m-1))                                                    */
/* ======================================================================== */
static double dihedral_theta(d3 a, d3 b, d3 c, d3 d) {
    d3 k = cross3(sub3(a, b), sub3(a, c)), n = cross3(sub3(b, d), sub3(b, c));
    double cost = dot3(k, n) / sqrt(dot3(k, k) * dot3(n, n));
    double sint = sqrt(fmax(1.0 - cost * cost, 0.0));
    sint = copysign(sint, dot3(sub3(k, n), sub3(d, a)));
    return atan2(sint, cost);
}

double orc_rbc_energy(const orc_rbc_par *p, int md, int nv, int nc, int nt, const int *tt, const int *adj0,
                      const int *adj1, const double *r, const float *ll, const float *aa, float l0, float a0,
                      int terms) {
    int m[6];
    double E = 0;
    double phi = p->phi / 180.0 * ORC_PI;
#define R(i) mk(r[3 * (i)], r[3 * (i) + 1], r[3 * (i) + 2])
    for (int c = 0; c < nc; c++) {
        double A = 0, V = 0;
        for (int t = 0; t < nt; t++) {
            d3 r0 = R(tt[3 * t] + c * nv), r1 = R(tt[3 * t + 1] + c * nv), r2 = R(tt[3 * t + 2] + c * nv);
            d3 n = cross3(sub3(r1, r0), sub3(r2, r0));
            A += 0.5 * sqrt(dot3(n, n));
            V += dot3(r0, cross3(r1, r2)) / 6.0;
        }
        if (terms & 1) E += p->ka * (A - p->totArea) * (A - p->totArea) / (2.0 * p->totArea);
        if (terms & 2) E += p->kv * (V - p->totVolume) * (V - p->totVolume) / (2.0 * p->totVolume);
    }
    for (int v = 0; v < nc * nv; v++)
        for (int j = 0; j < md; j++) {
            int i = v * md + j;
            if (!orc_adj_map(md, nv, adj0, adj1, i, m)) continue;
            int s = i % (md * nv);
            double L0 = ll ? ll[s] : l0, A0 = aa ? aa[s] : a0;
            d3 r0 = R(m[0]), r1 = R(m[1]), r2 = R(m[2]);
            if (terms & 4) { /* local area: each triangle from its 3 corners */
                d3 n = cross3(sub3(r1, r0), sub3(r2, r0));
                double a = 0.5 * sqrt(dot3(n, n));
                E += p->kd * (a - A0) * (a - A0) / (2.0 * A0) / 3.0;
            }
            if (terms & 8) { /* springs: each edge from both ends */
                double lmax = L0 / p->x0, x = sqrt(dot3(sub3(r1, r0), sub3(r1, r0))) / lmax, mm = p->mpow;
                double C = wlc(lmax, p->ks, L0) * pow(L0, mm + 1);
                double e = p->ks * lmax / 4.0 * (3 * x * x - 2 * x * x * x) / (1 - x) +
                           C * pow(x * lmax, 1 - mm) / (mm - 1);
                E += e / 2.0;
            }
            if (terms & 16) { /* bending: each dihedral from both wings */
                double th = dihedral_theta(r0, r2, r1, R(m[4]));
                E += p->kb * (1 - cos(th - phi)) / 2.0;
            }
        }
#undef R
    return E;
}

/* stretching: a constant force per vertex of every cell
 * (src/rbc/stretch/dev/main.h:1-17) */
void orc_rbc_stretch(int nc, int nv, const float *f, float *ff) {
    for (int i = 0; i < nc * nv; i++) {
        int k = i % nv;
        ff[3 * i] += f[3 * k];
        ff[3 * i + 1] += f[3 * k + 1];
        ff[3 * i + 2] += f[3 * k + 2];
    }
}
