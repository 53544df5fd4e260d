/*
 * CPU oracle of the Juelicher bending forces (src/rbc/force/bending/juelicher).
 *
 * TEST INFRASTRUCTURE ONLY (see udx_oracle.h).
 *
 * Per cell, with vertex areas A_v (a third of each incident triangle's area),
 * edge lengths l_e, dihedral angles theta_e and L_v = sum over the edges at v
 * of l_e theta_e, kb = 1 and H0 = -1/2 as the reference hard-codes them
 * (src/rbc/force/bending/juelicher/imp/forces.h:49-53) and
 * c_e = (L_j / (4 A_j) - H0) + (L_k / (4 A_k) - H0):
 *   f   = -sum_e c_e (theta_e grad l_e + l_e D_e)
 *         + sum_t sum_{v in t} (L_v^2 / (8 A_v^2) - 2 H0^2) / 3 grad_v area_t,
 *   fad = -(C/4) sum_e theta_e grad l_e - (C/2) sum_e l_e D_e,
 *   C = (L_tot / 4 - H0 A_tot) (4 kad pi / A_tot),  kad = 2 kb / pi,
 * all as the reference writes them (dev/main.h:283-400): D_e, its dihedral
 * gradient formula, is minus the gradient of its angle, and each triangle's
 * area gradient carries the coefficient of the vertex it moves, so f is not
 * minus the gradient of E = sum_v 2 A_v (L_v / (4 A_v) - H0)^2.  The
 * geometry is double (Kahan's triangle area, the dihedral of
 * src/math/tri/dev.h), every term rounds to float once and accumulates in
 * float.
 *
 * The reference accumulates with float atomics; the order here is defined
 * (and the HIP path follows it): per vertex, the area and L terms in
 * ascending triangle / edge order; the cell totals in ascending vertex order;
 * the force terms edge-length part (ascending edge), dihedral part
 * (ascending edge), area part (ascending triangle); fad edge-length part then
 * dihedral part; the vertex force += f + fad.  Edges are numbered by their
 * first appearance walking triangles in order, half-edges x->y, y->z, z->x;
 * the first half-edge j->k of an edge gives (i, j, k, l) with i the third
 * vertex of its triangle and l the third vertex of the triangle of k->j.
 */
#include "udx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define JUL_PI 3.141592653589793

int orc_jul_edges(int nv, int nt, const int *tt, int *dih) {
    /* half-edge a->b -> (triangle, third vertex) */
    int ne = 0;
    int *seen = (int *)malloc(sizeof(int) * 3 * (size_t)nt);
    for (int k = 0; k < 3 * nt; k++) seen[k] = 0;
    for (int t = 0; t < nt; t++)
        for (int q = 0; q < 3; q++) {
            if (seen[3 * t + q]) continue;
            const int j = tt[3 * t + q], k = tt[3 * t + (q + 1) % 3], i = tt[3 * t + (q + 2) % 3];
            int l = -1;
            for (int u = 0; u < nt && l < 0; u++)
                for (int w = 0; w < 3; w++)
                    if (tt[3 * u + w] == k && tt[3 * u + (w + 1) % 3] == j) {
                        l = tt[3 * u + (w + 2) % 3];
                        seen[3 * u + w] = 1;
                        break;
                    }
            seen[3 * t + q] = 1;
            dih[4 * ne] = i;
            dih[4 * ne + 1] = j;
            dih[4 * ne + 2] = k;
            dih[4 * ne + 3] = l;
            ne++;
        }
    free(seen);
    (void)nv;
    return ne;
}

/* ---- geometry (src/math/tri/dev.h, juelicher/dev/main.h) --------------- */
static double vabs3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static void vsub(const double a[3], const double b[3], double c[3]) {
    c[0] = a[0] - b[0];
    c[1] = a[1] - b[1];
    c[2] = a[2] - b[2];
}
static double vdot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void vcross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - b[1] * a[2];
    c[1] = b[0] * a[2] - a[0] * b[2];
    c[2] = a[0] * b[1] - b[0] * a[1];
}

static double kahan_area(const double r0[3], const double r1[3], const double r2[3]) {
    double u[3], a, b, c, t;
    vsub(r0, r1, u);
    a = vabs3(u);
    vsub(r1, r2, u);
    b = vabs3(u);
    vsub(r2, r0, u);
    c = vabs3(u);
    /* c <= b <= a */
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    return sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c))) / 4;
}

static double orient3d(const double a[3], const double b[3], const double c[3], const double d[3]) {
    const double ax = a[0] - d[0], ay = a[1] - d[1], az = a[2] - d[2];
    const double bx = b[0] - d[0], by = b[1] - d[1], bz = b[2] - d[2];
    const double cx = c[0] - d[0], cy = c[1] - d[1], cz = c[2] - d[2];
    return ax * (by * cz - bz * cy) + bx * (cy * az - cz * ay) + cx * (ay * bz - az * by);
}

/* (a - c) x (b - c) by the 2-D orientations */
static void ac_bc_cross(const double a[3], const double b[3], const double c[3], double r[3]) {
    r[0] = (a[1] - c[1]) * (b[2] - c[2]) - (a[2] - c[2]) * (b[1] - c[1]);
    r[1] = (a[2] - c[2]) * (b[0] - c[0]) - (a[0] - c[0]) * (b[2] - c[2]);
    r[2] = (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0]);
}

double orc_jul_theta(const double a[3], const double b[3], const double c[3], const double d[3]) {
    double k[3], l[3];
    ac_bc_cross(b, c, a, k);
    ac_bc_cross(c, b, d, l);
    const double bcx = b[0] - c[0], bcy = b[1] - c[1], bcz = b[2] - c[2];
    const double bc = sqrt(bcx * bcx + bcy * bcy + bcz * bcz);
    const double y = -orient3d(a, b, c, d) * bc;
    const double x = k[0] * l[0] + k[1] * l[1] + k[2] * l[2];
    return -atan2(y, x);
}

static int small(double s) { return !(s > 1e-12 || s < -1e-12); }

static void vnorm(const double a[3], double b[3]) {
    const double s = vabs3(a);
    if (!small(s)) {
        b[0] = (1 / s) * a[0];
        b[1] = (1 / s) * a[1];
        b[2] = (1 / s) * a[2];
    } else {
        b[0] = a[0];
        b[1] = a[1];
        b[2] = a[2];
    }
}

static void tri_normal(const double a[3], const double b[3], const double c[3], double e[3]) {
    double u[3], v[3], n[3];
    vsub(b, a, u);
    vsub(c, a, v);
    vcross(u, v, n);
    vnorm(n, e);
}

static double tri_cot(const double a[3], const double b[3], const double c[3]) {
    double u[3], v[3];
    const double y = 2 * kahan_area(a, b, c);
    vsub(a, b, u);
    vsub(c, b, v);
    return vdot(u, v) / y;
}

/* gradients of the dihedral angle (dev/main.h:104-128) */
static void ddih(const double a[3], const double b[3], const double c[3], const double d[3], double g[4][3]) {
    double n[3], k[3];
    tri_normal(a, b, c, n);
    tri_normal(c, b, d, k);
    double u[3];
    vsub(b, c, u);
    const double e = vabs3(u);
    const double An = kahan_area(a, b, c), Ak = kahan_area(c, b, d);
    const double sa = e / (2 * An), sd = e / (2 * Ak);
    const double cn = tri_cot(b, c, a), bn = tri_cot(a, b, c), bk = tri_cot(c, b, d), ck = tri_cot(d, c, b);
    for (int x = 0; x < 3; x++) {
        g[0][x] = sa * n[x];
        g[3][x] = sd * k[x];
        g[2][x] = (-bn / e) * n[x] + (-bk / e) * k[x];
        g[1][x] = (-cn / e) * n[x] + (-ck / e) * k[x];
    }
}

/* gradient of |c - b| at b and c (dev/main.h:80-87) */
static void dedg(const double b[3], const double c[3], double db[3], double dc[3]) {
    double u[3], e[3];
    vsub(c, b, u);
    vnorm(u, e);
    for (int x = 0; x < 3; x++) {
        dc[x] = e[x];
        db[x] = -e[x];
    }
}

/* gradient of the triangle area at a, b, c (dev/main.h:136-145) */
static void dtri_area(const double a[3], const double b[3], const double c[3], double g[3][3]) {
    double n[3], n2[3], ab[3], bc[3], ca[3];
    tri_normal(a, b, c, n);
    for (int x = 0; x < 3; x++) n2[x] = 0.5 * n[x];
    vsub(b, a, ab);
    vsub(c, b, bc);
    vsub(a, c, ca);
    vcross(n2, bc, g[0]);
    vcross(n2, ca, g[1]);
    vcross(n2, ab, g[2]);
}

static void getv(const float *pp, int v, double r[3]) {
    r[0] = pp[6 * (size_t)v];
    r[1] = pp[6 * (size_t)v + 1];
    r[2] = pp[6 * (size_t)v + 2];
}

/* the forces of nc cells: f (bending) and fad (area difference), 3 floats per
 * vertex each, written (not added) */
void orc_jul_forces(int nc, int nv, int nt, const int *tt, int ne, const int *dih, const float *pp, float *f,
                    float *fad) {
    const float H0 = -1.0f / 2.0f, kb = 1.0f, pi = (float)JUL_PI, kad = 2 * kb / pi;
    float *area = (float *)malloc(sizeof(float) * (size_t)nv);
    float *lt = (float *)malloc(sizeof(float) * (size_t)nv);
    float *theta = (float *)malloc(sizeof(float) * (size_t)ne);
    double *t0 = (double *)malloc(sizeof(double) * (size_t)ne);
    for (int c = 0; c < nc; c++) {
        const float *p = pp + 6 * (size_t)c * nv;
        float *fc = f + 3 * (size_t)c * nv, *gc = fad + 3 * (size_t)c * nv;
        for (int v = 0; v < nv; v++) area[v] = lt[v] = 0;
        /* vertex areas: each vertex adds its triangles in ascending order */
        for (int t = 0; t < nt; t++) {
            double a[3], b[3], cc[3];
            getv(p, tt[3 * t], a);
            getv(p, tt[3 * t + 1], b);
            getv(p, tt[3 * t + 2], cc);
            const float a0 = (float)(kahan_area(a, b, cc) / 3);
            for (int q = 0; q < 3; q++) area[tt[3 * t + q]] += a0;
        }
        for (int e = 0; e < ne; e++) {
            const int *d = dih + 4 * e;
            double a[3], b[3], cc[3], dd[3], u[3];
            getv(p, d[0], a);
            getv(p, d[1], b);
            getv(p, d[2], cc);
            getv(p, d[3], dd);
            t0[e] = orc_jul_theta(a, b, cc, dd);
            theta[e] = (float)t0[e];
            vsub(cc, b, u);
            const float lt0 = (float)(vabs3(u) * t0[e]);
            lt[d[1]] += lt0;
            lt[d[2]] += lt0;
        }
        float A = 0, Lt = 0;
        for (int v = 0; v < nv; v++) {
            A += area[v];
            Lt += lt[v];
        }
        const float C = (Lt / 4 - H0 * A) * (4 * kad * pi / A);
        for (int k = 0; k < 3 * nv; k++) fc[k] = gc[k] = 0;
        /* edge-length part */
        for (int e = 0; e < ne; e++) {
            const int *d = dih + 4 * e;
            const int j = d[1], k = d[2];
            double b[3], cc[3], db[3], dc[3];
            getv(p, j, b);
            getv(p, k, cc);
            dedg(b, cc, db, dc);
            double coef = -((lt[j] / area[j] / 4 - H0) + (lt[k] / area[k] / 4 - H0)) * theta[e];
            for (int x = 0; x < 3; x++) {
                fc[3 * j + x] += (float)(coef * db[x]);
                fc[3 * k + x] += (float)(coef * dc[x]);
            }
            coef = -C / 4 * theta[e];
            for (int x = 0; x < 3; x++) {
                gc[3 * j + x] += (float)(coef * db[x]);
                gc[3 * k + x] += (float)(coef * dc[x]);
            }
        }
        /* dihedral part */
        for (int e = 0; e < ne; e++) {
            const int *d = dih + 4 * e;
            double r[4][3], g[4][3], u[3];
            for (int q = 0; q < 4; q++) getv(p, d[q], r[q]);
            ddih(r[0], r[1], r[2], r[3], g);
            vsub(r[2], r[1], u);
            const double len0 = vabs3(u);
            const int j = d[1], k = d[2];
            double coef = -((lt[j] / area[j] / 4 - H0) + (lt[k] / area[k] / 4 - H0)) * len0;
            for (int q = 0; q < 4; q++)
                for (int x = 0; x < 3; x++) fc[3 * d[q] + x] += (float)(coef * g[q][x]);
            coef = -C / 2.0 * len0;
            for (int q = 0; q < 4; q++)
                for (int x = 0; x < 3; x++) gc[3 * d[q] + x] += (float)(coef * g[q][x]);
        }
        /* area part */
        for (int t = 0; t < nt; t++) {
            double r[3][3], g[3][3];
            for (int q = 0; q < 3; q++) getv(p, tt[3 * t + q], r[q]);
            dtri_area(r[0], r[1], r[2], g);
            for (int q = 0; q < 3; q++) {
                const int v = tt[3 * t + q];
                const double coef = (1.0 / 3.0) * (lt[v] * lt[v] / 8.0 / area[v] / area[v] - 2.0 * H0 * H0);
                for (int x = 0; x < 3; x++) fc[3 * v + x] += (float)(coef * g[q][x]);
            }
        }
    }
    free(area);
    free(lt);
    free(theta);
    free(t0);
}
