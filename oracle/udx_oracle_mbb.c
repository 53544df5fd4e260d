/*
 * CPU oracle of the solvent bounce-back on moving membrane meshes
 * (src/mesh_bounce) on one periodic domain.
 *
 * TEST INFRASTRUCTURE ONLY (see udx_oracle.h).
 *
 * Per step (src/sim/objects/imp/bounce.h:185-203 for membranes):
 *  - every triangle moves linearly from its vertices' positions at the start
 *    of the step (rp) to those after the membrane advance (rc); a solvent
 *    particle moves linearly from pp0 (start of step) to pp (after the move);
 *  - meshes near the domain boundary also act through their periodic images
 *    (the mesh halo of src/exch/mesh: a mesh whose bounding box at rc leaves
 *    the domain cropped by 1 is copied, shifted by a period, once per face,
 *    edge and corner direction of src/exch/map/dev.h:21-83);
 *  - every triangle instance tests the solvent of the cells overlapping its
 *    start-of-step bounding box grown by 0.1 (src/mesh_bounce/dev/main.h:1-86)
 *    for a crossing in [0, dt] (Fedosov's cubic, dev/intersection.h);
 *  - each particle keeps its earliest crossing; it is reflected with the
 *    triangle velocity at the crossing point, v' = 2 u_w - v, moved for the
 *    rest of the step and pushed 0.01 off the triangle (dev/main.h:130-210);
 *  - the momentum given to each triangle is collected into the velocities of
 *    its three vertices through the triangle's inertia (dev/collect.h:84-103,
 *    dev/utils.h:110-213).
 *
 * Mixed precision is the reference's: float particle and vertex data, double
 * plane and root computations (dev/utils.h:1-2, dev/type.h:4-5), float
 * results.  Where the reference relies on atomics the order here is defined
 * (and the HIP path follows it):
 *  - a particle's crossing is the one with the smallest time, ties to the
 *    smaller triangle instance index (the reference keeps the smallest time
 *    among at most 4 recorded candidates, the first recorded on ties);
 *  - instances: meshes 0..nm-1 unshifted, then images by fragment id
 *    (0..25) of the direction the mesh leaves toward, meshes ascending; the
 *    image toward d is shifted by -d L;
 *  - a triangle's momentum sums its collisions (all its instances) in
 *    ascending particle index; a vertex adds the velocity changes of its
 *    triangles in ascending triangle index.
 * Other definitions: a triangle whose start-of-step bounding box is wider
 * than 8 (or not finite) is not searched; 1/sqrt for rsqrtf in the push; the "nonzero" test of
 * src/mesh_bounce/dev/utils.h:94-99 (every component of P and L above 1e-6
 * in magnitude) on a triangle's total momentum; the angular momentum change
 * as the reference writes it (dev/utils.h:76-84).
 */
#include "udx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MBB_PI 3.14159265358979323846

typedef struct { float r[3], v[3]; } mpart;

/* ---- roots (src/mesh_bounce/dev/roots.h) -------------------------------- */
static int valid_time(double dt, double t) { return t >= 0 && t <= dt; }

static int quadratic(double a, double b, double c, double *h0, double *h1) {
    const int sgnb = b > 0 ? 1 : -1;
    const double D = b * b - 4 * a * c;
    if (D < 0) return 0;
    *h0 = (-b - sgnb * sqrt(D)) / (2 * a);
    *h1 = c / (a * *h0);
    if (*h0 > *h1) {
        double t = *h0;
        *h0 = *h1;
        *h1 = t;
    }
    return 1;
}

/* real roots of x^3 + a x^2 + b x + c, ascending (the GSL method the
 * reference restates, roots.h:18-97) */
static int cubic(double a, double b, double c, double *x0, double *x1, double *x2) {
    const double q = a * a - 3 * b;
    const double r = 2 * a * a * a - 9 * a * b + 27 * c;
    const double Q = q / 9;
    const double R = r / 54;
    const double Q3 = Q * Q * Q;
    const double R2 = R * R;
    const double CR2 = 729 * r * r;
    const double CQ3 = 2916 * q * q * q;
    if (R == 0 && Q == 0) {
        *x0 = *x1 = *x2 = -a / 3;
        return 3;
    } else if (CR2 == CQ3) {
        const double sqrtQ = sqrt(Q);
        if (R > 0) {
            *x0 = -2 * sqrtQ - a / 3;
            *x1 = sqrtQ - a / 3;
            *x2 = sqrtQ - a / 3;
        } else {
            *x0 = -sqrtQ - a / 3;
            *x1 = -sqrtQ - a / 3;
            *x2 = 2 * sqrtQ - a / 3;
        }
        return 3;
    } else if (R2 < Q3) {
        const double sgnR = R >= 0 ? 1 : -1;
        const double ratio = sgnR * sqrt(R2 / Q3);
        const double theta = acos(ratio);
        const double norm = -2 * sqrt(Q);
        double t;
        *x0 = norm * cos(theta / 3) - a / 3;
        *x1 = norm * cos((theta + 2.0 * MBB_PI) / 3) - a / 3;
        *x2 = norm * cos((theta - 2.0 * MBB_PI) / 3) - a / 3;
        if (*x0 > *x1) { t = *x0; *x0 = *x1; *x1 = t; }
        if (*x1 > *x2) {
            t = *x1; *x1 = *x2; *x2 = t;
            if (*x0 > *x1) { t = *x0; *x0 = *x1; *x1 = t; }
        }
        return 3;
    } else {
        const double sgnR = R >= 0 ? 1 : -1;
        const double A = -sgnR * pow(fabs(R) + sqrt(R2 - Q3), 1.0 / 3.0);
        const double B = Q / A;
        *x0 = A + B - a / 3;
        return 1;
    }
}

/* the first root of a t^3 + b t^2 + c t + d in [0, dt] (intersection.h:1-28) */
static int cubic_root(double dt, double a, double b, double c, double d, double *h) {
    const double eps = 1e-8;
    double h1, h2, h3;
    if (fabs(a) > eps) {
        b /= a;
        c /= a;
        d /= a;
        const int nsol = cubic(b, c, d, &h1, &h2, &h3);
        if (valid_time(dt, h1)) { *h = h1; return 1; }
        if (nsol > 1 && valid_time(dt, h2)) { *h = h2; return 1; }
        if (nsol > 2 && valid_time(dt, h3)) { *h = h3; return 1; }
    } else if (fabs(b) > eps) {
        if (!quadratic(b, c, d, &h1, &h2)) return 0;
        if (valid_time(dt, h1)) { *h = h1; return 1; }
        if (valid_time(dt, h2)) { *h = h2; return 1; }
    } else if (fabs(c) > eps) {
        h1 = -d / c;
        if (valid_time(dt, h1)) { *h = h1; return 1; }
    }
    return 0;
}

static double ddot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void dcross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Fedosov's moving-triangle crossing (src/mesh_bounce/dev/intersection.h:
 * 30-127): triangle s1 s2 s3 moving with vs1 vs2 vs3 (float), particle
 * p0 = (r, v) (float); dt in double.  Returns ORC_BB_* and (h, u, v, s). */
int orc_mbb_intersect(float dt_f, const float s1[3], const float s2[3], const float s3[3], const float vs1[3],
                      const float vs2[3], const float vs3[3], const float pr[3], const float pv[3], float out[4]) {
    const double dt = dt_f;
    double a1[3], a2[3], at1[3], at2[3], n0[3], nt[3], ntt[3], dr0[3];
    for (int c = 0; c < 3; c++) {
        a1[c] = (double)(s2[c] - s1[c]);
        a2[c] = (double)(s3[c] - s1[c]);
        at1[c] = (double)(vs2[c] - vs1[c]);
        at2[c] = (double)(vs3[c] - vs1[c]);
    }
    dcross(a1, a2, n0);
    dcross(at1, at2, ntt);
    nt[0] = a1[1] * at2[2] - a1[2] * at2[1] + at1[1] * a2[2] - at1[2] * a2[1];
    nt[1] = a1[2] * at2[0] - a1[0] * at2[2] + at1[2] * a2[0] - at1[0] * a2[2];
    nt[2] = a1[0] * at2[1] - a1[1] * at2[0] + at1[0] * a2[1] - at1[1] * a2[0];
    for (int c = 0; c < 3; c++) dr0[c] = (double)(pr[c] - s1[c]);
    double s;
    {
        double r1[3], s11[3], n1[3], dr1[3];
        for (int c = 0; c < 3; c++) {
            r1[c] = (double)pr[c] + dt * (double)pv[c];
            s11[c] = (double)s1[c] + dt * (double)vs1[c];
            n1[c] = n0[c] + dt * (nt[c] + dt * ntt[c]);
        }
        for (int c = 0; c < 3; c++) dr1[c] = r1[c] - s11[c];
        const double b0 = ddot(dr0, n0), b1 = ddot(dr1, n1);
        s = b0 > 0 ? 1 : -1;
        out[3] = (float)s;
        if (b0 * b1 > 0) return ORC_BB_NOCROSS;
    }
    double hl;
    {
        double dv[3];
        for (int c = 0; c < 3; c++) dv[c] = (double)(pv[c] - vs1[c]);
        const double a = ddot(ntt, dv);
        const double b = ddot(ntt, dr0) + ddot(nt, dv);
        const double c = ddot(nt, dr0) + ddot(n0, dv);
        const double d = ddot(n0, dr0);
        if (!cubic_root(dt, a, b, c, d, &hl)) return ORC_BB_HFAIL;
    }
    double rwl[3], g[3], a1_[3], a2_[3];
    for (int c = 0; c < 3; c++) {
        rwl[c] = (double)pr[c] + hl * (double)pv[c];
        g[c] = rwl[c] - (double)s1[c] - hl * (double)vs1[c];
        a1_[c] = a1[c] + hl * at1[c];
        a2_[c] = a2[c] + hl * at2[c];
    }
    const double ga1 = ddot(g, a1_), ga2 = ddot(g, a2_);
    const double a11 = ddot(a1_, a1_), a12 = ddot(a1_, a2_), a22 = ddot(a2_, a2_);
    const double fac = 1.0 / (a11 * a22 - a12 * a12);
    const double ul = (ga1 * a22 - ga2 * a12) * fac;
    const double vl = (ga2 * a11 - ga1 * a12) * fac;
    if ((ul < 0) || (vl < 0) || (ul + vl > 1)) return ORC_BB_WTRIANGLE;
    out[0] = (float)hl;
    out[1] = (float)ul;
    out[2] = (float)vl;
    return ORC_BB_SUCCESS;
}

/* ---- mesh instances (src/exch/mesh/imp/map.h, src/exch/map/dev.h) ------- */
static int frag_d2i(int x, int y, int z) { return ((x + 2) % 3) + 3 * ((y + 2) % 3) + 9 * ((z + 2) % 3); }

/* the directions a box code (x, y, z) sends to: every non-zero sub-vector */
static int box_dests(int x, int y, int z, int fids[7]) {
    int n = 0;
    const int d[3] = {x, y, z};
    for (int m = 1; m < 8; m++) {
        int e[3], ok = 1;
        for (int c = 0; c < 3; c++) {
            e[c] = (m >> c & 1) ? d[c] : 0;
            if ((m >> c & 1) && d[c] == 0) ok = 0;
        }
        if (ok) fids[n++] = frag_d2i(e[0], e[1], e[2]);
    }
    return n;
}

static void frag_i2d(int fid, int d[3]) {
    const int t[3] = {fid % 3, (fid / 3) % 3, fid / 9};
    for (int c = 0; c < 3; c++) d[c] = t[c] == 0 ? 1 : (t[c] == 1 ? -1 : 0);
}

int orc_mbb_instances(const int L[3], int nm, int nv, const float *rc, int *inst_mesh, float *inst_shift) {
    int ni = 0;
    int *codes = (int *)malloc(sizeof(int) * 3 * (nm > 0 ? nm : 1));
    for (int m = 0; m < nm; m++) {
        float lo[3], hi[3];
        for (int c = 0; c < 3; c++) lo[c] = hi[c] = rc[6 * (size_t)m * nv + c];
        for (int v = 1; v < nv; v++)
            for (int c = 0; c < 3; c++) {
                const float x = rc[6 * ((size_t)m * nv + v) + c];
                lo[c] = x < lo[c] ? x : lo[c];
                hi[c] = x > hi[c] ? x : hi[c];
            }
        for (int c = 0; c < 3; c++) {
            const int h = (L[c] - 2) / 2;   /* the domain cropped by 1, src/exch/mesh/imp/map.h:1-4 */
            codes[3 * m + c] = -1 + (lo[c] >= -h) + (hi[c] >= h);
        }
        inst_mesh[ni] = m;
        inst_shift[3 * ni] = inst_shift[3 * ni + 1] = inst_shift[3 * ni + 2] = 0;
        ni++;
    }
    for (int fid = 0; fid < 26; fid++) {
        int d[3];
        frag_i2d(fid, d);
        for (int m = 0; m < nm; m++) {
            int fids[7];
            const int nd = box_dests(codes[3 * m], codes[3 * m + 1], codes[3 * m + 2], fids);
            int hit = 0;
            for (int k = 0; k < nd; k++) hit |= fids[k] == fid;
            if (!hit) continue;
            inst_mesh[ni] = m;
            for (int c = 0; c < 3; c++) inst_shift[3 * ni + c] = (float)(-d[c] * L[c]);
            ni++;
        }
    }
    free(codes);
    return ni;
}

/* ---- collision search, bounce, momentum ---------------------------------- */
/* vertex j of instance k: r = rc + shift, v = (r - (rp + shift)) / dt
 * (src/mesh_bounce/dev/type.h:29-44 after the halo shift of
 * src/exch/common/dev.h:49-57) */
static void fetch_vertex(float dt, const float *rp, const float *rc, const float *shift, int j, mpart *o) {
    for (int c = 0; c < 3; c++) {
        const float pc = rc[6 * (size_t)j + c] + shift[c];
        const float pp = rp[3 * (size_t)j + c] + shift[c];
        o->r[c] = pc;
        o->v[c] = (pc - pp) / dt;
    }
}

static void fetch_tri(float dt, int gid, int nt, int nv, const int *tt, const int *inst_mesh, const float *inst_shift,
                      const float *rp, const float *rc, mpart *A, mpart *B, mpart *C) {
    const int k = gid / nt, t = gid % nt;
    const int base = inst_mesh[k] * nv;
    const float *sh = inst_shift + 3 * k;
    fetch_vertex(dt, rp, rc, sh, base + tt[3 * t], A);
    fetch_vertex(dt, rp, rc, sh, base + tt[3 * t + 1], B);
    fetch_vertex(dt, rp, rc, sh, base + tt[3 * t + 2], C);
}

static void revert(float h, mpart *p) {
    for (int c = 0; c < 3; c++) p->r[c] = p->r[c] - p->v[c] * h;
}

/* the particle's start and its straight-line velocity (dev/utils.h:43-48) */
static void particle_path(float dt, const float *pp, const float *pp0, int i, mpart *p0) {
    for (int c = 0; c < 3; c++) {
        const float r0 = pp0[6 * (size_t)i + c], r1 = pp[6 * (size_t)i + c];
        p0->r[c] = r0;
        p0->v[c] = (float)((double)(r1 - r0) / (double)dt);
    }
}

static int cell_of(int L, float x) {
    int c = (int)floor((double)x + L / 2);
    return c < 0 ? 0 : (c > L - 1 ? L - 1 : c);
}

static uint64_t make_key(float h, int gid) {
    uint32_t b;
    memcpy(&b, &h, 4);
    if (h == 0) b = 0;
    return ((uint64_t)b << 32) | (uint32_t)gid;
}

void orc_mbb_find(float dt, const int L[3], int ninst, const int *inst_mesh, const float *inst_shift, int nt, int nv,
                  const int *tt, const float *rp, const float *rc, const int *starts, const float *pp,
                  const float *pp0, uint64_t *keys) {
    const float tol = 0.1f;
    for (int gid = 0; gid < ninst * nt; gid++) {
        mpart A, B, C;
        fetch_tri(dt, gid, nt, nv, tt, inst_mesh, inst_shift, rp, rc, &A, &B, &C);
        revert(dt, &A);
        revert(dt, &B);
        revert(dt, &C);
        int lo[3], hi[3], skip = 0;
        for (int c = 0; c < 3; c++) {
            float mn = A.r[c], mx = A.r[c];
            mn = B.r[c] < mn ? B.r[c] : mn;
            mn = C.r[c] < mn ? C.r[c] : mn;
            mx = B.r[c] > mx ? B.r[c] : mx;
            mx = C.r[c] > mx ? C.r[c] : mx;
            if (!(mx - mn <= 8.0f)) skip = 1;   /* no membrane triangle: not searched */
            lo[c] = cell_of(L[c], mn - tol);
            hi[c] = cell_of(L[c], mx + tol);
        }
        if (skip) continue;
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++)
                for (int ix = lo[0]; ix <= hi[0]; ix++) {
                    const int cid = ix + L[0] * (iy + L[1] * iz);
                    for (int i = starts[cid]; i < starts[cid + 1]; i++) {
                        mpart p0;
                        particle_path(dt, pp, pp0, i, &p0);
                        float o[4];
                        if (orc_mbb_intersect(dt, A.r, B.r, C.r, A.v, B.v, C.v, p0.r, p0.v, o) != ORC_BB_SUCCESS)
                            continue;
                        const uint64_t k = make_key(o[0], gid);
                        if (k < keys[i]) keys[i] = k;
                    }
                }
    }
}

/* bounce of particle i off instance triangle gid (dev/main.h:130-210);
 * writes the new particle and its momentum change (P, L) */
static void perform_one(float dt, float mass, int i, int gid, int nt, int nv, const int *tt, const int *inst_mesh,
                        const float *inst_shift, const float *rp, const float *rc, const float *pp0, float *pp,
                        float dm[6]) {
    mpart p1, p0, A, B, C;
    for (int c = 0; c < 3; c++) {
        p1.r[c] = pp[6 * (size_t)i + c];
        p1.v[c] = pp[6 * (size_t)i + 3 + c];
    }
    particle_path(dt, pp, pp0, i, &p0);
    fetch_tri(dt, gid, nt, nv, tt, inst_mesh, inst_shift, rp, rc, &A, &B, &C);
    /* the crossing again, from the start-of-step triangle */
    float d[4];
    {
        mpart a = A, b = B, c = C;
        revert(dt, &a);
        revert(dt, &b);
        revert(dt, &c);
        (void)orc_mbb_intersect(dt, a.r, b.r, c.r, a.v, b.v, c.v, p0.r, p0.v, d);
    }
    /* collision point and wall velocity (dev/main.h:110-128) */
    float rw[3], vw[3];
    {
        mpart a = A, b = B, c = C;
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
    /* bounce_back (dev/utils.h:93-102), dt in double */
    mpart pn;
    for (int k = 0; k < 3; k++) {
        pn.v[k] = 2 * vw[k] - p0.v[k];
        pn.r[k] = (float)((double)rw[k] + ((double)dt - (double)d[0]) * (double)pn.v[k]);
    }
    /* push 0.01 along the side's normal at the current triangle (dev/main.h:130-150) */
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
    for (int k = 0; k < 3; k++) {
        pp[6 * (size_t)i + k] = pn.r[k];
        pp[6 * (size_t)i + 3 + k] = pn.v[k];
    }
    /* momentum given to the triangle (dev/utils.h:104-118), angular part
     * about the current triangle's centre */
    float rcen[3], r0[3], r1[3];
    for (int k = 0; k < 3; k++) {
        rcen[k] = (float)(0.333333 * (double)(A.r[k] + B.r[k] + C.r[k]));
        r0[k] = p1.r[k] - rcen[k];
        r1[k] = pn.r[k] - rcen[k];
        dm[k] = -mass * (pn.v[k] - p1.v[k]);
    }
    const float *v0 = p1.v, *v1 = pn.v;
    dm[3] = -mass * (r1[1] * v1[2] - r1[2] * v1[1] - r0[1] * v0[2] + r0[2] - v0[1]);
    dm[4] = -mass * (r1[2] * v1[0] - r1[0] * v1[2] - r0[2] * v0[0] + r0[0] - v0[2]);
    dm[5] = -mass * (r1[0] * v1[1] - r1[1] * v1[0] - r0[0] * v0[1] + r0[1] - v0[0]);
}

/* ---- triangle inertia and the velocity change of its vertices ----------- */
enum { XX, XY, XZ, YY, YZ, ZZ };

void orc_mbb_inertia(float mass, const float a[3], const float b[3], const float c[3], float I[6]) {
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

void orc_mbb_inverse(const float A[6], float I[6]) {
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

static int nz(float a) { return fabsf(a) > 1e-6f; }

int orc_mbb_nonzero(const float m[6]) {
    return nz(m[0]) && nz(m[1]) && nz(m[2]) && nz(m[3]) && nz(m[4]) && nz(m[5]);
}

/* rbc_M2v (dev/utils.h:172-213): velocity changes of the vertices a, b, c */
void orc_mbb_m2v(float mass, const float m[6], const float a_[3], const float b_[3], const float c_[3], float va[3],
                 float vb[3], float vc[3]) {
    float I[6], Ii[6], om[3], v[3], com[3], a[3], b[3], c[3];
    const float one_third = 1.f / 3.f;
    orc_mbb_inertia(mass, a_, b_, c_, I);
    orc_mbb_inverse(I, Ii);
    const float *L = m + 3;
    om[0] = one_third * (Ii[XX] * L[0] + Ii[XY] * L[1] + Ii[XZ] * L[2]);
    om[1] = one_third * (Ii[XY] * L[0] + Ii[YY] * L[1] + Ii[YZ] * L[2]);
    om[2] = one_third * (Ii[XZ] * L[0] + Ii[YZ] * L[1] + Ii[ZZ] * L[2]);
    const float inv_mass = (float)(1.0 / (double)mass);
    for (int k = 0; k < 3; k++) {
        v[k] = one_third * m[k] * inv_mass;
        com[k] = one_third * (a_[k] + b_[k] + c_[k]);
    }
    for (int k = 0; k < 3; k++) {
        a[k] = a_[k] - com[k];
        b[k] = b_[k] - com[k];
        c[k] = c_[k] - com[k];
    }
    const float *rr[3] = {a, b, c};
    float *out[3] = {va, vb, vc};
    for (int j = 0; j < 3; j++) {
        const float *r = rr[j];
        out[j][0] = v[0] + r[1] * om[2] - r[2] * om[1];
        out[j][1] = v[1] + r[2] * om[0] - r[0] * om[2];
        out[j][2] = v[2] + r[0] * om[1] - r[1] * om[0];
    }
}

/* ---- the whole bounce ---------------------------------------------------- */
long orc_mbb_bounce(float dt, float flu_mass, float rbc_mass, const int L[3], int nm, int nt, int nv, const int *tt,
                    const float *rp, float *rc, const int *starts, int n, const float *pp0, float *pp, float *mm,
                    int *col_tri) {
    int *inst_mesh = (int *)malloc(sizeof(int) * 8 * (nm > 0 ? nm : 1));
    float *inst_shift = (float *)malloc(sizeof(float) * 24 * (nm > 0 ? nm : 1));
    const int ninst = orc_mbb_instances(L, nm, nv, rc, inst_mesh, inst_shift);
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) keys[i] = UINT64_MAX;
    orc_mbb_find(dt, L, ninst, inst_mesh, inst_shift, nt, nv, tt, rp, rc, starts, pp, pp0, keys);
    const long ntri = (long)nm * nt;
    memset(mm, 0, sizeof(float) * 6 * (ntri > 0 ? ntri : 1));
    long ncol = 0;
    for (int i = 0; i < n; i++) {
        col_tri[i] = -1;
        if (keys[i] == UINT64_MAX) continue;
        const int gid = (int)(keys[i] & 0xffffffffu);
        float dm[6];
        perform_one(dt, flu_mass, i, gid, nt, nv, tt, inst_mesh, inst_shift, rp, rc, pp0, pp, dm);
        const int tri = inst_mesh[gid / nt] * nt + gid % nt;
        col_tri[i] = tri;
        for (int k = 0; k < 6; k++) mm[6 * (size_t)tri + k] += dm[k];   /* ascending particle index */
        ncol++;
    }
    /* collect_rbc_mom (dev/collect.h:84-103): ascending triangle per vertex */
    for (long t = 0; t < ntri; t++) {
        const float *m = mm + 6 * t;
        if (!orc_mbb_nonzero(m)) continue;
        const int cell = (int)(t / nt), tl = (int)(t % nt);
        const int j[3] = {cell * nv + tt[3 * tl], cell * nv + tt[3 * tl + 1], cell * nv + tt[3 * tl + 2]};
        float dv[3][3];
        orc_mbb_m2v(rbc_mass, m, rc + 6 * (size_t)j[0], rc + 6 * (size_t)j[1], rc + 6 * (size_t)j[2], dv[0], dv[1],
                    dv[2]);
        for (int q = 0; q < 3; q++)
            for (int k = 0; k < 3; k++) rc[6 * (size_t)j[q] + 3 + k] += dv[q][k];
    }
    free(keys);
    free(inst_mesh);
    free(inst_shift);
    return ncol;
}
