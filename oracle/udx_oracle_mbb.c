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

typedef struct { float r[3], v[3]; } mpart;

/* ---- crossing time -------------------------------------------------------
 * The crossing test of src/mesh_bounce/dev/intersection.h:30-127, derived
 * here.  Over the step every vertex moves linearly, X(t) = X0 + t Xv, and so
 * does the particle, p(t) = p0 + t pv.  With the edges e1(t) = B(t) - A(t)
 * and e2(t) = C(t) - A(t) the plane normal N(t) = e1 x e2 is the quadratic
 *   N(t) = N0 + t N1 + t^2 N2,
 *   N0 = e1 x e2,  N1 = e1 x e2v + e1v x e2,  N2 = e1v x e2v,
 * and the particle's offset from A is g(t) = g0 + t gv.  The signed volume
 * phi(t) = g . N is the cubic k0 + k1 t + k2 t^2 + k3 t^3 with
 *   k0 = g0.N0,  k1 = gv.N0 + g0.N1,  k2 = gv.N1 + g0.N2,  k3 = gv.N2,
 * and the particle passes through the plane at the first root of phi in
 * [0, dt] where phi changes sign.  The root: [0, dt] is cut at the turning
 * points of phi (roots of phi' = k1 + 2 k2 t + 3 k3 t^2, by the quadratic
 * formula in its cancellation-free form), phi is monotone on each piece, the
 * first piece whose ends differ in sign holds the root, and a Newton
 * iteration kept inside that bracket (bisection when a step leaves it)
 * converges to it.  The coefficients are used as they are (the reference
 * drops a leading coefficient below 1e-8 and solves the lower degree), and a
 * root where phi touches zero without changing sign is not a crossing. */
static double cub_eval(const double k[4], double t) { return ((k[3] * t + k[2]) * t + k[1]) * t + k[0]; }
static double cub_slope(const double k[4], double t) { return (3 * k[3] * t + 2 * k[2]) * t + k[1]; }

/* turning points of the cubic strictly inside (0, T), ascending */
static int cub_turning(const double k[4], double T, double out[2]) {
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
    if (m == 2 && r[1] < r[0]) { const double t = r[0]; r[0] = r[1]; r[1] = t; }
    for (int i = 0; i < m; i++)
        if (r[i] > 0 && r[i] < T) out[n++] = r[i];
    return n;
}

/* the root in (lo, hi) of a cubic monotone there, flo = value at lo (the
 * value at hi has the other sign) */
static double cub_bracketed(const double k[4], double lo, double hi, double flo) {
    double t = 0.5 * (lo + hi);
    for (int it = 0; it < 200; it++) {
        const double f = cub_eval(k, t);
        if (f == 0) return t;
        if ((f < 0) == (flo < 0)) { lo = t; flo = f; } else hi = t;
        const double d = cub_slope(k, t);
        double tn = d != 0 ? t - f / d : lo;
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == t || !(hi - lo > 4e-16 * hi)) return tn;
        t = tn;
    }
    return t;
}

/* first sign-changing root of the cubic in [0, T] */
static int cub_first_root(const double k[4], double T, double *h) {
    double s[4];
    int m = 0;
    s[m++] = 0;
    m += cub_turning(k, T, s + 1);
    s[m++] = T;
    double fprev = cub_eval(k, s[0]);
    if (fprev == 0) { *h = s[0]; return 1; }
    for (int i = 1; i < m; i++) {
        const double f = cub_eval(k, s[i]);
        if (f == 0) { *h = s[i]; return 1; }
        if ((f < 0) != (fprev < 0)) { *h = cub_bracketed(k, s[i - 1], s[i], fprev); return 1; }
        fprev = f;
    }
    return 0;
}

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Triangle A B C (vertex positions XA.., velocities VA..; float) and a
 * particle (pr, pv; float) over [0, dt].  Returns ORC_BB_* and (h, u, v, s):
 * crossing time, barycentric coordinates of the crossing point, and the
 * side of the plane the particle starts on (+1: phi(0) > 0). */
int orc_mbb_intersect(float dt_f, const float XA[3], const float XB[3], const float XC[3], const float VA[3],
                      const float VB[3], const float VC[3], const float pr[3], const float pv[3], float out[4]) {
    const double T = dt_f;
    double e1[3], e2[3], e1v[3], e2v[3], g0[3], gv[3];
    for (int c = 0; c < 3; c++) {
        e1[c] = (double)(XB[c] - XA[c]);
        e2[c] = (double)(XC[c] - XA[c]);
        e1v[c] = (double)(VB[c] - VA[c]);
        e2v[c] = (double)(VC[c] - VA[c]);
        g0[c] = (double)(pr[c] - XA[c]);
        gv[c] = (double)(pv[c] - VA[c]);
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
    if (phi0 * phiT > 0) return ORC_BB_NOCROSS;
    const double k[4] = {phi0, dot3(gv, N0) + dot3(g0, N1), dot3(gv, N1) + dot3(g0, N2), dot3(gv, N2)};
    double h;
    if (!cub_first_root(k, T, &h)) return ORC_BB_HFAIL;
    /* the crossing point on the triangle at time h: g = u e1 + v e2 in the
     * least-squares sense (normal equations of the two edge directions) */
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
    if (u < 0 || v < 0 || u + v > 1) return ORC_BB_WTRIANGLE;
    out[0] = (float)h;
    out[1] = (float)u;
    out[2] = (float)v;
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
