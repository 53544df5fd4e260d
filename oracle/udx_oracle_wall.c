/*
 * CPU oracle of the SDF walls (test infrastructure only; see udx_oracle.h).
 * Build with -ffp-contract=off: fused multiply-adds are written out.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "udx_oracle.h"

/* ======================================================================== */
/* grid transforms: src/math/tform/imp/main.h                               */
/* ======================================================================== */
/* o, s with b = s a + o mapping a0 -> a1 and b0 -> b1 (tform_vector, :13-31) */
static void os1(float a0, float a1, float b0, float b1, float *o, float *s) {
    *o = (a1 * b0 - a0 * b1) / (b0 - a0);
    *s = (b1 - a1) / (b0 - a0);
}

static void tvec(const float a0[3], const float a1[3], const float b0[3], const float b1[3], float o[3], float s[3]) {
    for (int c = 0; c < 3; c++) os1(a0[c], a1[c], b0[c], b1[c], &o[c], &s[c]);
}

static void tconv(const float o[3], const float s[3], const float a[3], float b[3]) {
    for (int c = 0; c < 3; c++) b[c] = s[c] * a[c] + o[c];   /* host: no contraction */
}

/* tform_to_grid / tform_from_grid (:71-84): cell-centred grid of n points */
static void to_grid(const float lo[3], const float hi[3], const int n[3], float o[3], float s[3]) {
    float a1[3], b1[3];
    for (int c = 0; c < 3; c++) { a1[c] = -0.5f; b1[c] = n[c] - 0.5f; }
    tvec(lo, a1, hi, b1, o, s);
}

static void from_grid(const float lo[3], const float hi[3], const int n[3], float o[3], float s[3]) {
    float a1[3], b1[3];
    for (int c = 0; c < 3; c++) { a1[c] = -0.5f; b1[c] = n[c] - 0.5f; }
    tvec(a1, lo, b1, hi, o, s);
}

/* tform_chain (:55-66) of from_grid(f) then to_grid(t): tform_grid2grid */
void orc_tform_grid2grid(const float flo[3], const float fhi[3], const int fn[3], const float tlo[3],
                         const float thi[3], const int tn[3], float o[3], float s[3]) {
    float o1[3], s1[3], o2[3], s2[3];
    float a0[3] = {0, 0, 0}, b0[3] = {1, 1, 1}, a1[3], b1[3], a2[3], b2[3];
    from_grid(flo, fhi, fn, o1, s1);
    to_grid(tlo, thi, tn, o2, s2);
    tconv(o1, s1, a0, a1);
    tconv(o1, s1, b0, b1);
    tconv(o2, s2, a1, a2);
    tconv(o2, s2, b1, b2);
    tvec(a0, a2, b0, b2, o, s);
}

/* ======================================================================== */
/* SDF texture: src/wall/sdf/imp/main.h:1-33, gen.h:1-80, field/imp/sample.h */
/* ======================================================================== */
static int ceiln(int a, int b) { return (a + b - 1) / b; }

void orc_sdf_texture_size(const int L[3], int T[3]) {
    enum { PAD = 16 };
    int M = ORC_WALL_MARGIN;
    int x = L[0] + 2 * M, y = L[1] + 2 * M, z = L[2] + 2 * M;
    T[0] = PAD * PAD;
    T[1] = PAD * ceiln(ceiln(y * T[0], x), PAD);
    T[2] = PAD * ceiln(ceiln(z * T[0], x), PAD);
}

/* cubic B-spline weight (src/wall/sdf/field/imp/sample.h:28-36) */
static float spl(float x) {
    return x <= 0 ? 0.0f
         : x <= 1 ? x * x * x / 6
         : x <= 2 ? (x * ((12 - 3 * x) * x - 12) + 4) / 6
         : x <= 3 ? (x * (x * (3 * x - 24) + 60) - 44) / 6
         : x <= 4 ? (x * ((12 - x) * x - 48) + 64) / 6
         : 0.0f;
}

/* periodic B-spline sample of field D (grid N) at grid coordinate r
 * (src/wall/sdf/field/imp/sample.h:38-70) */
static float spl_get(const int N[3], const float *D, const float r[3]) {
    int anchor[3], g[3];
    float w[3][4], tmp[4][4], partial[4], s, val;
    for (int c = 0; c < 3; c++) anchor[c] = (int)floorf(r[c]);
    for (int c = 0; c < 3; c++)
        for (int i = 0; i < 4; i++) w[c][i] = spl(r[c] - (float)(anchor[c] - 1 + i) + 2);
    for (int sz = 0; sz < 4; sz++)
        for (int sy = 0; sy < 4; sy++) {
            s = 0;
            for (int sx = 0; sx < 4; sx++) {
                int l[3] = {sx, sy, sz};
                for (int c = 0; c < 3; c++) g[c] = (l[c] - 1 + anchor[c] + N[c]) % N[c];
                s += w[0][sx] * D[g[0] + N[0] * (g[1] + N[1] * g[2])];
            }
            tmp[sz][sy] = s;
        }
    for (int sz = 0; sz < 4; sz++) {
        s = 0;
        for (int sy = 0; sy < 4; sy++) s += w[1][sy] * tmp[sz][sy];
        partial[sz] = s;
    }
    val = 0;
    for (int sz = 0; sz < 4; sz++) val += w[2][sz] * partial[sz];
    return val;
}

void orc_sdf_gen(const int L[3], const float ext[3], const int N[3], const float *D0, orc_sdf *q) {
    int M = ORC_WALL_MARGIN, T[3], n0 = N[0] * N[1] * N[2];
    orc_sdf_texture_size(L, T);
    for (int c = 0; c < 3; c++) { q->L[c] = L[c]; q->T[c] = T[c]; }
    /* field_scale by xdomain / ext[x] (src/wall/sdf/imp/gen.h:36-46) */
    float sc = (float)L[0] / ext[0];
    float *D = (float *)malloc(sizeof(float) * (size_t)n0);
    for (int i = 0; i < n0; i++) D[i] = D0[i] * sc;
    /* texture grid -> field grid (tex2sdf_ini, src/wall/sdf/tform/imp.cpp:74-88) */
    float tlo[3], thi[3], slo[3], shi[3], o[3], s[3];
    for (int c = 0; c < 3; c++) {
        tlo[c] = (float)(0 - M);
        thi[c] = (float)(L[c] + M);
        slo[c] = 0;
        shi[c] = (float)L[c];
    }
    orc_tform_grid2grid(tlo, thi, T, slo, shi, N, o, s);
    for (int iz = 0; iz < T[2]; iz++)
        for (int iy = 0; iy < T[1]; iy++)
            for (int ix = 0; ix < T[0]; ix++) {
                float a[3] = {(float)ix, (float)iy, (float)iz}, r[3];
                tconv(o, s, a, r);
                q->tex[ix + T[0] * ((long)iy + T[1] * (long)iz)] = spl_get(N, D, r);
            }
    free(D);
    /* local coordinates -> texture grid (sub2tex_ini, :101-117) */
    float lo[3], hi[3];
    for (int c = 0; c < 3; c++) {
        lo[c] = (float)(-L[c] / 2 - M);
        hi[c] = (float)(L[c] / 2 + M);
    }
    to_grid(lo, hi, T, q->o, q->s);
    q->cheap_threshold = -sqrtf(3.f) * ((float)(L[0] + 2 * M) / (float)T[0]);
}

/* the scaled field sampled at the cell centres of the box (the wall.h5 dump,
 * src/wall/sdf/field/imp/main.h:1-8 with out2sdf_ini) */
void orc_sdf_sample_box(const int L[3], const float ext[3], const int N[3], const float *D0, float *out) {
    int n0 = N[0] * N[1] * N[2];
    float sc = (float)L[0] / ext[0];
    float *D = (float *)malloc(sizeof(float) * (size_t)n0);
    for (int i = 0; i < n0; i++) D[i] = D0[i] * sc;
    float lo[3] = {0, 0, 0}, hi[3], o[3], s[3];
    for (int c = 0; c < 3; c++) hi[c] = (float)L[c];
    orc_tform_grid2grid(lo, hi, L, lo, hi, N, o, s);
    for (int iz = 0; iz < L[2]; iz++)
        for (int iy = 0; iy < L[1]; iy++)
            for (int ix = 0; ix < L[0]; ix++) {
                float a[3] = {(float)ix, (float)iy, (float)iz}, r[3];
                tconv(o, s, a, r);
                out[ix + L[0] * (iy + L[1] * iz)] = spl_get(N, D, r);
            }
    free(D);
}

/* point fetch with wrap addressing at integer texel coordinates */
static float fetch(const orc_sdf *q, int i, int j, int k) {
    const int *T = q->T;
    i %= T[0]; if (i < 0) i += T[0];
    j %= T[1]; if (j < 0) j += T[1];
    k %= T[2]; if (k < 0) k += T[2];
    return q->tex[i + T[0] * ((long)j + T[1] * (long)k)];
}

/* texture coordinate of a local position: s r + o, fused (device code) */
static void to_tex(const orc_sdf *q, float x, float y, float z, float t[3]) {
    t[0] = fmaf(q->s[0], x, q->o[0]);
    t[1] = fmaf(q->s[1], y, q->o[1]);
    t[2] = fmaf(q->s[2], z, q->o[2]);
}

/* trilinear interpolation at texel centres (src/wall/sdf/dev.h:79-112) */
float orc_sdf_eval(const orc_sdf *q, float x, float y, float z) {
    float t[3], l[3];
    int c0[3];
    to_tex(q, x, y, z, t);
    for (int c = 0; c < 3; c++) {
        c0[c] = (int)t[c];
        l[c] = t[c] - (float)c0[c];
    }
#define TX(a, b, c) fetch(q, c0[0] + (a), c0[1] + (b), c0[2] + (c))
    float s000 = TX(0, 0, 0), s001 = TX(1, 0, 0), s010 = TX(0, 1, 0), s011 = TX(1, 1, 0);
    float s100 = TX(0, 0, 1), s101 = TX(1, 0, 1), s110 = TX(0, 1, 1), s111 = TX(1, 1, 1);
#undef TX
    /* A (1 - p) + p B, the first product fused into the sum */
#define WAV(A, B, p) fmaf((A), 1 - (p), (p) * (B))
    float s00x = WAV(s000, s001, l[0]), s01x = WAV(s010, s011, l[0]);
    float s10x = WAV(s100, s101, l[0]), s11x = WAV(s110, s111, l[0]);
    float s0yx = WAV(s00x, s01x, l[1]), s1yx = WAV(s10x, s11x, l[1]);
    return WAV(s0yx, s1yx, l[2]);
#undef WAV
}

/* forward-difference gradient (src/wall/sdf/dev.h:32-46) */
void orc_sdf_grad(const orc_sdf *q, float x, float y, float z, float g[3]) {
    float t[3];
    int c0[3];
    to_tex(q, x, y, z, t);
    for (int c = 0; c < 3; c++) c0[c] = (int)t[c];
    float v = fetch(q, c0[0], c0[1], c0[2]);
    g[0] = q->s[0] * (fetch(q, c0[0] + 1, c0[1], c0[2]) - v);
    g[1] = q->s[1] * (fetch(q, c0[0], c0[1] + 1, c0[2]) - v);
    g[2] = q->s[2] * (fetch(q, c0[0], c0[1], c0[2] + 1) - v);
}

static int iround(float x) { return (x > 0.5f) ? (int)(x + 0.5f) : (int)(x - 0.5f); }

/* nearest-texel value (src/wall/sdf/dev.h:61-66) */
float orc_sdf_cheap(const orc_sdf *q, float x, float y, float z) {
    float t[3];
    to_tex(q, x, y, z, t);
    return fetch(q, iround(t[0]), iround(t[1]), iround(t[2]));
}

int orc_sdf_far(const orc_sdf *q, float x, float y, float z) {
    return orc_sdf_cheap(q, x, y, z) <= q->cheap_threshold - 1;
}

/* ======================================================================== */
/* wall freeze: src/wall/imp/generate.h:1-60, src/wall/sdf/imp/split.h:1-33, */
/* src/wall/exch/imp.cpp:14-120                                            */
/* ======================================================================== */
/* labels: 0 bulk (s < 0), 1 wall (0 <= s <= 2), 2 deep (s > 2);
 * pp keeps the bulk particles in order, wall gets the frozen ones in order */
int orc_wall_freeze(const orc_sdf *q, int *n, float *pp, float *wall) {
    int is = 0, iw = 0;
    for (int i = 0; i < *n; i++) {
        float *p = pp + 6 * i;
        float s = orc_sdf_eval(q, p[0], p[1], p[2]);
        if (s < 0) memmove(pp + 6 * is++, p, 6 * sizeof(float));
        else if (s <= 2) memcpy(wall + 6 * iw++, p, 6 * sizeof(float));
    }
    *n = is;
    return iw;
}

/* frozen particles followed by their periodic images inside the margin box:
 * the bag received from fragment j holds the images shifted by -d_j L
 * (what the opposite fragment sent), bags in fragment order */
int orc_wall_images(const int L[3], int nw, const float *wall, float *out) {
    int k = 0, M = ORC_WALL_MARGIN;
    memcpy(out, wall, sizeof(float) * 6 * (size_t)nw);
    k = nw;
    for (int j = 0; j < 26; j++) {
        int d[3];
        orc_frag_i2d3(orc_frag_anti(j), d);
        for (int i = 0; i < nw; i++) {
            float r[3];
            int inside = 1;
            for (int c = 0; c < 3; c++) {
                r[c] = wall[6 * i + c] + (float)(d[c] * L[c]);
                double lw = 0.5 * (L[c] + 2 * M);
                if (!((double)r[c] >= -lw && (double)r[c] < lw)) inside = 0;
            }
            if (!inside) continue;
            float *o = out + 6 * k++;
            o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
            o[3] = wall[6 * i + 3]; o[4] = wall[6 * i + 4]; o[5] = wall[6 * i + 5];
        }
    }
    return k;
}

/* ======================================================================== */
/* wall velocity: src/wall/wvel/dev.h (one domain: centre coordinates are   */
/* the local ones, src/coords/dev.h:17-33)                                  */
/* ======================================================================== */
void orc_wvel_at(const orc_wvel *w, const float r[3], float v[3]) {
    v[0] = v[1] = v[2] = 0;
    if (!w) return;
    if (w->type == ORC_WVEL_CSTE) {
        v[0] = w->u[0]; v[1] = w->u[1]; v[2] = w->u[2];
    } else if (w->type == ORC_WVEL_SHEAR) {
        float d = r[w->gdir];
        if (w->half) d = d < 0.f ? d : 0.f;
        v[w->vdir] = d * w->gdot;
    } else if (w->type == ORC_WVEL_HS) {
        float r2inv = 1.f / (r[0] * r[0] + r[1] * r[1]);
        float hfac = w->hu * (1.f - r[2] * r[2] * 4 / (w->hh * w->hh));
        hfac = hfac > 0.f ? hfac : 0.f;
        v[0] = r2inv * hfac * r[0];
        v[1] = r2inv * hfac * r[1];
        v[2] = 0;
    }
}

float orc_wvel_sin_rate(float gdot, float w, float t) { return gdot * sinf(w * t); }

/* v = 2 u_wall(rw) - v (src/wall/wvel/dev.h:57-66) */
static void bounce_vel(const orc_wvel *w, const float rw[3], float v[3]) {
    float u[3];
    orc_wvel_at(w, rw, u);
    for (int c = 0; c < 3; c++) v[c] = 2 * u[c] - v[c];
}

/* ======================================================================== */
/* solvent-wall DPD: src/wall/force/dev/main.h:1-50, map/ini.h:1-50         */
/* ======================================================================== */
/* pp: n cell-ordered solvent particles; wpp: nw wall particles cell-ordered
 * on the margin box (L + 2 M) with wstarts; wall particle velocities from
 * wv at their positions (NULL: resting walls).  Forces are added to ff.
 * Returns the number of pairs. */
long orc_wall_forces(const int L[3], int n, const float *pp, const orc_sdf *q, int nw, const float *wpp,
                     const int *wstarts, orc_dpd p, float seed, float *ff, const orc_wvel *wv) {
    int M = ORC_WALL_MARGIN, W[3] = {L[0] + 2 * M, L[1] + 2 * M, L[2] + 2 * M};
    long npairs = 0;
    (void)nw;
    for (int a = 0; a < n; a++) {
        const float *pa = pp + 6 * a;
        if (orc_sdf_far(q, pa[0], pa[1], pa[2])) continue;
        int b0[3];
        for (int c = 0; c < 3; c++) {
            int v = (int)floor((double)pa[c] + 0.5 * W[c]);
            b0[c] = v < 1 ? 1 : (v > W[c] - 2 ? W[c] - 2 : v);
        }
        float acc[3] = {0, 0, 0};
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++) {
                int c0 = (b0[0] - 1) + W[0] * ((b0[1] + dy) + W[1] * (b0[2] + dz));
                for (int b = wstarts[c0]; b < wstarts[c0 + 3]; b++) {
                    const float *pw = wpp + 6 * b;
                    float wb[6] = {pw[0], pw[1], pw[2], 0, 0, 0};
                    orc_wvel_at(wv, pw, wb + 3);
                    float f[3];
                    float rnd = orc_mean0var1ii(seed, a, b);
                    if (orc_pair_force(&p, pa, wb, rnd, f)) {
                        acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
                        npairs++;
                    }
                }
            }
        ff[3 * a] += acc[0]; ff[3 * a + 1] += acc[1]; ff[3 * a + 2] += acc[2];
    }
    return npairs;
}

/* ======================================================================== */
/* bounce-back: src/wall/sdf/bounce/dev/main.h:1-114                       */
/* ======================================================================== */
static void ugrad(const orc_sdf *q, const float r[3], float g[3]) {
    orc_sdf_grad(q, r[0], r[1], r[2], g);
    float mag = sqrtf(fmaf(g[2], g[2], fmaf(g[0], g[0], g[1] * g[1])));
    if (mag > 1e-6f) { g[0] /= mag; g[1] /= mag; g[2] /= mag; }
}

static void axpy(float a, const float x[3], float y[3]) {
    for (int c = 0; c < 3; c++) y[c] = fmaf(a, x[c], y[c]);
}

static void bounce_1p(const orc_sdf *q, const orc_wvel *wv, float dt, float currsdf, float r[3], float v[3]) {
    float r0[3] = {r[0], r[1], r[2]};
    axpy(-dt, v, r0);
    if (orc_sdf_eval(q, r0[0], r0[1], r0[2]) >= 0) {
        /* rescue: step back along the gradient (:27-43) */
        float g[3];
        ugrad(q, r, g);
        float sdf0 = currsdf;
        axpy(-sdf0, g, r);
        for (int l = 8; l >= 1; --l) {
            if (orc_sdf_eval(q, r[0], r[1], r[2]) < 0) {
                bounce_vel(wv, r, v);
                return;
            }
            float jump = 1.1f * sdf0 / (float)(1 << l);
            axpy(-jump, g, r);
        }
        return;
    }
    float t = 0;
    for (int l = 0; l < 2; l++) {
        float rc[3] = {r[0], r[1], r[2]}, g[3];
        axpy(t, v, rc);
        float phi = orc_sdf_eval(q, rc[0], rc[1], rc[2]);
        orc_sdf_grad(q, rc[0], rc[1], rc[2], g);
        float dphi = fmaf(v[2], g[2], fmaf(v[0], g[0], v[1] * g[1]));
        if (fabsf(dphi) < 1e-6f) break;
        t -= phi / dphi;
        if (t < -dt) t = -dt;
        if (t > 0) t = 0;
    }
    float rw[3] = {r[0], r[1], r[2]};
    axpy(t, v, rw);
    bounce_vel(wv, rw, v);
    r[0] = rw[0]; r[1] = rw[1]; r[2] = rw[2];
    axpy(-t, v, r);
    if (orc_sdf_eval(q, r[0], r[1], r[2]) >= 0) { r[0] = r0[0]; r[1] = r0[1]; r[2] = r0[2]; }
}

void orc_wall_bounce(const orc_sdf *q, float dt, int n, float *pp, const orc_wvel *wv) {
    for (int i = 0; i < n; i++) {
        float *r = pp + 6 * i, *v = r + 3;
        if (orc_sdf_cheap(q, r[0], r[1], r[2]) >= q->cheap_threshold) {
            float s = orc_sdf_eval(q, r[0], r[1], r[2]);
            if (s >= 0) bounce_1p(q, wv, dt, s, r, v);
        }
    }
}

/* kick-drift with the constant body force, bounce-back, periodic wrap:
 * the solvent update of a wall-bounded step (src/sim/imp/run.h:43-54) */
void orc_integrate_wall_ex(const int L[3], int n, float dt, float mass, const float accel[3], const float *ff,
                           float *pp, const orc_sdf *q, const orc_wvel *wv, int wrap) {
    for (int i = 0; i < n; i++) {
        float *r = pp + 6 * i, *v = r + 3;
        const float *f = ff + 3 * i;
        for (int k = 0; k < 3; k++) {
            float ft = fmaf(mass, accel[k], f[k]);
            v[k] = fmaf(ft / mass, dt, v[k]);
            r[k] = fmaf(v[k], dt, r[k]);
        }
        if (q && orc_sdf_cheap(q, r[0], r[1], r[2]) >= q->cheap_threshold) {
            float s = orc_sdf_eval(q, r[0], r[1], r[2]);
            if (s >= 0) bounce_1p(q, wv, dt, s, r, v);
        }
        if (!wrap) continue;
        for (int k = 0; k < 3; k++) {
            float Lk = (float)L[k], h = 0.5f * Lk;
            if (r[k] < -h) r[k] = r[k] + Lk;
            if (r[k] >= h) r[k] = r[k] - Lk;
        }
    }
}

void orc_integrate_wall(const int L[3], int n, float dt, float mass, const float accel[3], const float *ff,
                        float *pp, const orc_sdf *q, const orc_wvel *wv) {
    orc_integrate_wall_ex(L, n, dt, mass, accel, ff, pp, q, wv, 1);
}

/* repulsion of object particles (src/wall/force/dev/main.h:52-91) */
void orc_wall_repulse(const orc_sdf *q, float lambda, int n, const float *pp, float *ff) {
    for (int i = 0; i < n; i++) {
        const float *r = pp + 6 * i;
        if (orc_sdf_far(q, r[0], r[1], r[2])) continue;
        float s = orc_sdf_eval(q, r[0], r[1], r[2]), g[3];
        ugrad(q, r, g);
        float val = expf(lambda * (s + 1.f)) - 1.f;
        float f0 = -(val < 1e4f ? val : 1e4f);
        ff[3 * i] += f0 * g[0]; ff[3 * i + 1] += f0 * g[1]; ff[3 * i + 2] += f0 * g[2];
    }
}
