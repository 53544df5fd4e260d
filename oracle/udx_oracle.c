/*
 * CPU oracle of the uDeviceX DPD per-timestep hot path (test infrastructure
 * only; see udx_oracle.h).  Build with -ffp-contract=off: every fused
 * multiply-add below is written out with fmaf() on purpose and nothing else
 * may be contracted.
 */
#include "udx_oracle.h"

#include <fenv.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================== */
/* KISS trunk: src/math/rnd/imp.cpp:9-34                                    */
/* ======================================================================== */
void orc_kiss_ini(orc_kiss *r, int x, int y, int z, int c) {
    r->x = (uint32_t)x; r->y = (uint32_t)y; r->z = (uint32_t)z; r->c = (uint32_t)c;
}

/* 32-bit LCG + xorshift + multiply-with-carry (src/math/rnd/imp.cpp:21-30) */
uint32_t orc_kiss_get_int(orc_kiss *r) {
    uint64_t t, a = 698769069ULL;
    r->x = 69069u * r->x + 12345u;
    r->y ^= (r->y << 13);
    r->y ^= (r->y >> 17);
    r->y ^= (r->y << 5);
    t = a * (uint64_t)r->z + (uint64_t)r->c;
    r->c = (uint32_t)(t >> 32);
    r->z = (uint32_t)t;
    return r->x + r->y + r->z;
}

/* uint32 -> float (round to nearest) divided by float(UINT32_MAX) = 2^32
 * (src/math/rnd/imp.cpp:32-34) */
float orc_kiss_get(orc_kiss *r) {
    uint32_t i = orc_kiss_get_int(r);
    float den = (float)UINT32_MAX;
    return (float)i / den;
}

void orc_kiss_fill(int x, int y, int z, int c, int n, float *out) {
    orc_kiss r;
    orc_kiss_ini(&r, x, y, z, c);
    for (int i = 0; i < n; i++) out[i] = orc_kiss_get(&r);
}

/* ======================================================================== */
/* Branch generator: src/math/rnd/dev.h:27-72                               */
/* ======================================================================== */
static const float R_GOLD   = 0.6180339887498948482f;
static const float R_SILVER = 0.4142135623730950488f;
static const float R_BRONZE = 0.00008877875787352212838853023f;
static const float R_TIN    = 0.00004602357186447026756768986f;
static const float R_SQRT2  = 1.41421356237309514547f;

/* fma rounded toward zero, computed exactly: the product of two floats is
 * exact in double, TwoSum gives the exact residual of the double sum, and the
 * truncation to float is corrected when the double sum is itself a float and
 * the residual points toward zero. */
float orc_fma_rz(float a, float b, float c) {
    double p = (double)a * (double)b;
    double cd = (double)c;
    double s = p + cd;
    double bv = s - p;
    double av = s - bv;
    double e = (p - av) + (cd - bv);           /* p + c == s + e exactly   */
    float f = (float)s;                         /* round to nearest          */
    if (fabs((double)f) > fabs(s)) f = nextafterf(f, 0.0f);  /* truncate   */
    if ((double)f == s && e != 0.0) {
        /* s is a float; the exact value is s + e */
        if ((s > 0 && e < 0) || (s < 0 && e > 0)) f = nextafterf(f, 0.0f);
    }
    return f;
}

/* fma under the hardware round-toward-zero mode: cross-check of the exact
 * emulation above (needs a correctly rounded fmaf honouring the mode). */
float orc_fma_rz_fenv(float a, float b, float c) {
    volatile float va = a, vb = b, vc = c;
    int old = fegetround();
    fesetround(FE_TOWARDZERO);
    volatile float r = fmaf(va, vb, vc);
    fesetround(old);
    return r;
}

static float rem1(float r) { return r - floorf(r); }

/* Nine rounds of the Chebyshev T4 map x <- 8x^4 - 8x^2 + 1 with RZ fma,
 * the square itself rounded to nearest (src/math/rnd/dev.h:42-47, N=18). */
static float logistic18(float x) {
    for (int k = 0; k < 9; k++) {
        float x2 = x * x;
        x = orc_fma_rz(orc_fma_rz(8.0f, x2, -8.0f), x2, 1.0f);
    }
    return x;
}

/* p = frac((u&1023)*gold + u*bronze + (v&1023)*silver + v*tin), left to
 * right, each multiply fused into the following addition and the second
 * product of the first addition rounded on its own:
 *   s = fma(u&1023, gold, u*bronze); s = fma(v&1023, silver, s);
 *   s = fma(v, tin, s).
 * (src/math/rnd/dev.h:57-60 under FMA contraction; the contraction the CUDA
 * compiler applied is not recorded in the reference, this is the one this
 * build defines.) */
static float low_discrepancy(float fu_lo, float fu, float fv_lo, float fv) {
    float s = fmaf(fu_lo, R_GOLD, fu * R_BRONZE);
    s = fmaf(fv_lo, R_SILVER, s);
    s = fmaf(fv, R_TIN, s);
    return rem1(s);
}

float orc_mean0var1ii(float seed, int u, int v) {
    float p = low_discrepancy((float)(u & 0x3FF), (float)u, (float)(v & 0x3FF), (float)v);
    float l = logistic18(seed - p);
    return l * R_SQRT2;
}

float orc_mean0var1uu(float seed, uint32_t u, uint32_t v) {
    float p = low_discrepancy((float)(u & 0x3FFu), (float)u, (float)(v & 0x3FFu), (float)v);
    float l = logistic18(seed - p);
    return l * R_SQRT2;
}

void orc_rng_ii_batch(float seed, int n, const int *u, const int *v, float *out) {
    for (int i = 0; i < n; i++) out[i] = orc_mean0var1ii(seed, u[i], v[i]);
}

void orc_rng_uu_batch(float seed, int n, const uint32_t *u, const uint32_t *v, float *out) {
    for (int i = 0; i < n; i++) out[i] = orc_mean0var1uu(seed, u[i], v[i]);
}

void orc_fma_rz_batch(int n, const float *a, const float *b, const float *c, float *out, int fenv) {
    for (int i = 0; i < n; i++)
        out[i] = fenv ? orc_fma_rz_fenv(a[i], b[i], c[i]) : orc_fma_rz(a[i], b[i], c[i]);
}

/* ======================================================================== */
/* Fragments: src/frag/dev.h:19-81                                          */
/* ======================================================================== */
void orc_frag_i2d3(int i, int d[3]) {
    d[0] = (i + 2) % 3 - 1;
    d[1] = (i / 3 + 2) % 3 - 1;
    d[2] = (i / 9 + 2) % 3 - 1;
}

int orc_frag_d2i(int x, int y, int z) {
    return ((x + 2) % 3) + 3 * ((y + 2) % 3) + 9 * ((z + 2) % 3);
}

int orc_frag_anti(int i) {
    int d[3];
    orc_frag_i2d3(i, d);
    return orc_frag_d2i(-d[0], -d[1], -d[2]);
}

static int frag_ncell(const int L[3], int fid) {
    int d[3];
    orc_frag_i2d3(fid, d);
    return (d[0] ? 1 : L[0]) * (d[1] ? 1 : L[1]) * (d[2] ? 1 : L[2]);
}

/* fragment box origin/extent (src/exch/flu/dev/map.h:1-11) */
static void frag_box(const int L[3], int fid, int org[3], int ext[3]) {
    int d[3];
    orc_frag_i2d3(fid, d);
    for (int c = 0; c < 3; c++) {
        org[c] = (d[c] == 1) ? L[c] - 1 : 0;
        ext[c] = (d[c] == 0) ? L[c] : 1;
    }
}

int orc_frag_table_size(const int L[3]) {
    int s = 0;
    for (int f = 0; f < 26; f++) s += frag_ncell(L, f) + 1;
    return s;
}

/* src/exch/flu/dev/map.h:22-98: counts gathered in fragment-cell order (x
 * fastest), then an exclusive scan per fragment */
void orc_frag_tables(const int L[3], const int *counts, int *fss, int *fss_off) {
    int off = 0;
    for (int f = 0; f < 26; f++) {
        int org[3], ext[3], nc = frag_ncell(L, f), acc = 0;
        frag_box(L, f, org, ext);
        fss_off[f] = off;
        for (int h = 0; h < nc; h++) {
            int x = org[0] + h % ext[0];
            int y = org[1] + (h / ext[0]) % ext[1];
            int z = org[2] + h / (ext[0] * ext[1]);
            int cid = x + L[0] * (y + L[1] * z);
            fss[off + h] = acc;
            acc += counts[cid];
        }
        fss[off + nc] = acc;
        off += nc + 1;
    }
    fss_off[26] = off;
}

/* halo trunks on one rank (src/fluforces/imp/ini.h:11-58): seed_base = 0 and
 * seed_offset from the direction with the is_plus sign convention. */
static int is_plus(const int d[3]) {
    int s = d[0] + d[1] + d[2];
    return (s > 0) || (s == 0 && d[0] > 0) || (d[0] == 0 && d[1] > 0) || (d[1] == 0 && d[2] > 0);
}

void orc_seeds_ini(orc_seeds *s) {
    for (int f = 0; f < 26; f++) {
        int d[3], v[3], sign, seed;
        orc_frag_i2d3(f, d);
        sign = 2 * is_plus(d) - 1;
        for (int c = 0; c < 3; c++) v[c] = 1 + sign * d[c];
        seed = v[0] + 3 * (v[1] + 3 * v[2]);
        orc_kiss_ini(&s->t[f], 390 + seed, seed + 615, 12309, 23094);
    }
    orc_kiss_ini(&s->t[26], 0, 0, 0, 0);
}

void orc_seeds_next(orc_seeds *s, float out[27]) {
    for (int f = 0; f < 27; f++) out[f] = orc_kiss_get(&s->t[f]);
}

void orc_seeds_table(int nsteps, float *out) {
    orc_seeds s;
    orc_seeds_ini(&s);
    for (int k = 0; k < nsteps; k++) orc_seeds_next(&s, out + 27 * k);
}

/* ======================================================================== */
/* gen_random: pre/units/gen_random.cpp.  srand() does not seed drand48, so  */
/* the generator starts from glibc's zero-initialised state (X0 = 0).        */
/* ======================================================================== */
static double drand48_next(uint64_t *x) {
    *x = (0x5DEECE66DULL * *x + 0xBULL) & ((1ULL << 48) - 1);
    return (double)*x / 281474976710656.0;
}

static float gen_rnd(uint64_t *st, int L) {
    return (float)(L * (-0.5f + 0.9999f * drand48_next(st)));
}

void orc_gen_random(int X, int Y, int Z, int n, int roundtrip, float *pp) {
    uint64_t st = 0;
    for (int i = 0; i < n; i++) {
        float *p = pp + 6 * i;
        p[0] = gen_rnd(&st, X);
        p[1] = gen_rnd(&st, Y);
        p[2] = gen_rnd(&st, Z);
        p[3] = gen_rnd(&st, 2);
        p[4] = gen_rnd(&st, 2);
        p[5] = gen_rnd(&st, 2);
        if (roundtrip) {
            char buf[64];
            for (int c = 0; c < 6; c++) {
                snprintf(buf, sizeof buf, "%f", p[c]);
                p[c] = strtof(buf, NULL);
            }
        }
    }
}

/* ======================================================================== */
/* Cell list: src/clist/dev/main.h:6-101                                    */
/* ======================================================================== */
/* Cell of a position over [-L/2, L/2) with unit cells, x fastest.  The
 * index is floor(r + L/2) evaluated exactly in double (the reference's
 * (int)((double)r + L/2), src/clist/dev/main.h:21-24, for even L).
 * Returns -1 for positions outside the box. */
int orc_cell_of(const int L[3], const float r[3], int c[3]) {
    for (int k = 0; k < 3; k++) {
        double v = floor((double)r[k] + 0.5 * (double)L[k]);
        if (!(v >= 0.0 && v < (double)L[k])) return -1;
        c[k] = (int)v;
    }
    return c[0] + L[0] * (c[1] + L[1] * c[2]);
}

int orc_slab_of(int L, int sub, float x) {
    double v = floor(((double)x + 0.5 * (double)L) * (double)sub);
    if (!(v >= 0.0 && v < (double)L * sub)) return -1;
    return (int)v;
}

/* Counting sort by slab id (cell * sub + slab): the order of the build's
 * cell lists.  The reference's in-cell order comes from atomicAdd
 * (src/clist/dev/main.h:64-65) and is not reproducible; here it is defined
 * by the x slab, then by input index. */
int orc_clist_build_sub(const int L[3], int sub, int n, const float *pp_in, int *starts, int *counts,
                        int *bstarts, float *pp_out, int *order) {
    int nc = L[0] * L[1] * L[2], nb = nc * sub;
    int *bid = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *bcnt = (int *)calloc((size_t)nb + 1, sizeof(int));
    int *bst = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
    for (int i = 0; i < n; i++) {
        int c[3];
        const float *r = pp_in + 6 * i;
        int cid = orc_cell_of(L, r, c);
        bid[i] = -1;
        if (cid < 0) continue;
        int sx = orc_slab_of(L[0], sub, r[0]);
        bid[i] = cid * sub + (sx - c[0] * sub);
        bcnt[bid[i]]++;
    }
    int acc = 0;
    for (int b = 0; b < nb; b++) { bst[b] = acc; acc += bcnt[b]; }
    bst[nb] = acc;
    for (int c = 0; c < nc; c++) {
        starts[c] = bst[c * sub];
        counts[c] = bst[(c + 1) * sub] - bst[c * sub];
    }
    starts[nc] = acc;
    if (bstarts) memcpy(bstarts, bst, sizeof(int) * ((size_t)nb + 1));
    int *fill = (int *)calloc((size_t)nb, sizeof(int));
    for (int i = 0; i < n; i++) {
        if (bid[i] < 0) continue;
        int k = bst[bid[i]] + fill[bid[i]]++;
        if (order) order[k] = i;
        if (pp_out) memcpy(pp_out + 6 * k, pp_in + 6 * i, 6 * sizeof(float));
    }
    free(fill);
    free(bst);
    free(bcnt);
    free(bid);
    return acc;
}

int orc_clist_build(const int L[3], int n, const float *pp_in, int *starts, int *counts,
                    float *pp_out, int *order) {
    return orc_clist_build_sub(L, 1, n, pp_in, starts, counts, NULL, pp_out, order);
}

/* ======================================================================== */
/* DPD pair force: src/pair/dev.h:12-47,110-139                             */
/* ======================================================================== */
float orc_dpd_sigma(float kBT, float g, float dt) {
    return sqrtf(2 * kBT * g / dt);
}

/* w_R = (1 - r)^spow; 0.25, 0.5 and 1 are evaluated with square roots */
static float wr_pow(float rm, float spow) {
    if (spow == 0.25f) return sqrtf(sqrtf(rm));
    if (spow == 0.5f)  return sqrtf(rm);
    if (spow == 1.0f)  return rm;
    return powf(rm, spow);
}

/* a = (x, y, z, vx, vy, vz) of the particle the force acts on */
int orc_pair_force(const orc_dpd *p, const float *a, const float *b, float rnd, float *f) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    float r2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
    f[0] = f[1] = f[2] = 0.0f;
    if (r2 >= 1.0f) return 0;
    float ex, ey, ez, r;
    if (r2 < 1e-6f) {
        ex = ey = ez = r = 0.0f;
    } else {
        /* the reference's r = r2 * rsqrtf(r2) evaluated as the correctly
         * rounded square root, 1/r correctly rounded */
        r = sqrtf(r2);
        float invr = 1.0f / r;
        ex = dx * invr; ey = dy * invr; ez = dz * invr;
    }
    float dvx = a[3] - b[3], dvy = a[4] - b[4], dvz = a[5] - b[5];
    float ev = fmaf(ez, dvz, fmaf(ex, dvx, ey * dvy));
    float rm = 1.0f - r;
    rm = rm > 0.0f ? rm : 0.0f;
    float wr = wr_pow(rm, p->spow);
    float f0 = fmaf(-p->g * wr, ev, p->s * rnd) * wr;
    f0 = fmaf(p->a, rm, f0);
    f[0] = f0 * ex; f[1] = f0 * ey; f[2] = f0 * ez;
    return 1;
}

/* ======================================================================== */
/* DPD forces: half-shell of src/fluforces/bulk/dev/main.h:23-97 plus the     */
/* one-sided halo of src/fluforces/halo/dev/main.h:7-83, restated as one      */
/* full-shell gather per particle in a fixed order: the pairs inside the box  */
/* first, then the pairs across the periodic boundary (as the reference adds  */
/* the halo forces after the bulk's).                                        */
/* ======================================================================== */
typedef struct {
    const int *L, *starts, *fss, *fss_off;
} orc_grid;

/* position of particle idx (cell coords c) in fragment fid's packed list
 * (src/exch/flu/dev/pack.h:2-40) */
static int frag_pos(const orc_grid *g, int fid, const int c[3], int idx) {
    int org[3], ext[3];
    frag_box(g->L, fid, org, ext);
    int h = (c[0] - org[0]) + ext[0] * ((c[1] - org[1]) + ext[1] * (c[2] - org[2]));
    int cid = c[0] + g->L[0] * (c[1] + g->L[1] * c[2]);
    return g->fss[g->fss_off[fid] + h] + (idx - g->starts[cid]);
}

/* so (may be NULL): the particle's stress, sum over its pairs of
 * 1/2 f (x) dr as (xx, xy, xz, yy, yz, zz) with s_ab = 1/2 f_a dr_b
 * (src/pair/dev.h:87-100) */
/* the pair force rounded to a multiple of 2^-32 (unchanged for |f| >= 2^-9,
 * whose float ulp is at least 2^-32); sums of such values below 2^21 are
 * exact in double, whatever the order of the additions */
static double quant32(float f) {
    return (double)(rintf(f * 4294967296.0f) * 2.3283064365386963e-10f);
}

/* q64 = 1: the forces are summed exactly (quant32 terms in double) and the
 * sum rounded to float once, the production kernel's definition; q64 = 2:
 * the float pair forces summed in double, added to fo64 (the reference value
 * every summation order approximates); 0: summed in float in the traversal
 * order */
static long gather_one(const orc_grid *g, int n, const float *pp, orc_dpd p, const float seeds[27],
                       int periodic, int i, float *fo, float *so, int q64, double *fo64) {
    const int *L = g->L;
    const float *pi = pp + 6 * i;
    int ci[3];
    long npairs = 0;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    double dacc[3] = {0.0, 0.0, 0.0};
    float sacc[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    (void)n;
    if (orc_cell_of(L, pi, ci) < 0) return 0;
    /* pass 0: the cells inside the box (the bulk, src/fluforces/bulk);
     * pass 1: the cells across the periodic boundary (the halo,
     * src/fluforces/halo), each pass in (dz, dy, dx) order */
    for (int pass = 0; pass < 2; pass++)
    for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
        int o[3] = {dx, dy, dz}, cj[3], w[3], skip = 0;
        for (int k = 0; k < 3; k++) {
            int c = ci[k] + o[k];
            w[k] = c < 0 ? -1 : (c >= L[k] ? 1 : 0);
            if (w[k] && !periodic) skip = 1;
            cj[k] = c - w[k] * L[k];
        }
        if (skip) continue;
        int wrapped = w[0] | w[1] | w[2];
        if ((wrapped != 0) != (pass == 1)) continue;
        int cid = cj[0] + L[0] * (cj[1] + L[1] * cj[2]);
        /* the local particle is moved into the neighbour's frame
         * (src/fluforces/halo/dev/main.h:40-48) */
        float a[6];
        memcpy(a, pi, sizeof a);
        for (int k = 0; k < 3; k++) if (w[k]) a[k] = a[k] - (float)(w[k] * L[k]);
        int fid = 0, fanti = 0, aid = 0, mask = 0;
        if (wrapped) {
            fid = orc_frag_d2i(w[0], w[1], w[2]);
            fanti = orc_frag_anti(fid);
            aid = frag_pos(g, fid, ci, i);
            mask = fid < fanti;
        }
        for (int j = g->starts[cid]; j < g->starts[cid + 1]; j++) {
            if (j == i) continue;
            const float *pj = pp + 6 * j;
            float ddx = a[0] - pj[0], ddy = a[1] - pj[1], ddz = a[2] - pj[2];
            float r2 = fmaf(ddz, ddz, fmaf(ddx, ddx, ddy * ddy));
            if (!(r2 < 1.0f)) continue;
            float rnd;
            if (!wrapped) {
                int hi = i > j ? i : j, lo = i > j ? j : i;
                rnd = orc_mean0var1ii(seeds[26], hi, lo);
            } else {
                int bid = frag_pos(g, fanti, cj, j);
                uint32_t a1 = (uint32_t)(mask ? aid : bid), a2 = (uint32_t)(mask ? bid : aid);
                rnd = orc_mean0var1uu(seeds[fid], a1, a2);
            }
            float f[3];
            orc_pair_force(&p, a, pj, rnd, f);
            if (q64 == 1) {
                dacc[0] += quant32(f[0]); dacc[1] += quant32(f[1]); dacc[2] += quant32(f[2]);
            } else if (q64 == 2) {
                dacc[0] += (double)f[0]; dacc[1] += (double)f[1]; dacc[2] += (double)f[2];
            } else {
                acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
            }
            if (so) {
                sacc[0] += 0.5f * f[0] * ddx;
                sacc[1] += 0.5f * f[0] * ddy;
                sacc[2] += 0.5f * f[0] * ddz;
                sacc[3] += 0.5f * f[1] * ddy;
                sacc[4] += 0.5f * f[1] * ddz;
                sacc[5] += 0.5f * f[2] * ddz;
            }
            npairs++;
        }
    }
    if (q64 == 2) {
        fo64[0] += dacc[0]; fo64[1] += dacc[1]; fo64[2] += dacc[2];
        return npairs;
    }
    if (q64 == 1) {
        acc[0] = (float)dacc[0]; acc[1] = (float)dacc[1]; acc[2] = (float)dacc[2];
    }
    fo[0] += acc[0]; fo[1] += acc[1]; fo[2] += acc[2];
    if (so)
        for (int k = 0; k < 6; k++) so[k] += sacc[k];
    return npairs;
}

long orc_dpd_forces_stress(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                           orc_dpd p, const float seeds[27], int periodic, int nthreads, float *ff, float *ss) {
    int tsz = orc_frag_table_size(L);
    int *fss = (int *)malloc(sizeof(int) * (size_t)tsz);
    int fss_off[27];
    orc_frag_tables(L, counts, fss, fss_off);
    orc_grid g = {L, starts, fss, fss_off};
    long tot = 0;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:tot) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++)
        tot += gather_one(&g, n, pp, p, seeds, periodic, i, ff + 3 * i, ss ? ss + 6 * (size_t)i : NULL, 0, NULL);
    free(fss);
    return tot;
}

long orc_dpd_forces(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                    orc_dpd p, const float seeds[27], int periodic, int nthreads, float *ff) {
    int tsz = orc_frag_table_size(L);
    int *fss = (int *)malloc(sizeof(int) * (size_t)tsz);
    int fss_off[27];
    orc_frag_tables(L, counts, fss, fss_off);
    orc_grid g = {L, starts, fss, fss_off};
    long tot = 0;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:tot) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++) tot += gather_one(&g, n, pp, p, seeds, periodic, i, ff + 3 * i, NULL, 0, NULL);
    free(fss);
    return tot;
}

/* The same pairs and pair forces with exact sums (the production kernel's
 * definition: each particle's force is the exact sum of its pair forces
 * rounded to multiples of 2^-32, rounded to float once; added to ff).  The
 * traversal order does not matter here: any order gives the same sum. */
long orc_dpd_forces_sum64(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                          orc_dpd p, const float seeds[27], int periodic, int nthreads, float *ff) {
    int tsz = orc_frag_table_size(L);
    int *fss = (int *)malloc(sizeof(int) * (size_t)tsz);
    int fss_off[27];
    orc_frag_tables(L, counts, fss, fss_off);
    orc_grid g = {L, starts, fss, fss_off};
    long tot = 0;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:tot) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++) tot += gather_one(&g, n, pp, p, seeds, periodic, i, ff + 3 * i, NULL, 1, NULL);
    free(fss);
    return tot;
}

/* The same pairs, the float pair forces summed in double and ADDED to ff64
 * (3 per particle): the value every float summation order approximates. */
long orc_dpd_forces_f64(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                        orc_dpd p, const float seeds[27], int periodic, int nthreads, double *ff64) {
    int tsz = orc_frag_table_size(L);
    int *fss = (int *)malloc(sizeof(int) * (size_t)tsz);
    int fss_off[27];
    orc_frag_tables(L, counts, fss, fss_off);
    orc_grid g = {L, starts, fss, fss_off};
    long tot = 0;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:tot) num_threads(nthreads)
#endif
    for (int i = 0; i < n; i++) tot += gather_one(&g, n, pp, p, seeds, periodic, i, NULL, NULL, 2, ff64 + 3 * i);
    free(fss);
    return tot;
}

/* Half-shell pass of particle i, the reference's own traversal
 * (src/fluforces/bulk/dev/main.h:23-97 for pairs inside the box: each
 * unordered pair once, from its larger index, the reaction added to the
 * smaller; src/fluforces/halo/dev/main.h:7-83 for pairs across the periodic
 * boundary: one-sided, each side with its own fragment noise).  Forces go
 * to the thread's buffer f. */
static long half_one(const orc_grid *g, const float *pp, orc_dpd p, const float seeds[27], int i, float *f,
                     double *fd) {
    const int *L = g->L;
    const float *pi = pp + 6 * i;
    int ci[3];
    long npairs = 0;
    if (orc_cell_of(L, pi, ci) < 0) return 0;
    const int cid_i = ci[0] + L[0] * (ci[1] + L[1] * ci[2]);
    for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
        int o[3] = {dx, dy, dz}, cj[3], w[3];
        for (int k = 0; k < 3; k++) {
            int c = ci[k] + o[k];
            w[k] = c < 0 ? -1 : (c >= L[k] ? 1 : 0);
            cj[k] = c - w[k] * L[k];
        }
        const int wrapped = w[0] | w[1] | w[2];
        const int cid = cj[0] + L[0] * (cj[1] + L[1] * cj[2]);
        if (!wrapped && cid > cid_i) continue;      /* the other half of the shell */
        float a[6];
        memcpy(a, pi, sizeof a);
        for (int k = 0; k < 3; k++) if (w[k]) a[k] = a[k] - (float)(w[k] * L[k]);
        int fid = 0, aid = 0, mask = 0;
        if (wrapped) {
            fid = orc_frag_d2i(w[0], w[1], w[2]);
            aid = frag_pos(g, fid, ci, i);
            mask = fid < orc_frag_anti(fid);
        }
        const int jend = (!wrapped && cid == cid_i) ? i : g->starts[cid + 1];
        for (int j = g->starts[cid]; j < jend; j++) {
            if (j == i) continue;
            const float *pj = pp + 6 * j;
            float ddx = a[0] - pj[0], ddy = a[1] - pj[1], ddz = a[2] - pj[2];
            float r2 = fmaf(ddz, ddz, fmaf(ddx, ddx, ddy * ddy));
            if (!(r2 < 1.0f)) continue;
            float rnd, fij[3];
            if (!wrapped) {
                rnd = orc_mean0var1ii(seeds[26], i, j);
            } else {
                int bid = frag_pos(g, orc_frag_anti(fid), cj, j);
                uint32_t a1 = (uint32_t)(mask ? aid : bid), a2 = (uint32_t)(mask ? bid : aid);
                rnd = orc_mean0var1uu(seeds[fid], a1, a2);
            }
            orc_pair_force(&p, a, pj, rnd, fij);
            if (fd) {
                /* exact sums: quantised terms in double */
                for (int k = 0; k < 3; k++) fd[3 * i + k] += quant32(fij[k]);
                if (!wrapped)
                    for (int k = 0; k < 3; k++) fd[3 * j + k] -= quant32(fij[k]);
            } else {
                for (int k = 0; k < 3; k++) f[3 * i + k] += fij[k];
                if (!wrapped)
                    for (int k = 0; k < 3; k++) f[3 * j + k] -= fij[k];
            }
            npairs += wrapped ? 1 : 2;
        }
    }
    return npairs;
}

long orc_dpd_forces_half(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                         orc_dpd p, const float seeds[27], int nthreads, float *ff) {
    int tsz = orc_frag_table_size(L);
    int *fss = (int *)malloc(sizeof(int) * (size_t)tsz);
    int fss_off[27];
    orc_frag_tables(L, counts, fss, fss_off);
    orc_grid g = {L, starts, fss, fss_off};
    if (nthreads < 1) nthreads = 1;
    const size_t m = 3 * (size_t)n;
    float *buf = (float *)calloc((size_t)nthreads * m + 1, sizeof(float));
    long tot = 0;
#ifdef _OPENMP
    #pragma omp parallel num_threads(nthreads) reduction(+:tot)
#endif
    {
#ifdef _OPENMP
        float *mine = buf + m * (size_t)omp_get_thread_num();
        #pragma omp for schedule(dynamic, 256)
#else
        float *mine = buf;
#endif
        for (int i = 0; i < n; i++) tot += half_one(&g, pp, p, seeds, i, mine, NULL);
    }
    /* the threads' buffers added in thread order */
#ifdef _OPENMP
    #pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (long k = 0; k < (long)m; k++) {
        float s = 0.0f;
        for (int t = 0; t < nthreads; t++) s += buf[(size_t)t * m + k];
        ff[k] += s;
    }
    free(buf);
    free(fss);
    return tot;
}

/* The half-shell traversal with exact sums: per-thread FP64 buffers of
 * quantised pair forces (each pair once inside the box, reaction to the
 * smaller index), added in double and rounded to float once; ADDED to ff.
 * Equal bit for bit to orc_dpd_forces_sum64 whatever the thread count, since
 * every sum is exact. */
long orc_dpd_forces_half_sum64(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                               orc_dpd p, const float seeds[27], int nthreads, float *ff) {
    int tsz = orc_frag_table_size(L);
    int *fss = (int *)malloc(sizeof(int) * (size_t)tsz);
    int fss_off[27];
    orc_frag_tables(L, counts, fss, fss_off);
    orc_grid g = {L, starts, fss, fss_off};
    if (nthreads < 1) nthreads = 1;
    const size_t m = 3 * (size_t)n;
    double *buf = (double *)calloc((size_t)nthreads * m + 1, sizeof(double));
    long tot = 0;
#ifdef _OPENMP
    #pragma omp parallel num_threads(nthreads) reduction(+:tot)
#endif
    {
#ifdef _OPENMP
        double *mine = buf + m * (size_t)omp_get_thread_num();
        #pragma omp for schedule(dynamic, 256)
#else
        double *mine = buf;
#endif
        for (int i = 0; i < n; i++) tot += half_one(&g, pp, p, seeds, i, NULL, mine);
    }
    for (size_t k = 0; k < m; k++) {
        double s = 0.0;
        for (int t = 0; t < nthreads; t++) s += buf[(size_t)t * m + k];
        ff[k] += (float)s;
    }
    free(buf);
    free(fss);
    return tot;
}

/* ======================================================================== */
/* Integrator: kick-drift of src/scheme/move/dev/vv.h:1-10, with the          */
/* constant body force f += m a of src/scheme/force/dev/main.h:1-4 folded     */
/* in, then the periodic wrap of the single-domain redistribution.           */
/* ======================================================================== */
static float wrap_coord(float x, float L) {
    float h = 0.5f * L;
    if (x < -h) x = x + L;
    if (x >= h) x = x - L;
    return x;
}

void orc_integrate_ex(const int L[3], int n, float dt, float mass, const float accel[3], const float *ff,
                      float *pp, int wrap) {
    for (int i = 0; i < n; i++) {
        float *r = pp + 6 * i, *v = r + 3;
        const float *f = ff + 3 * i;
        for (int k = 0; k < 3; k++) {
            float ft = fmaf(mass, accel[k], f[k]);
            v[k] = fmaf(ft / mass, dt, v[k]);
            r[k] = fmaf(v[k], dt, r[k]);
            if (wrap) r[k] = wrap_coord(r[k], (float)L[k]);
        }
    }
}

void orc_integrate(const int L[3], int n, float dt, float mass, const float accel[3],
                   const float *ff, float *pp) {
    orc_integrate_ex(L, n, dt, mass, accel, ff, pp, 1);
}

void orc_wrap(const int L[3], int n, float *pp) {
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) pp[6 * i + k] = wrap_coord(pp[6 * i + k], (float)L[k]);
}

/* kick-drift without wrap, for membrane vertices (src/scheme/move/dev/vv.h:1-10) */
void orc_move(int n, float dt, float mass, const float *ff, float *pp) {
    for (int i = 0; i < n; i++) {
        float *r = pp + 6 * i, *v = r + 3;
        const float *f = ff + 3 * i;
        for (int k = 0; k < 3; k++) {
            v[k] = fmaf(f[k] / mass, dt, v[k]);
            r[k] = fmaf(v[k], dt, r[k]);
        }
    }
}

/* ======================================================================== */
/* Diagnostics: src/io/diag/part/imp/main.h:37-63 in double                 */
/* ======================================================================== */
void orc_diag(int n, const float *pp, double out[5]) {
    double m[3] = {0, 0, 0}, ke = 0, mx = 0;
    for (int i = 0; i < n; i++) {
        const float *v = pp + 6 * i + 3;
        double v2 = 0;
        for (int k = 0; k < 3; k++) { m[k] += v[k]; v2 += (double)v[k] * v[k]; }
        ke += 0.5 * v2;
        if (v2 > mx) mx = v2;
    }
    out[0] = m[0]; out[1] = m[1]; out[2] = m[2];
    out[3] = n > 0 ? ke / (1.5 * n) : 0.0;
    out[4] = mx;
}

long orc_flu_step(const int L[3], int n, float *pp, float *scratch_pp, int *starts, int *counts,
                  float *ff, orc_dpd p, const float seeds[27], float dt, float mass,
                  const float accel[3], int nthreads) {
    int nout = orc_clist_build_sub(L, ORC_SOLVENT_SUB, n, pp, starts, counts, NULL, scratch_pp, NULL);
    memcpy(pp, scratch_pp, sizeof(float) * 6 * (size_t)nout);
    memset(ff, 0, sizeof(float) * 3 * (size_t)nout);
    long np = orc_dpd_forces(L, nout, pp, starts, counts, p, seeds, 1, nthreads, ff);
    orc_integrate(L, nout, dt, mass, accel, ff, pp);
    return np / 2;
}

/* ======================================================================== */
/* Exhaustive check of the kernels' correctly rounded square root           */
/* ======================================================================== */
static float f_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bits_f(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float dev_sqrt_fix(float x, float s) {
    const float sd = f_bits(bits_f(s) - 1u), su = f_bits(bits_f(s) + 1u);
    const float rd = fmaf(-sd, s, x), ru = fmaf(-su, s, x);
    const float t = rd <= 0.f ? sd : s;
    return ru > 0.f ? su : t;
}

long orc_verify_sqrt_rn(float lo, float hi, int nthreads, int correct) {
    const uint32_t a = bits_f(lo), b = bits_f(hi);
    long bad = 0;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
    #pragma omp parallel for schedule(static, 1 << 16) reduction(+:bad) num_threads(nthreads)
#endif
    for (long u = a; u <= (long)b; u++) {
        const float x = f_bits((uint32_t)u), want = sqrtf(x);
        for (int d = -1; d <= 1; d++) {
            const float s = f_bits(bits_f(want) + (uint32_t)d);
            if ((correct ? dev_sqrt_fix(x, s) : s) != want) bad++;
        }
    }
    return bad;
}
