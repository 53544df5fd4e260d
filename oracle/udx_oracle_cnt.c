/*
 * CPU oracle of the membrane-membrane contact forces (src/cnt) for one
 * object type on one periodic domain.
 *
 * TEST INFRASTRUCTURE ONLY (see udx_oracle.h).
 *
 * Pair force: DPD plus the capped repulsive Lennard-Jones term of
 * src/pair/dev.h:34-83 (f_LJ = 24 e / r (s/r)^6 (2 (s/r)^6 - 1) clipped to
 * [0, 1e4]).  Cells: floor(x) + L/2 clamped into the domain (vertices outside
 * it are projected onto the boundary cells, src/cnt/imp/main.h:33-44).
 *  - bulk (src/cnt/dev/bulk.h): every vertex pair a > b whose cells differ by
 *    at most one per axis, noise mean0var1ii(seed_bulk, a, b), +f on a, -f
 *    on b;
 *  - halo (src/cnt/dev/halo.h): the vertices outside the domain, shifted by a
 *    period into the neighbour fragment (the object halo of src/exch/obj:
 *    receive fragment i = 0..25, vertices ascending; index h in that list),
 *    against every vertex whose cell lies within one cell of the image's
 *    (unclamped) cell, noise mean0var1ii(seed_halo, h, b), +f on the imaged
 *    vertex, -f on b.
 * The reference adds with float atomics; here each vertex sums, in one
 * accumulator: its bulk pairs (its 3x3x3 cell block in z, y, x order,
 * ascending vertex within a cell), then the images acting on it (the same
 * block of an image cell list, ascending image within a cell), then its own
 * images' pairs (its images in list order, each over the image's block), and
 * adds the sum to its force once.
 */
#include "udx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static float cap(float x, float lo, float hi) { return x > hi ? hi : (x < lo ? lo : x); }

/* DPD + LJ force on a from b (positions and velocities, 6 floats each) */
int orc_cnt_pair(const orc_dpd_lj *p, const float *a, const float *b, float rnd, float *f) {
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    const float r2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
    f[0] = f[1] = f[2] = 0.0f;
    if (r2 >= 1.0f) return 0;
    float ex, ey, ez, r, invr;
    if (r2 < 1e-6f) {
        ex = ey = ez = r = invr = 0.0f;
    } else {
        r = sqrtf(r2);
        invr = 1.0f / r;
        ex = dx * invr;
        ey = dy * invr;
        ez = dz * invr;
    }
    const float dvx = a[3] - b[3], dvy = a[4] - b[4], dvz = a[5] - b[5];
    const float ev = fmaf(ez, dvz, fmaf(ex, dvx, ey * dvy));
    float rm = 1.0f - r;
    rm = rm > 0.0f ? rm : 0.0f;
    const float wr = p->spow == 0.25f ? sqrtf(sqrtf(rm)) : (p->spow == 0.5f ? sqrtf(rm) : (p->spow == 1.0f ? rm : powf(rm, p->spow)));
    float f0 = fmaf(-p->g * wr, ev, p->s * rnd) * wr;
    f0 = fmaf(p->a, rm, f0);
    const float t2 = p->ljs * p->ljs * invr * invr;
    const float t4 = t2 * t2, t6 = t4 * t2;
    const float flj = cap(p->lje * 24 * invr * t6 * (2 * t6 - 1), 0, 1e4f);
    f0 = f0 + flj;
    f[0] = f0 * ex;
    f[1] = f0 * ey;
    f[2] = f0 * ez;
    return 1;
}

/* floor(x) + L/2 clamped into the domain (vertices) or into [-2, L + 1]
 * (image centres: a block that far out reaches no cell of the domain) */
static int ccell(float x, int L, int clampit) {
    const float f = floorf(x);
    const int lo = clampit ? 0 : -2, hi = clampit ? L - 1 : L + 1;
    if (!(f >= (float)(lo - L / 2))) return lo;
    if (!(f <= (float)(hi - L / 2))) return hi;
    return L / 2 + (int)f;
}

/* counting sort of points by clamped cell, ascending index within a cell */
static void cell_list(const int L[3], int n, const float *r, int stride, int *starts, int *order) {
    const int nc = L[0] * L[1] * L[2];
    int *cnt = (int *)calloc((size_t)nc + 1, sizeof(int));
    int *cid = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        const float *p = r + (size_t)stride * i;
        cid[i] = ccell(p[0], L[0], 1) + L[0] * (ccell(p[1], L[1], 1) + L[1] * ccell(p[2], L[2], 1));
        cnt[cid[i] + 1]++;
    }
    for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
    memcpy(starts, cnt, sizeof(int) * ((size_t)nc + 1));
    for (int i = 0; i < n; i++) order[cnt[cid[i]]++] = i;
    free(cnt);
    free(cid);
}

static int emap_code(const int L[3], const float *r, int d[3]) {
    for (int c = 0; c < 3; c++) d[c] = -1 + (r[c] >= (float)(-L[c] / 2)) + (r[c] >= (float)(L[c] / 2));
    return d[0] | d[1] | d[2];
}

static int sends_to(const int d[3], const int s[3]) {
    int nz = 0;
    for (int c = 0; c < 3; c++) {
        if (s[c] != 0 && s[c] != d[c]) return 0;
        nz += s[c] != 0;
    }
    return nz > 0;
}

/* the halo list: images (6 floats) and their vertices; returns the count */
int orc_cnt_images(const int L[3], int n, const float *pp, float *img, int *img_v) {
    int h = 0;
    for (int i = 0; i < 26; i++) {
        int di[3], s[3];
        orc_frag_i2d3(i, di);
        for (int c = 0; c < 3; c++) s[c] = -di[c];
        for (int v = 0; v < n; v++) {
            int d[3];
            if (!emap_code(L, pp + 6 * (size_t)v, d) || !sends_to(d, s)) continue;
            memcpy(img + 6 * (size_t)h, pp + 6 * (size_t)v, 6 * sizeof(float));
            for (int c = 0; c < 3; c++) img[6 * (size_t)h + c] = pp[6 * (size_t)v + c] + (float)(L[c] * di[c]);
            img_v[h++] = v;
        }
    }
    return h;
}

/* forces of n vertices added to ff (3 floats each); returns the pair count */
long orc_cnt_forces(const int L[3], int n, const float *pp, orc_dpd_lj p, float seed_bulk, float seed_halo,
                    float *ff) {
    const int ncell = L[0] * L[1] * L[2];
    int *starts = (int *)malloc(sizeof(int) * ((size_t)ncell + 1));
    int *order = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    cell_list(L, n, pp, 6, starts, order);
    float *img = (float *)malloc(sizeof(float) * 6 * 26 * (size_t)(n > 0 ? n : 1));
    int *img_v = (int *)malloc(sizeof(int) * 26 * (size_t)(n > 0 ? n : 1));
    const int nh = orc_cnt_images(L, n, pp, img, img_v);
    int *hstarts = (int *)malloc(sizeof(int) * ((size_t)ncell + 1));
    int *horder = (int *)malloc(sizeof(int) * (size_t)(nh > 0 ? nh : 1));
    cell_list(L, nh, img, 6, hstarts, horder);
    long npair = 0;
    for (int v = 0; v < n; v++) {
        const float *rv = pp + 6 * (size_t)v;
        int cv[3];
        for (int c = 0; c < 3; c++) cv[c] = ccell(rv[c], L[c], 1);
        float acc[3] = {0, 0, 0};
        /* bulk pairs */
        for (int z = cv[2] - 1; z <= cv[2] + 1; z++)
            for (int y = cv[1] - 1; y <= cv[1] + 1; y++)
                for (int x = cv[0] - 1; x <= cv[0] + 1; x++) {
                    if (x < 0 || x >= L[0] || y < 0 || y >= L[1] || z < 0 || z >= L[2]) continue;
                    const int cid = x + L[0] * (y + L[1] * z);
                    for (int s = starts[cid]; s < starts[cid + 1]; s++) {
                        const int u = order[s];
                        if (u == v) continue;
                        const int hi = u > v ? u : v, lo = u > v ? v : u;
                        float f[3];
                        if (!orc_cnt_pair(&p, pp + 6 * (size_t)hi, pp + 6 * (size_t)lo,
                                          orc_mean0var1ii(seed_bulk, hi, lo), f))
                            continue;
                        if (v == hi) {
                            acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
                            npair++;
                        } else {
                            acc[0] -= f[0]; acc[1] -= f[1]; acc[2] -= f[2];
                        }
                    }
                }
        /* images acting on v: v's cell within one of the image's unclamped cell */
        for (int z = cv[2] - 1; z <= cv[2] + 1; z++)
            for (int y = cv[1] - 1; y <= cv[1] + 1; y++)
                for (int x = cv[0] - 1; x <= cv[0] + 1; x++) {
                    if (x < 0 || x >= L[0] || y < 0 || y >= L[1] || z < 0 || z >= L[2]) continue;
                    const int cid = x + L[0] * (y + L[1] * z);
                    for (int s = hstarts[cid]; s < hstarts[cid + 1]; s++) {
                        const int h = horder[s];
                        const float *ri = img + 6 * (size_t)h;
                        int ok = 1;
                        for (int c = 0; c < 3; c++) {
                            const int ch = ccell(ri[c], L[c], 0);
                            if (cv[c] < ch - 1 || cv[c] > ch + 1) ok = 0;
                        }
                        if (!ok) continue;
                        float f[3];
                        if (!orc_cnt_pair(&p, ri, rv, orc_mean0var1ii(seed_halo, h, v), f)) continue;
                        acc[0] -= f[0]; acc[1] -= f[1]; acc[2] -= f[2];
                    }
                }
        /* v's own images against the vertices of their blocks */
        for (int h = 0; h < nh; h++) {
            if (img_v[h] != v) continue;
            const float *ri = img + 6 * (size_t)h;
            int ch[3];
            for (int c = 0; c < 3; c++) ch[c] = ccell(ri[c], L[c], 0);
            for (int z = ch[2] - 1; z <= ch[2] + 1; z++)
                for (int y = ch[1] - 1; y <= ch[1] + 1; y++)
                    for (int x = ch[0] - 1; x <= ch[0] + 1; x++) {
                        if (x < 0 || x >= L[0] || y < 0 || y >= L[1] || z < 0 || z >= L[2]) continue;
                        const int cid = x + L[0] * (y + L[1] * z);
                        for (int s = starts[cid]; s < starts[cid + 1]; s++) {
                            const int b = order[s];
                            float f[3];
                            if (!orc_cnt_pair(&p, ri, pp + 6 * (size_t)b, orc_mean0var1ii(seed_halo, h, b), f))
                                continue;
                            acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
                            npair++;
                        }
                    }
        }
        float *o = ff + 3 * (size_t)v;
        o[0] += acc[0];
        o[1] += acc[1];
        o[2] += acc[2];
    }
    free(starts);
    free(order);
    free(img);
    free(img_v);
    free(hstarts);
    free(horder);
    return npair;
}
