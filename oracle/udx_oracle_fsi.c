/*
 * CPU oracle of solvent-membrane interactions (FSI) on one periodic domain.
 *
 * TEST INFRASTRUCTURE ONLY (see udx_oracle.h).
 *
 * The reference pairs every membrane vertex with the solvent of its 3x3x3
 * cell block clipped to the domain (src/fsi/dev/bulk.h, map/bulk.h), then
 * sends the vertices lying outside the domain to the neighbour domains
 * (src/exch/obj, src/exch/map/dev.h); on one periodic domain those come back
 * as images shifted by a period and pair with the solvent the same way
 * (src/fsi/dev/halo.h), with their own seed and their position in the halo
 * list as the noise index.  Forces are symmetric: +f on the vertex, -f on
 * the solvent particle.
 *
 * The reference sums with atomics in arrival order and fills the halo list
 * in atomic order; the order here is defined (and the HIP path follows it):
 *  - halo list: receive fragment i = 0..25 (the vertices outside the domain
 *    toward -d_i, shifted by +d_i L), vertices ascending within a fragment;
 *  - vertex force: += the vertex's own sum (z planes 0, 1, 2, rows y - 1..
 *    y + 1, solvent in cell order), then += each image's sum in halo order;
 *  - solvent force: += one sum over the interacting instances (vertices,
 *    then images) in the order of their cells on the (L+2)^3 grid of cells
 *    [-1, L] (dz, dy, dx around the solvent's cell), ascending instance
 *    index within a cell.
 */
#include "udx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* per-axis -1 / 0 / +1: outside below, inside, outside above
 * (emap_code, src/exch/map/dev.h:8-18) */
static int emap_code(const int L[3], const float *r, int d[3]) {
    for (int c = 0; c < 3; c++) d[c] = -1 + (r[c] >= (float)(-L[c] / 2)) + (r[c] >= (float)(L[c] / 2));
    return d[0] | d[1] | d[2];
}

/* is direction s among the destinations of code d: faces, edges and the
 * corner spanned by the non-zero axes (emap_decode, src/exch/map/dev.h:43-82) */
static int sends_to(const int d[3], const int s[3]) {
    int nz = 0;
    for (int c = 0; c < 3; c++) {
        if (s[c] != 0 && s[c] != d[c]) return 0;
        nz += s[c] != 0;
    }
    return nz > 0;
}

/* the three x-runs (rows y - 1, y, y + 1) of one z plane of the clipped cell
 * block around r (src/fsi/dev/map/bulk.h:1-62); returns 0 if the plane has
 * no cells */
typedef struct { int org[3], cnt[3]; } fsi_rows;
static int rows_of(const int L[3], const int *starts, int zplane, const float *r, fsi_rows *m) {
    int xc = L[0] / 2 + (int)floorf(r[0]);
    int yc = L[1] / 2 + (int)floorf(r[1]);
    int zc = L[2] / 2 + (int)floorf(r[2]);
    int xs = xc - 1 > 0 ? xc - 1 : 0;
    int xe = xc + 2 < L[0] ? xc + 2 : L[0];
    if (xc - 1 >= L[0] || xc + 2 <= 0) return 0;
    int z = zc - 1 + zplane;
    for (int k = 0; k < 3; k++) {
        int y = yc - 1 + k;
        m->org[k] = 0;
        m->cnt[k] = 0;
        if (z >= 0 && z < L[2] && y >= 0 && y < L[1]) {
            int cid = xs + L[0] * (y + L[1] * z);
            m->org[k] = starts[cid];
            m->cnt[k] = starts[cid + (xe - xs)] - starts[cid];
        }
    }
    return 1;
}

typedef struct { float r[6]; int vertex, lid, halo; } fsi_inst;

static int build_instances(const int L[3], int no, const float *opp, fsi_inst **out) {
    int n = no, cap = no * 8 + 1;
    fsi_inst *a = (fsi_inst *)malloc(sizeof(fsi_inst) * (size_t)cap);
    for (int v = 0; v < no; v++) {
        memcpy(a[v].r, opp + 6 * v, sizeof a[v].r);
        a[v].vertex = v;
        a[v].lid = v;
        a[v].halo = 0;
    }
    int aid = 0;
    for (int i = 0; i < 26; i++) {
        int di[3], s[3];
        orc_frag_i2d3(i, di);
        for (int c = 0; c < 3; c++) s[c] = -di[c];
        for (int v = 0; v < no; v++) {
            int d[3];
            if (!emap_code(L, opp + 6 * v, d) || !sends_to(d, s)) continue;
            fsi_inst *q = &a[n++];
            memcpy(q->r, opp + 6 * v, sizeof q->r);
            for (int c = 0; c < 3; c++) q->r[c] = q->r[c] + (float)(L[c] * di[c]);
            q->vertex = v;
            q->lid = aid++;
            q->halo = 1;
        }
    }
    *out = a;
    return n;
}

static int inst_sum(const int L[3], const float *spp, const int *starts, const fsi_inst *q, orc_dpd p, float seed,
                    int zplane, float f[3]) {
    fsi_rows m;
    int np = 0;
    if (!rows_of(L, starts, zplane, q->r, &m)) return 0;
    for (int k = 0; k < 3; k++)
        for (int b = m.org[k]; b < m.org[k] + m.cnt[k]; b++) {
            float g[3];
            float rnd = orc_mean0var1uu(seed, (uint32_t)q->lid, (uint32_t)b);
            np += orc_pair_force(&p, q->r, spp + 6 * b, rnd, g);
            f[0] += g[0]; f[1] += g[1]; f[2] += g[2];
        }
    return np;
}

long orc_fsi_forces(const int L[3], int ns, const float *spp, const int *starts, int no, const float *opp,
                    orc_dpd p, const float seeds[2], float *fo, float *fs) {
    fsi_inst *inst;
    int ni = build_instances(L, no, opp, &inst);
    long npairs = 0;
    (void)ns;
    /* vertex side: one sum per instance over its three z planes; the vertex
     * itself first, then its images (instances no.. are in halo order) */
    for (int k = 0; k < ni; k++) {
        float s[3] = {0, 0, 0};
        for (int zp = 0; zp < 3; zp++) npairs += inst_sum(L, spp, starts, &inst[k], p, seeds[k < no ? 0 : 1], zp, s);
        float *o = fo + 3 * inst[k].vertex;
        o[0] += s[0]; o[1] += s[1]; o[2] += s[2];
    }
    /* solvent side: instances on the (L+2)^3 grid of cells -1..L */
    int W[3] = {L[0] + 2, L[1] + 2, L[2] + 2}, nw = W[0] * W[1] * W[2];
    int *cell = (int *)malloc(sizeof(int) * (size_t)(ni > 0 ? ni : 1));
    int *wst = (int *)calloc((size_t)nw + 1, sizeof(int));
    int *idx = (int *)malloc(sizeof(int) * (size_t)(ni > 0 ? ni : 1));
    for (int k = 0; k < ni; k++) {
        int c3[3], ok = 1;
        for (int c = 0; c < 3; c++) {
            c3[c] = L[c] / 2 + (int)floorf(inst[k].r[c]) + 1;
            if (c3[c] < 0 || c3[c] >= W[c]) ok = 0;
        }
        cell[k] = ok ? c3[0] + W[0] * (c3[1] + W[1] * c3[2]) : -1;
        if (ok) wst[cell[k] + 1]++;
    }
    for (int c = 0; c < nw; c++) wst[c + 1] += wst[c];
    {
        int *fill = (int *)calloc((size_t)nw, sizeof(int));
        for (int k = 0; k < ni; k++)
            if (cell[k] >= 0) idx[wst[cell[k]] + fill[cell[k]]++] = k;
        free(fill);
    }
    for (int c = 0; c < L[0] * L[1] * L[2]; c++) {
        int cx = c % L[0], cy = (c / L[0]) % L[1], cz = c / (L[0] * L[1]);
        for (int b = starts[c]; b < starts[c + 1]; b++) {
            float acc[3] = {0, 0, 0};
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int w = (cx + dx + 1) + W[0] * ((cy + dy + 1) + W[1] * (cz + dz + 1));
                        for (int t = wst[w]; t < wst[w + 1]; t++) {
                            const fsi_inst *q = &inst[idx[t]];
                            float g[3];
                            float rnd = orc_mean0var1uu(seeds[q->halo], (uint32_t)q->lid, (uint32_t)b);
                            if (orc_pair_force(&p, q->r, spp + 6 * b, rnd, g)) {
                                acc[0] -= g[0]; acc[1] -= g[1]; acc[2] -= g[2];
                            }
                        }
                    }
            fs[3 * b] += acc[0]; fs[3 * b + 1] += acc[1]; fs[3 * b + 2] += acc[2];
        }
    }
    free(cell);
    free(wst);
    free(idx);
    free(inst);
    return npairs;
}

/* FSI trunk of rank 0: KISS(1908, 1409, 290, 12968), two draws per step with
 * one membrane type (bulk, then halo) (src/fsi/imp/main.h:1-8,
 * src/fsi/imp/bulk.h:8, src/fsi/imp/halo.h:7) */
void orc_fsi_seeds(int nsteps, float *out) {
    orc_kiss_fill(1908, 1409, 290, 12968, 2 * nsteps, out);
}
