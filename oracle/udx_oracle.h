/*
 * CPU oracle of the uDeviceX DPD per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the reference the HIP path is
 * compared with:
 * it is loaded by tests/, by the smoke check of __graft_entry__.py and by the
 * cpu_baseline leg of bench.py, never by the product library (ctc_amd/).
 *
 * Every function restates the reference algorithm from its source; the
 * reference file:line is cited next to each declaration.  Arithmetic is IEEE
 * FP32 with the operation order spelled out (compile with -ffp-contract=off),
 * so that the HIP kernels, which use the same order, can be compared bit for
 * bit where the operations are correctly rounded.
 */
#ifndef UDX_ORACLE_H
#define UDX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- KISS trunk generator: src/math/rnd/imp.cpp:9-34 ------------------- */
typedef struct { uint32_t x, y, z, c; } orc_kiss;
void     orc_kiss_ini(orc_kiss *r, int x, int y, int z, int c);
uint32_t orc_kiss_get_int(orc_kiss *r);
float    orc_kiss_get(orc_kiss *r);
void     orc_kiss_fill(int x, int y, int z, int c, int n, float *out);

/* ---- branch generator: src/math/rnd/dev.h:27-72 ------------------------- */
float orc_fma_rz(float a, float b, float c);      /* exact emulation       */
float orc_fma_rz_fenv(float a, float b, float c); /* fesetround + fmaf     */
float orc_mean0var1ii(float seed, int u, int v);
float orc_mean0var1uu(float seed, uint32_t u, uint32_t v);
void  orc_rng_ii_batch(float seed, int n, const int *u, const int *v, float *out);
void  orc_rng_uu_batch(float seed, int n, const uint32_t *u, const uint32_t *v, float *out);
void  orc_fma_rz_batch(int n, const float *a, const float *b, const float *c, float *out, int fenv);

/* ---- per-step seeds ----------------------------------------------------- */
/* 27 trunks: index fid (0..25) are the halo fragment trunks of
 * src/fluforces/imp/ini.h:22-58 on a single rank, index 26 is the bulk trunk
 * (0,0,0,0) of src/fluforces/imp/ini.h:6.  One draw per trunk per step
 * (src/fluforces/bulk/imp/main.h:34, src/fluforces/imp/main.h:27-31). */
typedef struct { orc_kiss t[27]; } orc_seeds;
void orc_seeds_ini(orc_seeds *s);
void orc_seeds_next(orc_seeds *s, float out[27]);
void orc_seeds_table(int nsteps, float *out /* nsteps*27 */);

/* ---- fragments: src/frag/dev.h:19-81 ------------------------------------ */
void orc_frag_i2d3(int fid, int d[3]);
int  orc_frag_d2i(int x, int y, int z);
int  orc_frag_anti(int fid);
/* per-fragment exclusive prefix of cell counts in fragment-cell order
 * (src/exch/flu/dev/map.h:1-98).  fss_off[27] gives each fragment's offset
 * into fss; fragment fid holds ncell(fid)+1 entries. */
int  orc_frag_table_size(const int L[3]);
void orc_frag_tables(const int L[3], const int *counts, int *fss, int *fss_off);

/* ---- gen_random: pre/units/gen_random.cpp:13-33 ------------------------- */
/* drand48 from glibc's zero initial state; text round trip through "%f"
 * (pre/units/common/dump.cpp:8-15) when roundtrip != 0. */
void orc_gen_random(int X, int Y, int Z, int n, int roundtrip, float *pp);

/* ---- cell list: src/clist/dev/main.h:17-101, src/clist/imp/main.h:75-84 -- */
/* Stable counting sort by cell (ties keep input order).  Particles outside
 * [-L/2, L/2) are dropped.  starts has ncells+1 entries.  Returns nout. */
int orc_clist_build(const int L[3], int n, const float *pp_in, int *starts, int *counts,
                    float *pp_out, int *order);
/* The same with every cell cut into `sub` slabs along x (sub = 1, 2, 4 or
 * 8): particles are ordered by cell, then by slab
 * floor((x + L/2) sub) mod sub, ties keep the input order.  bstarts (may be
 * NULL) receives the slab starts, ncells * sub + 1 entries, slab s of cell c
 * at c * sub + s.  The solvent's cell lists use sub = SOLVENT_SUB. */
#define ORC_SOLVENT_SUB 4
int orc_clist_build_sub(const int L[3], int sub, int n, const float *pp_in, int *starts, int *counts,
                        int *bstarts, float *pp_out, int *order);
int orc_cell_of(const int L[3], const float r[3], int c[3]);
/* slab index floor((x + L/2) sub) in [0, L sub), exact in double; -1 outside */
int orc_slab_of(int L, int sub, float x);

/* ---- DPD pair force: src/pair/dev.h:12-47,110-139 ------------------------ */
typedef struct { float a, g, s, spow; } orc_dpd;
float orc_dpd_sigma(float kBT, float g, float dt);  /* src/pair/imp/main.h:54-62 */
/* returns 1 if r2 < 1 (a pair), 0 otherwise; f is written in both cases */
int orc_pair_force(const orc_dpd *p, const float *a, const float *b, float rnd, float *f);

/* ---- DPD forces (full-shell gather, defined order) ---------------------- */
/* pp: cell-ordered AoS particles (r[3], v[3]); seeds[27] as above.
 * Each particle sums its pairs in (dz, dy, dx) cell order and cell order
 * inside a cell, first over the cells inside the box, then over the cells
 * across the periodic boundary.
 * periodic=1: neighbour cells wrap, wrapped pairs use the halo RNG of
 * src/fluforces/halo/dev/main.h:7-12 with fragment positions;
 * periodic=0: cells outside the box are skipped (bulk-only unit harness,
 * src/u/bulk/main.cu). Forces are ADDED to ff.  Returns the number of ordered
 * pairs (each unordered pair counts twice). */
long orc_dpd_forces(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                    orc_dpd p, const float seeds[27], int periodic, int nthreads, float *ff);

/* orc_dpd_forces that also ADDS each particle's stress to ss (6 per
 * particle: xx, xy, xz, yy, yz, zz): the sum over its pairs of 1/2 f_a dr_b
 * with f the force on the particle and dr = r_particle - r_other
 * (src/pair/dev.h:87-100; both particles of a pair get the same stress) */
long orc_dpd_forces_stress(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                           orc_dpd p, const float seeds[27], int periodic, int nthreads, float *ff, float *ss);

/* The same pairs with exact sums, the production (UDX_MODE_FAST) kernel's
 * definition: each particle's force is the exact sum of its pair forces, each
 * component rounded to a multiple of 2^-32, rounded to float once and ADDED
 * to ff (exact in double while the sums stay below 2^21; independent of the
 * order of the additions).  Returns ordered pairs as orc_dpd_forces. */
long orc_dpd_forces_sum64(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                          orc_dpd p, const float seeds[27], int periodic, int nthreads, float *ff);

/* The same pairs, the float pair forces summed in double and ADDED to ff64
 * (3 doubles per particle): the value that every float summation order of
 * the reference approximates. */
long orc_dpd_forces_f64(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                        orc_dpd p, const float seeds[27], int periodic, int nthreads, double *ff64);

/* The same forces with the reference's traversal: a half shell inside the
 * box (each pair once, reaction to the smaller index), one-sided pairs across
 * the periodic boundary; per-thread force buffers added in thread order.
 * The CPU baseline's force pass (not a parity reference: its FP32 sums are
 * in another order).  Returns ordered pairs as orc_dpd_forces. */
long orc_dpd_forces_half(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                         orc_dpd p, const float seeds[27], int nthreads, float *ff);

/* orc_dpd_forces_half with exact sums (quantised pair forces in double, each
 * pair inside the box once with its reaction): equal bit for bit to
 * orc_dpd_forces_sum64, which checks that the production kernel's half-shell
 * evaluation and the gather give the same forces. */
long orc_dpd_forces_half_sum64(const int L[3], int n, const float *pp, const int *starts, const int *counts,
                               orc_dpd p, const float seeds[27], int nthreads, float *ff);

/* ---- integrator: src/scheme/move/dev/vv.h:1-10 + body force + wrap ------ */
void orc_integrate(const int L[3], int n, float dt, float mass, const float accel[3],
                   const float *ff, float *pp);
void orc_move(int n, float dt, float mass, const float *ff, float *pp);   /* no wrap */
/* orc_integrate with the wrap optional (the membrane bounce-back wraps after
 * the bounce), and the wrap on its own */
void orc_integrate_ex(const int L[3], int n, float dt, float mass, const float accel[3], const float *ff,
                      float *pp, int wrap);
void orc_wrap(const int L[3], int n, float *pp);

/* ---- diagnostics: src/io/diag/part/imp/main.h:37-63 (FP64) -------------- */
/* out: [0..2] momentum (sum v), [3] kBT = 0.5*sum|v|^2/(1.5 n), [4] max |v|^2 */
void orc_diag(int n, const float *pp, double out[5]);

/* ---- one full step: clist, forces, integrate --------------------------- */
/* pp (n particles, any order) is replaced by the cell-ordered, integrated
 * state; returns the number of unordered pairs of the step. */
long orc_flu_step(const int L[3], int n, float *pp, float *scratch_pp, int *starts, int *counts,
                  float *ff, orc_dpd p, const float seeds[27], float dt, float mass,
                  const float accel[3], int nthreads);

/* ---- SDF walls (udx_oracle_wall.c) -------------------------------------- */
/* wall margin in cells (XWM/YWM/ZWM, src/inc/def.h:15-18) */
#define ORC_WALL_MARGIN 6
typedef struct {
    int L[3], T[3];          /* box, texture grid                          */
    float o[3], s[3];        /* local position -> texture coordinate        */
    float cheap_threshold;
    float *tex;              /* T[0] T[1] T[2] values, x fastest (caller)   */
} orc_sdf;
void  orc_tform_grid2grid(const float flo[3], const float fhi[3], const int fn[3], const float tlo[3],
                          const float thi[3], const int tn[3], float o[3], float s[3]);
/* texture grid of src/wall/sdf/imp/main.h:1-18 */
void  orc_sdf_texture_size(const int L[3], int T[3]);
/* sdf_gen: scale the sdf.dat field (extents ext, grid N) to the box and
 * resample it with a periodic cubic B-spline onto the texture grid
 * (src/wall/sdf/imp/gen.h, src/wall/sdf/field/imp/sample.h) */
void  orc_sdf_gen(const int L[3], const float ext[3], const int N[3], const float *D, orc_sdf *q);
void  orc_sdf_sample_box(const int L[3], const float ext[3], const int N[3], const float *D, float *out);
float orc_sdf_eval(const orc_sdf *q, float x, float y, float z);          /* src/wall/sdf/dev.h:79-112 */
void  orc_sdf_grad(const orc_sdf *q, float x, float y, float z, float g[3]); /* :32-46 */
float orc_sdf_cheap(const orc_sdf *q, float x, float y, float z);         /* :61-66 */
int   orc_sdf_far(const orc_sdf *q, float x, float y, float z);           /* :70-75 */
/* freeze (src/wall/imp/generate.h:1-60): solvent keeps s < 0, returns the
 * number of frozen wall particles (0 <= s <= 2) written to wall */
int   orc_wall_freeze(const orc_sdf *q, int *n, float *pp, float *wall);
/* frozen particles + periodic images in the margin box (src/wall/exch/imp.cpp) */
int   orc_wall_images(const int L[3], int nw, const float *wall, float *out);
/* wall velocity (src/wall/wvel/dev.h, src/wall/wvel/imp/type.h): resting,
 * constant u, shear (v[vdir] = gdot r[gdir], r[gdir] clipped at 0 when half)
 * or Hele-Shaw (radial inflow hu, channel height hh); a time-dependent shear
 * rate is resolved by the caller.  NULL means resting walls. */
enum { ORC_WVEL_NONE = 0, ORC_WVEL_CSTE = 1, ORC_WVEL_SHEAR = 2, ORC_WVEL_HS = 3 };
typedef struct { int type; float u[3]; float gdot; int vdir, gdir, half; float hu, hh; } orc_wvel;
void  orc_wvel_at(const orc_wvel *w, const float r[3], float v[3]);
/* the "shear sin" rate at time t, gdot sin(w t) (src/wall/wvel/imp/main.h:71-84) */
float orc_wvel_sin_rate(float gdot, float w, float t);
/* solvent-wall DPD (src/wall/force/dev/main.h:1-50); wall clist on the
 * margin box L + 2 margin */
long  orc_wall_forces(const int L[3], int n, const float *pp, const orc_sdf *q, int nw, const float *wpp,
                      const int *wstarts, orc_dpd p, float seed, float *ff, const orc_wvel *wv);
/* src/wall/sdf/bounce/dev/main.h */
void  orc_wall_bounce(const orc_sdf *q, float dt, int n, float *pp, const orc_wvel *wv);
/* kick-drift (+ body force), bounce-back when q != NULL, periodic wrap */
void  orc_integrate_wall(const int L[3], int n, float dt, float mass, const float accel[3], const float *ff,
                         float *pp, const orc_sdf *q, const orc_wvel *wv);
/* the same with the wrap optional (the membrane bounce-back wraps after it) */
void  orc_integrate_wall_ex(const int L[3], int n, float dt, float mass, const float accel[3], const float *ff,
                            float *pp, const orc_sdf *q, const orc_wvel *wv, int wrap);
void  orc_wall_repulse(const orc_sdf *q, float lambda, int n, const float *pp, float *ff);

/* ---- RBC membrane (udx_oracle_rbc.c) ------------------------------------ */
/* parameters of src/rbc/params/imp.h:3-9 */
typedef struct {
    float gammaC, gammaT, kBT, kb, phi, ks, x0, mpow, ka, kd, kv, totArea, totVolume;
} orc_rbc_par;
int    orc_mesh_md(int nv, int nt, const int *tt);                 /* src/io/mesh_read/imp/main.h:109-129 */
/* ring adjacency, md*nv entries each (src/rbc/adj/imp/ini.h:1-45) */
int    orc_adj_ini(int md, int nt, int nv, const int *tt, int *adj0, int *adj1);
/* (vertex, slot) -> i0 i1 i2 i3 i4 cell; 0 for an empty slot (src/rbc/adj/dev.h:2-42) */
int    orc_adj_map(int md, int nv, const int *adj0, const int *adj1, int i, int m[6]);
/* per cell [area, volume] in FP64 (src/rbc/force/area_volume/dev/main.h) */
void   orc_rbc_area_volume(int nt, int nv, int nc, const int *tt, const float *pp, double *av);
/* stress-free edge lengths and triangle areas per slot (src/rbc/shape/imp/main.h) */
void   orc_rbc_shape(int md, int nv, const int *adj0, const int *adj1, const float *rr, float *ll, float *aa);
void   orc_rbc_stressful(int nt, float totArea, float *l0, float *a0);  /* src/rbc/force/imp/main.h:50-59 */
/* area + volume + spring + viscous forces ADDED to ff (double, 3 per
 * vertex); ll/aa NULL selects the stress-full l0/a0
 * (src/rbc/force/dev/main.h:48-77, common.h:1-108) */
void   orc_rbc_forces(const orc_rbc_par *p, int md, int nv, int nc, const int *adj0, const int *adj1,
                      const float *pp, const double *av, const float *ll, const float *aa, float l0, float a0,
                      double *ff);
/* Kantor bending ADDED to ff (src/rbc/force/bending/kantor/dev/main.h:23-66) */
void   orc_rbc_bending(const orc_rbc_par *p, int md, int nv, int nc, const int *adj0, const int *adj1,
                       const float *pp, double *ff);
/* membrane energy at double positions r (3 per vertex); terms bit mask:
 * 1 global area, 2 volume, 4 local area, 8 springs, 16 bending */
double orc_rbc_energy(const orc_rbc_par *p, int md, int nv, int nc, int nt, const int *tt, const int *adj0,
                      const int *adj1, const double *r, const float *ll, const float *aa, float l0, float a0,
                      int terms);
void   orc_rbc_stretch(int nc, int nv, const float *f, float *ff);     /* src/rbc/stretch/dev/main.h */

/* membrane adjoint (udx_oracle_adj.c): for a cotangent g (3 per vertex),
 * gx = (dF/dx)^T g = (dF/dx) g (F = -grad E, symmetric Jacobian) of the
 * conservative membrane force, terms[k] the unit-coefficient term forces
 * (k: ka, kd, kv, ks, kb; NULL to skip) and gp[k] = g . terms[k] = dL/dk,
 * at double positions x (3 per vertex) */
void   orc_rbc_adjoint(const orc_rbc_par *p, int md, int nv, int nc, int nt, const int *tt, const int *adj0,
                       const int *adj1, const double *x, const float *ll, const float *aa, float l0, float a0,
                       const double *g, double *gx, double *terms, double gp[5]);

/* ---- solvent-membrane interactions (udx_oracle_fsi.c) ------------------- */
/* spp: cell-ordered solvent with starts (ncells + 1); opp: no membrane
 * vertices (any position); seeds[0] bulk, seeds[1] halo.  Forces ADDED to fo
 * (vertices) and fs (solvent); returns the number of pairs (src/fsi/dev/bulk.h, halo.h) */
long   orc_fsi_forces(const int L[3], int ns, const float *spp, const int *starts, int no, const float *opp,
                      orc_dpd p, const float seeds[2], float *fo, float *fs);
void   orc_fsi_seeds(int nsteps, float *out /* 2 per step: bulk, halo */);

/* ---- membrane bounce-back: src/mesh_bounce (udx_oracle_mbb.c) ---------- */
enum { ORC_BB_SUCCESS = 0, ORC_BB_NOCROSS = 1, ORC_BB_WTRIANGLE = 2, ORC_BB_HFAIL = 3 };
/* out = (h, u, v, side) */
int   orc_mbb_intersect(float dt, const float s1[3], const float s2[3], const float s3[3], const float vs1[3],
                        const float vs2[3], const float vs3[3], const float pr[3], const float pv[3], float out[4]);
/* meshes + periodic images: returns the instance count (at most 8 nm) */
int   orc_mbb_instances(const int L[3], int nm, int nv, const float *rc, int *inst_mesh, float *inst_shift);
void  orc_mbb_find(float dt, const int L[3], int ninst, const int *inst_mesh, const float *inst_shift, int nt, int nv,
                   const int *tt, const float *rp, const float *rc, const int *starts, const float *pp,
                   const float *pp0, uint64_t *keys);
void  orc_mbb_inertia(float mass, const float a[3], const float b[3], const float c[3], float I[6]);
void  orc_mbb_inverse(const float A[6], float I[6]);
int   orc_mbb_nonzero(const float m[6]);
void  orc_mbb_m2v(float mass, const float m[6], const float a[3], const float b[3], const float c[3], float va[3],
                  float vb[3], float vc[3]);
/* rp: start-of-step vertex positions (3 per vertex); rc: vertices (6 per
 * vertex) after the advance, velocities updated in place; pp0/pp: solvent at
 * the start of the step / after the move (unwrapped), pp bounced in place;
 * mm: 6 per triangle (P, L); col_tri: the triangle each particle bounced off
 * (-1: none).  Returns the number of bounced particles. */
long  orc_mbb_bounce(float dt, float flu_mass, float rbc_mass, const int L[3], int nm, int nt, int nv, const int *tt,
                     const float *rp, float *rc, const int *starts, int n, const float *pp0, float *pp, float *mm,
                     int *col_tri);

/* ---- Juelicher bending: src/rbc/force/bending/juelicher (udx_oracle_jul.c) */
/* edges as (i, j, k, l): 4 ints per edge; returns the edge count */
int    orc_jul_edges(int nv, int nt, const int *tt, int *dih);
double orc_jul_theta(const double a[3], const double b[3], const double c[3], const double d[3]);
void   orc_jul_forces(int nc, int nv, int nt, const int *tt, int ne, const int *dih, const float *pp, float *f,
                      float *fad);

/* ---- membrane contact: src/cnt (udx_oracle_cnt.c) ------------------------ */
typedef struct { float a, g, s, spow, ljs, lje; } orc_dpd_lj;
int   orc_cnt_pair(const orc_dpd_lj *p, const float *a, const float *b, float rnd, float *f);
int   orc_cnt_images(const int L[3], int n, const float *pp, float *img, int *img_v);
long  orc_cnt_forces(const int L[3], int n, const float *pp, orc_dpd_lj p, float seed_bulk, float seed_halo,
                     float *ff);

/* ---- correctly rounded square root from the 1-ulp hardware root --------
 * The GPU kernels take the hardware square root s (within 1 ulp of the
 * exact value) and correct it with two fma residuals:
 *   s- = s - 1 ulp, s+ = s + 1 ulp;
 *   t = (x - s- s <= 0) ? s- : s;  result = (x - s+ s > 0) ? s+ : t
 * This checks every float x in [lo, hi] with every value the hardware may
 * give (the correctly rounded root and its two neighbours) against the
 * host's correctly rounded sqrtf.  Returns the number of failures.
 * correct = 0 checks the uncorrected hardware values instead (a control:
 * two failures per float). */
long orc_verify_sqrt_rn(float lo, float hi, int nthreads, int correct);

#ifdef __cplusplus
}
#endif
#endif
