/*
 * udx: MI355X (gfx950) HIP implementation of the uDeviceX per-timestep DPD
 * hot path behind a C ABI.
 *
 * Conventions shared by every entry point:
 *  - particle arrays at the boundary are the reference's AoS layout
 *    Particle{float r[3], v[3]} (24 B, src/inc/type.h:6-8) and forces are
 *    Force{float f[3]} (12 B, src/inc/type.h:20-22), as raw device pointers;
 *  - forces are ACCUMULATED (+=) as in the reference; the caller clears them;
 *  - every call returns 0 on success and a non-zero status on failure, with
 *    a message in udx_last_error() (the reference aborts through ERR/UC,
 *    src/utils/error.h:1-17; a caller that wants that behaviour exits on a
 *    non-zero status);
 *  - all work is enqueued on the stream bound to the udx_ctx, nothing
 *    synchronises unless the call returns a host value.
 * The box is [-L/2, L/2)^3 with unit cells, x fastest, one periodic domain.
 */
#ifndef UDX_H
#define UDX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { UDX_OK = 0, UDX_EINVAL = 1, UDX_EHIP = 2, UDX_ECAP = 3, UDX_ESTATE = 4 };

/* message of the last failed call on this thread */
const char *udx_last_error(void);

/* ---- KISS trunk generator: replaces rnd_ini/rnd_get/rnd_fin
 *      (src/math/rnd/imp.h:2-5; algorithm src/math/rnd/imp.cpp:9-34) ---- */
typedef struct udx_rnd { uint32_t x, y, z, c; } udx_rnd;
void  udx_rnd_ini(int x, int y, int z, int c, udx_rnd *r);
float udx_rnd_get(udx_rnd *r);

/* ---- DPD pair parameters: replaces pair_set_dpd / pair_compute_dpd_sigma /
 *      pair_get_view_dpd (src/pair/imp.h:11-35, grey single-colour view) ---- */
typedef struct udx_pair_dpd { float a, g, s, spow; } udx_pair_dpd;
/* DPD + Lennard-Jones (PairDPDLJ, src/pair/type.h:22-25): the membrane contact pair */
typedef struct udx_pair_dpd_lj { float a, g, s, spow, ljs, lje; } udx_pair_dpd_lj;
/* s = sqrt(2 kBT g / dt) (src/pair/imp/main.h:54-62) */
void udx_pair_compute_dpd_sigma(float kBT, float dt, udx_pair_dpd *p);

/* ---- context: one per host thread, bound to a HIP stream ---- */
typedef struct udx_ctx udx_ctx;
int udx_ctx_ini(void *hip_stream /* hipStream_t, NULL = default */, udx_ctx **ctx);
int udx_ctx_fin(udx_ctx *ctx);
int udx_ctx_sync(udx_ctx *ctx);

/* ---- branch generator on the GPU: rnd::mean0var1ii / mean0var1uu
 *      (src/math/rnd/dev.h:57-72) for n index pairs (device arrays);
 *      uu != 0 selects the unsigned variant ---- */
int udx_rnd_branch(udx_ctx *ctx, float seed, int n, const int *u, const int *v, int uu, float *out);

/* ---- self-check of the device arithmetic (no counterpart in the reference):
 *      over every float x in [lo, hi] (0 <= lo <= hi, fewer than 2^31 floats)
 *      compares the library's device routine with the correctly rounded
 *      result and returns the number of floats that differ (synchronises).
 *      UDX_CHECK_RCP: the reciprocal of the pair force (for x in [2^-24, 1]);
 *      UDX_CHECK_SQRT: its square root; UDX_CHECK_FRACT: x - floor(x) of the
 *      branch generator; UDX_CHECK_RCP_ESTIMATE: the hardware reciprocal
 *      estimate alone (not correctly rounded: shows that the check counts). ---- */
enum { UDX_CHECK_RCP = 0, UDX_CHECK_SQRT = 1, UDX_CHECK_FRACT = 2, UDX_CHECK_RCP_ESTIMATE = 3 };
int udx_selfcheck(udx_ctx *ctx, int what, float lo, float hi, unsigned long long *mismatches);

/* ---- cell list: replaces clist_ini / clist_ini_map / clist_build /
 *      clists_get_ss / clists_get_cc (src/clist/imp.h:5-33) ----
 * Stable counting sort: particles are ordered by cell, ties keep their
 * input order; particles outside the box are dropped. */
typedef struct udx_clist udx_clist;
int udx_clist_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int maxp, udx_clist **c);
int udx_clist_fin(udx_clist *c);
/* pp: n input particles (device AoS); ppout: cell-ordered output (device AoS,
 * capacity maxp); *nout (host) receives the number kept (synchronises). */
int udx_clist_build(udx_clist *c, int n, const float *pp, float *ppout, int *nout);
const int *udx_clist_starts(const udx_clist *c);   /* device, ncells + 1 */
const int *udx_clist_counts(const udx_clist *c);   /* device, ncells     */
const int *udx_clist_order(const udx_clist *c);    /* device, input index per output slot */
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  udx_clist_build without a handle: the caller passes every
 * output (ppout: n x 6 floats or NULL, starts: ncells + 1, counts: ncells,
 * order: n ints; device), same order and tables as udx_clist_build.  The
 * working arrays live in a static device buffer that each GPU holds for the
 * process (about 17 MB), sized for UDX_NOALLOC_MAXCELLS cells and
 * UDX_NOALLOC_MAXP particles; larger calls return UDX_ECAP.  Calls that
 * overlap on one GPU share that buffer and race.  Does not synchronise: the
 * number kept is starts[ncells] on the device. */
#define UDX_NOALLOC_MAXCELLS (1 << 18)
#define UDX_NOALLOC_MAXP (1 << 20)
int udx_clist_build_noalloc(udx_ctx *ctx, int Lx, int Ly, int Lz, int n, const float *pp, float *ppout,
                            int *starts, int *counts, int *order);

/* ---- exclusive scan: replaces scan_ini / scan_apply / scan_fin
 *      (src/algo/scan/imp.h:4-6) ----
 * output[i] = input[0] + ... + input[i - 1] for 0 <= i < size (device int
 * arrays, on the context's stream; output[size] is not written).  size above
 * the capacity given to udx_scan_ini returns UDX_ECAP, as the reference's
 * scan_apply aborts.  The reference compresses each count to 8 bits (counts
 * above 255 overflow); here every count is a full int. */
typedef struct udx_scan udx_scan;
int udx_scan_ini(udx_ctx *ctx, int size, udx_scan **w);
int udx_scan_apply(udx_scan *w, const int *input, int size, int *output);
int udx_scan_fin(udx_scan *w);

/* ---- fluid-fluid DPD forces: replaces fluforces_bulk_ini / prepare / apply
 *      and fluforces_halo_ini / prepare / apply (src/fluforces/imp.h:8-25) ----
 * One periodic domain: pairs across the box boundary are computed in the
 * same pass with the halo RNG of src/fluforces/halo/dev/main.h:7-12. */
typedef struct udx_fluforces udx_fluforces;
int udx_fluforces_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int maxp, udx_fluforces **f);
int udx_fluforces_fin(udx_fluforces *f);
/* AoS -> internal float4 position/velocity streams (the reference's
 * zipped_pp conversion, src/fluforces/imp/main.h:1-8) */
int udx_fluforces_prepare(udx_fluforces *f, int n, const float *pp);
/* pp given to prepare must be cell ordered by the clist whose starts are
 * passed.  periodic = 0 skips cells outside the box (the bulk-only unit
 * harness, src/u/bulk/main.cu).  Draws one seed per trunk on the host
 * (bulk trunk (0,0,0,0), 26 halo trunks) and adds the forces into ff. */
int udx_fluforces_apply(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic,
                        float *ff);
/* same with explicit seeds: seeds[fid] for halo fragment fid (0..25),
 * seeds[26] for the bulk */
int udx_fluforces_apply_seeds(udx_fluforces *f, const udx_pair_dpd *p, const int *starts,
                              int periodic, const float seeds[27], float *ff);
/* forces and the per-particle stress: the reference's FoSArray path
 * (src/fluforces/bulk/imp/main.h:42-53 dispatch_stress, src/pair/dev.h:87-100):
 * ss (device, 6 floats per particle: xx, xy, xz, yy, yz, zz) gets, per pair
 * of the particle, 1/2 f_a dr_b with f the force on it and dr = r - r_other
 * (both particles of a pair get the same), ADDED like the forces.  seeds as
 * udx_fluforces_apply_seeds, or NULL for the next draws of the trunks.
 * Correctly rounded arithmetic (the mode setting does not apply). */
int udx_fluforces_apply_stress(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic,
                               const float seeds[27], float *ff, float *ss);
/* force kernel modes.  FAST and EXACT gather the full shell of each particle
 * in a fixed order (no atomics) with correctly rounded reciprocals and roots,
 * so both give the same bits, equal to the CPU restatement; EXACT is kept so
 * callers written against the two-mode interface keep working.  HALF (solvent
 * forces only: udx_fluforces, udx_flu, udx_ens) evaluates each pair inside the
 * box once, as the reference's half-shell bulk kernel, and sums every
 * particle's pair forces exactly (each component rounded to a multiple of
 * 2^-32, added in FP64, rounded to float once): deterministic, independent of
 * the order of the additions, equal to the CPU restatement's exact sums. */
#define UDX_MODE_FAST 0    /* default */
#define UDX_MODE_EXACT 1
#define UDX_MODE_HALF 2
int udx_fluforces_set_mode(udx_fluforces *f, int mode);
/* The gather's pairs across the periodic boundary (the halo) are evaluated in
 * the force kernel, or, when a launch's boxes hold at least n particles in
 * boundary cells (estimated from the box sizes; default 2^20), by a second
 * launch over those particles only, which keeps its lanes busy when there are
 * many.  The forces are the same bits either way.  Process-wide; no
 * counterpart in the reference (its halo is a separate set of kernels). */
int udx_dpd_set_split_edges(double n);
/* unordered pair interactions (r^2 < 1) counted by apply calls since the last reset */
int udx_fluforces_pairs(udx_fluforces *f, unsigned long long *npairs, int reset);

/* ---- integrator: replaces scheme_move_apply (src/scheme/move/imp.h:5-6) and
 *      bforce_apply for a constant body force (src/scheme/force/imp.h) ---- */
/* kick-drift v += f/m dt; r += v dt (src/scheme/move/dev/vv.h:1-10) */
int udx_scheme_move_apply(udx_ctx *ctx, float dt, float mass, int n, const float *ff, float *pp);
/* f += m a (src/scheme/force/dev/main.h:1-4,65-80) */
int udx_bforce_apply_cste(udx_ctx *ctx, float mass, const float accel[3], int n, float *ff);
/* periodic wrap of positions into [-L/2, L/2) (single-domain redistribution) */
int udx_wrap_pp(udx_ctx *ctx, int Lx, int Ly, int Lz, int n, float *pp);

/* ---- diagnostics (src/io/diag/part/imp/main.h:37-63), FP64 on the GPU ----
 * out[0..2] = sum v, out[3] = kBT = 0.5 sum|v|^2 / (1.5 n), out[4] = max |v|^2
 * (synchronises) */
int udx_diag_part(udx_ctx *ctx, int n, const float *pp, double out[5]);

/* ---- fluid solver: the per-step pipeline of src/sim/imp/run.h:8-68 for a
 *      solvent-only box (cell list, DPD forces, body force, kick-drift,
 *      periodic wrap) on internal float4 state ---- */
typedef struct udx_flu_conf {
    int L[3];
    int maxp;
    udx_pair_dpd dpd;
    float dt, mass;
    float accel[3];     /* constant body force per unit mass */
    int mode;           /* UDX_MODE_* of the solvent force kernel */
} udx_flu_conf;
typedef struct udx_flu udx_flu;
int udx_flu_ini(udx_ctx *ctx, const udx_flu_conf *conf, udx_flu **f);
int udx_flu_fin(udx_flu *f);
int udx_flu_set_pp(udx_flu *f, int n, const float *pp);       /* device AoS in  */
int udx_flu_get_pp(udx_flu *f, float *pp, int *n);            /* device AoS out */
int udx_flu_get_ff(udx_flu *f, float *ff);                    /* last step forces, AoS */
/* advance nsteps steps; seeds come from the internal trunks unless a table
 * is given (seeds_host: nsteps x 28 floats: [0..26] as in
 * udx_fluforces_apply_seeds, [27] the wall seed) */
int udx_flu_run(udx_flu *f, int nsteps, const float *seeds_host);
int udx_flu_pairs(udx_flu *f, unsigned long long *npairs, int reset);
/* solvent-wall pairs counted since the last reset of udx_flu_pairs */
int udx_flu_wall_pairs(udx_flu *f, unsigned long long *npairs);

/* ---- SDF walls: replace sdf_ini / sdf_gen / sdf_get_view / sdf_bounce
 *      (src/wall/sdf/imp.h:9-25), wall_gen / wall_interact / wall_bounce
 *      (src/sim/walls/imp.h:14-31) and wall_force / wall_repulse
 *      (src/wall/imp.h:42-48); resting walls (wall velocity 0) ---- */
typedef struct udx_sdf udx_sdf;
int udx_sdf_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, udx_sdf **q);
int udx_sdf_fin(udx_sdf *q);
/* field D (host, N[0] N[1] N[2] floats, x fastest) over the box with
 * extents ext: scaled by Lx / ext[0] and resampled with a periodic cubic
 * B-spline onto the texture grid (src/wall/sdf/imp/gen.h) */
int udx_sdf_gen(udx_sdf *q, const float ext[3], const int N[3], const float *D);
int udx_sdf_gen_file(udx_sdf *q, const char *path);            /* sdf.dat format */
int udx_sdf_size(const udx_sdf *q, int T[3], float *cheap_threshold);
const float *udx_sdf_texture(const udx_sdf *q);                /* device, T[0] T[1] T[2] */
/* r3: n positions (device, 3 floats each); what = 0: sdf_eval, 1: nearest
 * texel, 2: gradient (3 floats per position) (src/wall/sdf/dev.h) */
int udx_sdf_probe(const udx_sdf *q, int n, const float *r3, int what, float *out);

typedef struct udx_wall udx_wall;
/* freeze (src/wall/imp/generate.h:1-60): pp (device AoS, *n particles) keeps
 * the solvent (sdf < 0) in order and *n is updated; particles with
 * 0 <= sdf <= 2 become wall particles, deeper ones are dropped; the wall
 * set holds them and their periodic images in the 6-cell margin, cell-sorted */
/* wall velocity (src/wall/wvel/imp.h, src/wall/wvel/imp/type.h): resting
 * (the default), constant u, shear v[vdir] = gdot r[gdir] (r[gdir] clipped
 * at 0 when half), shear with the rate gdot sin(w t), Hele-Shaw radial
 * inflow (hs_u, channel height hs_h).  Used for the wall particles' velocity
 * in the wall DPD and in the bounce-back, 2 u_wall - v. */
enum { UDX_WVEL_NONE = 0, UDX_WVEL_CSTE = 1, UDX_WVEL_SHEAR = 2, UDX_WVEL_SHEAR_SIN = 3, UDX_WVEL_HS = 4 };
typedef struct udx_wall_vel {
    int type;
    float u[3];
    float gdot;
    int vdir, gdir, half;
    float w;
    float hs_u, hs_h;
} udx_wall_vel;
int udx_wall_gen(udx_ctx *ctx, const udx_sdf *q, int *n, float *pp, udx_wall **w);
int udx_wall_fin(udx_wall *w);
int udx_wall_counts(const udx_wall *w, int *nfrozen, int *ntotal);
const float *udx_wall_pp(const udx_wall *w);       /* device AoS, cell order on the margin box */
const int *udx_wall_starts(const udx_wall *w);     /* device, margin-box cells + 1 */
int udx_wall_set_mode(udx_wall *w, int mode);      /* UDX_MODE_EXACT or not */
/* solvent-wall DPD on n cell-ordered solvent particles (device AoS), forces
 * added to ff; the seed is one draw of the wall trunk, or given */
int udx_wall_force(udx_wall *w, const udx_pair_dpd *p, int n, const float *pp, float *ff);
int udx_wall_force_seed(udx_wall *w, const udx_pair_dpd *p, int n, const float *pp, float seed, float *ff);
int udx_wall_pairs(udx_wall *w, unsigned long long *npairs, int reset);
/* bounce-back of particles inside the walls (src/wall/sdf/bounce/dev/main.h) */
int udx_wall_bounce(const udx_sdf *q, float dt, int n, float *pp);
/* wall_vel_set (wvel_set_cste/shear/shear_sin/hs, src/wall/wvel/imp.h) and
 * the time of the "shear sin" rate (wvel_get_step, src/wall/wvel/imp.h);
 * udx_flu_run sets the time of its walls from its own clock */
int udx_wall_set_vel(udx_wall *w, const udx_wall_vel *v);
int udx_wall_set_time(udx_wall *w, float t);
/* bounce-back with the wall's velocity at its time */
int udx_wall_bounce_wall(const udx_wall *w, float dt, int n, float *pp);
/* wall repulsion of object particles, f += -min(1e4, e^{l(s+1)} - 1) grad s / |grad s| */
int udx_wall_repulse(const udx_sdf *q, float lambda, int n, const float *pp, float *ff);
/* walls in the solvent step: wall DPD after the solvent forces, bounce-back
 * after the drift (NULL removes them) */
int udx_flu_set_wall(udx_flu *f, udx_wall *w);
/* the step clock: time of the next step (advanced by dt per step; the walls'
 * velocity is evaluated at the step's time) */
int udx_flu_set_time(udx_flu *f, float t);
/* enable != 0: udx_flu_run captures the step pair (even, odd) into a HIP
 * graph on the solver's own stream and replays it (the same results; one
 * launch per two steps instead of about fifteen per step).  Runs with
 * membranes attached, a time-dependent wall velocity or timing enabled stay
 * eager.  No counterpart in the reference (its steps are launched from the
 * host one kernel at a time). */
int udx_flu_set_graph(udx_flu *f, int enable);
/* per-step kernel timing of the last udx_flu_run (ms, HIP events on the
 * bound stream): [0] cell list + reorder + fragment tables, [1] solvent DPD,
 * [2] wall DPD, [3] integrate + bounce-back, [4] their sum.  Without objects
 * the solvent DPD kernel also evaluates the wall forces and moves the
 * particles: [1] holds all three and [2], [3] are the gaps between them. */
int udx_flu_timing(udx_flu *f, int enable, float out[5]);

/* ---- RBC membrane: replaces rbc_force_ini / rbc_force_set_conf /
 *      rbc_force_apply / rbc_force_fin (src/rbc/force/imp.h:9-28),
 *      area_volume_ini / area_volume_compute (src/rbc/force/area_volume/imp.h:6-10),
 *      bending_apply with the Kantor model (src/rbc/force/bending/imp.h:16)
 *      and rbc_stretch_apply (src/rbc/stretch/imp.h) ----
 * Cells are nc consecutive copies of one template mesh (nv vertices, nt
 * triangles); pp holds nc * nv Particles, cell-major. */
typedef struct udx_rbc_params {        /* src/rbc/params/imp.h:3-9 */
    float gammaC, gammaT, kBT;
    float kb, phi;                     /* bending rigidity, spontaneous angle (degrees) */
    float ks, x0, mpow;                /* WLC + POW springs */
    float ka, kd, kv;                  /* global area, local area, volume */
    float totArea, totVolume;
} udx_rbc_params;
typedef struct udx_rbc udx_rbc;
/* tt: host, 3 * nt vertex indices; builds the ring adjacency */
int udx_rbc_ini(udx_ctx *ctx, int nv, int nt, const int *tt, int max_cells, udx_rbc **r);
int udx_rbc_fin(udx_rbc *r);
int udx_rbc_md(const udx_rbc *r, int *md);                 /* maximum vertex degree */
/* host copies of the ring adjacency (md * nv each, src/rbc/adj/imp.h) */
int udx_rbc_adjacency(const udx_rbc *r, int *adj0, int *adj1);
/* stress-full references l0, a0 from the total area (src/rbc/force/imp/main.h:50-59) */
int udx_rbc_set_stressful(udx_rbc *r, float totArea);
/* stress-free references from a shape (host, 3 * nv floats; src/rbc/force/imp/main.h:61-92) */
int udx_rbc_set_stressfree(udx_rbc *r, const float *rr);
int udx_rbc_set_bending(udx_rbc *r, int kantor);           /* 1: "kantor", 0: "none" */
/* rbc_force_set_bending (src/rbc/force/imp/main.h:120-130): "none",
 * "kantor" or "juelicher" (src/rbc/force/bending/juelicher: kb = 1 and
 * H0 = -1/2 fixed as there, the params' kb unused; area-difference term
 * included).  The adjoint covers Kantor bending only. */
enum { UDX_BENDING_NONE = 0, UDX_BENDING_KANTOR = 1, UDX_BENDING_JUELICHER = 2 };
int udx_rbc_set_bending_type(udx_rbc *r, int type);
/* per cell [area, volume] in FP64 into av (device, 2 * nc; NULL: the
 * handle's own buffer, see udx_rbc_av) */
int udx_rbc_area_volume(udx_rbc *r, int nc, const float *pp, double *av);
const double *udx_rbc_av(const udx_rbc *r);                /* device, 2 * max_cells */
/* area/volume, bending and membrane forces ADDED to ff (nc * nv Forces);
 * the membrane's thermal noise (rnd = true) is not supported */
int udx_rbc_force_apply(udx_rbc *r, const udx_rbc_params *p, float dt, int nc, const float *pp, float *ff);
/* Membrane adjoint.  No counterpart in the reference; for calibrating the
 * membrane coefficients.  For a cotangent g (device, 3 floats per vertex) of
 * the conservative membrane force F (area, volume, springs, bending; the
 * viscous term is left out, i.e. zero relative velocities), writes
 * gx = (dF/dx)^T g (device, 3 floats per vertex) and gp = g . dF/dk for
 * k = ka, kd, kv, ks, kb (host, summed over all cells).  Nothing is kept
 * from the forward call: the geometry is recomputed. */
int udx_rbc_force_adjoint(udx_rbc *r, const udx_rbc_params *p, int nc, const float *pp, const float *g,
                          float *gx, double gp[5]);
/* whole-cell periodic redistribution on one domain (drbc_build_map / pack /
 * unpack, src/distr/rbc/imp/map.h, src/distr/rbc/dev.h:3-20,
 * src/distr/map/dev.h:17-26, src/distr/rbc/imp/unpack.h:30-51): a cell whose
 * bounding-box centre left [-L/2, L/2) is shifted by -L d and moved after
 * the kept cells; kept cells first, then fragments 0..25, ascending within
 * each (the reference's order within a fragment comes from atomics).
 * pp: device AoS, in place. */
int udx_rbc_distribute(udx_rbc *r, int Lx, int Ly, int Lz, int nc, float *pp);
/* ff[i] += f[i % nv] for nc * nv vertices (f: device, 3 * nv) */
int udx_rbc_stretch_apply(udx_ctx *ctx, int nc, int nv, const float *f, float *ff);

/* ---- solvent-membrane interactions: replaces fsi_ini / fsi_bind_solvent /
 *      fsi_bulk / fsi_halo for one membrane type (src/fsi/imp.h:6-13) ----
 * One periodic domain with even sides: the vertices outside the domain act
 * through their periodic images, with the second seed of the step and their
 * position in the halo list (fragment-major, vertex order within a
 * fragment) as noise index.  Both sides are gathers in a fixed order (no
 * atomics); +f goes to the vertex, -f to the solvent particle. */
typedef struct udx_fsi udx_fsi;
int udx_fsi_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int max_vertices, udx_fsi **f);
int udx_fsi_fin(udx_fsi *f);
int udx_fsi_set_mode(udx_fsi *f, int mode);      /* UDX_MODE_FAST or UDX_MODE_EXACT */
/* pp: cell-ordered solvent (AoS) with the starts of its cell list (ncells + 1);
 * ff: solvent forces (Force) the interactions are ADDED to */
int udx_fsi_bind_solvent(udx_fsi *f, const float *pp, float *ff, int n, const int *starts);
/* nv membrane vertices opp (AoS, any position) with forces off (ADDED);
 * draws the step's two seeds (bulk, halo) from the FSI trunk
 * (1908, 1409, 290, 12968) (src/fsi/imp/main.h:1-8) */
int udx_fsi_apply(udx_fsi *f, const udx_pair_dpd *p, int nv, const float *opp, float *off);
int udx_fsi_apply_seeds(udx_fsi *f, const udx_pair_dpd *p, int nv, const float *opp, float *off,
                        const float seeds[2]);
/* vertex-solvent pairs (r^2 < 1) since the last reset */
int udx_fsi_pairs(udx_fsi *f, unsigned long long *npairs, int reset);
/* Membranes in the solvent step (src/sim/imp/run.h:8-68): nc cells of r's
 * mesh whose vertices pp (device AoS, nc * nv particles, owned by the caller)
 * the step advances in place.  Per step, after the solvent cell list: whole
 * cells are redistributed (a cell whose bounding-box centre left the box is
 * shifted by a box length and moved after the kept cells, kept cells first,
 * then fragments 0..25, ascending within each; src/distr/rbc/dev.h,
 * src/distr/map/dev.h:17-26); after the solvent and wall forces, FSI with the
 * pair fp (fsi: buffers and the seed trunk, two draws per step) adds +f to the
 * vertices and -f to the solvent; after the solvent kick-drift, each of the
 * `substeps` membrane sub-steps starts the vertex force from the FSI force,
 * adds the membrane force rp (udx_rbc_force_apply) and kick-drifts the
 * vertices with `mass` and dt / substeps (src/sim/objects/imp/main.h:40-61).
 * Replaces objects_distribute, forces_objects and objects_advance of the
 * reference's step.  r = NULL detaches the membranes.  Not combined with walls. */
int udx_flu_set_rbc(udx_flu *f, udx_rbc *r, udx_fsi *fsi, const udx_rbc_params *rp, const udx_pair_dpd *fp,
                    float mass, int substeps, int nc, float *pp);

/* ---- membrane contact: replaces cnt_ini / cnt_fin / cnt_build_cells /
 *      cnt_bulk / cnt_halo (src/cnt/imp.h:5-13) for one object type on one
 *      periodic domain ----
 * DPD plus the capped repulsive Lennard-Jones term between membrane vertices
 * (pairs a > b whose cells, clamped into the domain, differ by at most one
 * per axis; noise mean0var1ii(seed_bulk, a, b)) and between the periodic
 * images of the vertices outside the domain and the vertices (noise
 * mean0var1ii(seed_halo, image, b), the image list ordered by receive
 * fragment, vertices ascending).  pp: n vertices (6 floats, device), ff: 3
 * floats per vertex, added to.  udx_cnt_apply draws (seed_bulk, seed_halo)
 * from the handle's trunk (7119, 187, 18278, 15674).  Each vertex sums its
 * terms in a fixed order: no dependence on scheduling. */
typedef struct udx_cnt udx_cnt;
int udx_cnt_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int max_n, udx_cnt **c);
int udx_cnt_fin(udx_cnt *c);
int udx_cnt_apply(udx_cnt *c, const udx_pair_dpd_lj *p, int n, const float *pp, float *ff);
int udx_cnt_apply_seed(udx_cnt *c, const udx_pair_dpd_lj *p, int n, const float *pp, float seed_bulk,
                       float seed_halo, float *ff);
int udx_cnt_pairs(udx_cnt *c, unsigned long long *npairs, int reset);
/* contact between the membranes of udx_flu_set_rbc inside udx_flu_run (with
 * the FSI force, src/sim/objinter/imp/main.h:42-80); NULL switches it off */
int udx_flu_set_rbc_contact(udx_flu *f, udx_cnt *c, const udx_pair_dpd_lj *p);
/* walls and membranes together (udx_flu_set_wall and udx_flu_set_rbc): the
 * vertices get the wall DPD with the FSI pair (seeded by the wall trunk after
 * the solvent's draw of each step; with a seed table the solvent's wall seed
 * is its column 27 and the membranes' the handle's next wall-trunk draw; pairs
 * counted by udx_wall_pairs), the wall repulsion with coefficient lambda
 * (wall_repulse, src/wall/force/imp.h:17-19; 0, the default, for "none") and
 * the wall bounce-back after the membrane advance */
int udx_flu_set_rbc_repulsion(udx_flu *f, float lambda);

/* ---- membrane bounce-back: replaces mesh_bounce_ini / _fin /
 *      _find_collisions / _select_collisions / _bounce /
 *      _collect_rbc_momentum (src/mesh_bounce/imp.h:17-31) and the mesh
 *      halo of src/exch/mesh on one periodic domain ----
 * Solvent particles that cross a membrane triangle during the step are
 * reflected with the triangle's velocity at the crossing point and pushed
 * 0.01 back to their side; the momentum they give each triangle changes the
 * velocities of its three vertices.  Triangles move linearly from rp (the
 * vertices at the start of the step) to rc (after the membrane advance);
 * particles from pp0 (start of the step) to pp (after the move, not
 * wrapped).  Meshes leaving the domain cropped by 1 also act through their
 * periodic images.  starts: the solvent cell list of pp0 (L^3 + 1 entries,
 * cells of unit size, x fastest); pp must be in that order.  All pointers
 * are device memory; rp and rc hold 6 floats per vertex (rc's velocities are
 * updated), mm (optional) receives P and L per triangle (6 floats).  A
 * particle takes its earliest crossing (ties: the lower triangle instance),
 * a triangle's momentum sums its crossings in particle order and a vertex
 * adds its triangles' changes in triangle order, so results do not depend
 * on scheduling. */
typedef struct udx_mbb udx_mbb;
int udx_mbb_ini(udx_ctx *ctx, int nv, int nt, const int *tt_host, int max_meshes, int maxp, udx_mbb **b);
int udx_mbb_fin(udx_mbb *b);
int udx_mbb_bounce(udx_mbb *b, float dt, float flu_mass, float rbc_mass, int Lx, int Ly, int Lz, int nm,
                   const float *rp, float *rc, const int *starts, int n, const float *pp0, float *pp, float *mm);
/* particles bounced by the last call */
int udx_mbb_count(udx_mbb *b, unsigned long long *ncol);
/* per particle of the last call: the triangle (mesh * nt + t) it bounced
 * off, -1 for none (device memory, n entries) */
const int *udx_mbb_collisions(const udx_mbb *b);
/* the list of (triangle, particle) pairs that pass the plane test and get
 * the full crossing test (default max(2 maxp, 2^20)); when a call has more,
 * the search runs again with the full test in place (same results, slower) */
int udx_mbb_set_pair_capacity(udx_mbb *b, int cap);
/* bounce-back inside udx_flu_run after the membrane advance
 * (objects_bounce, src/sim/objects/imp/bounce.h:236-248); NULL switches it
 * off.  Needs the membranes of udx_flu_set_rbc; the same template. */
int udx_flu_set_rbc_bounce(udx_flu *f, udx_mbb *b);

/* ---- configuration: replaces conf_ini / conf_read / conf_lookup_* /
 *      conf_set_* (src/conf/imp.h) and the parameter readers pair_set_conf
 *      (src/pair/conf.cpp:17-57) and rbc_params_set_conf
 *      (src/rbc/params/conf.cpp:11-34) ----
 * The libconfig subset of the reference's .cfg files.  Four layers are
 * searched in order: values set by the program, the argument string, the
 * optional file, the default file.  Lookups are typed ("2" is an int, "2.0"
 * a float); a missing key or a wrong type is an error (UDX_EINVAL) with the
 * reference's message.  Host only: no GPU work. */
typedef struct udx_conf udx_conf;
int udx_conf_ini(udx_conf **c);
int udx_conf_fin(udx_conf *c);
int udx_conf_set_include_dir(udx_conf *c, const char *dir);   /* for @include */
/* any argument may be NULL; args is settings text ("a = 1 b = { c = 2.0 }") */
int udx_conf_read(udx_conf *c, const char *default_file, const char *opt_file, const char *args);
int udx_conf_has(const udx_conf *c, const char *desc);        /* 1 if found in any layer */
int udx_conf_lookup_int(const udx_conf *c, const char *desc, int *a);
int udx_conf_lookup_float(const udx_conf *c, const char *desc, float *a);
int udx_conf_lookup_bool(const udx_conf *c, const char *desc, int *a);
/* the string stays valid while the udx_conf lives */
int udx_conf_lookup_string(const udx_conf *c, const char *desc, const char **a);
int udx_conf_lookup_vint(const udx_conf *c, const char *desc, int maxn, int *n, int *a);
int udx_conf_lookup_vfloat(const udx_conf *c, const char *desc, int maxn, int *n, float *a);
int udx_conf_lookup_vstring(const udx_conf *c, const char *desc, int maxn, int *n, const char **a);
int udx_conf_lookup_int3(const udx_conf *c, const char *desc, int a[3]);
int udx_conf_lookup_float3(const udx_conf *c, const char *desc, float a[3]);
int udx_conf_set_int(udx_conf *c, const char *desc, int a);
int udx_conf_set_float(udx_conf *c, const char *desc, float a);
int udx_conf_set_bool(udx_conf *c, const char *desc, int a);
int udx_conf_set_string(udx_conf *c, const char *desc, const char *a);
/* <ns>.{dpd, a, g, spow} -> grey DPD parameters with sigma from kBT and dt */
int udx_pair_set_conf(const udx_conf *c, const char *ns, float kBT, float dt, udx_pair_dpd *p);
/* <ns>.{gammaC, gammaT, kBT, kb, phi, ks, x0, mpow, ka, kd, kv, totArea, totVolume} */
int udx_rbc_params_set_conf(const udx_conf *c, const char *ns, udx_rbc_params *p);
/* wvel.type ("constant", "shear", "shear sin", "hele shaw") and its keys
 * (wvel_set_conf, src/wall/wvel/conf.cpp:10-47) */
int udx_wall_vel_set_conf(const udx_conf *c, udx_wall_vel *v);
/* the wall repulsion coefficient of an object type for
 * udx_flu_set_rbc_repulsion (ini_repulsion_params,
 * src/sim/objects/imp/ini.h:69-75, and wall_repulse_prm_ini_conf,
 * src/wall/imp/repulse.h:8-12): <ns>.repulsion is "none" (lambda = 0) or the
 * name of a group whose key lambda is read */
int udx_wall_repulse_set_conf(const udx_conf *c, const char *ns, float *lambda);

/* ---- restart files: replaces restart_write_pp / restart_read_pp /
 *      restart_write_ii / restart_read_ii (src/io/restart/imp.h:14-32),
 *      flu_strt_dump / flu_strt_quants (src/flu/imp/start.h:20-41) and the
 *      wall template (src/wall/imp/strt.h:1-60) ----
 * Files BASE/CODE/ID.bop (text header) and BASE/CODE/ID.values (n x 6
 * floats "x y z vx vy vz", or n ints "i"), ID = %05d, "templ" or "final";
 * the directories must exist.  The reader takes the binary float/int and the
 * ascii/iascii value formats and fails (UDX_ECAP) above maxn entries.
 * udx_restart_* are host only (host buffers, no GPU work). */
enum { UDX_RESTART_TEMPL = -1, UDX_RESTART_FINAL = -2 };
int udx_restart_write_pp(const char *base, const char *code, int id, long n, const float *pp);
int udx_restart_read_pp(const char *base, const char *code, int id, int maxn, int *n, float *pp);
int udx_restart_write_ii(const char *base, const char *code, int id, long n, const int *ii);
int udx_restart_read_ii(const char *base, const char *code, int id, int maxn, int *n, int *ii);
/* the solvent's particles as CODE "flu.pp" (cell order of the last step) */
int udx_flu_strt_dump(udx_flu *f, const char *base, int id);
int udx_flu_strt_quants(udx_flu *f, const char *base, int id);
/* the frozen wall particles as CODE "wall.pp", ID templ; udx_wall_strt_gen
 * builds the wall (images, cell list) from them instead of freezing */
int udx_wall_strt_dump_templ(const udx_wall *w, const char *base);
int udx_wall_strt_gen(udx_ctx *ctx, const udx_sdf *q, const char *base, int maxn, udx_wall **w);

/* ---- ensemble: many independent periodic boxes advanced together ----
 * No counterpart in the reference (each box there is a separate run of the
 * solvent step, src/sim/imp/run.h:8-68).  Every kernel of the step is
 * launched once for all boxes; per box the arithmetic and its order equal a
 * single-box udx_flu run with the same seeds, so results are identical.
 * Box b's default trunks are the solvent trunks with the first KISS word
 * offset by 7919 b (box 0 = the reference's). */
typedef struct udx_ens_box {
    int L[3];
    int n;                     /* particles */
    udx_pair_dpd dpd;          /* a, gamma, sigma, spow (spow shared by all boxes) */
    float dt, mass;
    float accel[3];
} udx_ens_box;
typedef struct udx_ens udx_ens;
int udx_ens_ini(udx_ctx *ctx, int nbox, const udx_ens_box *boxes, udx_ens **e);
int udx_ens_fin(udx_ens *e);
int udx_ens_set_mode(udx_ens *e, int mode);      /* UDX_MODE_FAST, UDX_MODE_EXACT or UDX_MODE_HALF */
/* pp: AoS particles of all boxes, box after box (sum of n); positions are
 * wrapped into their box's [-L/2, L/2) (each box keeps all its particles) */
int udx_ens_set_pp(udx_ens *e, const float *pp);
int udx_ens_get_pp(udx_ens *e, float *pp);       /* cell order within each box */
/* seeds: NULL (the boxes' trunks) or nsteps x nbox x 27 floats (26 halo
 * fragments then the bulk, as udx_fluforces_apply_seeds) */
int udx_ens_run(udx_ens *e, int nsteps, const float *seeds);
/* replay each pair of steps from one captured hipGraph */
int udx_ens_set_graph(udx_ens *e, int enable);
/* time the force kernel's tile variants (threads per block, queued pairs
 * per lane) over `steps` steps each and keep the fastest; ms[v] per step */
int udx_ens_autotune(udx_ens *e, int steps, int *best, float *ms);
int udx_ens_set_variant(udx_ens *e, int variant);
int udx_ens_pairs(udx_ens *e, unsigned long long *npairs, int reset);   /* unordered pairs */

#ifdef __cplusplus
}
#endif
#endif
