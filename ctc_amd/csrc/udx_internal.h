// Host-side internals of libudx: contexts, error reporting, device buffers and
// the kernel launchers each module exports to the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/udx.h"
#include "udx_dev.h"

namespace udx {

void set_error(const char *fmt, ...);

#define UDX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::udx::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #call,               \
                             hipGetErrorString(e_));                                   \
            return UDX_EHIP;                                                           \
        }                                                                              \
    } while (0)

#define UDX_CHECK(cond, code, ...)                                                     \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::udx::set_error(__VA_ARGS__);                                             \
            return (code);                                                             \
        }                                                                              \
    } while (0)

#define UDX_TRY(call)                                                                  \
    do {                                                                               \
        int s_ = (call);                                                               \
        if (s_ != UDX_OK) return s_;                                                   \
    } while (0)

// launch-error check after a kernel launch
#define UDX_LAUNCHED() UDX_HIP(hipGetLastError())

template <typename T>
int dev_alloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    UDX_HIP(hipMalloc((void **)p, n * sizeof(T)));
    return UDX_OK;
}

template <typename T>
void dev_free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------------ grid
struct Box {
    int L[3];
    int ncells() const { return L[0] * L[1] * L[2]; }
};

// ------------------------------------------------------------------ scan
// exclusive scan of n ints into out[0..n] (out[n] = total); scratch holds
// scan_scratch_size(n) ints
int scan_scratch_size(int n);
int scan_exclusive(hipStream_t s, const int *in, int n, int *out, int *scratch);
// the same, clearing in[] as it is read
int scan_exclusive_consume(hipStream_t s, int *in, int n, int *out, int *scratch);
// out[0..n) only (no total at out[n])
int scan_exclusive_n(hipStream_t s, const int *in, int n, int *out, int *scratch);

// ------------------------------------------------------------------ cell list
// Particles are sorted by cell and, inside a cell, by x slab: every cell is
// cut into `sub` slabs of width 1/sub along x (slab id = cell * sub +
// floor((x + L/2) sub) mod sub), ties keeping the input order.  The slab
// starts let a force kernel read only the part of a row of cells that lies
// within the cutoff along x.
struct ClistBufs {
    Box box;
    int maxp;
    int sub;         // x slabs per cell (1: plain cells)
    int *counts;     // ncells
    int *starts;     // ncells + 1 (starts[ncells] = number kept)
    int *bcounts;    // ncells * sub (the same array as counts when sub == 1)
    int *bstarts;    // ncells * sub + 1 (the same array as starts when sub == 1)
    int2 *cellslot;  // maxp: (slab, provisional slot)
    int *order;      // maxp: input index of each output slot
    int *scan_tmp;
};
// x slabs per cell of the solvent's cell lists
constexpr int SOLVENT_SUB = 4;
int clist_alloc(ClistBufs *c, const Box &b, int maxp, int sub = 1);
void clist_free(ClistBufs *c);
// pos4 (input, n particles): counting sort into cell order; order[] and
// starts[] filled; the number kept is starts[ncells] on the device
int clist_sort(hipStream_t s, ClistBufs *c, int n, const float4 *pos4);
int clist_sort_aos(hipStream_t s, ClistBufs *c, int n, const float *pp);
// the two halves of a sort for callers that fill bcounts[] and cellslot[]
// themselves: zero the slab counts; scan, scatter, the stable in-slab order
// and the cell tables
int clist_clear(hipStream_t s, ClistBufs *c);
int clist_finish(hipStream_t s, ClistBufs *c, int n);
// the counting half alone (clist_finish completes the sort)
int clist_count_only(hipStream_t s, ClistBufs *c, int n, const float4 *pos4);
// out[k] = in[order[k]] for k < starts[ncells]; n bounds the launch
int gather_f4(hipStream_t s, const ClistBufs *c, int n, const float4 *pin, const float4 *vin,
              float4 *pout, float4 *vout);
int gather_aos(hipStream_t s, const ClistBufs *c, int n, const float *pin, float *pout);

// ------------------------------------------------------------------ forces
struct FragTables {
    int *cnt;        // cell counts of the 26 fragments, concatenated (one 0 after each)
    int *fss;        // exclusive scan of cnt: fragment f's offsets are fss[off[f] + h] - fss[off[f]]
    int *scan_tmp;
    int off[27];     // start of each fragment in cnt
    int size;
};
struct FragOff {
    int off[27];
};
int frag_alloc(FragTables *t, const Box &b);
void frag_free(FragTables *t);
// ecount (optional): a box's edge list counter, cleared here for the step
int frag_build(hipStream_t s, const FragTables *t, const Box &b, const int *starts, int *ecount = nullptr);

struct DpdLaunch {
    const float4 *pos, *vel;
    const int *starts;      // ncells + 1, starts[ncells] = n
    const int *bstarts = nullptr;   // x-slab starts (ncells * sub + 1) of a slab-sorted list, or null
    int sub = 1;
    const FragTables *frag;
    Box box;
    int nmax;               // launch bound (>= particles)
    int periodic;
    udx_pair_dpd par;
    const float *seeds;     // device, 27 floats
    float *ff;              // AoS Force (stride 3) or float4 (stride 4)
    int ff_stride;
    int accumulate;
    unsigned long long *pairs;   // ordered-pair counter or null
    int i0 = 0;             // first particle of the box in pos/vel (noise indices are relative)
    float *ss = nullptr;    // per-particle stress, 6 floats added (null: forces only)
    // FP64 accumulators, 3 per particle, zero on entry and left zero: with one,
    // UDX_MODE_HALF evaluates each pair once (half shell) and sums exactly
    double *acc = nullptr;
    // edge list of a periodic box (entries from i0, nmax capacity; *ecount
    // cleared by frag_build for the step): with one the gather's halo pass
    // runs as its own launch over the listed particles only
    int *ecount = nullptr, *eidx = nullptr;
    float4 *ef = nullptr;
    // with one, the gather also moves the particles (dpd_move_fused)
    const struct FusedMove *move = nullptr;
};
// The move fused into the gather, for the solvent step without objects: after
// its sum each particle that is not sdf_far adds its wall force (the gather's
// wall pass: what wall_forces computes), stores the total in ff, kick-drifts
// with the body force, bounces off the walls, wraps, writes itself to
// opos/ovel at its own index and counts itself into next's x slabs: what
// wall_forces and move_f4_wall do after the pass, with the same operations in
// the same order.
struct FusedMove {
    float4 *opos, *ovel;
    float dt, mass, accel[3];
    int walls;         // the fields below are set
    SdfView sdf;
    WallVel wv;
    const float4 *wpos;        // frozen wall particles (nwall) in the cell order of the margin grid W
    const int *wstarts;
    int W[3], nwall;
    const float *wseed;        // device, the step's wall noise seed
    unsigned long long *wpairs;   // solvent-wall pair counter or null
    ClistBufs *next;
    int slab0 = 0;             // the box's first slab in next (boxes end to end)
    int *step = nullptr;       // device step counter advanced once per step (batches)
};
// whether dpd_forces_mode can run a with a.move set: the gather (FAST or
// EXACT) in one launch (no split halo) writing float4 forces
bool dpd_move_fused(const DpdLaunch &a, int mode);
int dpd_forces(hipStream_t s, const DpdLaunch &a);
// mode: UDX_MODE_FAST or UDX_MODE_EXACT
int dpd_forces_mode(hipStream_t s, const DpdLaunch &a, int mode);
// the gather's halo pass runs as its own launch when the expected boundary
// particles of a launch (all boxes of a batch) reach n
int dpd_set_split_edges(double n);

// many boxes in one launch (one descriptor per box, all sharing spow)
struct DpdBatch {
    void *table = nullptr;     // device: per-box kernel descriptors
    int *bblock[2] = {nullptr, nullptr};   // device: first block of each box for 128 / 256 threads
    int nblocks[2] = {0, 0};
    int nbox = 0;
    float spow = 0.25f;
    // the half-shell pass (UDX_MODE_FAST): FP64 sums shared by the boxes, and
    // the float forces they are rounded into (all boxes' particles in one array)
    double *acc = nullptr;
    float *ff = nullptr;
    int ntot = 0;
    const int *n_end = nullptr;   // device: end of the last box's cell list
    int split = 0;                // every box has an edge list: bulk and halo passes launched apart
    // the boxes' moves fused into the gather (FusedMove of every box): the
    // device step counter the first launch of a step advances, or null
    int *step = nullptr;
};
// tile variants of the batched force kernel: threads per block and queued
// pairs per lane; variant 0 is the default
enum { DPD_NVARIANT = 6 };
void dpd_variant_shape(int v, int *threads, int *qcap);
int dpd_batch_build(DpdBatch *b, const DpdLaunch *boxes, int nbox);
void dpd_batch_free(DpdBatch *b);
int dpd_forces_batch(hipStream_t s, const DpdBatch &b, int mode, int variant);

// ------------------------------------------------------------------ scheme
int aos_to_f4(hipStream_t s, int n, const float *pp, float4 *pos, float4 *vel);
int f4_to_aos(hipStream_t s, int n, const int *n_dev, const float4 *pos, const float4 *vel, float *pp);
int f4_to_force_aos(hipStream_t s, int n, const int *n_dev, const float4 *ff, float *out);
// next (with wrap): also count the moved particles into next's x slabs, the
// counting pass of its next sort (clist_finish then completes it)
int move_f4(hipStream_t s, int nmax, const int *n_dev, const Box &b, float dt, float mass,
            const float accel[3], const float4 *ff, float4 *pos, float4 *vel, int wrap = 1,
            ClistBufs *next = nullptr);
int move_aos(hipStream_t s, int n, float dt, float mass, const float *ff, float *pp);
int bforce_aos(hipStream_t s, int n, float mass, const float accel[3], float *ff);
int wrap_aos(hipStream_t s, const Box &b, int n, float *pp);
int rnd_branch(hipStream_t s, float seed, int n, const int *u, const int *v, int uu, float *out);

// ------------------------------------------------------------------ walls
struct SdfField {
    int L[3], T[3];
    float *tex;      // T[0] T[1] T[2] texels, x fastest
    SdfView view;    // device view (tex pointer, grid, transform, threshold)
    int ready;
    // runs [x, y) of consecutive box cells that can hold a particle that is
    // not sdf_far (every other cell's particles are far from the walls)
    int2 *segs = nullptr;
    int nseg = 0;
};
struct WallSet {
    ClistBufs cl;    // cell list of the wall particles on the margin box
    float4 *pos;     // cell-ordered wall positions
    float *pp_aos;   // the same in AoS (velocity 0)
    float *frozen;   // the frozen particles in freeze order (AoS), the restart template
    int n, nfrozen;  // all wall particles (frozen + periodic images), frozen ones
};
void sdf_texture_size(const int L[3], int T[3]);
int sdf_build(hipStream_t s, SdfField *q, const float ext[3], const int N[3], const float *D_host);
int sdf_probe(hipStream_t s, const SdfField *q, int n, const float *r3, int what, float *out);
int wall_build(hipStream_t s, WallSet *w, const SdfField *q, int *n, float *pp);
int wall_build_frozen(hipStream_t s, WallSet *w, const SdfField *q, int nw, float *frozen);
void wall_free(WallSet *w);
// assign: the particles that are not sdf_far get ff = their wall force (the
// others are left alone) instead of ff += it.  cell_starts: the particles are
// in the cell order of the SDF's box with these starts, so only the runs of
// cells near a wall are visited.
int wall_forces(hipStream_t s, const SdfField *q, const WallSet *w, int nmax, const int *n_dev, const float4 *pos,
                const float4 *vel, const udx_pair_dpd &p, const float *seed_dev, float *ff, int ff_stride,
                unsigned long long *pairs, int exact, const WallVel &wv, int assign = 0,
                const int *cell_starts = nullptr);
int wall_bounce_aos(hipStream_t s, const SdfField *q, float dt, int n, float *pp, const WallVel &wv);
// the wall velocity of v at time t: shear sin resolved to its rate gdot sin(w t)
// (src/wall/wvel/imp/main.h:71-84)
WallVel wall_vel_view(const udx_wall_vel &v, float t);
int wall_repulse_aos(hipStream_t s, const SdfField *q, float lambda, int n, const float *pp, float *ff);
int move_f4_wall(hipStream_t s, int nmax, const int *n_dev, const Box &b, float dt, float mass, const float accel[3],
                 const float4 *ff, float4 *pos, float4 *vel, const SdfField *q, const WallVel &wv, int wrap = 1,
                 ClistBufs *next = nullptr);
// membrane bounce-back inside the solvent step: rp, rc the vertices (AoS) at
// the start of the step and after the membrane advance (velocities updated),
// pos0 the solvent positions at the start of the step; applies the periodic
// wrap to every particle
// Juelicher bending of one template (juelicher.hip): build from the
// triangles, add the forces of nc cells to ff (3 floats per vertex)
struct JulTemplate;
int jul_build(hipStream_t s, int nv, int nt, const int *tt_host, int max_cells, JulTemplate **out);
int jul_apply(hipStream_t s, const JulTemplate *j, int nc, const float *pp, float *ff);
void jul_free(JulTemplate *j);
// membrane contact forces (cnt.hip) and the handle's seed draws
udx_ctx *cnt_ctx(const udx_cnt *c);
int cnt_forces(hipStream_t s, udx_cnt *c, const udx_pair_dpd_lj &pr, int n, const float *pp, float seed_bulk,
               float seed_halo, float *ff);
float cnt_draw(udx_cnt *c);
udx_ctx *mbb_ctx(const udx_mbb *b);
int mbb_nv(const udx_mbb *b);
int mbb_step(hipStream_t s, udx_mbb *b, float dt, float flu_mass, float rbc_mass, const Box &box, int nm,
             const float *rp, float *rc, const int *starts, int nmax, const int *n_dev, float4 *pos, float4 *vel,
             const float4 *pos0);
int diag_aos(hipStream_t s, int n, const float *pp, double *partials_dev, double out[5]);

// solvent trunk f of one rank: the halo fragments 0..25
// (src/fluforces/imp/ini.h:11-58, seed_base = 0) and the bulk trunk at 26
// (src/fluforces/imp/ini.h:6)
void flu_trunk_ini(int f, udx_rnd *r);

// ------------------------------------------------------------------ FSI
struct FsiBufs {
    Box box;
    int max_obj;
    int *flag, *aid, *scan_tmp;   // 26 * max_obj halo flags and their exclusive scan
    float *inst, *inst_sorted;    // vertex and image instances (AoS), 8 * max_obj
    int2 *meta, *meta_sorted;     // (noise index, image?)
    float *sinst;                 // per-instance force sums, 3 * 8 * max_obj
    ClistBufs cl;                 // instances on the (L+2)^3 grid of cells -1..L
};
int fsi_alloc(FsiBufs *f, const Box &b, int max_obj);
void fsi_free(FsiBufs *f);
// +f on the vertices (fo, stride 3) and -f on the solvent (fs, fs_stride),
// both ADDED; seeds_dev = [bulk, halo]
int fsi_forces(hipStream_t s, FsiBufs *f, int ns, const float4 *spos, const float4 *svel, const int *starts,
               int no, const float *opp, const udx_pair_dpd &p, const float *seeds_dev, float *fo, float *fs,
               int fs_stride, unsigned long long *pairs, int exact);
FsiBufs *fsi_bufs(udx_fsi *q);
udx_ctx *fsi_ctx(const udx_fsi *q);
float fsi_draw(udx_fsi *q);               // next seed of the FSI trunk
unsigned long long *fsi_pairs_dev(udx_fsi *q);
udx_ctx *rbc_ctx(const udx_rbc *r);
int rbc_nv(const udx_rbc *r);
int rbc_max_cells(const udx_rbc *r);
// whole-cell periodic redistribution of nc cells of nv vertices (AoS pp,
// in place; tmp: 6 nc nv floats, cfid and cnew: nc ints)
int rbc_redistribute(hipStream_t s, const Box &b, int nc, int nv, float *pp, float *tmp, int *cfid, int *cnew);

}  // namespace udx

struct udx_ctx {
    hipStream_t stream;
    int device;
};
