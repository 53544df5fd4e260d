// Ensemble of independent periodic DPD boxes advanced together on MI355X.
//
// The boxes (edges, densities, gamma, kBT and seeds of their own) are laid
// end to end: particles box-major, cells box-major, fragment tables
// box-major.  One step is the same sequence of kernels as one box's step —
// cell list, reorder, fragment tables, forces, kick-drift + wrap — each
// launched once for all boxes: the per-particle kernels find their box by a
// binary search over the box offsets, and the force kernel gives every box
// its own run of blocks (no block spans two boxes) with the box's own
// parameters, seeds and tables.  Per box the arithmetic and its order are
// exactly those of a single-box run, so results are identical to launching
// each box on its own.
//
// The step pair (even/odd ping-pong) is captured once into a hipGraph and
// replayed; a device-side step counter selects each step's seeds, so the
// graph needs no host updates.  The force kernel's tile (threads per block,
// queued pairs per lane) is picked by timing the variants on the ensemble
// itself (udx_ens_autotune).
#include <vector>

#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {
namespace {

__device__ __forceinline__ int find_box(const int *off, int nbox, int i) {
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The box of particle i: one search per wave from its first active lane's
// index (wave-uniform, so it runs on scalar loads), then each lane steps over
// the box starts between that particle and its own (none or one for boxes of
// more than 64 particles).  Active lanes hold consecutive indices from the
// first active one upward.
__device__ __forceinline__ int find_box_wave(const int *off, int nbox, int i) {
    int b = find_box(off, nbox, __builtin_amdgcn_readfirstlane(i));
    while (b + 1 < nbox && off[b + 1] <= i) b++;
    return b;
}

// x slab (global index: cell * SOLVENT_SUB + slab) and provisional slot of
// each particle; consecutive lanes in one slab take their slots with one
// atomic (as k_clist_count)
__global__ __launch_bounds__(256) void k_ens_count(int n, int nbox, const int *poff, const int *coff,
                                                   const int4 *boxL, const float4 *pos, int *counts,
                                                   int2 *cellslot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int b = 0;
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) {
        b = find_box_wave(poff, nbox, i);
        r = pos[i];
    }
    const int4 L = boxL[b];
    clist_count_one(i < n, i, make_float3(r.x, r.y, r.z), L.x, L.y, L.z, SOLVENT_SUB, counts, cellslot,
                    coff[b] * SOLVENT_SUB);
}

// cell counts of the 26 fragments of every box (as k_frag_counts); the
// first threads also copy the step's seeds of every box from the run's table
// and clear the boxes' edge list counters
__global__ __launch_bounds__(256) void k_ens_frag_counts(int total, int nbox, const int *fbase, const int *foff,
                                                         const int *coff, const int4 *boxL, const int *starts,
                                                         int *cnt, const float *table, const int *step,
                                                         float *seeds, int *ecount) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < 27 * nbox) seeds[g] = table[(long)*step * 27 * nbox + g];
    if (g < nbox) ecount[g] = 0;
    if (g >= total) return;
    const int b = find_box(fbase, nbox, g);
    const int4 L = boxL[b];
    const int *off = foff + 27 * b;
    const int gl = g - fbase[b];
    int f = 0;
    while (f < 25 && gl >= off[f + 1]) f++;
    const int h = gl - off[f];
    const int dx = (f + 2) % 3 - 1, dy = (f / 3 + 2) % 3 - 1, dz = (f / 9 + 2) % 3 - 1;
    const int ex = dx ? 1 : L.x, ey = dy ? 1 : L.y, ez = dz ? 1 : L.z;
    int v = 0;
    if (h < ex * ey * ez) {
        const int x = (dx == 1 ? L.x - 1 : 0) + h % ex;
        const int y = (dy == 1 ? L.y - 1 : 0) + (h / ex) % ey;
        const int z = (dz == 1 ? L.z - 1 : 0) + h / (ex * ey);
        const int cid = coff[b] + x + L.x * (y + L.y * z);
        v = starts[cid + 1] - starts[cid];
    }
    cnt[g] = v;
}

struct MoveBox {
    float Lx, Ly, Lz, dt;
    float mass, ax, ay, az;
};

// kick-drift + wrap, then the count of the next step's slabs (as
// k_ens_count) and the step counter
__global__ __launch_bounds__(256) void k_ens_move(int n, int nbox, const int *poff, const MoveBox *mb,
                                                  const float4 *ff, float4 *pos, float4 *vel, const int *coff,
                                                  const int4 *boxL, int *counts, int2 *cellslot, int *step) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *step += 1;
    int b = 0;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) {
        b = find_box_wave(poff, nbox, i);
        const MoveBox m = mb[b];
        p = pos[i];
        float4 v = vel[i];
        const float4 f = ff[i];
        kick_drift(m.dt, m.mass, m.ax, f.x, &p.x, &v.x);
        kick_drift(m.dt, m.mass, m.ay, f.y, &p.y, &v.y);
        kick_drift(m.dt, m.mass, m.az, f.z, &p.z, &v.z);
        p.x = wrap_coord(p.x, m.Lx);
        p.y = wrap_coord(p.y, m.Ly);
        p.z = wrap_coord(p.z, m.Lz);
        pos[i] = p;
        vel[i] = v;
    }
    const int4 L = boxL[b];
    clist_count_one(i < n, i, make_float3(p.x, p.y, p.z), L.x, L.y, L.z, SOLVENT_SUB, counts, cellslot,
                    coff[b] * SOLVENT_SUB);
}

// periodic wrap of the particles into their boxes (set_pp): the boxes keep
// every particle, so a position rounded onto +L/2 must come back inside
__global__ __launch_bounds__(256) void k_ens_wrap(int n, int nbox, const int *poff, const MoveBox *mb, float4 *pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MoveBox m = mb[find_box_wave(poff, nbox, i)];
    float4 p = pos[i];
    p.x = wrap_coord(p.x, m.Lx);
    p.y = wrap_coord(p.y, m.Ly);
    p.z = wrap_coord(p.z, m.Lz);
    pos[i] = p;
}

__global__ void k_ens_unpack(int n, const float4 *pos, const float4 *vel, float *pp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i], v = vel[i];
    float *o = pp + 6 * (long)i;
    o[0] = p.x; o[1] = p.y; o[2] = p.z;
    o[3] = v.x; o[4] = v.y; o[5] = v.z;
}

}  // namespace
}  // namespace udx

using namespace udx;

struct udx_ens {
    udx_ctx *ctx;
    hipStream_t s;                 // own stream (graph capture needs a non-null stream)
    hipEvent_t done;
    int nbox, ntot, ctot, ftot, mode, variant, cur, use_graph, seed_cap;
    int counted;                   // the last move counted pos[cur] into the slabs
    std::vector<udx_ens_box> box;
    std::vector<int> poff_h, coff_h, fbase_h, foff_h;
    std::vector<udx_rnd> trunk;    // 27 per box
    int *poff, *coff, *fbase, *foff;
    int4 *boxL;
    MoveBox *mb;
    float4 *pos[2], *vel[2], *ff;
    double *acc;                   // FP64 force sums of the fast pass (kept zero between steps)
    int *ecount, *eidx;            // edge lists of the gather's halo pass (a counter per box)
    float4 *ef;
    ClistBufs cl;
    int *fcnt, *fss, *fscan;
    float *seeds_table, *seeds_cur, *seeds_host;
    int *step;
    unsigned long long *pairs;
    std::vector<FragTables> frag;  // per-box views into fss
    DpdBatch batch[2];             // descriptors for the two ping-pong parities
    hipGraph_t graph;
    hipGraphExec_t exec;
    int graph_variant;
};

// Both parities' tables carry the boxes' moves: in FAST and EXACT mode the
// gather moves the particles into the other pair of streams and counts them
// into the next step's slabs (the half-shell mode moves them in k_ens_move).
static int ens_build_batches(udx_ens *e) {
    for (int par = 0; par < 2; par++) {
        std::vector<DpdLaunch> d(e->nbox);
        std::vector<FusedMove> mv(e->nbox);
        for (int b = 0; b < e->nbox; b++) {
            FusedMove &m = mv[b];
            m = FusedMove{};
            m.opos = e->pos[1 - par];
            m.ovel = e->vel[1 - par];
            m.dt = e->box[b].dt;
            m.mass = e->box[b].mass;
            for (int c = 0; c < 3; c++) m.accel[c] = e->box[b].accel[c];
            m.walls = 0;
            m.next = &e->cl;
            m.slab0 = e->coff_h[b] * SOLVENT_SUB;
            m.step = e->step;
            DpdLaunch &a = d[b];
            a.move = &m;
            a.pos = e->pos[par];
            a.vel = e->vel[par];
            a.starts = e->cl.starts + e->coff_h[b];
            a.bstarts = e->cl.bstarts + (size_t)e->coff_h[b] * SOLVENT_SUB;
            a.sub = SOLVENT_SUB;
            a.frag = &e->frag[b];
            a.box = Box{{e->box[b].L[0], e->box[b].L[1], e->box[b].L[2]}};
            a.nmax = e->box[b].n;
            a.periodic = 1;
            a.par = e->box[b].dpd;
            a.seeds = e->seeds_cur + 27 * b;
            a.ff = (float *)e->ff;
            a.ff_stride = 4;
            a.accumulate = 0;
            a.pairs = e->pairs;
            a.i0 = e->poff_h[b];
            a.acc = e->acc;
            a.ecount = e->ecount + b;
            a.eidx = e->eidx;
            a.ef = e->ef;
        }
        UDX_TRY(dpd_batch_build(&e->batch[par], d.data(), e->nbox));
    }
    return UDX_OK;
}

// one step: state in pos/vel[cur] -> pos/vel[1 - cur]
static int ens_count(udx_ens *e, int cur) {
    const int n = e->ntot;
    UDX_TRY(clist_clear(e->s, &e->cl));
    hipLaunchKernelGGL(k_ens_count, dim3(ceil_div(n, 256)), dim3(256), 0, e->s, n, e->nbox, e->poff, e->coff,
                       e->boxL, e->pos[cur], e->cl.bcounts, e->cl.cellslot);
    UDX_LAUNCHED();
    return UDX_OK;
}

// The move counts the next step's slabs, so a step starts from the scan
// (ens_count first when the positions were set since).  *next: the parity
// that holds the state after the step (the gather's move writes it back to
// cur; k_ens_move moves it in 1 - cur).
static int ens_step(udx_ens *e, int cur, int *next) {
    hipStream_t s = e->s;
    const int a = cur, b = 1 - cur, n = e->ntot;
    if (!e->counted) UDX_TRY(ens_count(e, a));
    UDX_TRY(clist_finish(s, &e->cl, n));
    UDX_TRY(gather_f4(s, &e->cl, n, e->pos[a], e->vel[a], e->pos[b], e->vel[b]));
    const int nf = e->ftot > 27 * e->nbox ? e->ftot : 27 * e->nbox;
    hipLaunchKernelGGL(k_ens_frag_counts, dim3(ceil_div(nf, 256)), dim3(256), 0, s, e->ftot, e->nbox, e->fbase,
                       e->foff, e->coff, e->boxL, e->cl.starts, e->fcnt, e->seeds_table, e->step, e->seeds_cur,
                       e->ecount);
    UDX_LAUNCHED();
    UDX_TRY(scan_exclusive(s, e->fcnt, e->ftot, e->fss, e->fscan));
    UDX_TRY(dpd_forces_batch(s, e->batch[b], e->mode, e->variant));
    const bool fused = e->mode != UDX_MODE_HALF;
    if (!fused) {
        hipLaunchKernelGGL(k_ens_move, dim3(ceil_div(n > 0 ? n : 1, 256)), dim3(256), 0, s, n, e->nbox, e->poff,
                           e->mb, e->ff, e->pos[b], e->vel[b], e->coff, e->boxL, e->cl.bcounts, e->cl.cellslot,
                           e->step);
        UDX_LAUNCHED();
    }
    e->counted = 1;
    *next = fused ? a : b;
    return UDX_OK;
}

static void ens_drop_graph(udx_ens *e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

// the even/odd step pair as one graph (state ends where it started)
static int ens_capture(udx_ens *e) {
    ens_drop_graph(e);
    // the captured steps start from counted positions (udx_ens_run counts
    // before the first replay when they are not)
    const int counted0 = e->counted;
    e->counted = 1;
    UDX_HIP(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
    int mid = e->cur, end = e->cur;
    int st = ens_step(e, e->cur, &mid);
    if (!st) st = ens_step(e, mid, &end);
    if (!st && end != e->cur) st = UDX_EINVAL;   // the pair ends where it started
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->s, &g);
    e->counted = counted0;
    if (st && g) (void)hipGraphDestroy(g);
    UDX_TRY(st);
    UDX_CHECK(ce == hipSuccess && g, UDX_EHIP, "ensemble: graph capture failed: %s", hipGetErrorString(ce));
    e->graph = g;
    UDX_HIP(hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
    e->graph_variant = e->variant * 2 + e->cur;
    return UDX_OK;
}

static int ens_seed_table(udx_ens *e, int nsteps, const float *seeds) {
    const size_t per = 27 * (size_t)e->nbox;
    if (nsteps > e->seed_cap) {
        UDX_HIP(hipStreamSynchronize(e->s));
        dev_free(e->seeds_table);
        if (e->seeds_host) (void)hipHostFree(e->seeds_host);
        e->seeds_host = nullptr;
        UDX_TRY(dev_alloc(&e->seeds_table, per * nsteps));
        UDX_HIP(hipHostMalloc((void **)&e->seeds_host, sizeof(float) * per * nsteps));
        e->seed_cap = nsteps;
        // the graph reads the table through a fixed pointer
        ens_drop_graph(e);
    } else {
        UDX_HIP(hipStreamSynchronize(e->s));
    }
    for (int k = 0; k < nsteps; k++)
        for (int b = 0; b < e->nbox; b++)
            for (int f = 0; f < 27; f++) {
                float v;
                if (seeds) v = seeds[((size_t)k * e->nbox + b) * 27 + f];
                else v = udx_rnd_get(&e->trunk[27 * b + f]);
                e->seeds_host[((size_t)k * e->nbox + b) * 27 + f] = v;
            }
    UDX_HIP(hipMemcpyAsync(e->seeds_table, e->seeds_host, sizeof(float) * per * nsteps, hipMemcpyHostToDevice,
                           e->s));
    UDX_HIP(hipMemsetAsync(e->step, 0, sizeof(int), e->s));
    return UDX_OK;
}

extern "C" {

int udx_ens_fin(udx_ens *e) {
    if (!e) return UDX_OK;
    if (e->s) (void)hipStreamSynchronize(e->s);
    ens_drop_graph(e);
    for (int k = 0; k < 2; k++) {
        dpd_batch_free(&e->batch[k]);
        dev_free(e->pos[k]);
        dev_free(e->vel[k]);
    }
    dev_free(e->ff);
    dev_free(e->acc);
    dev_free(e->ecount);
    dev_free(e->eidx);
    dev_free(e->ef);
    dev_free(e->poff);
    dev_free(e->coff);
    dev_free(e->fbase);
    dev_free(e->foff);
    dev_free(e->boxL);
    dev_free(e->mb);
    dev_free(e->fcnt);
    dev_free(e->fss);
    dev_free(e->fscan);
    dev_free(e->seeds_table);
    dev_free(e->seeds_cur);
    dev_free(e->step);
    dev_free(e->pairs);
    if (e->seeds_host) (void)hipHostFree(e->seeds_host);
    clist_free(&e->cl);
    if (e->done) (void)hipEventDestroy(e->done);
    if (e->s) (void)hipStreamDestroy(e->s);
    delete e;
    return UDX_OK;
}

int udx_ens_ini(udx_ctx *ctx, int nbox, const udx_ens_box *boxes, udx_ens **out) {
    UDX_CHECK(ctx && boxes && out && nbox > 0, UDX_EINVAL, "udx_ens_ini: bad arguments");
    udx_ens *e = new udx_ens();
    e->ctx = ctx;
    e->nbox = nbox;
    e->box.assign(boxes, boxes + nbox);
    e->poff_h.assign(nbox + 1, 0);
    e->coff_h.assign(nbox + 1, 0);
    e->fbase_h.assign(nbox + 1, 0);
    e->foff_h.assign(27 * nbox, 0);
    e->trunk.resize(27 * (size_t)nbox);
    std::vector<int4> L(nbox);
    std::vector<MoveBox> mb(nbox);
    int st = UDX_OK;
    for (int b = 0; b < nbox && !st; b++) {
        const udx_ens_box &x = boxes[b];
        if (x.L[0] < 3 || x.L[1] < 3 || x.L[2] < 3 || x.n < 0 || x.mass <= 0) {
            set_error("udx_ens_ini: box %d: sides must be >= 3, n >= 0, mass > 0", b);
            st = UDX_EINVAL;
            break;
        }
        e->poff_h[b + 1] = e->poff_h[b] + x.n;
        e->coff_h[b + 1] = e->coff_h[b] + x.L[0] * x.L[1] * x.L[2];
        int off = 0;
        for (int f = 0; f < 26; f++) {
            const int dx = (f + 2) % 3 - 1, dy = (f / 3 + 2) % 3 - 1, dz = (f / 9 + 2) % 3 - 1;
            e->foff_h[27 * b + f] = off;
            off += (dx ? 1 : x.L[0]) * (dy ? 1 : x.L[1]) * (dz ? 1 : x.L[2]) + 1;
        }
        e->foff_h[27 * b + 26] = off;
        e->fbase_h[b + 1] = e->fbase_h[b] + off;
        L[b] = make_int4(x.L[0], x.L[1], x.L[2], 0);
        mb[b] = MoveBox{(float)x.L[0], (float)x.L[1], (float)x.L[2], x.dt, x.mass, x.accel[0], x.accel[1],
                        x.accel[2]};
        // box b's trunks: the reference's, first KISS word offset by 7919 b
        for (int f = 0; f < 27; f++) {
            flu_trunk_ini(f, &e->trunk[27 * b + f]);
            e->trunk[27 * b + f].x += 7919u * (uint32_t)b;
        }
    }
    e->ntot = e->poff_h[nbox];
    e->ctot = e->coff_h[nbox];
    e->ftot = e->fbase_h[nbox];
    if (!st && (long)e->ntot >= (1L << 27)) {
        set_error("udx_ens_ini: %d particles exceed the 2^27 queue index range", e->ntot);
        st = UDX_ECAP;
    }
    if (!st && hipStreamCreate(&e->s) != hipSuccess) st = UDX_EHIP;
    if (!st && hipEventCreateWithFlags(&e->done, hipEventDisableTiming) != hipSuccess) st = UDX_EHIP;
    auto up = [&](auto *dst, const void *src, size_t bytes) {
        if (!st && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) st = UDX_EHIP;
    };
    if (!st) st = dev_alloc(&e->poff, (size_t)nbox + 1);
    if (!st) st = dev_alloc(&e->coff, (size_t)nbox + 1);
    if (!st) st = dev_alloc(&e->fbase, (size_t)nbox + 1);
    if (!st) st = dev_alloc(&e->foff, 27 * (size_t)nbox);
    if (!st) st = dev_alloc(&e->boxL, (size_t)nbox);
    if (!st) st = dev_alloc(&e->mb, (size_t)nbox);
    for (int k = 0; k < 2 && !st; k++) {
        st = dev_alloc(&e->pos[k], (size_t)e->ntot);
        if (!st) st = dev_alloc(&e->vel[k], (size_t)e->ntot);
    }
    if (!st) st = dev_alloc(&e->ff, (size_t)e->ntot);
    if (!st) st = dev_alloc(&e->acc, 3 * (size_t)(e->ntot > 0 ? e->ntot : 1));
    if (!st && hipMemset(e->acc, 0, sizeof(double) * 3 * (size_t)(e->ntot > 0 ? e->ntot : 1)) != hipSuccess)
        st = UDX_EHIP;
    if (!st) st = dev_alloc(&e->ecount, (size_t)nbox);
    if (!st) st = dev_alloc(&e->eidx, (size_t)(e->ntot > 0 ? e->ntot : 1));
    if (!st) st = dev_alloc(&e->ef, (size_t)(e->ntot > 0 ? e->ntot : 1));
    if (!st) st = clist_alloc(&e->cl, Box{{e->ctot, 1, 1}}, e->ntot > 0 ? e->ntot : 1, SOLVENT_SUB);
    if (!st) st = dev_alloc(&e->fcnt, (size_t)e->ftot);
    if (!st) st = dev_alloc(&e->fss, (size_t)e->ftot + 1);
    if (!st) st = dev_alloc(&e->fscan, (size_t)scan_scratch_size(e->ftot));
    if (!st) st = dev_alloc(&e->seeds_cur, 27 * (size_t)nbox);
    if (!st) st = dev_alloc(&e->step, 1);
    if (!st) st = dev_alloc(&e->pairs, 1);
    up(e->poff, e->poff_h.data(), sizeof(int) * (nbox + 1));
    up(e->coff, e->coff_h.data(), sizeof(int) * (nbox + 1));
    up(e->fbase, e->fbase_h.data(), sizeof(int) * (nbox + 1));
    up(e->foff, e->foff_h.data(), sizeof(int) * 27 * nbox);
    up(e->boxL, L.data(), sizeof(int4) * nbox);
    up(e->mb, mb.data(), sizeof(MoveBox) * nbox);
    if (!st && hipMemset(e->pairs, 0, sizeof(unsigned long long)) != hipSuccess) st = UDX_EHIP;
    if (!st) {
        e->frag.resize(nbox);
        for (int b = 0; b < nbox; b++) {
            FragTables &t = e->frag[b];
            t = FragTables{};
            t.fss = e->fss + e->fbase_h[b];
            for (int f = 0; f < 27; f++) t.off[f] = e->foff_h[27 * b + f];
            t.size = e->foff_h[27 * b + 26];
        }
        st = ens_build_batches(e);
    }
    if (st) {
        if (st == UDX_EHIP) set_error("udx_ens_ini: HIP allocation or copy failed");
        udx_ens_fin(e);
        return st;
    }
    *out = e;
    return UDX_OK;
}

int udx_ens_set_mode(udx_ens *e, int mode) {
    UDX_CHECK(e && mode >= UDX_MODE_FAST && mode <= UDX_MODE_HALF, UDX_EINVAL, "udx_ens_set_mode: bad mode");
    if (mode != e->mode) ens_drop_graph(e);
    e->mode = mode;
    return UDX_OK;
}

int udx_ens_set_variant(udx_ens *e, int v) {
    UDX_CHECK(e && v >= 0 && v < DPD_NVARIANT, UDX_EINVAL, "udx_ens_set_variant: variant %d", v);
    if (v != e->variant) ens_drop_graph(e);
    e->variant = v;
    return UDX_OK;
}

int udx_ens_set_graph(udx_ens *e, int enable) {
    UDX_CHECK(e, UDX_EINVAL, "udx_ens_set_graph: null handle");
    e->use_graph = enable ? 1 : 0;
    return UDX_OK;
}

int udx_ens_set_pp(udx_ens *e, const float *pp) {
    UDX_CHECK(e && pp, UDX_EINVAL, "udx_ens_set_pp: null argument");
    // wait for work on the caller's stream that produced pp
    UDX_HIP(hipEventRecord(e->done, e->ctx->stream));
    UDX_HIP(hipStreamWaitEvent(e->s, e->done, 0));
    UDX_TRY(aos_to_f4(e->s, e->ntot, pp, e->pos[0], e->vel[0]));
    if (e->ntot > 0)
        hipLaunchKernelGGL(k_ens_wrap, dim3(ceil_div(e->ntot, 256)), dim3(256), 0, e->s, e->ntot, e->nbox, e->poff,
                           e->mb, e->pos[0]);
    UDX_LAUNCHED();
    e->cur = 0;
    e->counted = 0;
    UDX_HIP(hipEventRecord(e->done, e->s));
    UDX_HIP(hipStreamWaitEvent(e->ctx->stream, e->done, 0));
    return UDX_OK;
}

int udx_ens_get_pp(udx_ens *e, float *pp) {
    UDX_CHECK(e && pp, UDX_EINVAL, "udx_ens_get_pp: null argument");
    if (e->ntot > 0)
        hipLaunchKernelGGL(k_ens_unpack, dim3(ceil_div(e->ntot, 256)), dim3(256), 0, e->s, e->ntot, e->pos[e->cur],
                           e->vel[e->cur], pp);
    UDX_LAUNCHED();
    UDX_HIP(hipEventRecord(e->done, e->s));
    UDX_HIP(hipStreamWaitEvent(e->ctx->stream, e->done, 0));
    return UDX_OK;
}

int udx_ens_run(udx_ens *e, int nsteps, const float *seeds) {
    UDX_CHECK(e && nsteps >= 0, UDX_EINVAL, "udx_ens_run: bad arguments");
    if (nsteps == 0 || e->ntot == 0) return UDX_OK;
    UDX_HIP(hipEventRecord(e->done, e->ctx->stream));
    UDX_HIP(hipStreamWaitEvent(e->s, e->done, 0));
    UDX_TRY(ens_seed_table(e, nsteps, seeds));
    int k = 0;
    if (e->use_graph && nsteps >= 2) {
        if (!e->counted) {
            UDX_TRY(ens_count(e, e->cur));
            e->counted = 1;
        }
        if (!e->exec || e->graph_variant != e->variant * 2 + e->cur) UDX_TRY(ens_capture(e));
        for (; k + 2 <= nsteps; k += 2) UDX_HIP(hipGraphLaunch(e->exec, e->s));
    }
    for (; k < nsteps; k++) {
        int next = e->cur;
        UDX_TRY(ens_step(e, e->cur, &next));
        e->cur = next;
    }
    UDX_HIP(hipEventRecord(e->done, e->s));
    UDX_HIP(hipStreamWaitEvent(e->ctx->stream, e->done, 0));
    return UDX_OK;
}

int udx_ens_pairs(udx_ens *e, unsigned long long *npairs, int reset) {
    UDX_CHECK(e && npairs, UDX_EINVAL, "udx_ens_pairs: null argument");
    UDX_HIP(hipMemcpyAsync(npairs, e->pairs, sizeof *npairs, hipMemcpyDeviceToHost, e->s));
    if (reset) UDX_HIP(hipMemsetAsync(e->pairs, 0, sizeof(unsigned long long), e->s));
    UDX_HIP(hipStreamSynchronize(e->s));
    *npairs /= 2;
    return UDX_OK;
}

// time each tile variant of the force kernel on `steps` steps of the
// ensemble (the state advances), keep the fastest
int udx_ens_autotune(udx_ens *e, int steps, int *best, float *ms) {
    UDX_CHECK(e && steps > 0, UDX_EINVAL, "udx_ens_autotune: bad arguments");
    hipEvent_t t0, t1;
    UDX_HIP(hipEventCreate(&t0));
    UDX_HIP(hipEventCreate(&t1));
    const int keep = e->use_graph;
    e->use_graph = 0;
    float bestms = 1e30f;
    int bv = 0, st = UDX_OK;
    for (int v = 0; v < DPD_NVARIANT && !st; v++) {
        e->variant = v;
        st = udx_ens_run(e, 1, nullptr);   // warm the variant
        if (st) break;
        if (hipEventRecord(t0, e->s) != hipSuccess) st = UDX_EHIP;
        if (!st) st = udx_ens_run(e, steps, nullptr);
        float t = 0;
        if (!st && (hipEventRecord(t1, e->s) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                    hipEventElapsedTime(&t, t0, t1) != hipSuccess))
            st = UDX_EHIP;
        if (ms) ms[v] = t / steps;
        if (!st && t < bestms) {
            bestms = t;
            bv = v;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->use_graph = keep;
    ens_drop_graph(e);
    e->variant = bv;
    if (best) *best = bv;
    return st;
}

}  // extern "C"
