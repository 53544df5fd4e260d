// C ABI of libudx (declared in include/udx.h): handles, host-side KISS seed
// draws, and the per-step pipeline of the solvent-only solver.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "udx_internal.h"

namespace udx {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// KISS trunk (src/math/rnd/imp.cpp:21-34)
static uint32_t kiss_int(udx_rnd *r) {
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

// the per-step trunks: halo fragments 0..25 (src/fluforces/imp/ini.h:11-58,
// one rank: seed_base = 0), the bulk trunk (0,0,0,0) at index 26 and the
// wall trunk (42,42,42,42) of src/wall/imp/ini.h:11 at index 27
constexpr int NSEED = 28;
struct Trunks {
    udx_rnd t[NSEED];
    void ini() {
        for (int f = 0; f < 27; f++) flu_trunk_ini(f, &t[f]);
        udx_rnd_ini(42, 42, 42, 42, &t[27]);
    }
    void next(float out[NSEED]) {
        for (int f = 0; f < NSEED; f++) out[f] = udx_rnd_get(&t[f]);
    }
};

static Box make_box(int Lx, int Ly, int Lz) {
    Box b;
    b.L[0] = Lx;
    b.L[1] = Ly;
    b.L[2] = Lz;
    return b;
}

void flu_trunk_ini(int f, udx_rnd *r) {
    if (f == 26) {
        udx_rnd_ini(0, 0, 0, 0, r);
        return;
    }
    int d[3] = {(f + 2) % 3 - 1, (f / 3 + 2) % 3 - 1, (f / 9 + 2) % 3 - 1};
    int sum = d[0] + d[1] + d[2];
    int plus = (sum > 0) || (sum == 0 && d[0] > 0) || (d[0] == 0 && d[1] > 0) || (d[1] == 0 && d[2] > 0);
    int sign = 2 * plus - 1, v[3];
    for (int c = 0; c < 3; c++) v[c] = 1 + sign * d[c];
    int seed = v[0] + 3 * (v[1] + 3 * v[2]);
    udx_rnd_ini(390 + seed, seed + 615, 12309, 23094, r);
}

}  // namespace udx

using namespace udx;

// ------------------------------------------------------------------ handles
struct udx_clist {
    udx_ctx *ctx;
    ClistBufs c;
};

struct udx_fluforces {
    udx_ctx *ctx;
    Box box;
    int maxp, n, mode;
    float4 *pos, *vel;
    float *seeds_dev;
    FragTables frag;
    Trunks trunks;
    unsigned long long *pairs_dev;
    double *acc;        // FP64 force sums of the fast pass (3 per particle, kept zero between calls)
    int *ecount, *eidx; // edge list of the gather's halo pass
    float4 *ef;
};

struct udx_sdf {
    udx_ctx *ctx;
    SdfField f;
};

struct udx_wall {
    udx_ctx *ctx;
    const udx_sdf *sdf;
    WallSet w;
    udx_rnd trunk;              // (42,42,42,42), src/wall/imp/ini.h:11
    int mode, cap;
    float *seed_dev;
    float4 *pos, *vel;          // AoS -> float4 staging for udx_wall_force
    unsigned long long *pairs_dev;
    udx_wall_vel wvel;          // resting unless udx_wall_set_vel (src/sim/walls/imp/main.h:114-116)
    float t;                    // time of the "shear sin" rate
};

static const SdfField *wall_sdf(const udx_wall *w) { return &w->sdf->f; }
static const WallSet *wall_set(const udx_wall *w) { return &w->w; }
static unsigned long long *wall_pairs_dev(udx_wall *w) { return w->pairs_dev; }

struct udx_flu {
    udx_ctx *ctx;
    udx_flu_conf conf;
    Box box;
    int n;
    float4 *pos[2], *vel[2], *ff;
    double *acc;        // FP64 force sums of the fast pass (3 per particle, kept zero between steps)
    int *ecount, *eidx; // edge list of the gather's halo pass
    float4 *ef;
    int cur;
    ClistBufs cl;
    // the last step's move counted pos[cur] into the cell list's x slabs, so
    // the next sort starts from its scan
    int counted = 0;
    FragTables frag;
    Trunks trunks;
    float *seeds_dev;   // seed_cap x NSEED
    int seed_cap;
    float *seeds_host;  // pinned staging
    unsigned long long *pairs_dev;   // [0] ordered solvent pairs, [1] solvent-wall pairs
    struct udx_wall *wall;
    float t;            // step clock, float like the reference's (src/scheme/time_line/imp/type.h:1-5)
    int timing;
    int fused;          // the last step moved the particles in the gather (walls before it)
    // timing: five events per step (clist, DPD, walls and objects, move,
    // end), recorded on the library's stream and read after the run, so the
    // timed steps run without host synchronisation
    std::vector<hipEvent_t> ev;
    hipEvent_t *evk = nullptr;   // the current step's five, null when not timing
    float tms[5];
    // graph replay of step pairs (udx_flu_set_graph): the even/odd step pair
    // captured on an own stream (capture needs one) and replayed; a device
    // step index selects each step's seeds from the run's table
    int use_graph = 0;
    hipStream_t gs = nullptr;       // the stream steps launch on while capturing and replaying
    hipEvent_t gdone = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    struct GraphKey {
        const float *table;
        int n, cur, mode;
        const void *wall;
    } gkey{};
    int *step_dev = nullptr;        // index of the current step in the seed table
    float *seeds_cur = nullptr;     // that step's NSEED seeds
    hipStream_t step_stream = nullptr;   // non-null: flu_step launches there
    // membranes in the step (udx_flu_set_rbc)
    struct Membranes {
        udx_rbc *r = nullptr;
        udx_fsi *fsi = nullptr;
        udx_rbc_params rp;
        udx_pair_dpd fp;
        float mass = 0;
        int substeps = 0, nc = 0, nv = 0, cap = 0;
        int ccap = 0;                                      // cells the per-cell arrays (cfid, cnew) hold
        float *pp = nullptr;                               // the caller's vertices (AoS), advanced in place
        float *fo = nullptr, *ffast = nullptr, *tmp = nullptr;   // FSI force, sub-step force, redistribution staging
        int *cfid = nullptr, *cnew = nullptr;
        // bounce-back (udx_flu_set_rbc_bounce): the vertices and the solvent
        // positions at the start of the step
        udx_mbb *bb = nullptr;
        float *start_pp = nullptr;
        float4 *pos0 = nullptr;
        // contact between membranes (udx_flu_set_rbc_contact)
        udx_cnt *cnt = nullptr;
        udx_pair_dpd_lj cnt_par{};
        // walls: the vertices as float4 streams for the wall DPD, their
        // count on the device, the wall repulsion coefficient (0: none)
        float4 *vpos = nullptr, *vvel = nullptr;
        int *nvtx_dev = nullptr;
        float wall_lambda = 0.f;
    } mb;
};

// FSI seeds of step k: after the NSEED x seed_cap solvent table; then the
// seeds of the membranes' wall DPD, one per step
static const float *flu_fsi_seeds(const udx_flu *f, int k) { return f->seeds_dev + NSEED * f->seed_cap + 2 * k; }
static const float *flu_mwall_seed(const udx_flu *f, int k) { return f->seeds_dev + (NSEED + 2) * f->seed_cap + k; }

// keep_bounce: a new membrane capacity keeps the bounce-back set-up
static void flu_free_membranes(udx_flu *f, bool keep_bounce = false) {
    dev_free(f->mb.fo);
    dev_free(f->mb.ffast);
    dev_free(f->mb.tmp);
    dev_free(f->mb.cfid);
    dev_free(f->mb.cnew);
    dev_free(f->mb.start_pp);
    dev_free(f->mb.vpos);
    dev_free(f->mb.vvel);
    dev_free(f->mb.nvtx_dev);
    const float wall_lambda = f->mb.wall_lambda;
    udx_mbb *bb = f->mb.bb;
    float4 *pos0 = f->mb.pos0;
    udx_cnt *cnt = f->mb.cnt;
    const udx_pair_dpd_lj cnt_par = f->mb.cnt_par;
    f->mb = udx_flu::Membranes{};
    if (keep_bounce) {
        f->mb.bb = bb;
        f->mb.pos0 = pos0;
        f->mb.cnt = cnt;
        f->mb.cnt_par = cnt_par;
        f->mb.wall_lambda = wall_lambda;
    } else {
        dev_free(pos0);
    }
}

extern "C" {

const char *udx_last_error(void) { return g_err; }

void udx_rnd_ini(int x, int y, int z, int c, udx_rnd *r) {
    r->x = (uint32_t)x;
    r->y = (uint32_t)y;
    r->z = (uint32_t)z;
    r->c = (uint32_t)c;
}

float udx_rnd_get(udx_rnd *r) {
    uint32_t i = kiss_int(r);
    return (float)i / (float)UINT32_MAX;
}

void udx_pair_compute_dpd_sigma(float kBT, float dt, udx_pair_dpd *p) { p->s = sqrtf(2 * kBT * p->g / dt); }

int udx_ctx_ini(void *stream, udx_ctx **ctx) {
    UDX_CHECK(ctx, UDX_EINVAL, "udx_ctx_ini: null output");
    int dev = 0;
    UDX_HIP(hipGetDevice(&dev));
    udx_ctx *c = new udx_ctx;
    c->stream = (hipStream_t)stream;
    c->device = dev;
    *ctx = c;
    return UDX_OK;
}

int udx_ctx_fin(udx_ctx *ctx) {
    delete ctx;
    return UDX_OK;
}

int udx_ctx_sync(udx_ctx *ctx) {
    UDX_HIP(hipStreamSynchronize(ctx->stream));
    return UDX_OK;
}

int udx_rnd_branch(udx_ctx *ctx, float seed, int n, const int *u, const int *v, int uu, float *out) {
    UDX_CHECK(ctx && n >= 0, UDX_EINVAL, "udx_rnd_branch: bad arguments");
    return rnd_branch(ctx->stream, seed, n, u, v, uu, out);
}

// ------------------------------------------------------------------ clist
int udx_clist_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int maxp, udx_clist **out) {
    UDX_CHECK(ctx && out, UDX_EINVAL, "udx_clist_ini: null argument");
    UDX_CHECK(Lx > 0 && Ly > 0 && Lz > 0 && maxp >= 0, UDX_EINVAL, "udx_clist_ini: bad box %d %d %d / maxp %d",
              Lx, Ly, Lz, maxp);
    udx_clist *c = new udx_clist;
    c->ctx = ctx;
    int s = clist_alloc(&c->c, make_box(Lx, Ly, Lz), maxp, SOLVENT_SUB);
    if (s) {
        clist_free(&c->c);
        delete c;
        return s;
    }
    *out = c;
    return UDX_OK;
}

int udx_clist_fin(udx_clist *c) {
    if (c) clist_free(&c->c);
    delete c;
    return UDX_OK;
}

int udx_clist_build(udx_clist *c, int n, const float *pp, float *ppout, int *nout) {
    UDX_CHECK(c, UDX_EINVAL, "udx_clist_build: null clist");
    hipStream_t s = c->ctx->stream;
    UDX_TRY(clist_sort_aos(s, &c->c, n, pp));
    if (ppout) UDX_TRY(gather_aos(s, &c->c, n, pp, ppout));
    if (nout) {
        UDX_HIP(hipMemcpyAsync(nout, c->c.starts + c->c.box.ncells(), sizeof(int), hipMemcpyDeviceToHost, s));
        UDX_HIP(hipStreamSynchronize(s));
    }
    return UDX_OK;
}

const int *udx_clist_starts(const udx_clist *c) { return c->c.starts; }
const int *udx_clist_counts(const udx_clist *c) { return c->c.counts; }
const int *udx_clist_order(const udx_clist *c) { return c->c.order; }

// ------------------------------------------------------------------ scan
// scan_ini / scan_apply / scan_fin (src/algo/scan/imp.h:4-6; the CUDA
// version src/algo/scan/cuda/imp.h:9-31 checks size against the workspace)
struct udx_scan {
    udx_ctx *ctx;
    int size;
    int *tmp;
};

int udx_scan_ini(udx_ctx *ctx, int size, udx_scan **w) {
    UDX_CHECK(ctx && w && size >= 0, UDX_EINVAL, "udx_scan_ini: bad arguments");
    udx_scan *q = new udx_scan{ctx, size, nullptr};
    const int st = dev_alloc(&q->tmp, (size_t)scan_scratch_size(size));
    if (st) {
        delete q;
        return st;
    }
    *w = q;
    return UDX_OK;
}

int udx_scan_apply(udx_scan *w, const int *input, int size, int *output) {
    UDX_CHECK(w && size >= 0, UDX_EINVAL, "udx_scan_apply: bad arguments");
    UDX_CHECK(size <= w->size, UDX_ECAP, "udx_scan_apply: size = %d > w->size = %d", size, w->size);
    if (size == 0) return UDX_OK;
    UDX_CHECK(input && output, UDX_EINVAL, "udx_scan_apply: null array");
    return scan_exclusive_n(w->ctx->stream, input, size, output, w->tmp);
}

int udx_scan_fin(udx_scan *w) {
    if (!w) return UDX_OK;
    dev_free(w->tmp);
    delete w;
    return UDX_OK;
}

// ------------------------------------------------------------------ fluforces
int udx_fluforces_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int maxp, udx_fluforces **out) {
    UDX_CHECK(ctx && out, UDX_EINVAL, "udx_fluforces_ini: null argument");
    UDX_CHECK(Lx > 0 && Ly > 0 && Lz > 0 && maxp >= 0, UDX_EINVAL, "udx_fluforces_ini: bad box or capacity");
    udx_fluforces *f = new udx_fluforces();
    f->ctx = ctx;
    f->box = make_box(Lx, Ly, Lz);
    f->maxp = maxp;
    f->trunks.ini();
    int s = UDX_OK;
    if (!s) s = dev_alloc(&f->pos, maxp);
    if (!s) s = dev_alloc(&f->vel, maxp);
    if (!s) s = dev_alloc(&f->seeds_dev, 27);
    if (!s) s = dev_alloc(&f->pairs_dev, 1);
    if (!s) s = frag_alloc(&f->frag, f->box);
    if (!s) s = dev_alloc(&f->acc, 3 * (size_t)(maxp > 0 ? maxp : 1));
    if (!s) s = dev_alloc(&f->ecount, 1);
    if (!s) s = dev_alloc(&f->eidx, (size_t)(maxp > 0 ? maxp : 1));
    if (!s) s = dev_alloc(&f->ef, (size_t)(maxp > 0 ? maxp : 1));
    if (!s && (hipMemset(f->pairs_dev, 0, sizeof(unsigned long long)) != hipSuccess ||
               hipMemset(f->acc, 0, sizeof(double) * 3 * (size_t)(maxp > 0 ? maxp : 1)) != hipSuccess)) {
        set_error("udx_fluforces_ini: memset failed");
        s = UDX_EHIP;
    }
    if (s) {
        udx_fluforces_fin(f);
        return s;
    }
    *out = f;
    return UDX_OK;
}

int udx_fluforces_fin(udx_fluforces *f) {
    if (!f) return UDX_OK;
    dev_free(f->pos);
    dev_free(f->vel);
    dev_free(f->seeds_dev);
    dev_free(f->pairs_dev);
    dev_free(f->acc);
    dev_free(f->ecount);
    dev_free(f->eidx);
    dev_free(f->ef);
    frag_free(&f->frag);
    delete f;
    return UDX_OK;
}

int udx_fluforces_set_mode(udx_fluforces *f, int mode) {
    UDX_CHECK(f, UDX_EINVAL, "udx_fluforces_set_mode: null handle");
    UDX_CHECK(mode >= UDX_MODE_FAST && mode <= UDX_MODE_HALF, UDX_EINVAL, "udx_fluforces_set_mode: mode %d",
              mode);
    f->mode = mode;
    return UDX_OK;
}

int udx_fluforces_prepare(udx_fluforces *f, int n, const float *pp) {
    UDX_CHECK(f, UDX_EINVAL, "udx_fluforces_prepare: null handle");
    UDX_CHECK(n >= 0 && n <= f->maxp, UDX_ECAP, "udx_fluforces_prepare: %d particles exceed capacity %d", n,
              f->maxp);
    f->n = n;
    return aos_to_f4(f->ctx->stream, n, pp, f->pos, f->vel);
}

static int fluforces_apply(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic,
                           const float seeds[27], float *ff, float *ss);

int udx_fluforces_apply_seeds(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic,
                              const float seeds[27], float *ff) {
    UDX_CHECK(f && p && starts && seeds, UDX_EINVAL, "udx_fluforces_apply_seeds: null argument");
    return fluforces_apply(f, p, starts, periodic, seeds, ff, nullptr);
}

int udx_fluforces_apply_stress(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic,
                               const float seeds[27], float *ff, float *ss) {
    UDX_CHECK(f && p && starts && ff && ss, UDX_EINVAL, "udx_fluforces_apply_stress: null argument");
    float own[NSEED];
    if (!seeds) {
        f->trunks.next(own);
        seeds = own;
    }
    return fluforces_apply(f, p, starts, periodic, seeds, ff, ss);
}

static int fluforces_apply(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic,
                           const float seeds[27], float *ff, float *ss) {
    hipStream_t s = f->ctx->stream;
    if (f->n == 0) return UDX_OK;
    UDX_HIP(hipMemcpyAsync(f->seeds_dev, seeds, 27 * sizeof(float), hipMemcpyHostToDevice, s));
    UDX_TRY(frag_build(s, &f->frag, f->box, starts, f->ecount));
    DpdLaunch a;
    a.pos = f->pos;
    a.vel = f->vel;
    a.starts = starts;
    a.frag = &f->frag;
    a.box = f->box;
    a.nmax = f->n;
    a.periodic = periodic;
    a.par = *p;
    a.seeds = f->seeds_dev;
    a.ff = ff;
    a.ff_stride = 3;
    a.accumulate = 1;
    a.pairs = f->pairs_dev;
    a.ss = ss;
    a.acc = f->acc;
    a.ecount = f->ecount;
    a.eidx = f->eidx;
    a.ef = f->ef;
    UDX_TRY(dpd_forces_mode(s, a, f->mode));
    // the seeds live on the host stack: wait for the copy before returning
    UDX_HIP(hipStreamSynchronize(s));
    return UDX_OK;
}

int udx_fluforces_apply(udx_fluforces *f, const udx_pair_dpd *p, const int *starts, int periodic, float *ff) {
    UDX_CHECK(f, UDX_EINVAL, "udx_fluforces_apply: null handle");
    float seeds[NSEED];
    f->trunks.next(seeds);
    return udx_fluforces_apply_seeds(f, p, starts, periodic, seeds, ff);
}

int udx_dpd_set_split_edges(double n) { return dpd_set_split_edges(n); }

int udx_fluforces_pairs(udx_fluforces *f, unsigned long long *npairs, int reset) {
    UDX_CHECK(f && npairs, UDX_EINVAL, "udx_fluforces_pairs: null argument");
    unsigned long long v = 0;
    UDX_HIP(hipMemcpyAsync(&v, f->pairs_dev, sizeof v, hipMemcpyDeviceToHost, f->ctx->stream));
    UDX_HIP(hipStreamSynchronize(f->ctx->stream));
    *npairs = v / 2;
    if (reset) UDX_HIP(hipMemsetAsync(f->pairs_dev, 0, sizeof v, f->ctx->stream));
    return UDX_OK;
}

// ------------------------------------------------------------------ scheme
int udx_scheme_move_apply(udx_ctx *ctx, float dt, float mass, int n, const float *ff, float *pp) {
    UDX_CHECK(ctx, UDX_EINVAL, "udx_scheme_move_apply: null context");
    return move_aos(ctx->stream, n, dt, mass, ff, pp);
}

int udx_bforce_apply_cste(udx_ctx *ctx, float mass, const float accel[3], int n, float *ff) {
    UDX_CHECK(ctx && accel, UDX_EINVAL, "udx_bforce_apply_cste: null argument");
    return bforce_aos(ctx->stream, n, mass, accel, ff);
}

int udx_wrap_pp(udx_ctx *ctx, int Lx, int Ly, int Lz, int n, float *pp) {
    UDX_CHECK(ctx, UDX_EINVAL, "udx_wrap_pp: null context");
    return wrap_aos(ctx->stream, make_box(Lx, Ly, Lz), n, pp);
}

int udx_diag_part(udx_ctx *ctx, int n, const float *pp, double out[5]) {
    UDX_CHECK(ctx && out, UDX_EINVAL, "udx_diag_part: null argument");
    double *part = nullptr;
    UDX_TRY(dev_alloc(&part, 5 * 1024 + 5));
    int s = diag_aos(ctx->stream, n, pp, part, out);
    dev_free(part);
    return s;
}

// ------------------------------------------------------------------ flu
int udx_flu_fin(udx_flu *f);
static void flu_drop_graph(udx_flu *f);

int udx_flu_ini(udx_ctx *ctx, const udx_flu_conf *conf, udx_flu **out) {
    UDX_CHECK(ctx && conf && out, UDX_EINVAL, "udx_flu_ini: null argument");
    const int *L = conf->L;
    UDX_CHECK(L[0] >= 3 && L[1] >= 3 && L[2] >= 3, UDX_EINVAL,
              "udx_flu_ini: a periodic box needs at least 3 cells per side (got %d %d %d)", L[0], L[1], L[2]);
    UDX_CHECK(conf->maxp > 0 && conf->mass > 0, UDX_EINVAL, "udx_flu_ini: bad capacity or mass");
    UDX_CHECK(conf->mode >= UDX_MODE_FAST && conf->mode <= UDX_MODE_HALF, UDX_EINVAL, "udx_flu_ini: mode %d",
              conf->mode);
    udx_flu *f = new udx_flu();
    f->ctx = ctx;
    f->conf = *conf;
    f->box = make_box(L[0], L[1], L[2]);
    f->trunks.ini();
    f->seed_cap = 0;
    int s = UDX_OK;
    for (int b = 0; b < 2 && !s; b++) {
        s = dev_alloc(&f->pos[b], conf->maxp);
        if (!s) s = dev_alloc(&f->vel[b], conf->maxp);
    }
    if (!s) s = dev_alloc(&f->ff, conf->maxp);
    if (!s) s = dev_alloc(&f->acc, 3 * (size_t)conf->maxp);
    if (!s) s = dev_alloc(&f->ecount, 1);
    if (!s) s = dev_alloc(&f->eidx, (size_t)conf->maxp);
    if (!s) s = dev_alloc(&f->ef, (size_t)conf->maxp);
    if (!s && hipMemset(f->acc, 0, sizeof(double) * 3 * (size_t)conf->maxp) != hipSuccess) {
        set_error("udx_flu_ini: memset failed");
        s = UDX_EHIP;
    }
    if (!s) s = clist_alloc(&f->cl, f->box, conf->maxp, SOLVENT_SUB);
    if (!s) s = frag_alloc(&f->frag, f->box);
    if (!s) s = dev_alloc(&f->pairs_dev, 2);
    if (!s && hipMemset(f->pairs_dev, 0, 2 * sizeof(unsigned long long)) != hipSuccess) {
        set_error("udx_flu_ini: memset failed");
        s = UDX_EHIP;
    }
    if (s) {
        udx_flu_fin(f);
        return s;
    }
    *out = f;
    return UDX_OK;
}

int udx_flu_fin(udx_flu *f) {
    if (!f) return UDX_OK;
    for (int b = 0; b < 2; b++) {
        dev_free(f->pos[b]);
        dev_free(f->vel[b]);
    }
    dev_free(f->ff);
    dev_free(f->acc);
    dev_free(f->ecount);
    dev_free(f->eidx);
    dev_free(f->ef);
    clist_free(&f->cl);
    frag_free(&f->frag);
    dev_free(f->seeds_dev);
    dev_free(f->pairs_dev);
    if (f->seeds_host) (void)hipHostFree(f->seeds_host);
    for (hipEvent_t e : f->ev) (void)hipEventDestroy(e);
    if (f->gs) (void)hipStreamSynchronize(f->gs);
    flu_drop_graph(f);
    if (f->gs) (void)hipStreamDestroy(f->gs);
    if (f->gdone) (void)hipEventDestroy(f->gdone);
    dev_free(f->step_dev);
    dev_free(f->seeds_cur);
    flu_free_membranes(f);
    delete f;
    return UDX_OK;
}

int udx_flu_set_pp(udx_flu *f, int n, const float *pp) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_pp: null handle");
    UDX_CHECK(n >= 0 && n <= f->conf.maxp, UDX_ECAP, "udx_flu_set_pp: %d particles exceed capacity %d", n,
              f->conf.maxp);
    if (n != f->n) flu_drop_graph(f);
    f->n = n;
    f->cur = 0;
    f->counted = 0;
    return aos_to_f4(f->ctx->stream, n, pp, f->pos[0], f->vel[0]);
}

int udx_flu_get_pp(udx_flu *f, float *pp, int *n) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_get_pp: null handle");
    if (n) *n = f->n;
    return f4_to_aos(f->ctx->stream, f->n, nullptr, f->pos[f->cur], f->vel[f->cur], pp);
}

int udx_flu_get_ff(udx_flu *f, float *ff) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_get_ff: null handle");
    return f4_to_force_aos(f->ctx->stream, f->n, nullptr, f->ff, ff);
}

static int flu_seed_table(udx_flu *f, int nsteps, const float *seeds_host) {
    hipStream_t s = f->ctx->stream;
    if (nsteps > f->seed_cap) {
        UDX_HIP(hipStreamSynchronize(s));
        dev_free(f->seeds_dev);
        if (f->seeds_host) (void)hipHostFree(f->seeds_host);
        f->seeds_host = nullptr;
        f->seed_cap = 0;
        UDX_TRY(dev_alloc(&f->seeds_dev, (NSEED + 3) * (size_t)nsteps));
        UDX_HIP(hipHostMalloc((void **)&f->seeds_host, (NSEED + 3) * sizeof(float) * (size_t)nsteps));
        f->seed_cap = nsteps;
    } else {
        // the staging buffer may still feed an earlier copy
        UDX_HIP(hipStreamSynchronize(s));
    }
    // the wall trunk draws for the solvent, then for the membranes
    // (wall_interact, src/sim/walls/imp/main.h:47-61): with a seed table the
    // solvent's comes from it and the membranes' from the wall trunk
    float *mw = f->seeds_host + (NSEED + 2) * f->seed_cap;
    // the reference skips wall_force and its draw when there are no vertices
    // (src/sim/walls/imp/main.h:54-58, src/wall/force/imp/main.h:12)
    const bool mwall = f->wall && f->mb.r && (long)f->mb.nc * f->mb.nv > 0;
    for (int k = 0; k < nsteps; k++) {
        if (seeds_host)
            memcpy(f->seeds_host + NSEED * k, seeds_host + NSEED * k, NSEED * sizeof(float));
        else
            f->trunks.next(f->seeds_host + NSEED * k);
        mw[k] = mwall ? udx_rnd_get(&f->trunks.t[27]) : 0.f;
    }
    // the FSI trunk draws twice per step, bulk then halo (src/fsi/imp/bulk.h:8, src/fsi/imp/halo.h:7)
    float *fsi = f->seeds_host + NSEED * f->seed_cap;
    for (int k = 0; k < nsteps; k++) {
        fsi[2 * k] = f->mb.r ? fsi_draw(f->mb.fsi) : 0.f;
        fsi[2 * k + 1] = f->mb.r ? fsi_draw(f->mb.fsi) : 0.f;
    }
    UDX_HIP(hipMemcpyAsync(f->seeds_dev, f->seeds_host, NSEED * sizeof(float) * (size_t)nsteps,
                           hipMemcpyHostToDevice, s));
    UDX_HIP(hipMemcpyAsync(f->seeds_dev + NSEED * f->seed_cap, fsi, 2 * sizeof(float) * (size_t)nsteps,
                           hipMemcpyHostToDevice, s));
    UDX_HIP(hipMemcpyAsync(f->seeds_dev + (NSEED + 2) * f->seed_cap, mw, sizeof(float) * (size_t)nsteps,
                           hipMemcpyHostToDevice, s));
    return UDX_OK;
}

// one step of src/sim/imp/run.h:8-68: solvent cell list + reorder, membrane
// redistribution, DPD forces, walls, FSI, body force + solvent kick-drift +
// bounce + periodic wrap, membrane sub-steps (FSI + internal forces, kick-drift)
static int flu_step(udx_flu *f, const float *seeds_dev, const float *fsi_seeds, const float *mwall_seed) {
    hipStream_t s = f->step_stream ? f->step_stream : f->ctx->stream;
    const int nc = f->box.ncells();
    const int a = f->cur, b = 1 - f->cur;
    udx_flu::Membranes &m = f->mb;
    const int nvtx = m.nc * m.nv;
    if (f->evk) UDX_HIP(hipEventRecord(f->evk[0], s));
    if (f->counted) UDX_TRY(clist_finish(s, &f->cl, f->n));
    else UDX_TRY(clist_sort(s, &f->cl, f->n, f->pos[a]));
    UDX_TRY(gather_f4(s, &f->cl, f->n, f->pos[a], f->vel[a], f->pos[b], f->vel[b]));
    UDX_TRY(frag_build(s, &f->frag, f->box, f->cl.starts, f->ecount));
    if (m.r) UDX_TRY(rbc_redistribute(s, f->box, m.nc, m.nv, m.pp, m.tmp, m.cfid, m.cnew));
    const bool bounce = m.r && m.bb;
    // objects_save_mesh (src/sim/imp/run.h:30-31): the vertices at the start of the step
    if (bounce && nvtx > 0)
        UDX_HIP(hipMemcpyAsync(m.start_pp, m.pp, 6 * sizeof(float) * (size_t)nvtx, hipMemcpyDeviceToDevice, s));
    if (f->evk) UDX_HIP(hipEventRecord(f->evk[1], s));
    DpdLaunch d;
    d.pos = f->pos[b];
    d.vel = f->vel[b];
    d.starts = f->cl.starts;
    d.bstarts = f->cl.bstarts;
    d.sub = f->cl.sub;
    d.frag = &f->frag;
    d.box = f->box;
    d.nmax = f->n;
    d.periodic = 1;
    d.par = f->conf.dpd;
    d.seeds = seeds_dev;
    d.ff = (float *)f->ff;
    d.ff_stride = 4;
    d.accumulate = 0;
    d.pairs = f->pairs_dev;
    d.acc = f->acc;
    d.ecount = f->ecount;
    d.eidx = f->eidx;
    d.ef = f->ef;
    const SdfField *sdf = nullptr;
    WallVel wv{};
    if (f->wall) {
        // wall_update_vel(t - tstart) at the start of the step (src/sim/imp/run.h:17-18)
        f->wall->t = f->t;
        wv = wall_vel_view(f->wall->wvel, f->t);
        sdf = wall_sdf(f->wall);
    }
    // Without objects the gather also evaluates the wall forces and moves the
    // particles (into the other pair of streams, which the next step's reorder
    // reads): the forces and the move are those of the passes one after the
    // other.
    f->fused = !m.r && dpd_move_fused(d, f->conf.mode);
    if (f->fused) {
        if (f->evk) UDX_HIP(hipEventRecord(f->evk[2], s));
        FusedMove mv{};
        mv.opos = f->pos[a];
        mv.ovel = f->vel[a];
        mv.dt = f->conf.dt;
        mv.mass = f->conf.mass;
        for (int c = 0; c < 3; c++) mv.accel[c] = f->conf.accel[c];
        mv.walls = sdf != nullptr;
        mv.sdf = sdf ? sdf->view : SdfView{};
        mv.wv = wv;
        if (sdf) {
            const WallSet *w = wall_set(f->wall);
            mv.wpos = w->pos;
            mv.wstarts = w->cl.starts;
            for (int c = 0; c < 3; c++) mv.W[c] = w->cl.box.L[c];
            mv.nwall = w->n;
            mv.wseed = seeds_dev + 27;
            mv.wpairs = f->pairs_dev + 1;
        }
        mv.next = &f->cl;
        d.move = &mv;
        UDX_TRY(dpd_forces_mode(s, d, f->conf.mode));
        f->counted = 1;
        if (f->evk) {
            UDX_HIP(hipEventRecord(f->evk[3], s));
            UDX_HIP(hipEventRecord(f->evk[4], s));
        }
        f->t += f->conf.dt;   // time_line_advance (src/scheme/time_line/imp/main.h:10-21)
        return UDX_OK;
    }
    UDX_TRY(dpd_forces_mode(s, d, f->conf.mode));
    if (f->evk) UDX_HIP(hipEventRecord(f->evk[2], s));
    if (f->wall)
        UDX_TRY(wall_forces(s, sdf, wall_set(f->wall), f->n, f->cl.starts + nc, f->pos[b], f->vel[b],
                            f->conf.dpd, seeds_dev + 27, (float *)f->ff, 4, f->pairs_dev + 1,
                            f->conf.mode == UDX_MODE_EXACT, wv, 0, f->cl.starts));
    if (m.r && nvtx > 0) {
        UDX_HIP(hipMemsetAsync(m.fo, 0, 3 * sizeof(float) * (size_t)nvtx, s));
        if (f->wall) {
            // the vertices' wall DPD with the FSI pair, then the wall
            // repulsion (forces_wall, src/sim/imp/force.h:37-59)
            UDX_TRY(aos_to_f4(s, nvtx, m.pp, m.vpos, m.vvel));
            UDX_TRY(wall_forces(s, sdf, wall_set(f->wall), nvtx, m.nvtx_dev, m.vpos, m.vvel, m.fp, mwall_seed, m.fo,
                                3, wall_pairs_dev(f->wall), f->conf.mode == UDX_MODE_EXACT, wv));
            if (m.wall_lambda > 0) UDX_TRY(wall_repulse_aos(s, sdf, m.wall_lambda, nvtx, m.pp, m.fo));
        }
        // FSI: +f on the vertices, -f on the solvent (src/sim/objinter/imp/main.h:42-91)
        UDX_TRY(fsi_forces(s, fsi_bufs(m.fsi), f->n, f->pos[b], f->vel[b], f->cl.starts, nvtx, m.pp, m.fp, fsi_seeds,
                           m.fo, (float *)f->ff, 4, fsi_pairs_dev(m.fsi), f->conf.mode == UDX_MODE_EXACT));
        // contact (forces_cnt, cnt_halo: src/sim/objinter/imp/main.h:22-26,79)
        if (m.cnt) {
            const float sb = cnt_draw(m.cnt), sh = cnt_draw(m.cnt);
            UDX_TRY(cnt_forces(s, m.cnt, m.cnt_par, nvtx, m.pp, sb, sh, m.fo));
        }
    } else if (m.r && m.cnt) {
        // the object interactions run for every object type, also one with
        // no cells, and draw their seeds (src/sim/objinter/imp/main.h:42-51)
        (void)cnt_draw(m.cnt);
        (void)cnt_draw(m.cnt);
    }
    if (f->evk) UDX_HIP(hipEventRecord(f->evk[3], s));
    // store_solvent (src/sim/imp/update.h:11-16): positions before the move
    if (bounce && f->n > 0)
        UDX_HIP(hipMemcpyAsync(m.pos0, f->pos[b], sizeof(float4) * (size_t)f->n, hipMemcpyDeviceToDevice, s));
    // without the membrane bounce-back (which moves particles after the
    // move and wraps them itself) the move also counts the next step's slabs
    UDX_TRY(move_f4_wall(s, f->n, f->cl.starts + nc, f->box, f->conf.dt, f->conf.mass, f->conf.accel, f->ff,
                         f->pos[b], f->vel[b], sdf, wv, bounce ? 0 : 1, bounce ? nullptr : &f->cl));
    f->counted = bounce ? 0 : 1;
    if (m.r && nvtx > 0) {
        // objects_advance (src/sim/objects/imp/main.h:40-61): per sub-step the
        // fast force starts as the FSI force, the membrane force is added and
        // the vertices kick-drift with their own mass
        const float sub_dt = f->conf.dt / (float)m.substeps;
        for (int k = 0; k < m.substeps; k++) {
            UDX_HIP(hipMemcpyAsync(m.ffast, m.fo, 3 * sizeof(float) * (size_t)nvtx, hipMemcpyDeviceToDevice, s));
            UDX_TRY(udx_rbc_force_apply(m.r, &m.rp, sub_dt, m.nc, m.pp, m.ffast));
            UDX_TRY(move_aos(s, nvtx, sub_dt, m.mass, m.ffast, m.pp));
        }
    }
    // the vertices' wall bounce-back (bounce_wall, src/sim/imp/update.h:41-53)
    if (f->wall && m.r && nvtx > 0) UDX_TRY(wall_bounce_aos(s, sdf, f->conf.dt, nvtx, m.pp, wv));
    // objects_bounce (src/sim/objects/imp/bounce.h:236-248); also the wrap
    // the move left out
    if (bounce)
        UDX_TRY(mbb_step(s, m.bb, f->conf.dt, f->conf.mass, m.mass, f->box, m.nc, m.start_pp, m.pp, f->cl.starts,
                         f->n, f->cl.starts + nc, f->pos[b], f->vel[b], m.pos0));
    if (f->evk) UDX_HIP(hipEventRecord(f->evk[4], s));
    f->cur = b;
    f->t += f->conf.dt;   // time_line_advance (src/scheme/time_line/imp/main.h:10-21)
    return UDX_OK;
}

namespace {
// the seeds of step *step from the run's table, then the counter advanced to
// the next step (one wave: every lane reads the counter before lane 0 writes
// it; nothing else in the step reads it)
__global__ void k_flu_seeds(const float *table, int *step, float *cur) {
    const int k = threadIdx.x;
    const int st = *step;
    if (k < NSEED) cur[k] = table[(long)st * NSEED + k];
    __syncthreads();
    if (k == 0) *step = st + 1;
}
}  // namespace

static void flu_drop_graph(udx_flu *f) {
    if (f->gexec) (void)hipGraphExecDestroy(f->gexec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    f->gexec = nullptr;
    f->graph = nullptr;
}

// steps that a captured graph can replay: everything the step does is on
// the device and the same from step to step (no membranes, whose seeds and
// sub-steps are host-driven; no time-dependent wall velocity; no timing)
static bool flu_graph_ok(const udx_flu *f) {
    return f->use_graph && !f->mb.r && !f->timing && !(f->wall && f->wall->wvel.type == UDX_WVEL_SHEAR_SIN);
}

// capture the step pair starting from the current parity on f->gs
static int flu_capture(udx_flu *f) {
    flu_drop_graph(f);
    const float t0 = f->t;
    const int cur0 = f->cur;
    const int counted0 = f->counted;
    f->counted = 1;   // replays start counted (flu_run_graph counts first when not)
    f->step_stream = f->gs;
    UDX_HIP(hipStreamBeginCapture(f->gs, hipStreamCaptureModeThreadLocal));
    int st = UDX_OK;
    for (int k = 0; k < 2 && !st; k++) {
        hipLaunchKernelGGL(k_flu_seeds, dim3(1), dim3(64), 0, f->gs, f->seeds_dev, f->step_dev, f->seeds_cur);
        st = hipGetLastError() == hipSuccess ? UDX_OK : UDX_EHIP;
        if (!st) st = flu_step(f, f->seeds_cur, nullptr, nullptr);
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(f->gs, &g);
    f->step_stream = nullptr;
    f->t = t0;
    f->cur = cur0;
    f->counted = counted0;
    if (st && g) (void)hipGraphDestroy(g);
    UDX_TRY(st);
    UDX_CHECK(ce == hipSuccess && g, UDX_EHIP, "udx_flu_run: graph capture failed: %s", hipGetErrorString(ce));
    f->graph = g;
    UDX_HIP(hipGraphInstantiate(&f->gexec, g, nullptr, nullptr, 0));
    f->gkey = udx_flu::GraphKey{f->seeds_dev, f->n, f->cur, f->conf.mode, f->wall};
    return UDX_OK;
}

// nsteps >= 2 steps by graph replay (pairs), the last one eagerly when odd
static int flu_run_graph(udx_flu *f, int nsteps) {
    hipStream_t s = f->ctx->stream;
    // the captured steps expect the current positions counted into the slabs
    if (!f->counted) {
        UDX_TRY(clist_count_only(s, &f->cl, f->n, f->pos[f->cur]));
        f->counted = 1;
    }
    UDX_HIP(hipMemsetAsync(f->step_dev, 0, sizeof(int), s));
    UDX_HIP(hipEventRecord(f->gdone, s));
    UDX_HIP(hipStreamWaitEvent(f->gs, f->gdone, 0));
    const udx_flu::GraphKey k{f->seeds_dev, f->n, f->cur, f->conf.mode, f->wall};
    if (!f->gexec || memcmp(&k, &f->gkey, sizeof k) != 0) UDX_TRY(flu_capture(f));
    const int pairs = nsteps / 2;
    for (int p = 0; p < pairs; p++) UDX_HIP(hipGraphLaunch(f->gexec, f->gs));
    UDX_HIP(hipEventRecord(f->gdone, f->gs));
    UDX_HIP(hipStreamWaitEvent(s, f->gdone, 0));
    f->t += (float)(2 * pairs) * f->conf.dt;
    f->counted = 1;   // the replayed moves counted
    if (nsteps & 1) UDX_TRY(flu_step(f, f->seeds_dev + NSEED * (nsteps - 1), flu_fsi_seeds(f, nsteps - 1),
                                     flu_mwall_seed(f, nsteps - 1)));
    return UDX_OK;
}

int udx_flu_run(udx_flu *f, int nsteps, const float *seeds_host) {
    UDX_CHECK(f && nsteps >= 0, UDX_EINVAL, "udx_flu_run: bad arguments");
    if (nsteps == 0) return UDX_OK;
    UDX_TRY(flu_seed_table(f, nsteps, seeds_host));
    if (nsteps >= 2 && flu_graph_ok(f)) return flu_run_graph(f, nsteps);
    if (f->timing) {
        while (f->ev.size() < 5 * (size_t)nsteps) {
            hipEvent_t e;
            UDX_HIP(hipEventCreate(&e));
            f->ev.push_back(e);
        }
    }
    for (int k = 0; k < nsteps; k++) {
        f->evk = f->timing ? f->ev.data() + 5 * (size_t)k : nullptr;
        const int st = flu_step(f, f->seeds_dev + NSEED * k, flu_fsi_seeds(f, k), flu_mwall_seed(f, k));
        f->evk = nullptr;
        UDX_TRY(st);
    }
    if (f->timing) {
        float acc[4] = {0, 0, 0, 0};
        UDX_HIP(hipEventSynchronize(f->ev[5 * (size_t)(nsteps - 1) + 4]));
        // a fused step records walls before the gather (with the move)
        const int sw = f->fused ? 1 : 0;
        for (int k = 0; k < nsteps; k++)
            for (int c = 0; c < 4; c++) {
                float t;
                UDX_HIP(hipEventElapsedTime(&t, f->ev[5 * (size_t)k + c], f->ev[5 * (size_t)k + c + 1]));
                acc[sw && (c == 1 || c == 2) ? 3 - c : c] += t;
            }
        for (int c = 0; c < 4; c++) f->tms[c] = acc[c] / nsteps;
        f->tms[4] = f->tms[0] + f->tms[1] + f->tms[2] + f->tms[3];
    }
    return UDX_OK;
}

int udx_flu_pairs(udx_flu *f, unsigned long long *npairs, int reset) {
    UDX_CHECK(f && npairs, UDX_EINVAL, "udx_flu_pairs: null argument");
    unsigned long long v[2] = {0, 0};
    UDX_HIP(hipMemcpyAsync(v, f->pairs_dev, sizeof v, hipMemcpyDeviceToHost, f->ctx->stream));
    UDX_HIP(hipStreamSynchronize(f->ctx->stream));
    npairs[0] = v[0] / 2;
    if (reset) UDX_HIP(hipMemsetAsync(f->pairs_dev, 0, sizeof v, f->ctx->stream));
    return UDX_OK;
}

int udx_flu_wall_pairs(udx_flu *f, unsigned long long *npairs) {
    UDX_CHECK(f && npairs, UDX_EINVAL, "udx_flu_wall_pairs: null argument");
    UDX_HIP(hipMemcpyAsync(npairs, f->pairs_dev + 1, sizeof *npairs, hipMemcpyDeviceToHost, f->ctx->stream));
    UDX_HIP(hipStreamSynchronize(f->ctx->stream));
    return UDX_OK;
}

// ------------------------------------------------------------------ sdf / walls
int udx_sdf_fin(udx_sdf *q) {
    if (!q) return UDX_OK;
    dev_free(q->f.tex);
    dev_free(q->f.segs);
    delete q;
    return UDX_OK;
}

int udx_sdf_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, udx_sdf **out) {
    UDX_CHECK(ctx && out && Lx > 0 && Ly > 0 && Lz > 0, UDX_EINVAL, "udx_sdf_ini: bad arguments");
    udx_sdf *q = new udx_sdf();
    q->ctx = ctx;
    q->f.L[0] = Lx;
    q->f.L[1] = Ly;
    q->f.L[2] = Lz;
    sdf_texture_size(q->f.L, q->f.T);
    const long nt = (long)q->f.T[0] * q->f.T[1] * q->f.T[2];
    int s = dev_alloc(&q->f.tex, nt);
    if (s) {
        udx_sdf_fin(q);
        return s;
    }
    q->f.view.tex = q->f.tex;
    for (int c = 0; c < 3; c++) q->f.view.T[c] = q->f.T[c];
    *out = q;
    return UDX_OK;
}

int udx_sdf_gen(udx_sdf *q, const float ext[3], const int N[3], const float *D) {
    UDX_CHECK(q && ext && N && D, UDX_EINVAL, "udx_sdf_gen: null argument");
    UDX_CHECK(N[0] > 0 && N[1] > 0 && N[2] > 0 && ext[0] > 0, UDX_EINVAL, "udx_sdf_gen: bad field");
    return sdf_build(q->ctx->stream, &q->f, ext, N, D);
}

// sdf.dat: "ex ey ez\n" "nx ny nz\n" then nx ny nz float32, x fastest
// (src/wall/sdf/field/imp/ini.h:19-43)
int udx_sdf_gen_file(udx_sdf *q, const char *path) {
    UDX_CHECK(q && path, UDX_EINVAL, "udx_sdf_gen_file: null argument");
    FILE *fp = fopen(path, "rb");
    UDX_CHECK(fp, UDX_EINVAL, "udx_sdf_gen_file: cannot open '%s'", path);
    char line[2048];
    float ext[3];
    int N[3], ok = 1;
    if (!fgets(line, sizeof line, fp) || sscanf(line, "%f %f %f", &ext[0], &ext[1], &ext[2]) != 3) ok = 0;
    if (ok && (!fgets(line, sizeof line, fp) || sscanf(line, "%d %d %d", &N[0], &N[1], &N[2]) != 3)) ok = 0;
    std::vector<float> D;
    if (ok && N[0] > 0 && N[1] > 0 && N[2] > 0) {
        D.resize((size_t)N[0] * N[1] * N[2]);
        if (fread(D.data(), sizeof(float), D.size(), fp) != D.size()) ok = 0;
    } else {
        ok = 0;
    }
    fclose(fp);
    UDX_CHECK(ok, UDX_EINVAL, "udx_sdf_gen_file: '%s' is not an sdf.dat file", path);
    return udx_sdf_gen(q, ext, N, D.data());
}

int udx_sdf_size(const udx_sdf *q, int T[3], float *cheap_threshold) {
    UDX_CHECK(q, UDX_EINVAL, "udx_sdf_size: null handle");
    for (int c = 0; c < 3; c++) T[c] = q->f.T[c];
    if (cheap_threshold) *cheap_threshold = q->f.view.cheap_threshold;
    return UDX_OK;
}

const float *udx_sdf_texture(const udx_sdf *q) { return q ? q->f.tex : nullptr; }

int udx_sdf_probe(const udx_sdf *q, int n, const float *r3, int what, float *out) {
    UDX_CHECK(q && q->f.ready && n >= 0 && what >= 0 && what <= 2, UDX_EINVAL, "udx_sdf_probe: bad arguments");
    return sdf_probe(q->ctx->stream, &q->f, n, r3, what, out);
}

int udx_wall_fin(udx_wall *w) {
    if (!w) return UDX_OK;
    wall_free(&w->w);
    dev_free(w->seed_dev);
    dev_free(w->pos);
    dev_free(w->vel);
    dev_free(w->pairs_dev);
    delete w;
    return UDX_OK;
}

int udx_wall_gen(udx_ctx *ctx, const udx_sdf *q, int *n, float *pp, udx_wall **out) {
    UDX_CHECK(ctx && q && q->f.ready && n && out, UDX_EINVAL, "udx_wall_gen: bad arguments");
    udx_wall *w = new udx_wall();
    w->ctx = ctx;
    w->sdf = q;
    udx_rnd_ini(42, 42, 42, 42, &w->trunk);
    int s = wall_build(ctx->stream, &w->w, &q->f, n, pp);
    if (!s) s = dev_alloc(&w->seed_dev, 1);
    if (!s) s = dev_alloc(&w->pairs_dev, 1);
    if (!s && hipMemset(w->pairs_dev, 0, sizeof(unsigned long long)) != hipSuccess) s = UDX_EHIP;
    if (s) {
        udx_wall_fin(w);
        return s;
    }
    *out = w;
    return UDX_OK;
}

// wall strt (src/wall/imp/strt.h:1-60): the frozen particles as template
// "wall.pp"/templ; reading it rebuilds the images and the wall cell list
int udx_wall_strt_dump_templ(const udx_wall *w, const char *base) {
    UDX_CHECK(w && base, UDX_EINVAL, "udx_wall_strt_dump_templ: null argument");
    const int n = w->w.nfrozen;
    std::vector<float> hst(6 * (size_t)(n > 0 ? n : 1));
    if (n > 0) {
        UDX_HIP(hipMemcpyAsync(hst.data(), w->w.frozen, 6 * sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                               w->ctx->stream));
        UDX_HIP(hipStreamSynchronize(w->ctx->stream));
    }
    return udx_restart_write_pp(base, "wall.pp", UDX_RESTART_TEMPL, n, hst.data());
}

int udx_wall_strt_gen(udx_ctx *ctx, const udx_sdf *q, const char *base, int maxn, udx_wall **out) {
    UDX_CHECK(ctx && q && q->f.ready && base && out && maxn >= 0, UDX_EINVAL, "udx_wall_strt_gen: bad arguments");
    std::vector<float> hst(6 * (size_t)(maxn > 0 ? maxn : 1));
    int n = 0;
    UDX_TRY(udx_restart_read_pp(base, "wall.pp", UDX_RESTART_TEMPL, maxn, &n, hst.data()));
    float *frozen = nullptr;
    UDX_TRY(dev_alloc(&frozen, 6 * (size_t)(n > 0 ? n : 1)));
    if (n > 0 && (hipMemcpyAsync(frozen, hst.data(), 6 * sizeof(float) * (size_t)n, hipMemcpyHostToDevice,
                                 ctx->stream) != hipSuccess)) {
        dev_free(frozen);
        set_error("udx_wall_strt_gen: copy failed");
        return UDX_EHIP;
    }
    udx_wall *w = new udx_wall();
    w->ctx = ctx;
    w->sdf = q;
    udx_rnd_ini(42, 42, 42, 42, &w->trunk);
    int s = wall_build_frozen(ctx->stream, &w->w, &q->f, n, frozen);
    if (!s) s = dev_alloc(&w->seed_dev, 1);
    if (!s) s = dev_alloc(&w->pairs_dev, 1);
    if (!s && hipMemset(w->pairs_dev, 0, sizeof(unsigned long long)) != hipSuccess) s = UDX_EHIP;
    if (s) {
        udx_wall_fin(w);
        return s;
    }
    *out = w;
    return UDX_OK;
}

int udx_wall_counts(const udx_wall *w, int *nfrozen, int *ntotal) {
    UDX_CHECK(w, UDX_EINVAL, "udx_wall_counts: null handle");
    if (nfrozen) *nfrozen = w->w.nfrozen;
    if (ntotal) *ntotal = w->w.n;
    return UDX_OK;
}

const float *udx_wall_pp(const udx_wall *w) { return w ? w->w.pp_aos : nullptr; }
const int *udx_wall_starts(const udx_wall *w) { return w ? w->w.cl.starts : nullptr; }

int udx_wall_set_mode(udx_wall *w, int mode) {
    UDX_CHECK(w, UDX_EINVAL, "udx_wall_set_mode: null handle");
    UDX_CHECK(mode >= UDX_MODE_FAST && mode <= UDX_MODE_EXACT, UDX_EINVAL, "udx_wall_set_mode: mode %d", mode);
    w->mode = mode;
    return UDX_OK;
}

int udx_wall_force_seed(udx_wall *w, const udx_pair_dpd *p, int n, const float *pp, float seed, float *ff) {
    UDX_CHECK(w && p && n >= 0, UDX_EINVAL, "udx_wall_force_seed: bad arguments");
    if (n == 0) return UDX_OK;
    hipStream_t s = w->ctx->stream;
    if (n > w->cap) {
        UDX_HIP(hipStreamSynchronize(s));
        dev_free(w->pos);
        dev_free(w->vel);
        w->cap = 0;
        UDX_TRY(dev_alloc(&w->pos, n));
        UDX_TRY(dev_alloc(&w->vel, n));
        w->cap = n;
    }
    UDX_TRY(aos_to_f4(s, n, pp, w->pos, w->vel));
    UDX_HIP(hipMemcpyAsync(w->seed_dev, &seed, sizeof seed, hipMemcpyHostToDevice, s));
    // n as the device-side count: the first element of a one-int buffer
    int *n_dev = nullptr;
    UDX_TRY(dev_alloc(&n_dev, 1));
    int st = UDX_OK;
    if (hipMemcpyAsync(n_dev, &n, sizeof n, hipMemcpyHostToDevice, s) != hipSuccess) st = UDX_EHIP;
    if (!st)
        st = wall_forces(s, &w->sdf->f, &w->w, n, n_dev, w->pos, w->vel, *p, w->seed_dev, ff, 3, w->pairs_dev,
                         w->mode == UDX_MODE_EXACT, wall_vel_view(w->wvel, w->t));
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = UDX_EHIP;
    dev_free(n_dev);
    if (st) set_error("udx_wall_force: launch failed");
    return st;
}

int udx_wall_force(udx_wall *w, const udx_pair_dpd *p, int n, const float *pp, float *ff) {
    UDX_CHECK(w, UDX_EINVAL, "udx_wall_force: null handle");
    return udx_wall_force_seed(w, p, n, pp, udx_rnd_get(&w->trunk), ff);
}

int udx_wall_pairs(udx_wall *w, unsigned long long *npairs, int reset) {
    UDX_CHECK(w && npairs, UDX_EINVAL, "udx_wall_pairs: null argument");
    UDX_HIP(hipMemcpyAsync(npairs, w->pairs_dev, sizeof *npairs, hipMemcpyDeviceToHost, w->ctx->stream));
    UDX_HIP(hipStreamSynchronize(w->ctx->stream));
    if (reset) UDX_HIP(hipMemsetAsync(w->pairs_dev, 0, sizeof *npairs, w->ctx->stream));
    return UDX_OK;
}

int udx_wall_bounce(const udx_sdf *q, float dt, int n, float *pp) {
    UDX_CHECK(q && q->f.ready, UDX_EINVAL, "udx_wall_bounce: sdf not generated");
    return wall_bounce_aos(q->ctx->stream, &q->f, dt, n, pp, WallVel{});
}

static int wall_vel_check(const udx_wall_vel *v) {
    UDX_CHECK(v, UDX_EINVAL, "udx_wall_set_vel: null argument");
    UDX_CHECK(v->type >= UDX_WVEL_NONE && v->type <= UDX_WVEL_HS, UDX_EINVAL, "udx_wall_set_vel: type %d", v->type);
    if (v->type == UDX_WVEL_SHEAR || v->type == UDX_WVEL_SHEAR_SIN)
        UDX_CHECK(v->vdir >= 0 && v->vdir < 3 && v->gdir >= 0 && v->gdir < 3, UDX_EINVAL,
                  "udx_wall_set_vel: vdir %d, gdir %d", v->vdir, v->gdir);
    if (v->type == UDX_WVEL_HS) UDX_CHECK(v->hs_h > 0, UDX_EINVAL, "udx_wall_set_vel: Hele-Shaw height %g", v->hs_h);
    return UDX_OK;
}

int udx_wall_set_vel(udx_wall *w, const udx_wall_vel *v) {
    UDX_CHECK(w, UDX_EINVAL, "udx_wall_set_vel: null handle");
    UDX_TRY(wall_vel_check(v));
    w->wvel = *v;
    return UDX_OK;
}

int udx_wall_set_time(udx_wall *w, float t) {
    UDX_CHECK(w, UDX_EINVAL, "udx_wall_set_time: null handle");
    w->t = t;
    return UDX_OK;
}

int udx_wall_bounce_wall(const udx_wall *w, float dt, int n, float *pp) {
    UDX_CHECK(w && n >= 0 && (n == 0 || pp), UDX_EINVAL, "udx_wall_bounce_wall: bad arguments");
    return wall_bounce_aos(w->ctx->stream, wall_sdf(w), dt, n, pp, wall_vel_view(w->wvel, w->t));
}

int udx_wall_repulse(const udx_sdf *q, float lambda, int n, const float *pp, float *ff) {
    UDX_CHECK(q && q->f.ready, UDX_EINVAL, "udx_wall_repulse: sdf not generated");
    return wall_repulse_aos(q->ctx->stream, &q->f, lambda, n, pp, ff);
}

int udx_flu_set_wall(udx_flu *f, udx_wall *w) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_wall: null handle");
    UDX_CHECK(!w || (w->sdf->f.L[0] == f->box.L[0] && w->sdf->f.L[1] == f->box.L[1] && w->sdf->f.L[2] == f->box.L[2]),
              UDX_EINVAL, "udx_flu_set_wall: wall and solvent boxes differ");
    if (w != f->wall) flu_drop_graph(f);
    f->wall = w;
    return UDX_OK;
}

int udx_flu_set_graph(udx_flu *f, int enable) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_graph: null handle");
    if (enable && !f->gs) {
        UDX_HIP(hipStreamCreate(&f->gs));
        UDX_HIP(hipEventCreateWithFlags(&f->gdone, hipEventDisableTiming));
        UDX_TRY(dev_alloc(&f->step_dev, 1));
        UDX_TRY(dev_alloc(&f->seeds_cur, NSEED));
    }
    if (!enable && f->gs) {
        UDX_HIP(hipStreamSynchronize(f->gs));
        flu_drop_graph(f);
    }
    f->use_graph = enable ? 1 : 0;
    return UDX_OK;
}

// flu_strt_dump / flu_strt_quants (src/flu/imp/start.h:20-41): code "flu.pp"
int udx_flu_strt_dump(udx_flu *f, const char *base, int id) {
    UDX_CHECK(f && base, UDX_EINVAL, "udx_flu_strt_dump: null argument");
    const int n = f->n;
    std::vector<float> hst(6 * (size_t)(n > 0 ? n : 1));
    if (n > 0) {
        float *d = nullptr;
        UDX_TRY(dev_alloc(&d, 6 * (size_t)n));
        int st = f4_to_aos(f->ctx->stream, n, nullptr, f->pos[f->cur], f->vel[f->cur], d);
        if (!st && (hipMemcpyAsync(hst.data(), d, 6 * sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                                   f->ctx->stream) != hipSuccess ||
                    hipStreamSynchronize(f->ctx->stream) != hipSuccess))
            st = UDX_EHIP;
        dev_free(d);
        UDX_CHECK(!st, st, "udx_flu_strt_dump: copy failed");
    }
    return udx_restart_write_pp(base, "flu.pp", id, n, hst.data());
}

int udx_flu_strt_quants(udx_flu *f, const char *base, int id) {
    UDX_CHECK(f && base, UDX_EINVAL, "udx_flu_strt_quants: null argument");
    std::vector<float> hst(6 * (size_t)f->conf.maxp);
    int n = 0;
    UDX_TRY(udx_restart_read_pp(base, "flu.pp", id, f->conf.maxp, &n, hst.data()));
    float *d = nullptr;
    UDX_TRY(dev_alloc(&d, 6 * (size_t)(n > 0 ? n : 1)));
    int st = UDX_OK;
    if (n > 0 && hipMemcpyAsync(d, hst.data(), 6 * sizeof(float) * (size_t)n, hipMemcpyHostToDevice,
                                f->ctx->stream) != hipSuccess)
        st = UDX_EHIP;
    if (!st) st = udx_flu_set_pp(f, n, d);
    if (!st && hipStreamSynchronize(f->ctx->stream) != hipSuccess) st = UDX_EHIP;
    dev_free(d);
    UDX_CHECK(!st, st, "udx_flu_strt_quants: copy failed");
    return UDX_OK;
}

int udx_flu_set_time(udx_flu *f, float t) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_time: null handle");
    f->t = t;
    return UDX_OK;
}

int udx_flu_set_rbc(udx_flu *f, udx_rbc *r, udx_fsi *fsi, const udx_rbc_params *rp, const udx_pair_dpd *fp,
                    float mass, int substeps, int nc, float *pp) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_rbc: null handle");
    if (!r) {
        UDX_HIP(hipStreamSynchronize(f->ctx->stream));
        flu_free_membranes(f);
        return UDX_OK;
    }
    UDX_CHECK(fsi && rp && fp && (nc == 0 || pp), UDX_EINVAL, "udx_flu_set_rbc: null argument");
    UDX_CHECK(rbc_ctx(r) == f->ctx && fsi_ctx(fsi) == f->ctx, UDX_EINVAL,
              "udx_flu_set_rbc: membranes, FSI and solvent must share one context");
    UDX_CHECK(mass > 0 && substeps >= 1, UDX_EINVAL, "udx_flu_set_rbc: bad mass or substeps");
    const int nv = rbc_nv(r);
    UDX_CHECK(nc >= 0 && nc <= rbc_max_cells(r), UDX_ECAP, "udx_flu_set_rbc: %d cells above the membrane capacity %d",
              nc, rbc_max_cells(r));
    UDX_CHECK((long)nc * nv <= fsi_bufs(fsi)->max_obj, UDX_ECAP,
              "udx_flu_set_rbc: %d vertices above the FSI capacity %d", nc * nv, fsi_bufs(fsi)->max_obj);
    const FsiBufs *fb = fsi_bufs(fsi);
    UDX_CHECK(fb->box.L[0] == f->box.L[0] && fb->box.L[1] == f->box.L[1] && fb->box.L[2] == f->box.L[2], UDX_EINVAL,
              "udx_flu_set_rbc: FSI and solvent boxes differ");
    udx_flu::Membranes &m = f->mb;
    const int n = nc * nv;
    // vertex arrays hold m.cap entries and the per-cell ones m.ccap: a new
    // mesh with fewer vertices per cell can bring more cells in the same
    // vertex capacity
    if (n > m.cap || nc > m.ccap) {
        UDX_HIP(hipStreamSynchronize(f->ctx->stream));
        flu_free_membranes(f, true);
        UDX_TRY(dev_alloc(&m.fo, 3 * (size_t)n));
        UDX_TRY(dev_alloc(&m.start_pp, 6 * (size_t)(n > 0 ? n : 1)));
        UDX_TRY(dev_alloc(&m.vpos, (size_t)(n > 0 ? n : 1)));
        UDX_TRY(dev_alloc(&m.vvel, (size_t)(n > 0 ? n : 1)));
        UDX_TRY(dev_alloc(&m.ffast, 3 * (size_t)n));
        UDX_TRY(dev_alloc(&m.tmp, 6 * (size_t)n));
        UDX_TRY(dev_alloc(&m.cfid, (size_t)(nc > 0 ? nc : 1)));
        UDX_TRY(dev_alloc(&m.cnew, (size_t)(nc > 0 ? nc : 1)));
        m.cap = n;
        m.ccap = nc;
    }
    m.r = r;
    m.fsi = fsi;
    m.rp = *rp;
    m.fp = *fp;
    m.mass = mass;
    m.substeps = substeps;
    m.nc = nc;
    m.nv = nv;
    m.pp = pp;
    if (!m.nvtx_dev) UDX_TRY(dev_alloc(&m.nvtx_dev, 1));
    {
        const int nvtx = nc * nv;
        UDX_HIP(hipStreamSynchronize(f->ctx->stream));
        UDX_HIP(hipMemcpy(m.nvtx_dev, &nvtx, sizeof nvtx, hipMemcpyHostToDevice));
    }
    if (m.bb && mbb_nv(m.bb) != nv) {
        m.bb = nullptr;
        UDX_CHECK(false, UDX_EINVAL, "udx_flu_set_rbc: the bounce-back template differs; bounce-back switched off");
    }
    return UDX_OK;
}

int udx_flu_set_rbc_repulsion(udx_flu *f, float lambda) {
    UDX_CHECK(f && lambda >= 0, UDX_EINVAL, "udx_flu_set_rbc_repulsion: bad arguments");
    f->mb.wall_lambda = lambda;
    return UDX_OK;
}

int udx_flu_set_rbc_contact(udx_flu *f, udx_cnt *c, const udx_pair_dpd_lj *p) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_rbc_contact: null handle");
    udx_flu::Membranes &m = f->mb;
    if (!c) {
        m.cnt = nullptr;
        return UDX_OK;
    }
    UDX_CHECK(p, UDX_EINVAL, "udx_flu_set_rbc_contact: null parameters");
    UDX_CHECK(m.r, UDX_EINVAL, "udx_flu_set_rbc_contact: no membranes (udx_flu_set_rbc first)");
    UDX_CHECK(cnt_ctx(c) == f->ctx, UDX_EINVAL, "udx_flu_set_rbc_contact: contact and solvent contexts differ");
    m.cnt = c;
    m.cnt_par = *p;
    return UDX_OK;
}

int udx_flu_set_rbc_bounce(udx_flu *f, udx_mbb *b) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_set_rbc_bounce: null handle");
    udx_flu::Membranes &m = f->mb;
    if (!b) {
        m.bb = nullptr;
        return UDX_OK;
    }
    UDX_CHECK(m.r, UDX_EINVAL, "udx_flu_set_rbc_bounce: no membranes (udx_flu_set_rbc first)");
    UDX_CHECK(mbb_ctx(b) == f->ctx, UDX_EINVAL, "udx_flu_set_rbc_bounce: bounce-back and solvent contexts differ");
    UDX_CHECK(mbb_nv(b) == m.nv, UDX_EINVAL, "udx_flu_set_rbc_bounce: template of %d vertices, membranes of %d",
              mbb_nv(b), m.nv);
    if (!m.pos0) UDX_TRY(dev_alloc(&m.pos0, (size_t)f->conf.maxp));
    m.bb = b;
    return UDX_OK;
}

int udx_flu_timing(udx_flu *f, int enable, float out[5]) {
    UDX_CHECK(f, UDX_EINVAL, "udx_flu_timing: null handle");
    f->timing = enable ? 1 : 0;
    if (out)
        for (int c = 0; c < 5; c++) out[c] = f->tms[c];
    return UDX_OK;
}

}  // extern "C"
