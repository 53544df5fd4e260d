// Solvent-solvent DPD forces for MI355X.
//
// Reference: the bulk kernel visits a half shell with float atomics on both
// particles (src/fluforces/bulk/dev/main.h:23-97) and 26 halo kernels add the
// one-sided forces of pairs across the periodic boundary with their own RNG
// (src/fluforces/halo/dev/main.h:7-83).  Here one pass gathers the full shell
// of each particle in a fixed order, so forces are deterministic and need no
// atomics: first the pairs inside the box, then (only for particles in a
// boundary cell of a periodic box) the pairs across the boundary, each
// computed exactly like the reference's halo: the local particle shifted by
// -L d into the neighbour frame and the noise drawn from the fragment's trunk
// with packed-list positions.
//
// Kernel shape: one thread per cell-ordered particle.  A candidate pass walks
// the 9 rows (dz, dy) of the 27-cell shell, each row a contiguous run of at
// most 3 cells cut to the particle's x window, tests the distance and appends
// hits to a per-lane queue in LDS.  When a queue fills (and at the end of a
// pass) the wave drains the queues together, so the expensive pair body (RNG
// + force) runs with nearly all lanes active instead of the ~25% of the
// candidates that pass the cutoff.
#include <type_traits>
#include <vector>

#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {

namespace {

// 128 threads: measured 5% faster per step than 256 on the 96^3, density-10 box
constexpr int DPD_THREADS = 128;
constexpr int QCAP = 32;   // queued pairs per lane (a power of two)
// staged candidates per wave and row (320 for batches of small boxes measured
// 4% slower than 160: more staging loads, fewer blocks per CU)
constexpr int SCAP = 160, SCAP_BATCH = 160;
constexpr int DRAIN_STEPS = 2;  // pair steps per lane when a queue is full
// bulk queue entries: generation (GBITS) and offset (OBITS) in 16 bits
constexpr int OBITS = 11, GBITS = 5, GTAB = 1 << GBITS;
constexpr unsigned OMASK = (1u << OBITS) - 1u;
static_assert(OBITS + GBITS == 16, "16-bit entries");

struct DpdK {
    const float4 *pos, *vel;
    const int *starts, *fss;
    const int *bstarts;       // x-slab starts (sub per cell) or null: whole cells
    int sub;
    int foff[27];
    int Lx, Ly, Lz, nmax, periodic;
    int i0;                   // first particle of the box; noise indices are relative to it
    DpdPar par;
    const float *seeds;
    float *ff;
    int ff_stride, accumulate;
    unsigned long long *pairs;
    float *ss;                // per-particle stress (6 floats, added) or null
    double *acc;              // FP64 force accumulators (half-shell pass) or null
    // edge list (null: the halo pass runs inside the force kernel): the
    // particles with pairs across the periodic boundary and their bulk sums,
    // box entries from i0, *ecount of them
    int *ecount, *eidx;
    float4 *ef;
    // the move fused into the pass (k_dpd with MOVE set; FusedMove)
    float4 *opos, *ovel;
    float mdt, mmass, max_, may, maz;
    SdfView sdf;
    WallVel wv;
    SlabCount sc;
    // the wall pass (MOVE == 2): frozen wall particles, their cell list on
    // the margin grid W, the wall noise seed and the wall pair counter
    const float4 *wpos;
    const int *wstarts;
    int Wx, Wy, Wz, nwall;
    const float *wseed;
    unsigned long long *wpairs;
};

// What the force pass reads of a box, read once at kernel entry.  The
// arrays are typed as global memory: pointers read from a table in memory
// (the batched kernel) are otherwise generic, and generic (flat) loads count
// against the LDS wait counter as well, so every LDS read of the staged
// candidate loops would also wait for the outstanding particle loads.  The
// scalars are copied so that the loops do not read them from the table again
// (a read inside a conditional block cannot be hoisted out of the loop).
#define UDX_GLOBAL __attribute__((address_space(1)))
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f3v __attribute__((ext_vector_type(3)));   // 16 B in memory; loads read 12

template <class T>
__device__ __forceinline__ UDX_GLOBAL T *gptr(T *p) {
    return (UDX_GLOBAL T *)p;
}

struct DpdP {
    const UDX_GLOBAL f4v *pos, *vel;
    const UDX_GLOBAL int *starts, *fss, *bstarts;
    const UDX_GLOBAL float *seeds;
    int Lx, Ly, Lz, nmax, periodic, i0, sub;
    DpdPar par;
    // x, y, z of a particle (w is not used)
    __device__ __forceinline__ float4 p(int j) const {
        const f3v v = *(const UDX_GLOBAL f3v *)(pos + j);
        return make_float4(v.x, v.y, v.z, 0.f);
    }
    __device__ __forceinline__ float4 v(int j) const {
        const f3v u = *(const UDX_GLOBAL f3v *)(vel + j);
        return make_float4(u.x, u.y, u.z, 0.f);
    }
};

__device__ __forceinline__ DpdP dpd_view(const DpdK &A) {
    return DpdP{(const UDX_GLOBAL f4v *)A.pos,
                (const UDX_GLOBAL f4v *)A.vel,
                gptr(A.starts),
                gptr(A.fss),
                gptr(A.bstarts),
                gptr(A.seeds),
                A.Lx,
                A.Ly,
                A.Lz,
                A.nmax,
                A.periodic,
                A.i0,
                A.sub,
                A.par};
}

// Runs of row (dy, dz) of a particle in column cx of a periodic box, in
// ascending dx order: seg 0 and seg 1 (two runs where x wraps, Lx >= 3);
// wx is the periodic shift of the run's cells seen from the particle.
__device__ __forceinline__ bool xrun(int sg, int cx, int Lx, int *x0, int *x1, int *wx) {
    if (sg == 0) {
        if (cx == 0) { *x0 = *x1 = Lx - 1; *wx = -1; return true; }
        *x0 = cx - 1;
        *x1 = cx + 1 < Lx ? cx + 1 : Lx - 1;
        *wx = 0;
        return true;
    }
    if (cx == 0) { *x0 = 0; *x1 = 1; *wx = 0; return true; }
    if (cx == Lx - 1) { *x0 = *x1 = 0; *wx = 1; return true; }
    return false;
}

// Queue entry of a candidate that passed the cutoff: the neighbour's index j
// (27 bits) and the periodic wrap w of its cell seen from the home cell,
// packed as (wx + 1) + 3 (wy + 1) + 9 (wz + 1); 13 means no wrap.  Entries of
// the bulk pass carry j only.
constexpr unsigned W_NONE = 13u;

__device__ __forceinline__ unsigned wcode(int wx, int wy, int wz) {
    return (unsigned)((wx + 1) + 3 * (wy + 1) + 9 * (wz + 1));
}

// inputs of one queued pair: separation, relative velocity and the logistic
// map argument seed - p of its noise
struct PairIn {
    float dx, dy, dz, r2, dvx, dvy, dvz, x;
};

__device__ __forceinline__ void pair_kin(PairIn &o, float ax, float ay, float az, const float4 &vi, const float4 &pj,
                                         const float4 &vj) {
    o.dx = ax - pj.x;
    o.dy = ay - pj.y;
    o.dz = az - pj.z;
    o.r2 = __builtin_fmaf(o.dz, o.dz, __builtin_fmaf(o.dx, o.dx, o.dy * o.dy));
    o.dvx = vi.x - vj.x;
    o.dvy = vi.y - vj.y;
    o.dvz = vi.z - vj.z;
}

// a pair inside the box: noise mean0var1ii(seed, max, min)
// (src/fluforces/bulk/dev/main.h:32-38)
__device__ __forceinline__ PairIn pair_in_bulk_j(const DpdP &P, int j, int i, const float4 &pi, const float4 &vi,
                                                 float seed_bulk) {
    const float4 pj = P.p(j), vj = P.v(j);
    const int u = (i > j ? i : j) - P.i0, v = (i > j ? j : i) - P.i0;
    PairIn o;
    pair_kin(o, pi.x, pi.y, pi.z, vi, pj, vj);
    o.x = seed_bulk - low_discrepancy((float)(u & 0x3FF), (float)u, (float)(v & 0x3FF), (float)v);
    return o;
}

__device__ __forceinline__ PairIn pair_in_bulk(const DpdP &P, unsigned ent, int i, const float4 &pi,
                                               const float4 &vi, float seed_bulk) {
    return pair_in_bulk_j(P, (int)(ent & 0x7FFFFFFu), i, pi, vi, seed_bulk);
}

// a pair across the periodic boundary: the local particle is moved into the
// neighbour frame (src/fluforces/halo/dev/main.h:40-48) and the noise is
// mean0var1uu(seed_d, ...) of fragment d = w with positions in the packed
// lists (src/fluforces/halo/dev/main.h:7-12, src/exch/flu/dev/pack.h:2-40).
// An entry with w = W_NONE (an idle slot) is evaluated as a bulk pair.
__device__ __forceinline__ PairIn pair_in_halo(const DpdK &A, const DpdP &P, unsigned ent, int i, const float4 &pi,
                                               const float4 &vi, int cx, int cy, int cz, int istart,
                                               float seed_bulk) {
    const unsigned w = ent >> 27;
    if (w == W_NONE) return pair_in_bulk(P, ent, i, pi, vi, seed_bulk);
    const int j = (int)(ent & 0x7FFFFFFu);
    const float4 pj = P.p(j), vj = P.v(j);
    const int Lx = P.Lx, Ly = P.Ly, Lz = P.Lz;
    const int wx = (int)(w % 3u) - 1, wy = (int)((w / 3u) % 3u) - 1, wz = (int)(w / 9u) - 1;
    const float ax = pi.x - (float)(wx * Lx), ay = pi.y - (float)(wy * Ly), az = pi.z - (float)(wz * Lz);
    const int fid = frag_d2i(wx, wy, wz), anti = frag_d2i(-wx, -wy, -wz);
    const int aid = P.fss[A.foff[fid] + frag_cell(wx, wy, wz, cx, cy, cz, Lx, Ly)] - P.fss[A.foff[fid]] + (i - istart);
    int jx = 0, jy = 0, jz = 0;
    cell_coord(pj.x, Lx, &jx);
    cell_coord(pj.y, Ly, &jy);
    cell_coord(pj.z, Lz, &jz);
    const int jstart = P.starts[jx + Lx * (jy + Ly * jz)];
    const int bid =
        P.fss[A.foff[anti] + frag_cell(-wx, -wy, -wz, jx, jy, jz, Lx, Ly)] - P.fss[A.foff[anti]] + (j - jstart);
    const bool mask = fid < anti;
    const int u = mask ? aid : bid, v = mask ? bid : aid;
    PairIn o;
    pair_kin(o, ax, ay, az, vi, pj, vj);
    o.x = P.seeds[fid] - low_discrepancy((float)(u & 0x3FF), (float)u, (float)(v & 0x3FF), (float)v);
    return o;
}

// a wrapped pair from its queue entry (j | w << 27) and its noise indices u, v
// (positions in the packed fragment lists, set at the push): the local
// particle moved into the neighbour frame, noise mean0var1uu(seed_w, u, v)
// (src/fluforces/halo/dev/main.h:7-12,40-48).  An idle entry (w = W_NONE, the
// particle itself) gives a zero force that is not added.
__device__ __forceinline__ PairIn pair_in_wrapped(const DpdP &P, unsigned ent, int u, int v, const float4 &pi,
                                                  const float4 &vi) {
    const unsigned w = ent >> 27;
    const int j = (int)(ent & 0x7FFFFFFu);
    const float4 pj = P.p(j), vj = P.v(j);
    const int wx = (int)(w % 3u) - 1, wy = (int)((w / 3u) % 3u) - 1, wz = (int)(w / 9u) - 1;
    const float ax = pi.x - (float)(wx * P.Lx), ay = pi.y - (float)(wy * P.Ly), az = pi.z - (float)(wz * P.Lz);
    PairIn o;
    pair_kin(o, ax, ay, az, vi, pj, vj);
    o.x = P.seeds[frag_d2i(wx, wy, wz)] - low_discrepancy((float)(u & 0x3FF), (float)u, (float)(v & 0x3FF), (float)v);
    return o;
}

// Candidate window along x.  With x-slab starts a row run [x0, x1] of cells
// (wrap wx) is cut to the slabs that can hold a particle within the cutoff:
// a candidate at y-z distance >= d from the particle has |dx| < sqrt(1 - d^2).
// The bound is conservative for the float pair test: the y-z gaps g are
// lowered and the half-width raised by WIN_MARGIN (above the rounding of
// the particle's periodic shift for boxes up to 2048 cells) and 1 + 1e-5
// stands for the cutoff, so every candidate the test would accept lies in
// the window.  hw < 0: no particle of the row is within the cutoff.
constexpr float WIN_MARGIN = 1.0f / 4096.0f;

__device__ __forceinline__ float row_halfwidth(int dy, int dz, float fy, float fz) {
    float gy = dy == 0 ? 0.f : (dy < 0 ? fy : 1.f - fy);
    float gz = dz == 0 ? 0.f : (dz < 0 ? fz : 1.f - fz);
    gy = fmaxf(gy - WIN_MARGIN, 0.f);
    gz = fmaxf(gz - WIN_MARGIN, 0.f);
    const float w2 = 1.00001f - gy * gy - gz * gz;
    return w2 < 0.f ? -1.f : __builtin_sqrtf(w2) + WIN_MARGIN;
}

// candidates of the run [x0, x1] of the row starting at cell `base`, for the
// particle at x = ax in the run's frame: index range [*j, *j + *n)
__device__ __forceinline__ void run_range(const DpdP &P, int base, int x0, int x1, float ax, float hw, int *j,
                                          int *n) {
    if (!P.bstarts) {
        *j = P.starts[base + x0];
        *n = P.starts[base + x1 + 1] - *j;
        return;
    }
    const int S = P.sub;
    const double h = 0.5 * (double)P.Lx;
    double lo = __builtin_floor(((double)ax - (double)hw + h) * (double)S);
    double hi = __builtin_floor(((double)ax + (double)hw + h) * (double)S);
    lo = fmax(lo, (double)(x0 * S));
    hi = fmin(hi, (double)(x1 * S + S - 1));
    if (!(hw >= 0.f) || lo > hi) {
        *j = 0;
        *n = 0;
        return;
    }
    const int sb = base * S;
    *j = P.bstarts[sb + (int)lo];
    *n = P.bstarts[sb + (int)hi + 1] - *j;
}

template <bool B>
using PassTag = std::integral_constant<bool, B>;

// what a launch of the gather does: both passes, or the bulk pass with the
// particles that have wrapped pairs handed to an edge list, or the halo pass
// of the listed particles (their sums continue from the listed bulk sums, so
// the order of the additions is that of one pass)
enum { PH_ALL = 0, PH_BULK = 1, PH_HALO = 2 };

// The force pass of one block of NT consecutive particles starting at i
// = first + threadIdx.x; QC (a power of two) queued pairs per lane.
//
// Bulk pass: the 9 rows (dz, dy) of the 27-cell shell clipped to the box,
// each a contiguous run of at most 3 cells cut to the particle's x window.
// The lanes are consecutive cell-ordered particles, so the union of the
// wave's runs of a row is one index range: staged in LDS with coalesced
// loads (or read from global memory when it does not fit) and tested four
// candidates per iteration.  Halo pass (periodic boxes, only waves with a
// particle in a boundary cell): the wrapped parts of the 9 rows, read from
// global memory.  Each lane sums its pairs in candidate order, bulk first.
//
// MOVE (one launch of a periodic box, float4 forces): 1 the particles are
// also moved (FusedMove), 2 the same with walls.
template <int WR, bool EXACT, int NT, int QC, int SC, bool STRESS = false, int PHASE = PH_ALL, int MOVE = 0>
__device__ __forceinline__ void dpd_block(const DpdK &A, int first) {
    static_assert(QC >= 8 && (QC & (QC - 1)) == 0, "the queue length is a power of two");
    // 16-bit bulk entries, one region of QC * 128 bytes per wave, which the
    // wave's halo rings reuse: the block's two waves run their passes
    // independently, so neither may touch the other's region
    __shared__ unsigned q[QC * NT / 2];
    __shared__ float4 sbuf[NT / 64][SC + 4];
    __shared__ int gtab[NT / 64][GTAB];    // candidate bases of the bulk entries' generations
    __shared__ unsigned long long lpairs[NT / 64];
    const int tid = threadIdx.x;
    const DpdP P = dpd_view(A);
    const int Lx = P.Lx, Ly = P.Ly, Lz = P.Lz;
    const int n = P.starts[Lx * Ly * Lz];
    int i = first + tid;
    bool active = i < n && i - P.i0 < P.nmax;
    float f0x = 0.f, f0y = 0.f, f0z = 0.f;   // the listed bulk sums (halo phase)
    if (PHASE == PH_HALO) {
        // entries first + tid of the box's edge list; a block past its end has nothing to do
        const int ne = *A.ecount;
        if (first - P.i0 >= ne) return;
        active = first + tid - P.i0 < ne;
        i = P.i0;
        if (active) {
            const int k = first + tid;
            i = A.eidx[k];
            const float4 f = A.ef[k];
            f0x = f.x;
            f0y = f.y;
            f0z = f.z;
        }
    }

    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi;
    int cx = 0, cy = 0, cz = 0, istart = 0;
    float ry = 0.f, rz = 0.f;   // position inside the cell along y and z
    if (active) {
        pi = P.p(i);
        vi = P.v(i);
        cell_coord(pi.x, Lx, &cx);
        cell_coord(pi.y, Ly, &cy);
        cell_coord(pi.z, Lz, &cz);
        istart = P.starts[cx + Lx * (cy + Ly * cz)];
        ry = (float)((double)pi.y + 0.5 * (double)Ly - (double)cy);
        rz = (float)((double)pi.z + 0.5 * (double)Lz - (double)cz);
    }
    const float seed_bulk = P.seeds[26];
    float fx = f0x, fy = f0y, fz = f0z;
    // stress 1/2 f_a dr_b, (xx, xy, xz, yy, yz, zz) (src/pair/dev.h:87-100)
    float sxx = 0.f, sxy = 0.f, sxz = 0.f, syy = 0.f, syz = 0.f, szz = 0.f;
    int npair = 0;

    // Each lane's queue is a ring of QC LDS slots (entry k at slot k mod QC):
    // head is the next entry to evaluate, tail the next free slot.  A pair
    // step evaluates the next two entries of every lane, so two noise chains
    // share each round-toward-zero window.  When a lane runs out of room the
    // wave takes DRAIN_STEPS steps only: lanes fill their queues at very
    // different rates (a lane near the -z face of its cell hits most in the
    // dz = -1 rows), and emptying every queue at that point would run the
    // expensive pair body with most lanes idle; deeper queues leave fewer
    // idle slots, so an entry is 16 bits: a generation g (GBITS) and an
    // offset (OBITS), the candidate being gtab[g mod GTAB] + offset.  Each
    // row stages one generation per 2^OBITS candidates of the wave's union,
    // and a generation's table slot is reused only after a full drain.  An
    // idle slot evaluates a discarded pair with an in-bounds particle (i
    // itself, or the box's first particle for a lane past the end).
    // head and tail count in units of STEP bytes (one slot of every lane of
    // the wave), so the LDS address of an entry in the wave's region is
    // (count & QMASK) | (lane address)
    constexpr unsigned STEP = 2u * 64u, QMASK = (unsigned)(QC - 1) * STEP;
    unsigned head = 0, tail = 0;
    char *const qw = (char *)q + (tid >> 6) * (QC * 128);
    const unsigned qlane = (unsigned)((tid & 63) * 2);
    auto slot = [&](unsigned k) { return (unsigned short *)(qw + ((k & QMASK) | qlane)); };
    int *const gt = gtab[tid >> 6];
    int gen = 0, gfull = 0;   // next generation; generations below gfull are drained
    const int idle_j = active ? i : P.i0;
    const unsigned idle = (unsigned)idle_j | (W_NONE << 27);
    auto add = [&](bool v, const float3 &f, const PairIn &pq) {
        if (v) {
            fx += f.x;
            fy += f.y;
            fz += f.z;
            if (STRESS) {
                sxx += 0.5f * f.x * pq.dx;
                sxy += 0.5f * f.x * pq.dy;
                sxz += 0.5f * f.x * pq.dz;
                syy += 0.5f * f.y * pq.dy;
                syz += 0.5f * f.y * pq.dz;
                szz += 0.5f * f.z * pq.dz;
            }
        }
    };
    auto decode = [&](unsigned e) { return gt[e >> OBITS] + (int)(e & OMASK); };
    auto pair_step = [&](auto) {
        const unsigned fl = tail - head;
        const bool va = fl > 0u, vb = fl > STEP;
        const int ja = va ? decode(*slot(head)) : idle_j;
        const int jb = vb ? decode(*slot(head + STEP)) : idle_j;
        const PairIn a = pair_in_bulk_j(P, ja, i, pi, vi, seed_bulk);
        const PairIn b = pair_in_bulk_j(P, jb, i, pi, vi, seed_bulk);
        float xa = a.x, xb = b.x;
        logistic18x2(&xa, &xb);
        const float3 f1 = dpd_pair<WR, EXACT>(P.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
        const float3 f2 = dpd_pair<WR, EXACT>(P.par, b.dx, b.dy, b.dz, b.r2, b.dvx, b.dvy, b.dvz, xb * R_SQRT2);
        add(va, f1, a);
        add(vb, f2, b);
        const unsigned took = fl < 2u * STEP ? fl : 2u * STEP;
        head += took;
        npair += (int)(took / STEP);
    };
    // room for `need` more entries in every lane (need <= 2 DRAIN_STEPS)
    auto make_room = [&](auto pass, unsigned need) {
        if (wave_any(tail - head + need * STEP > (unsigned)QC * STEP))
#pragma unroll 1
            for (int k = 0; k < DRAIN_STEPS; k++) pair_step(pass);
    };
    auto drain = [&](auto pass) {
        while (wave_any(tail != head)) pair_step(pass);
        gfull = gen;
    };
    // the slot at tail is free whenever there is room (make_room first)
    auto push = [&](bool hit, unsigned e) {
        unsigned short *const w = slot(tail);
        if (hit) *w = (unsigned short)e;
        tail += hit ? STEP : 0u;
    };
    // generations for the candidates [lo, lo + nch 2^OBITS) of a row: the
    // first one returned; table slots are free once every entry older than
    // GTAB generations is drained
    auto new_gens = [&](int lo, int nch) {
        if (gen + nch - gfull > GTAB) drain(PassTag<true>{});
        __builtin_amdgcn_wave_barrier();
        const int k = tid & 63;
        if (k < nch) gt[(gen + k) & (GTAB - 1)] = lo + (k << OBITS);
        __builtin_amdgcn_wave_barrier();
        const int g0 = gen;
        gen += nch;
        return g0;
    };

    // ---- bulk pass: rows 0-8; only the particle's own row (4) can hold i
    float4 *sb = sbuf[tid >> 6];
    auto bulk_row = [&](int row, auto self) {
        constexpr bool SELF = decltype(self)::value;
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        const int yy = cy + dy, zz = cz + dz;
        int j0 = 0, n0 = 0;
        if (active && yy >= 0 && yy < Ly && zz >= 0 && zz < Lz) {
            const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < Lx ? cx + 1 : Lx - 1;
            run_range(P, Lx * (yy + Ly * zz), x0, x1, pi.x, row_halfwidth(dy, dz, ry, rz), &j0, &n0);
        }
        // the union of the wave's candidate ranges of this row
        int lo = n0 > 0 ? j0 : 0x7fffffff, hi = n0 > 0 ? j0 + n0 : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
        }
        if (hi <= lo) return;
        // candidate k of the lane's range is i itself (own row only)
        const int kself = SELF ? i - j0 : -1;
        if (hi - lo <= SC) {
            // one generation: entry = g << OBITS | (j - lo)
            const unsigned e0 = ((unsigned)(new_gens(lo, 1) & (GTAB - 1)) << OBITS) + (unsigned)(n0 > 0 ? j0 - lo : 0);
            __builtin_amdgcn_wave_barrier();
            for (int k = (tid & 63); k < hi - lo; k += 64) sb[k] = P.p(lo + k);
            __builtin_amdgcn_wave_barrier();
            const float4 *sp = sb + (n0 > 0 ? j0 - lo : 0);
            for (int t = 0; wave_any(t < n0); t += 4) {
                make_room(PassTag<true>{}, 4u);
                const float4 c0 = sp[t], c1 = sp[t + 1], c2 = sp[t + 2], c3 = sp[t + 3];
#define UDX_TEST(C, U)                                                                             \
    {                                                                                              \
        const float ddx = pi.x - C.x, ddy = pi.y - C.y, ddz = pi.z - C.z;                          \
        const float r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddx, ddx, ddy * ddy));            \
        push(t + U < n0 && (!SELF || t + U != kself) && r2 < 1.0f, e0 + (unsigned)(t + U));        \
    }
                UDX_TEST(c0, 0)
                UDX_TEST(c1, 1)
                UDX_TEST(c2, 2)
                UDX_TEST(c3, 3)
#undef UDX_TEST
            }
            __builtin_amdgcn_wave_barrier();
        } else {
            // the union read from global memory in pieces of GTAB generations
            // (one per 2^OBITS candidates; a piece takes only its candidates)
            const int nch = ((hi - 1 - lo) >> OBITS) + 1;
            for (int c0 = 0; c0 < nch; c0 += GTAB) {
                const int pl = lo + (c0 << OBITS);
                const int np = nch - c0 < GTAB ? nch - c0 : GTAB;
                const int ph = pl + (np << OBITS);
                const unsigned g0 = (unsigned)new_gens(pl, np);
                auto ent = [&](int j) {
                    const int o = j - pl;
                    return (((g0 + (unsigned)(o >> OBITS)) & (GTAB - 1)) << OBITS) | (unsigned)(o & OMASK);
                };
                for (int t = 0; wave_any(t < n0); t += 2) {
                    make_room(PassTag<true>{}, 2u);
                    const int ja = j0 + t, jb = j0 + t + 1;
                    const bool oka = t < n0 && ja >= pl && ja < ph, okb = t + 1 < n0 && jb >= pl && jb < ph;
                    const float4 pa = P.p(oka ? ja : P.i0), pb = P.p(okb ? jb : P.i0);
                    const float dxa = pi.x - pa.x, dya = pi.y - pa.y, dza = pi.z - pa.z;
                    const float dxb = pi.x - pb.x, dyb = pi.y - pb.y, dzb = pi.z - pb.z;
                    const float r2a = __builtin_fmaf(dza, dza, __builtin_fmaf(dxa, dxa, dya * dya));
                    const float r2b = __builtin_fmaf(dzb, dzb, __builtin_fmaf(dxb, dxb, dyb * dyb));
                    push(oka && (!SELF || t != kself) && r2a < 1.0f, ent(ja));
                    push(okb && (!SELF || t + 1 != kself) && r2b < 1.0f, ent(jb));
                }
            }
        }
    };
    if (PHASE != PH_HALO) {
        for (int row = 0; row < 4; row++) bulk_row(row, PassTag<false>{});
        bulk_row(4, PassTag<true>{});
        for (int row = 5; row < 9; row++) bulk_row(row, PassTag<false>{});
        drain(PassTag<true>{});
    }

    // ---- halo pass: the cells across the periodic boundary.  The noise
    // indices of a wrapped pair are positions in packed fragment lists: the
    // home particle's in fragment w is one value per run, and the candidate's
    // in the opposite fragment is B + j with B one value per run (consecutive
    // cells of a row are consecutive in a fragment's list), so both are set at
    // the push and queued with the entry (three rings of HQC slots per lane in
    // the memory of the bulk queue and the staging buffer).
    const bool edge =
        active && P.periodic && (cx == 0 || cx == Lx - 1 || cy == 0 || cy == Ly - 1 || cz == 0 || cz == Lz - 1);
    bool listed = false;
    if (PHASE == PH_BULK) {
        // hand the particles with wrapped pairs to the edge list, one slot
        // range per wave (their order in the list does not matter)
        const unsigned long long em = __ballot(edge);
        int base = 0;
        if ((tid & 63) == 0 && em) base = atomicAdd(A.ecount, __popcll(em));
        base = __shfl(base, 0, 64);
        if (edge) {
            const int k = P.i0 + base +
                          (int)__builtin_amdgcn_mbcnt_hi((unsigned)(em >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)em, 0u));
            A.eidx[k] = i;
            A.ef[k] = make_float4(fx, fy, fz, 0.f);
        }
        listed = edge;
    }
    if (PHASE != PH_BULK && wave_any(edge)) {
        constexpr int HQC = QC / 4 < 8 ? QC / 4 : 8;
        static_assert(2 * HQC * 256 <= QC * 128 && HQC * 256 <= (SC + 4) * 16,
                      "halo rings fit the wave's queue region and staging buffer");
        constexpr unsigned HSTEP = 4u * 64u, HMASK = (unsigned)(HQC - 1) * HSTEP;
        // rings j|w and u in the wave's queue region, ring v in the wave's
        // staging buffer (slot k of every lane of the wave: HSTEP bytes)
        unsigned *const rj = (unsigned *)qw, *const ru = (unsigned *)(qw + HQC * HSTEP);
        char *const rv = (char *)&sbuf[tid >> 6][0];
        const unsigned hlane = (unsigned)((tid & 63) * 4);
        auto hs = [&](unsigned *base, unsigned k) { return (unsigned *)((char *)base + ((k & HMASK) | hlane)); };
        auto hv = [&](unsigned k) { return (unsigned *)(rv + ((k & HMASK) | hlane)); };
        unsigned hh = 0, ht = 0;
        auto hstep = [&]() {
            const unsigned fl = ht - hh;
            const bool va = fl > 0u, vb = fl > HSTEP;
            const unsigned ea = va ? *hs(rj, hh) : idle, eb = vb ? *hs(rj, hh + HSTEP) : idle;
            const unsigned ua = va ? *hs(ru, hh) : 0u, ub = vb ? *hs(ru, hh + HSTEP) : 0u;
            const unsigned xa_ = va ? *hv(hh) : 0u, xb_ = vb ? *hv(hh + HSTEP) : 0u;
            const PairIn a = pair_in_wrapped(P, ea, (int)ua, (int)xa_, pi, vi);
            const PairIn b = pair_in_wrapped(P, eb, (int)ub, (int)xb_, pi, vi);
            float xa = a.x, xb = b.x;
            logistic18x2(&xa, &xb);
            const float3 f1 = dpd_pair<WR, EXACT>(P.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
            const float3 f2 = dpd_pair<WR, EXACT>(P.par, b.dx, b.dy, b.dz, b.r2, b.dvx, b.dvy, b.dvz, xb * R_SQRT2);
            add(va, f1, a);
            add(vb, f2, b);
            const unsigned took = fl < 2u * HSTEP ? fl : 2u * HSTEP;
            hh += took;
            npair += (int)(took / HSTEP);
        };
        auto hpush = [&](bool hit, unsigned e, int u, int v) {
            if (hit) {
                *hs(rj, ht) = e;
                *hs(ru, ht) = (unsigned)u;
                *hv(ht) = (unsigned)v;
            }
            ht += hit ? HSTEP : 0u;
        };
        __builtin_amdgcn_wave_barrier();
        for (int row = 0; row < 9; row++) {
            const int dz = row / 3 - 1, dy = row % 3 - 1;
            int yy = cy + dy, zz = cz + dz;
            const int wy = yy < 0 ? -1 : (yy >= Ly ? 1 : 0), wz = zz < 0 ? -1 : (zz >= Lz ? 1 : 0);
            yy -= wy * Ly;
            zz -= wz * Lz;
            const bool yzwrap = (wy | wz) != 0;
            int j0 = 0, n0 = 0, j1 = 0, n1 = 0, wx0 = 0, wx1 = 0, x0, x1;
            // per run: the home particle's list position aid, the candidates' B,
            // and which of the two is the first noise index (mask)
            int aid0 = 0, b0 = 0, aid1 = 0, b1 = 0;
            bool m0 = false, m1 = false;
            auto runtab = [&](int wx, int xs, int *aid, int *bb, bool *mk) {
                const int fid = frag_d2i(wx, wy, wz), anti = frag_d2i(-wx, -wy, -wz);
                const int fo = A.foff[fid], ao = A.foff[anti];
                *aid = P.fss[fo + frag_cell(wx, wy, wz, cx, cy, cz, Lx, Ly)] - P.fss[fo] + (i - istart);
                *bb = P.fss[ao + frag_cell(-wx, -wy, -wz, xs, yy, zz, Lx, Ly)] - P.fss[ao] -
                      P.starts[xs + Lx * (yy + Ly * zz)];
                *mk = fid < anti;
            };
            if (edge) {
                const float hw = row_halfwidth(dy, dz, ry, rz);
                const int base = Lx * (yy + Ly * zz);
                if (xrun(0, cx, Lx, &x0, &x1, &wx0) && (yzwrap || wx0 != 0)) {
                    run_range(P, base, x0, x1, pi.x - (float)(wx0 * Lx), hw, &j0, &n0);
                    if (n0 > 0) runtab(wx0, x0, &aid0, &b0, &m0);
                }
                if (xrun(1, cx, Lx, &x0, &x1, &wx1) && (yzwrap || wx1 != 0)) {
                    run_range(P, base, x0, x1, pi.x - (float)(wx1 * Lx), hw, &j1, &n1);
                    if (n1 > 0) runtab(wx1, x0, &aid1, &b1, &m1);
                }
            }
            const float ay = pi.y - (float)(wy * Ly), az = pi.z - (float)(wz * Lz);
            const float ax0 = pi.x - (float)(wx0 * Lx), ax1 = pi.x - (float)(wx1 * Lx);
            const unsigned w0 = wcode(wx0, wy, wz) << 27, w1 = wcode(wx1, wy, wz) << 27;
            const int len = n0 + n1;
            for (int t = 0; wave_any(t < len); t += 2) {
                if (wave_any(ht - hh + 2u * HSTEP > (unsigned)HQC * HSTEP))
#pragma unroll 1
                    for (int k = 0; k < DRAIN_STEPS; k++) hstep();
                const int ta = t, tb = t + 1;
                const bool oka = ta < len, okb = tb < len;
                const bool sa = ta >= n0, sb = tb >= n0;
                const int ja = sa ? j1 + (ta - n0) : j0 + ta;
                const int jb = sb ? j1 + (tb - n0) : j0 + tb;
                const float4 pa = P.p(oka ? ja : P.i0);
                const float4 pb = P.p(okb ? jb : P.i0);
                const float axa = sa ? ax1 : ax0, axb = sb ? ax1 : ax0;
                const float dxa = axa - pa.x, dya = ay - pa.y, dza = az - pa.z;
                const float dxb = axb - pb.x, dyb = ay - pb.y, dzb = az - pb.z;
                const float r2a = __builtin_fmaf(dza, dza, __builtin_fmaf(dxa, dxa, dya * dya));
                const float r2b = __builtin_fmaf(dzb, dzb, __builtin_fmaf(dxb, dxb, dyb * dyb));
                const int aa = sa ? aid1 : aid0, ba = (sa ? b1 : b0) + ja;
                const int ab = sb ? aid1 : aid0, bbb = (sb ? b1 : b0) + jb;
                const bool ma = sa ? m1 : m0, mb = sb ? m1 : m0;
                hpush(oka && r2a < 1.0f, (unsigned)ja | (sa ? w1 : w0), ma ? aa : ba, ma ? ba : aa);
                hpush(okb && r2b < 1.0f, (unsigned)jb | (sb ? w1 : w0), mb ? ab : bbb, mb ? bbb : ab);
            }
        }
        while (wave_any(ht != hh)) hstep();
        __builtin_amdgcn_wave_barrier();
    }

    // ---- wall pass (MOVE == 2): the particles that are not sdf_far and the
    // frozen wall particles in the 27 cells of the wall cell list around
    // them, as k_wall_force evaluates them (candidate order, noise
    // mean0var1ii(seed_wall, i, wall index), wall velocity at the wall
    // particle); summed apart and added after the solvent sum.  The ring is
    // the wave's queue region (its bulk and halo queues are empty here).
    float wsx = 0.f, wsy = 0.f, wsz = 0.f;
    bool wnear = false;
    if (MOVE == 2) {
        wnear = active && !sdf_far(A.sdf, pi.x, pi.y, pi.z);
        if (A.nwall > 0 && wave_any(wnear)) {
            constexpr unsigned WQ = (unsigned)QC / 2u;   // 32-bit slots per lane in QC * 128 bytes
            const int Wx = A.Wx, Wy = A.Wy, Wz = A.Wz;
            const UDX_GLOBAL f4v *const wp = (const UDX_GLOBAL f4v *)A.wpos;
            const UDX_GLOBAL int *const ws = gptr(A.wstarts);
            int b0x = 1, b0y = 1, b0z = 1;
            if (wnear) {
                auto clampc = [](float x, int W) {
                    const int qq = (int)__builtin_floor((double)x + 0.5 * (double)W);
                    return qq < 1 ? 1 : (qq > W - 2 ? W - 2 : qq);
                };
                b0x = clampc(pi.x, Wx);
                b0y = clampc(pi.y, Wy);
                b0z = clampc(pi.z, Wz);
            }
            const float wseed = *A.wseed;
            const float fa_lo = (float)(i & 0x3FF), fa_hi = (float)i;
            const unsigned wlane = (unsigned)((tid & 63) * 4);
            auto wslot = [&](unsigned k) { return (unsigned *)(qw + (((k & (WQ - 1u)) << 8) | wlane)); };
            unsigned wh = 0, wt = 0;
            int nw = 0;
            auto wstep = [&]() {
                const unsigned fl = wt - wh;
                const bool ok0 = fl > 0u, ok1 = fl > 1u;
                const unsigned e0 = ok0 ? *wslot(wh) : 0u;   // wall particle 0 exists (nwall > 0)
                const unsigned e1 = ok1 ? *wslot(wh + 1u) : e0;
                const f4v w0 = wp[e0], w1 = wp[e1];
                const float dx0 = pi.x - w0.x, dy0 = pi.y - w0.y, dz0 = pi.z - w0.z;
                const float dx1 = pi.x - w1.x, dy1 = pi.y - w1.y, dz1 = pi.z - w1.z;
                const float r20 = __builtin_fmaf(dz0, dz0, __builtin_fmaf(dx0, dx0, dy0 * dy0));
                const float r21 = __builtin_fmaf(dz1, dz1, __builtin_fmaf(dx1, dx1, dy1 * dy1));
                // mean0var1ii(seed, i, wall) (src/math/rnd/dev.h:57-63)
                float x0 = wseed - low_discrepancy(fa_lo, fa_hi, (float)((int)e0 & 0x3FF), (float)(int)e0);
                float x1 = wseed - low_discrepancy(fa_lo, fa_hi, (float)((int)e1 & 0x3FF), (float)(int)e1);
                logistic18x2(&x0, &x1);
                float u0[3], u1[3];
                wvel_at(A.wv, w0.x, w0.y, w0.z, u0);
                wvel_at(A.wv, w1.x, w1.y, w1.z, u1);
                const float3 f0 = dpd_pair<WR, EXACT>(P.par, dx0, dy0, dz0, r20, vi.x - u0[0], vi.y - u0[1],
                                                      vi.z - u0[2], x0 * R_SQRT2);
                const float3 f1 = dpd_pair<WR, EXACT>(P.par, dx1, dy1, dz1, r21, vi.x - u1[0], vi.y - u1[1],
                                                      vi.z - u1[2], x1 * R_SQRT2);
                if (ok0) {
                    wsx += f0.x;
                    wsy += f0.y;
                    wsz += f0.z;
                }
                if (ok1) {
                    wsx += f1.x;
                    wsy += f1.y;
                    wsz += f1.z;
                }
                const unsigned took = fl < 2u ? fl : 2u;
                wh += took;
                nw += (int)took;
            };
            __builtin_amdgcn_wave_barrier();
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++) {
                    int c0 = 0, len = 0;
                    if (wnear) {
                        const int c = (b0x - 1) + Wx * ((b0y + dy) + Wy * (b0z + dz));
                        c0 = ws[c];
                        len = ws[c + 3] - c0;
                    }
                    for (int t = 0; wave_any(t < len); t += 4) {
                        if (wave_any(wt - wh + 4u > WQ))
#pragma unroll 1
                            for (int k = 0; k < DRAIN_STEPS; k++) wstep();
                        f4v pw[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) pw[u] = t + u < len ? wp[c0 + t + u] : f4v{1e30f, 1e30f, 1e30f, 0.f};
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const float dx = pi.x - pw[u].x, dyy = pi.y - pw[u].y, dzz = pi.z - pw[u].z;
                            const float r2 = __builtin_fmaf(dzz, dzz, __builtin_fmaf(dx, dx, dyy * dyy));
                            const bool hit = t + u < len && r2 < 1.0f;
                            if (hit) *wslot(wt) = (unsigned)(c0 + t + u);
                            wt += hit ? 1u : 0u;
                        }
                    }
                }
            while (wave_any(wt != wh)) wstep();
            __builtin_amdgcn_wave_barrier();
            if (A.wpairs) {
                unsigned long long c = (unsigned long long)nw;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
                if ((tid & 63) == 0 && c) atomicAdd(A.wpairs, c);
            }
        }
    }

    if (MOVE || A.opos) {
        // the step's total force (the wall sum added for the particles near a
        // wall: 0 + sum, then + wall, as ff = sum and ff += wall), the
        // kick-drift, the bounce-back and the wrap of move_f4_wall, then the
        // count into the next step's x slabs (MOVE: the single-box kernel;
        // A.opos without MOVE: a batch without walls, where a particle
        // listed for the halo launch is moved there)
        const bool mine = active && !listed;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (mine) {
            UDX_GLOBAL f4v *f4 = (UDX_GLOBAL f4v *)A.ff + i;
            float gx = 0.f + fx, gy = 0.f + fy, gz = 0.f + fz;
            if (MOVE == 2 && wnear) {
                gx = gx + wsx;
                gy = gy + wsy;
                gz = gz + wsz;
            }
            *f4 = f4v{gx, gy, gz, 0.f};
            p = pi;
            float4 v = vi;
            kick_drift(A.mdt, A.mmass, A.max_, gx, &p.x, &v.x);
            kick_drift(A.mdt, A.mmass, A.may, gy, &p.y, &v.y);
            kick_drift(A.mdt, A.mmass, A.maz, gz, &p.z, &v.z);
            if (MOVE == 2) {
                float r[3] = {p.x, p.y, p.z}, u[3] = {v.x, v.y, v.z};
                if (sdf_bounce(A.sdf, A.wv, A.mdt, r, u)) {
                    p.x = r[0]; p.y = r[1]; p.z = r[2];
                    v.x = u[0]; v.y = u[1]; v.z = u[2];
                }
            }
            p.x = wrap_coord(p.x, (float)Lx);
            p.y = wrap_coord(p.y, (float)Ly);
            p.z = wrap_coord(p.z, (float)Lz);
            *((UDX_GLOBAL f4v *)A.opos + i) = f4v{p.x, p.y, p.z, 0.f};
            *((UDX_GLOBAL f4v *)A.ovel + i) = f4v{v.x, v.y, v.z, 0.f};
        }
        clist_count_one(mine, i, make_float3(p.x, p.y, p.z), A.sc.L[0], A.sc.L[1], A.sc.L[2], A.sc.sub,
                        A.sc.counts, A.sc.cellslot, A.sc.slab0);
    } else if (active && !listed) {
        if (A.ff_stride == 4) {
            UDX_GLOBAL f4v *f4 = (UDX_GLOBAL f4v *)A.ff + i;
            f4v o = {0.f, 0.f, 0.f, 0.f};
            if (A.accumulate) o = *f4;
            *f4 = f4v{o.x + fx, o.y + fy, o.z + fz, 0.f};
        } else {
            UDX_GLOBAL float *f3 = gptr(A.ff) + 3 * (long)i;
            if (A.accumulate) {
                f3[0] += fx; f3[1] += fy; f3[2] += fz;
            } else {
                f3[0] = fx; f3[1] = fy; f3[2] = fz;
            }
        }
        if (STRESS) {
            UDX_GLOBAL float *s6 = gptr(A.ss) + 6 * (long)i;
            s6[0] += sxx; s6[1] += sxy; s6[2] += sxz;
            s6[3] += syy; s6[4] += syz; s6[5] += szz;
        }
    }
    if (A.pairs) {
        unsigned long long c = (unsigned long long)npair;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if ((tid & 63) == 0) lpairs[tid >> 6] = c;
        __syncthreads();
        if (tid == 0) {
            unsigned long long s = 0;
            for (int w = 0; w < NT / 64; w++) s += lpairs[w];
            if (s) atomicAdd(A.pairs, s);
        }
    }
}

// at most 102 VGPRs: five waves per SIMD (measured 10% faster on the 96^3 box
// than four waves at 104 VGPRs; six spill).  Blocks in launch order: the
// hardware deals them to the XCDs round robin, which spreads the waves near
// the walls (the first and last planes of a channel) over every XCD; giving
// each XCD a contiguous eighth of the cells fetched fewer bytes but left the
// wall work on two XCDs (1% slower with the wall pass fused).
template <int WR, bool EXACT, int PHASE, int MOVE = 0>
__global__ __launch_bounds__(DPD_THREADS) __attribute__((amdgpu_waves_per_eu(5))) void k_dpd(DpdK A) {
    dpd_block<WR, EXACT, DPD_THREADS, QCAP, SCAP, false, PHASE, MOVE>(A, blockIdx.x * DPD_THREADS);
}

// with the per-particle stress (the reference's FoSArray path,
// src/fluforces/bulk/imp/main.h:42-53)
template <int WR>
__global__ __launch_bounds__(DPD_THREADS) void k_dpd_stress(DpdK A) {
    dpd_block<WR, true, DPD_THREADS, QCAP, SCAP, true>(A, blockIdx.x * DPD_THREADS);
}

// Many boxes in one launch: block b belongs to the box k with
// bblock[k] <= b < bblock[k + 1]; each box's blocks start at its first
// particle, so no block spans two boxes.
// (five waves per SIMD: the 16-bit queues of every tile variant leave room)
template <int WR, bool EXACT, int NT, int QC, int PHASE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_dpd_batch(
    const DpdK *boxes, const int *bblock, int nbox, int *step) {
    // the fused moves' step counter (the step's seeds were read before it)
    if (PHASE != PH_HALO && step && blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
    // blocks in launch order (the hardware deals them to the XCDs round
    // robin): boxes of different densities given to the XCDs in contiguous
    // eighths left some XCDs with more work (8% slower on 64 boxes)
    const int blk = blockIdx.x;
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bblock[mid] <= blk) lo = mid;
        else hi = mid;
    }
    const DpdK &A = boxes[lo];
    dpd_block<WR, EXACT, NT, QC, SCAP_BATCH, false, PHASE>(A, A.i0 + (blk - bblock[lo]) * NT);
}

// ---------------------------------------------------------------- half shell
// The production pass (UDX_MODE_FAST with FP64 accumulators): each pair
// inside the box is evaluated once, by its higher-indexed particle, as the
// reference's bulk kernel does (src/fluforces/bulk/dev/main.h:23-97), and
// its force is added to one particle and subtracted from the other.
//
// Sums are exact: every pair force component is rounded to a multiple of
// 2^-32 (nothing changes for |f| >= 2^-9, whose float ulp is at least that)
// and added in FP64, where sums of such multiples stay exact while their
// magnitude is below 2^21.  The sum is therefore independent of the order of
// the additions (lanes, LDS and global atomics), the same from run to run,
// and the float force is that sum rounded once.  The pair force itself is
// antisymmetric bit for bit (negated differences give negated results), so
// which particle of a pair evaluates it does not matter either.
//
// A wave owns 64 consecutive cell-ordered particles.  The half shell of a
// particle is the part of its full shell with lower indices: the rows
// (dz, dy) = (-1,-1), (-1,0), (-1,1), (0,-1) and its own row (0,0) up to
// itself.  For each row the union of the wave's candidate runs is staged in
// LDS (positions, velocities and FP64 reaction accumulators, in chunks of
// HS_SC).  Candidates that pass the cutoff go to one queue per wave (slots
// from ballots); whenever 128 are queued every lane evaluates two of them,
// so the pair body runs with all lanes busy however unevenly the lanes hit.
// The force goes to the home particle's LDS accumulator and, negated, to the
// staged particle's; at the end of a chunk the staged reactions are added to
// the global accumulators, at the end of the wave the home sums.  Pairs
// across the periodic boundary keep the reference's one-sided halo noise
// (the full-shell gather's halo pass) and add to the home sum only.
constexpr int HS_SC = 128;    // staged candidates per wave and chunk of a row
constexpr int HS_QN = 512;    // wave queue ring: entries owner lane | slot << 6
constexpr int HS_HQC = 16;    // halo pass: queued pairs per lane (a power of two)

__device__ __forceinline__ double quant32(float f) {
    return (double)(__builtin_rintf(f * 4294967296.0f) * 2.3283064365386963e-10f);
}

struct HsWave {
    float4 hp[64], hv[64];            // home particles, read by the lane that evaluates a pair
    double hacc[3][64];               // home sums
    union {
        struct {
            float4 sp[HS_SC + 4];     // staged positions (4 reads past a lane's run)
            float4 sv[HS_SC];         // staged velocities
        } st;
        unsigned ring[HS_HQC * 64];   // halo pass: per-lane ring queues
    } u;
    double racc[3][HS_SC];            // reactions on the staged particles
    unsigned short q[HS_QN];
};

template <class T>
__device__ __forceinline__ void lds_add(T *p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ void gacc_add(UDX_GLOBAL double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int WR, int NT>
__device__ __forceinline__ void dpd_half_block(const DpdK &A, int first, HsWave *lds) {
    const int tid = threadIdx.x, lane = tid & 63;
    HsWave &W = lds[tid >> 6];
    const DpdP P = dpd_view(A);
    const int Lx = P.Lx, Ly = P.Ly, Lz = P.Lz;
    const int n = P.starts[Lx * Ly * Lz];
    const int wfirst = first + (tid & ~63);
    const int i = first + tid;
    const bool active = i < n && i - P.i0 < P.nmax;
    UDX_GLOBAL double *const gacc = (UDX_GLOBAL double *)A.acc;

    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi;
    int cx = 0, cy = 0, cz = 0, istart = 0;
    float ry = 0.f, rz = 0.f;
    if (active) {
        pi = P.p(i);
        vi = P.v(i);
        cell_coord(pi.x, Lx, &cx);
        cell_coord(pi.y, Ly, &cy);
        cell_coord(pi.z, Lz, &cz);
        istart = P.starts[cx + Lx * (cy + Ly * cz)];
        ry = (float)((double)pi.y + 0.5 * (double)Ly - (double)cy);
        rz = (float)((double)pi.z + 0.5 * (double)Lz - (double)cz);
    }
    W.hp[lane] = pi;
    W.hv[lane] = vi;
    W.hacc[0][lane] = W.hacc[1][lane] = W.hacc[2][lane] = 0.0;
    const float seed_bulk = P.seeds[26];
    const int ubase = wfirst - P.i0;   // noise index of the wave's first particle
    constexpr unsigned QM = HS_QN - 1;
    unsigned head = 0, tail = 0;       // wave-uniform queue counters
    unsigned bulk = 0;                 // pairs evaluated in the bulk pass

    // k = 1 or 2 queued pairs per lane; c0 is the chunk's first particle
    auto batch = [&](auto two, int c0) {
        constexpr bool TWO = decltype(two)::value;
        const unsigned avail = tail - head;
        const unsigned e0 = W.q[head & QM];
        const bool va = (unsigned)lane < avail, vb = TWO && (unsigned)lane + 64u < avail;
        const unsigned ea = va ? W.q[(head + lane) & QM] : e0;
        const unsigned eb = vb ? W.q[(head + 64u + lane) & QM] : e0;
        auto in = [&](unsigned e, PairIn &o) {
            const int ow = (int)(e & 63u), k = (int)(e >> 6);
            const float4 a = W.hp[ow], av = W.hv[ow], b = W.u.st.sp[k], bv = W.u.st.sv[k];
            pair_kin(o, a.x, a.y, a.z, av, b, bv);
            const int u = ubase + ow, v = c0 + k - P.i0;   // u > v: the home particle has the higher index
            o.x = seed_bulk - low_discrepancy((float)(u & 0x3FF), (float)u, (float)(v & 0x3FF), (float)v);
        };
        auto out = [&](bool ok, unsigned e, const float3 &f) {
            if (ok) {
                const int ow = (int)(e & 63u), k = (int)(e >> 6);
                const double qx = quant32(f.x), qy = quant32(f.y), qz = quant32(f.z);
                lds_add(&W.hacc[0][ow], qx);
                lds_add(&W.hacc[1][ow], qy);
                lds_add(&W.hacc[2][ow], qz);
                lds_add(&W.racc[0][k], -qx);
                lds_add(&W.racc[1][k], -qy);
                lds_add(&W.racc[2][k], -qz);
            }
        };
        PairIn a, b;
        in(ea, a);
        if (TWO) {
            in(eb, b);
            float xa = a.x, xb = b.x;
            logistic18x2(&xa, &xb);
            const float3 fa = dpd_pair<WR, false>(P.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
            const float3 fb = dpd_pair<WR, false>(P.par, b.dx, b.dy, b.dz, b.r2, b.dvx, b.dvy, b.dvz, xb * R_SQRT2);
            out(va, ea, fa);
            out(vb, eb, fb);
        } else {
            const float xa = logistic18(a.x);
            const float3 fa = dpd_pair<WR, false>(P.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
            out(va, ea, fa);
        }
        const unsigned took = avail < (TWO ? 128u : 64u) ? avail : (TWO ? 128u : 64u);
        head += took;
        bulk += took;
    };

    // ---- bulk pass over the half shell
    for (int row = 0; row < 5; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        const int yy = cy + dy, zz = cz + dz;
        int j0 = 0, n0 = 0;
        if (active && yy >= 0 && yy < Ly && zz >= 0) {
            const int x0 = cx > 0 ? cx - 1 : 0;
            const int x1 = row == 4 ? cx : (cx + 1 < Lx ? cx + 1 : Lx - 1);
            run_range(P, Lx * (yy + Ly * zz), x0, x1, pi.x, row_halfwidth(dy, dz, ry, rz), &j0, &n0);
            if (row == 4) n0 = max(0, min(n0, i - j0));   // own row: lower indices only
        }
        int lo = n0 > 0 ? j0 : 0x7fffffff, hi = n0 > 0 ? j0 + n0 : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
        }
        for (int c0 = lo; c0 < hi; c0 += HS_SC) {
            const int len = min(HS_SC, hi - c0);
            __builtin_amdgcn_wave_barrier();
            for (int k = lane; k < len; k += 64) {
                const f4v p = P.pos[c0 + k], v = P.vel[c0 + k];
                W.u.st.sp[k] = make_float4(p.x, p.y, p.z, 0.f);
                W.u.st.sv[k] = make_float4(v.x, v.y, v.z, 0.f);
                W.racc[0][k] = W.racc[1][k] = W.racc[2][k] = 0.0;
            }
            __builtin_amdgcn_wave_barrier();
            // this lane's candidates in the chunk: slots [s0, s0 + m)
            const int ja = max(j0, c0), jb = min(j0 + n0, c0 + len);
            const int m = n0 > 0 && jb > ja ? jb - ja : 0;
            const int s0 = m > 0 ? ja - c0 : 0;
            const float4 *sp = W.u.st.sp + s0;
            for (int t = 0; wave_any(t < m); t += 4) {
                const float4 c[4] = {sp[t], sp[t + 1], sp[t + 2], sp[t + 3]};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float ddx = pi.x - c[u].x, ddy = pi.y - c[u].y, ddz = pi.z - c[u].z;
                    const float r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddx, ddx, ddy * ddy));
                    const bool hit = t + u < m && r2 < 1.0f;
                    const unsigned long long mk = __ballot(hit);
                    const unsigned pos =
                        tail + __builtin_amdgcn_mbcnt_hi((unsigned)(mk >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mk, 0u));
                    if (hit) W.q[pos & QM] = (unsigned short)(lane | ((s0 + t + u) << 6));
                    tail += (unsigned)__popcll(mk);
                }
                while (tail - head >= 128u) batch(PassTag<true>{}, c0);
            }
            if (tail - head > 64u) batch(PassTag<true>{}, c0);
            else if (tail != head) batch(PassTag<false>{}, c0);
            __builtin_amdgcn_wave_barrier();
            // the chunk's reactions into the global sums
            for (int k = lane; k < len; k += 64) {
                const double rx = W.racc[0][k], rY = W.racc[1][k], rZ = W.racc[2][k];
                if (rx != 0.0 || rY != 0.0 || rZ != 0.0) {
                    UDX_GLOBAL double *g = gacc + 3 * (long)(c0 + k);
                    gacc_add(g, rx);
                    gacc_add(g + 1, rY);
                    gacc_add(g + 2, rZ);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }

    // ---- halo pass: pairs across the periodic boundary, one-sided (the
    // full-shell gather's halo pass with the sums exact)
    double hx = 0.0, hy = 0.0, hz = 0.0;
    int nhalo = 0;
    const bool edge =
        active && P.periodic && (cx == 0 || cx == Lx - 1 || cy == 0 || cy == Ly - 1 || cz == 0 || cz == Lz - 1);
    if (wave_any(edge)) {
        constexpr unsigned STEP = 4u * 64u, RMASK = (unsigned)(HS_HQC - 1) * STEP;
        unsigned rh = 0, rt = 0;
        const unsigned qlane = (unsigned)(lane * 4);
        auto slot = [&](unsigned k) { return (unsigned *)((char *)W.u.ring + ((k & RMASK) | qlane)); };
        const unsigned idle = (unsigned)(active ? i : P.i0) | (W_NONE << 27);
        auto step = [&]() {
            const unsigned fl = rt - rh;
            const bool va = fl > 0u, vb = fl > STEP;
            const unsigned ea = va ? *slot(rh) : idle;
            const unsigned eb = vb ? *slot(rh + STEP) : idle;
            const PairIn a = pair_in_halo(A, P, ea, i, pi, vi, cx, cy, cz, istart, seed_bulk);
            const PairIn b = pair_in_halo(A, P, eb, i, pi, vi, cx, cy, cz, istart, seed_bulk);
            float xa = a.x, xb = b.x;
            logistic18x2(&xa, &xb);
            const float3 fa = dpd_pair<WR, false>(P.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
            const float3 fb = dpd_pair<WR, false>(P.par, b.dx, b.dy, b.dz, b.r2, b.dvx, b.dvy, b.dvz, xb * R_SQRT2);
            if (va) {
                hx += quant32(fa.x);
                hy += quant32(fa.y);
                hz += quant32(fa.z);
            }
            if (vb) {
                hx += quant32(fb.x);
                hy += quant32(fb.y);
                hz += quant32(fb.z);
            }
            const unsigned took = fl < 2u * STEP ? fl : 2u * STEP;
            rh += took;
            nhalo += (int)(took / STEP);
        };
        auto push = [&](bool hit, unsigned e) {
            unsigned *const w = slot(rt);
            if (hit) *w = e;
            rt += hit ? STEP : 0u;
        };
        __builtin_amdgcn_wave_barrier();
        for (int row = 0; row < 9; row++) {
            const int dz = row / 3 - 1, dy = row % 3 - 1;
            int yy = cy + dy, zz = cz + dz;
            const int wy = yy < 0 ? -1 : (yy >= Ly ? 1 : 0), wz = zz < 0 ? -1 : (zz >= Lz ? 1 : 0);
            yy -= wy * Ly;
            zz -= wz * Lz;
            const bool yzwrap = (wy | wz) != 0;
            int j0 = 0, n0 = 0, j1 = 0, n1 = 0, wx0 = 0, wx1 = 0, x0, x1;
            if (edge) {
                const float hw = row_halfwidth(dy, dz, ry, rz);
                const int base = Lx * (yy + Ly * zz);
                if (xrun(0, cx, Lx, &x0, &x1, &wx0) && (yzwrap || wx0 != 0))
                    run_range(P, base, x0, x1, pi.x - (float)(wx0 * Lx), hw, &j0, &n0);
                if (xrun(1, cx, Lx, &x0, &x1, &wx1) && (yzwrap || wx1 != 0))
                    run_range(P, base, x0, x1, pi.x - (float)(wx1 * Lx), hw, &j1, &n1);
            }
            const float ay = pi.y - (float)(wy * Ly), az = pi.z - (float)(wz * Lz);
            const float ax0 = pi.x - (float)(wx0 * Lx), ax1 = pi.x - (float)(wx1 * Lx);
            const unsigned w0 = wcode(wx0, wy, wz) << 27, w1 = wcode(wx1, wy, wz) << 27;
            const int len = n0 + n1;
            for (int t = 0; wave_any(t < len); t += 2) {
                if (wave_any(rt - rh + 2u * STEP > (unsigned)HS_HQC * STEP))
#pragma unroll 1
                    for (int k = 0; k < DRAIN_STEPS; k++) step();
                const int ta = t, tb = t + 1;
                const bool oka = ta < len, okb = tb < len;
                const int ja = ta >= n0 ? j1 + (ta - n0) : j0 + ta;
                const int jb = tb >= n0 ? j1 + (tb - n0) : j0 + tb;
                const float4 pa = P.p(oka ? ja : P.i0);
                const float4 pb = P.p(okb ? jb : P.i0);
                const float axa = ta >= n0 ? ax1 : ax0, axb = tb >= n0 ? ax1 : ax0;
                const float dxa = axa - pa.x, dya = ay - pa.y, dza = az - pa.z;
                const float dxb = axb - pb.x, dyb = ay - pb.y, dzb = az - pb.z;
                const float r2a = __builtin_fmaf(dza, dza, __builtin_fmaf(dxa, dxa, dya * dya));
                const float r2b = __builtin_fmaf(dzb, dzb, __builtin_fmaf(dxb, dxb, dyb * dyb));
                push(oka && r2a < 1.0f, (unsigned)ja | (ta >= n0 ? w1 : w0));
                push(okb && r2b < 1.0f, (unsigned)jb | (tb >= n0 ? w1 : w0));
            }
        }
        while (wave_any(rt != rh)) step();
        __builtin_amdgcn_wave_barrier();
    }

    if (active) {
        UDX_GLOBAL double *g = gacc + 3 * (long)i;
        gacc_add(g, W.hacc[0][lane] + hx);
        gacc_add(g + 1, W.hacc[1][lane] + hy);
        gacc_add(g + 2, W.hacc[2][lane] + hz);
    }
    if (A.pairs) {
        // ordered pairs, as the gather counts them: each bulk pair twice
        unsigned long long c = (unsigned long long)nhalo;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        c += 2ull * bulk;
        if (lane == 0 && c) atomicAdd(A.pairs, c);
    }
}

template <int WR>
__global__ __launch_bounds__(DPD_THREADS) void k_dpd_half(DpdK A) {
    __shared__ HsWave lds[DPD_THREADS / 64];
    // blocks in launch order (dealt to the XCDs round robin): 4% faster on the
    // 96^3 wall box than contiguous eighths of the cells per XCD
    dpd_half_block<WR, DPD_THREADS>(A, blockIdx.x * DPD_THREADS, lds);
}

template <int WR>
__global__ __launch_bounds__(DPD_THREADS) void k_dpd_half_batch(const DpdK *boxes, const int *bblock, int nbox) {
    __shared__ HsWave lds[DPD_THREADS / 64];
    const int blk = blockIdx.x;   // launch order, as k_dpd_batch
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bblock[mid] <= blk) lo = mid;
        else hi = mid;
    }
    const DpdK &A = boxes[lo];
    dpd_half_block<WR, DPD_THREADS>(A, A.i0 + (blk - bblock[lo]) * DPD_THREADS, lds);
}

// float forces from the FP64 sums (rounded once), which are cleared for the
// next pass: particles i0 + k, k < nmax, below *n_dev (the end of the cell list)
__global__ __launch_bounds__(256) void k_acc_to_ff(int nmax, const int *n_dev, int i0, double *acc, float *ff,
                                                     int stride, int accumulate) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmax || i0 + k >= *n_dev) return;
    const long i = i0 + k;
    double *a = acc + 3 * i;
    const float fx = (float)a[0], fy = (float)a[1], fz = (float)a[2];
    a[0] = a[1] = a[2] = 0.0;
    if (stride == 4) {
        float4 *f4 = (float4 *)ff + i;
        float4 o = accumulate ? *f4 : make_float4(0.f, 0.f, 0.f, 0.f);
        *f4 = make_float4(o.x + fx, o.y + fy, o.z + fz, 0.f);
    } else {
        float *f3 = ff + 3 * i;
        if (accumulate) {
            f3[0] += fx; f3[1] += fy; f3[2] += fz;
        } else {
            f3[0] = fx; f3[1] = fy; f3[2] = fz;
        }
    }
}

// Counts of the cells of the 26 fragments in fragment-cell order (x
// fastest, src/exch/flu/dev/map.h:22-42), one extra 0 after each fragment.
// An exclusive scan over this array gives every fragment's packed-list
// offsets (src/exch/flu/dev/map.h:44-98) as differences G[off + h] - G[off].
// The box's edge list counter (if any) starts the step at zero here.
__global__ __launch_bounds__(256) void k_frag_counts(const int *starts, int Lx, int Ly, int Lz, FragOff fo,
                                                     int *cnt, int *ecount) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0 && ecount) *ecount = 0;
    if (g >= fo.off[26]) return;
    int f = 0;
    while (f < 25 && g >= fo.off[f + 1]) f++;
    const int h = g - fo.off[f];
    const int dx = (f + 2) % 3 - 1, dy = (f / 3 + 2) % 3 - 1, dz = (f / 9 + 2) % 3 - 1;
    const int ex = dx ? 1 : Lx, ey = dy ? 1 : Ly, ez = dz ? 1 : Lz;
    int v = 0;
    if (h < ex * ey * ez) {
        const int x = (dx == 1 ? Lx - 1 : 0) + h % ex;
        const int y = (dy == 1 ? Ly - 1 : 0) + (h / ex) % ey;
        const int z = (dz == 1 ? Lz - 1 : 0) + h / (ex * ey);
        const int cid = x + Lx * (y + Ly * z);
        v = starts[cid + 1] - starts[cid];
    }
    cnt[g] = v;
}

DpdK make_k(const DpdLaunch &a) {
    DpdK k;
    k.pos = a.pos;
    k.vel = a.vel;
    k.starts = a.starts;
    k.bstarts = a.bstarts;
    k.sub = a.sub;
    k.fss = a.frag->fss;
    for (int f = 0; f < 27; f++) k.foff[f] = a.frag->off[f];
    k.Lx = a.box.L[0];
    k.Ly = a.box.L[1];
    k.Lz = a.box.L[2];
    k.nmax = a.nmax;
    k.periodic = a.periodic;
    k.i0 = a.i0;
    k.par = DpdPar{a.par.a, a.par.g, a.par.s, a.par.spow};
    k.seeds = a.seeds;
    k.ff = a.ff;
    k.ff_stride = a.ff_stride;
    k.accumulate = a.accumulate;
    k.pairs = a.pairs;
    k.ss = a.ss;
    k.acc = a.acc;
    k.ecount = a.ecount;
    k.eidx = a.eidx;
    k.ef = a.ef;
    k.opos = k.ovel = nullptr;
    k.mdt = k.mmass = k.max_ = k.may = k.maz = 0.f;
    k.sdf = SdfView{};
    k.wv = WallVel{};
    k.sc = SlabCount{};
    k.wpos = nullptr;
    k.wstarts = nullptr;
    k.Wx = k.Wy = k.Wz = k.nwall = 0;
    k.wseed = nullptr;
    k.wpairs = nullptr;
    if (const FusedMove *m = a.move) {
        k.opos = m->opos;
        k.ovel = m->ovel;
        k.mdt = m->dt;
        k.mmass = m->mass;
        k.max_ = m->accel[0];
        k.may = m->accel[1];
        k.maz = m->accel[2];
        if (m->walls) {
            k.sdf = m->sdf;
            k.wv = m->wv;
            k.wpos = m->wpos;
            k.wstarts = m->wstarts;
            k.Wx = m->W[0];
            k.Wy = m->W[1];
            k.Wz = m->W[2];
            k.nwall = m->nwall;
            k.wseed = m->wseed;
            k.wpairs = m->wpairs;
        }
        for (int c = 0; c < 3; c++) k.sc.L[c] = a.box.L[c];
        k.sc.sub = m->next->sub;
        k.sc.counts = m->next->bcounts;
        k.sc.cellslot = m->next->cellslot;
        k.sc.slab0 = m->slab0;
    }
    return k;
}

// Expected particles in the boundary cells of a periodic box (the ones with
// pairs across the boundary), for nmax particles spread uniformly.
double edge_estimate(const Box &b, int nmax) {
    double inner = 1.0;
    for (int c = 0; c < 3; c++) inner *= b.L[c] > 2 ? (double)(b.L[c] - 2) / b.L[c] : 0.0;
    return (double)nmax * (1.0 - inner);
}

// The halo pass as its own launch over the listed particles keeps all its
// lanes busy, but its chains of dependent loads are exposed unless there are
// many of them: measured on MI355X, a 96^3 box (about 5e5 boundary particles)
// runs faster with the halo pass inside the force kernel, where other waves'
// bulk work hides them, and 64 boxes of 8..48 cells (1.4e6) faster split.
double split_edges = 1048576.0;   // udx_dpd_set_split_edges

// with an edge list (periodic boxes with many boundary particles) the bulk
// pass and the halo pass of the listed particles are two launches of the same
// grid; results are identical either way
bool split_halo(const DpdLaunch &a) { return a.eidx && a.periodic && edge_estimate(a.box, a.nmax) >= split_edges; }

template <int WR, bool EXACT>
int launch_dpd(hipStream_t s, const DpdK &k, int nmax) {
    const dim3 g(ceil_div(nmax, DPD_THREADS)), b(DPD_THREADS);
    const Box box{{k.Lx, k.Ly, k.Lz}};
    if (k.opos) {
        // the fused move (dpd_move_fused): both force modes give the same bits
        if (k.sdf.tex) hipLaunchKernelGGL((k_dpd<WR, false, PH_ALL, 2>), g, b, 0, s, k);
        else hipLaunchKernelGGL((k_dpd<WR, false, PH_ALL, 1>), g, b, 0, s, k);
    } else if (k.eidx && k.periodic && edge_estimate(box, nmax) >= split_edges) {
        hipLaunchKernelGGL((k_dpd<WR, EXACT, PH_BULK>), g, b, 0, s, k);
        UDX_LAUNCHED();
        hipLaunchKernelGGL((k_dpd<WR, EXACT, PH_HALO>), g, b, 0, s, k);
    } else {
        hipLaunchKernelGGL((k_dpd<WR, EXACT, PH_ALL>), g, b, 0, s, k);
    }
    UDX_LAUNCHED();
    return UDX_OK;
}


}  // namespace

int frag_alloc(FragTables *t, const Box &b) {
    *t = FragTables{};
    int off = 0;
    for (int f = 0; f < 26; f++) {
        const int dx = (f + 2) % 3 - 1, dy = (f / 3 + 2) % 3 - 1, dz = (f / 9 + 2) % 3 - 1;
        t->off[f] = off;
        off += (dx ? 1 : b.L[0]) * (dy ? 1 : b.L[1]) * (dz ? 1 : b.L[2]) + 1;
    }
    t->off[26] = off;
    t->size = off;
    UDX_TRY(dev_alloc(&t->cnt, off));
    UDX_TRY(dev_alloc(&t->fss, off + 1));
    UDX_TRY(dev_alloc(&t->scan_tmp, scan_scratch_size(off)));
    return UDX_OK;
}

void frag_free(FragTables *t) {
    dev_free(t->cnt);
    dev_free(t->fss);
    dev_free(t->scan_tmp);
}

int frag_build(hipStream_t s, const FragTables *t, const Box &b, const int *starts, int *ecount) {
    FragOff fo;
    for (int f = 0; f < 27; f++) fo.off[f] = t->off[f];
    hipLaunchKernelGGL(k_frag_counts, dim3(ceil_div(t->size, 256)), dim3(256), 0, s, starts, b.L[0], b.L[1],
                       b.L[2], fo, t->cnt, ecount);
    UDX_LAUNCHED();
    return scan_exclusive(s, t->cnt, t->size, t->fss, t->scan_tmp);
}

// mode UDX_MODE_EXACT selects correctly rounded sqrt/division in the pair force
int dpd_forces_mode(hipStream_t s, const DpdLaunch &a, int mode) {
    const bool exact = mode == UDX_MODE_EXACT;
    UDX_CHECK(!a.bstarts || (a.box.L[0] <= 2048 && a.box.L[1] <= 2048 && a.box.L[2] <= 2048), UDX_EINVAL,
              "dpd: x-slab windows need box sides of at most 2048 cells");
    UDX_CHECK(!a.periodic || (a.box.L[0] >= 3 && a.box.L[1] >= 3 && a.box.L[2] >= 3), UDX_EINVAL,
              "dpd: a periodic box needs at least 3 cells per side (got %d %d %d)", a.box.L[0], a.box.L[1],
              a.box.L[2]);
    UDX_CHECK(a.nmax < (1 << 27), UDX_ECAP, "dpd: %d particles exceed the 2^27 queue index range", a.nmax);
    UDX_CHECK(!a.move || (dpd_move_fused(a, mode) && a.move->next), UDX_EINVAL,
              "dpd: the fused move needs the gather in one launch with float4 forces");
    if (a.nmax <= 0) return UDX_OK;
    const DpdK k = make_k(a);
    const float sp = a.par.spow;
    if (a.ss) {
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(ceil_div(a.nmax, DPD_THREADS)), dim3(DPD_THREADS), 0, s, k);
            return hipGetLastError();
        };
        hipError_t e;
        if (sp == 0.25f) e = go(k_dpd_stress<WR_QUARTER>);
        else if (sp == 0.5f) e = go(k_dpd_stress<WR_HALF>);
        else if (sp == 1.0f) e = go(k_dpd_stress<WR_ONE>);
        else e = go(k_dpd_stress<WR_POW>);
        UDX_CHECK(e == hipSuccess, UDX_EHIP, "dpd: stress kernel launch: %s", hipGetErrorString(e));
        return UDX_OK;
    }
    if (exact) {
        if (sp == 0.25f) return launch_dpd<WR_QUARTER, true>(s, k, a.nmax);
        if (sp == 0.5f) return launch_dpd<WR_HALF, true>(s, k, a.nmax);
        if (sp == 1.0f) return launch_dpd<WR_ONE, true>(s, k, a.nmax);
        return launch_dpd<WR_POW, true>(s, k, a.nmax);
    }
    if (mode == UDX_MODE_HALF) {
        UDX_CHECK(a.acc, UDX_EINVAL, "dpd: the half-shell mode needs FP64 accumulators");
        // half shell with exact FP64 sums, then the float forces
        const dim3 g(ceil_div(a.nmax, DPD_THREADS)), b(DPD_THREADS);
        if (sp == 0.25f) hipLaunchKernelGGL(k_dpd_half<WR_QUARTER>, g, b, 0, s, k);
        else if (sp == 0.5f) hipLaunchKernelGGL(k_dpd_half<WR_HALF>, g, b, 0, s, k);
        else if (sp == 1.0f) hipLaunchKernelGGL(k_dpd_half<WR_ONE>, g, b, 0, s, k);
        else hipLaunchKernelGGL(k_dpd_half<WR_POW>, g, b, 0, s, k);
        UDX_LAUNCHED();
        const int nc = a.box.ncells();
        hipLaunchKernelGGL(k_acc_to_ff, dim3(ceil_div(a.nmax, 256)), dim3(256), 0, s, a.nmax, a.starts + nc, a.i0,
                           a.acc, a.ff, a.ff_stride, a.accumulate);
        UDX_LAUNCHED();
        return UDX_OK;
    }
    if (sp == 0.25f) return launch_dpd<WR_QUARTER, false>(s, k, a.nmax);
    if (sp == 0.5f) return launch_dpd<WR_HALF, false>(s, k, a.nmax);
    if (sp == 1.0f) return launch_dpd<WR_ONE, false>(s, k, a.nmax);
    return launch_dpd<WR_POW, false>(s, k, a.nmax);
}

int dpd_forces(hipStream_t s, const DpdLaunch &a) { return dpd_forces_mode(s, a, 0); }

bool dpd_move_fused(const DpdLaunch &a, int mode) {
    return (mode == UDX_MODE_FAST || mode == UDX_MODE_EXACT) && a.periodic && !split_halo(a) && a.ff_stride == 4 &&
           !a.accumulate && !a.ss;
}

int dpd_set_split_edges(double n) {
    UDX_CHECK(n >= 0.0, UDX_EINVAL, "udx_dpd_set_split_edges: %g", n);
    split_edges = n;
    return UDX_OK;
}

// ---------------------------------------------------------------- batches
namespace {
constexpr int VAR_SHAPE[DPD_NVARIANT][2] = {{128, 16}, {128, 8}, {128, 32}, {256, 16}, {256, 8}, {256, 32}};


template <int WR, bool EXACT, int NT, int QC>
int launch_batch(hipStream_t s, const DpdBatch &b) {
    const int w = NT == 128 ? 0 : 1;
    if (b.nblocks[w] == 0) return UDX_OK;
    const DpdK *tab = (const DpdK *)b.table;
    if (b.split) {
        hipLaunchKernelGGL((k_dpd_batch<WR, EXACT, NT, QC, PH_BULK>), dim3(b.nblocks[w]), dim3(NT), 0, s, tab,
                           b.bblock[w], b.nbox, b.step);
        UDX_LAUNCHED();
        hipLaunchKernelGGL((k_dpd_batch<WR, EXACT, NT, QC, PH_HALO>), dim3(b.nblocks[w]), dim3(NT), 0, s, tab,
                           b.bblock[w], b.nbox, b.step);
    } else {
        hipLaunchKernelGGL((k_dpd_batch<WR, EXACT, NT, QC, PH_ALL>), dim3(b.nblocks[w]), dim3(NT), 0, s, tab,
                           b.bblock[w], b.nbox, b.step);
    }
    UDX_LAUNCHED();
    return UDX_OK;
}

template <int WR, bool EXACT>
int launch_batch_wr(hipStream_t s, const DpdBatch &b) {
    return launch_batch<WR, EXACT, 128, 16>(s, b);
}
}  // namespace

void dpd_variant_shape(int v, int *threads, int *qcap) {
    *threads = VAR_SHAPE[v][0];
    *qcap = VAR_SHAPE[v][1];
}

int dpd_batch_build(DpdBatch *b, const DpdLaunch *boxes, int nbox) {
    UDX_CHECK(nbox > 0, UDX_EINVAL, "dpd: empty batch");
    std::vector<DpdK> ks(nbox);
    std::vector<int> bb[2] = {std::vector<int>(nbox + 1), std::vector<int>(nbox + 1)};
    for (int k = 0; k < nbox; k++) {
        const DpdLaunch &a = boxes[k];
        UDX_CHECK(!a.periodic || (a.box.L[0] >= 3 && a.box.L[1] >= 3 && a.box.L[2] >= 3), UDX_EINVAL,
                  "dpd: box %d: a periodic box needs at least 3 cells per side", k);
        UDX_CHECK(a.par.spow == boxes[0].par.spow, UDX_EINVAL, "dpd: the boxes of a batch share spow");
        UDX_CHECK(!a.bstarts || (a.box.L[0] <= 2048 && a.box.L[1] <= 2048 && a.box.L[2] <= 2048), UDX_EINVAL,
                  "dpd: box %d: x-slab windows need box sides of at most 2048 cells", k);
        ks[k] = make_k(a);
        for (int w = 0; w < 2; w++) bb[w][k + 1] = bb[w][k] + ceil_div(a.nmax, w == 0 ? 128 : 256);
    }
    UDX_CHECK((long)boxes[nbox - 1].i0 + boxes[nbox - 1].nmax < (1L << 27), UDX_ECAP,
              "dpd: the batch exceeds the 2^27 queue index range");
    dpd_batch_free(b);
    UDX_HIP(hipMalloc(&b->table, sizeof(DpdK) * nbox));
    UDX_HIP(hipMemcpy(b->table, ks.data(), sizeof(DpdK) * nbox, hipMemcpyHostToDevice));
    for (int w = 0; w < 2; w++) {
        UDX_TRY(dev_alloc(&b->bblock[w], (size_t)nbox + 1));
        UDX_HIP(hipMemcpy(b->bblock[w], bb[w].data(), sizeof(int) * (nbox + 1), hipMemcpyHostToDevice));
        b->nblocks[w] = bb[w][nbox];
    }
    b->nbox = nbox;
    b->spow = boxes[0].par.spow;
    b->acc = boxes[0].acc;
    b->ff = boxes[0].ff;
    for (int k = 0; k < nbox; k++)
        UDX_CHECK(boxes[k].acc == b->acc && boxes[k].ff == b->ff && boxes[k].ff_stride == 4 && !boxes[k].accumulate,
                  UDX_EINVAL, "dpd: the boxes of a batch share one float4 force array and one FP64 sum array");
    b->ntot = boxes[nbox - 1].i0 + boxes[nbox - 1].nmax;
    b->n_end = boxes[nbox - 1].starts + boxes[nbox - 1].box.ncells();
    b->step = boxes[0].move ? boxes[0].move->step : nullptr;
    for (int k = 0; k < nbox; k++)
        UDX_CHECK(!boxes[k].move == !boxes[0].move && (!boxes[k].move || !boxes[k].move->walls), UDX_EINVAL,
                  "dpd: the boxes of a batch are all moved in the gather (without walls) or none");
    double edges = 0.0;
    b->split = 1;
    for (int k = 0; k < nbox; k++) {
        if (!boxes[k].eidx || !boxes[k].periodic) b->split = 0;
        edges += edge_estimate(boxes[k].box, boxes[k].nmax);
    }
    if (edges < split_edges) b->split = 0;
    return UDX_OK;
}

void dpd_batch_free(DpdBatch *b) {
    if (b->table) (void)hipFree(b->table);
    b->table = nullptr;
    dev_free(b->bblock[0]);
    dev_free(b->bblock[1]);
}

int dpd_forces_batch(hipStream_t s, const DpdBatch &b, int mode, int variant) {
    const float sp = b.spow;
    if (mode == UDX_MODE_HALF) {
        UDX_CHECK(b.acc, UDX_EINVAL, "dpd: the half-shell mode needs FP64 accumulators");
        if (b.nblocks[0] > 0) {
            const dim3 g(b.nblocks[0]), t(DPD_THREADS);
            const DpdK *tab = (const DpdK *)b.table;
            if (sp == 0.25f) hipLaunchKernelGGL(k_dpd_half_batch<WR_QUARTER>, g, t, 0, s, tab, b.bblock[0], b.nbox);
            else if (sp == 0.5f) hipLaunchKernelGGL(k_dpd_half_batch<WR_HALF>, g, t, 0, s, tab, b.bblock[0], b.nbox);
            else if (sp == 1.0f) hipLaunchKernelGGL(k_dpd_half_batch<WR_ONE>, g, t, 0, s, tab, b.bblock[0], b.nbox);
            else hipLaunchKernelGGL(k_dpd_half_batch<WR_POW>, g, t, 0, s, tab, b.bblock[0], b.nbox);
            UDX_LAUNCHED();
        }
        if (b.ntot > 0) {
            hipLaunchKernelGGL(k_acc_to_ff, dim3(ceil_div(b.ntot, 256)), dim3(256), 0, s, b.ntot, b.n_end, 0, b.acc,
                               b.ff, 4, 0);
            UDX_LAUNCHED();
        }
        (void)variant;
        return UDX_OK;
    }
    if (mode == UDX_MODE_EXACT) {
        if (sp == 0.25f) return launch_batch_wr<WR_QUARTER, true>(s, b);
        if (sp == 0.5f) return launch_batch_wr<WR_HALF, true>(s, b);
        if (sp == 1.0f) return launch_batch_wr<WR_ONE, true>(s, b);
        return launch_batch_wr<WR_POW, true>(s, b);
    }
    if (sp == 0.25f) {
        switch (variant) {
            case 1: return launch_batch<WR_QUARTER, false, 128, 8>(s, b);
            case 2: return launch_batch<WR_QUARTER, false, 128, 32>(s, b);
            case 3: return launch_batch<WR_QUARTER, false, 256, 16>(s, b);
            case 4: return launch_batch<WR_QUARTER, false, 256, 8>(s, b);
            case 5: return launch_batch<WR_QUARTER, false, 256, 32>(s, b);
            default: return launch_batch<WR_QUARTER, false, 128, 16>(s, b);
        }
    }
    if (sp == 0.5f) return launch_batch_wr<WR_HALF, false>(s, b);
    if (sp == 1.0f) return launch_batch_wr<WR_ONE, false>(s, b);
    return launch_batch_wr<WR_POW, false>(s, b);
}

}  // namespace udx
