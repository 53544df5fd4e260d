// Solvent-solvent DPD forces for MI355X.
//
// Reference: the bulk kernel visits a half shell with float atomics on both
// particles (src/fluforces/bulk/dev/main.h:23-97) and 26 halo kernels add the
// one-sided forces of pairs across the periodic boundary with their own RNG
// (src/fluforces/halo/dev/main.h:7-83).  Here one pass gathers the full shell
// of each particle in a fixed order, so forces are deterministic and need no
// atomics, and each wrapped pair is computed exactly like the reference's
// halo: the local particle shifted by -L d into the neighbour frame and the
// noise drawn from the fragment's trunk with packed-list positions.
//
// Kernel shape: one thread per cell-ordered particle.  A candidate pass walks
// the 9 rows (dz, dy) of the 27-cell shell, each row a contiguous run of at
// most 3 cells (two runs where x wraps), tests the distance and appends hits
// to a per-lane queue in LDS.  When a queue fills (and at the end) the wave
// drains the queues together, so the expensive pair body (RNG + force) runs
// with nearly all lanes active instead of the ~15% that pass the cutoff.
#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {

namespace {

constexpr int DPD_THREADS = 256;
constexpr int QCAP = 32;

struct DpdK {
    const float4 *pos, *vel;
    const int *starts, *fss;
    int foff[27];
    int Lx, Ly, Lz, nmax, periodic;
    DpdPar par;
    const float *seeds;
    float *ff;
    int ff_stride, accumulate;
    unsigned long long *pairs;
};

// x-runs of row (dy, dz) for a particle in column cx: seg 0 and seg 1 in
// ascending dx order (dx = -1, 0, +1 across the wrap)
__device__ __forceinline__ bool xrun(int sg, int cx, int Lx, int periodic, int *x0, int *x1, int *wx) {
    if (sg == 0) {
        if (periodic && cx == 0) { *x0 = *x1 = Lx - 1; *wx = -1; return true; }
        *x0 = cx > 0 ? cx - 1 : 0;
        *x1 = cx + 1 < Lx ? cx + 1 : Lx - 1;
        *wx = 0;
        return true;
    }
    if (periodic && cx == 0) { *x0 = 0; *x1 = Lx > 1 ? 1 : 0; *wx = 0; return true; }
    if (periodic && cx == Lx - 1) { *x0 = *x1 = 0; *wx = 1; return true; }
    return false;
}

__device__ __forceinline__ int wrap_dir(int c, int L) { return c < 0 ? -1 : (c >= L ? 1 : 0); }

template <int WR, bool EXACT>
__global__ __launch_bounds__(DPD_THREADS) void k_dpd(DpdK A) {
    __shared__ unsigned q[QCAP * DPD_THREADS];
    __shared__ unsigned long long lpairs[DPD_THREADS / 64];
    const int tid = threadIdx.x;
    const int Lx = A.Lx, Ly = A.Ly, Lz = A.Lz;
    const int n = A.starts[Lx * Ly * Lz];
    const int i = blockIdx.x * DPD_THREADS + tid;
    const bool active = i < n && i < A.nmax;

    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi;
    int cx = 0, cy = 0, cz = 0, istart = 0;
    if (active) {
        pi = A.pos[i];
        vi = A.vel[i];
        cell_coord(pi.x, Lx, &cx);
        cell_coord(pi.y, Ly, &cy);
        cell_coord(pi.z, Lz, &cz);
        istart = A.starts[cx + Lx * (cy + Ly * cz)];
    }
    const float seed_bulk = A.seeds[26];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    int npair = 0, fill = 0;

    // drain the per-lane queues: pair body for every queued neighbour
    auto drain = [&]() {
        for (int e = 0; wave_any(e < fill); e++) {
            if (e < fill) {
                const unsigned ent = q[e * DPD_THREADS + tid];
                const int j = (int)(ent & 0x7FFFFFFu);
                const int code = (int)(ent >> 27);
                const int row = code / 3, wx = code % 3 - 1;
                const int dz = row / 3 - 1, dy = row % 3 - 1;
                const int yy = cy + dy, zz = cz + dz;
                const int wy = wrap_dir(yy, Ly), wz = wrap_dir(zz, Lz);
                const float4 pj = A.pos[j], vj = A.vel[j];
                const float ax = pi.x - (float)(wx * Lx);
                const float ay = pi.y - (float)(wy * Ly);
                const float az = pi.z - (float)(wz * Lz);
                const float ddx = ax - pj.x, ddy = ay - pj.y, ddz = az - pj.z;
                const float r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddx, ddx, ddy * ddy));
                float rnd;
                if ((wx | wy | wz) == 0) {
                    const int hi = i > j ? i : j, lo = i > j ? j : i;
                    rnd = mean0var1ii(seed_bulk, hi, lo);
                } else {
                    const int fid = frag_d2i(wx, wy, wz), anti = frag_d2i(-wx, -wy, -wz);
                    const int aid = A.fss[A.foff[fid] + frag_cell(wx, wy, wz, cx, cy, cz, Lx, Ly)] + (i - istart);
                    int jx = 0;
                    cell_coord(pj.x, Lx, &jx);
                    const int jy = yy - wy * Ly, jz = zz - wz * Lz;
                    const int jstart = A.starts[jx + Lx * (jy + Ly * jz)];
                    const int bid =
                        A.fss[A.foff[anti] + frag_cell(-wx, -wy, -wz, jx, jy, jz, Lx, Ly)] + (j - jstart);
                    const bool mask = fid < anti;
                    rnd = mean0var1uu(A.seeds[fid], (uint32_t)(mask ? aid : bid), (uint32_t)(mask ? bid : aid));
                }
                const float3 f = dpd_pair<WR, EXACT>(A.par, ddx, ddy, ddz, r2, vi.x - vj.x, vi.y - vj.y,
                                                     vi.z - vj.z, rnd);
                fx += f.x;
                fy += f.y;
                fz += f.z;
                npair++;
            }
        }
        fill = 0;
    };

    int k = 0, j = 0, je = 0, code = 0;
    float xa = 0.f, ya = 0.f, za = 0.f;
    for (;;) {
        // next non-empty run of candidate cells
        while (active && j >= je && k < 18) {
            const int row = k >> 1, sg = k & 1;
            k++;
            const int dz = row / 3 - 1, dy = row % 3 - 1;
            int yy = cy + dy, zz = cz + dz;
            const int wy = wrap_dir(yy, Ly), wz = wrap_dir(zz, Lz);
            if ((wy | wz) && !A.periodic) continue;
            int x0, x1, wx;
            if (!xrun(sg, cx, Lx, A.periodic, &x0, &x1, &wx)) continue;
            yy -= wy * Ly;
            zz -= wz * Lz;
            const int base = Lx * (yy + Ly * zz);
            j = A.starts[base + x0];
            je = A.starts[base + x1 + 1];
            code = row * 3 + (wx + 1);
            xa = pi.x - (float)(wx * Lx);
            ya = pi.y - (float)(wy * Ly);
            za = pi.z - (float)(wz * Lz);
        }
        const bool has = active && j < je;
        if (!wave_any(has)) break;
        if (has) {
            if (j != i) {
                const float4 pj = A.pos[j];
                const float ddx = xa - pj.x, ddy = ya - pj.y, ddz = za - pj.z;
                const float r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddx, ddx, ddy * ddy));
                if (r2 < 1.0f) {
                    q[fill * DPD_THREADS + tid] = (unsigned)j | ((unsigned)code << 27);
                    fill++;
                }
            }
            j++;
        }
        if (wave_any(fill == QCAP)) drain();
    }
    drain();

    if (active) {
        if (A.ff_stride == 4) {
            float4 *f4 = (float4 *)A.ff + i;
            float4 o = A.accumulate ? *f4 : make_float4(0.f, 0.f, 0.f, 0.f);
            *f4 = make_float4(o.x + fx, o.y + fy, o.z + fz, 0.f);
        } else {
            float *f3 = A.ff + 3 * (long)i;
            if (A.accumulate) {
                f3[0] += fx; f3[1] += fy; f3[2] += fz;
            } else {
                f3[0] = fx; f3[1] = fy; f3[2] = fz;
            }
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
            for (int w = 0; w < DPD_THREADS / 64; w++) s += lpairs[w];
            if (s) atomicAdd(A.pairs, s);
        }
    }
}

// per-fragment exclusive scan of cell counts in fragment-cell order
// (src/exch/flu/dev/map.h:22-98); one block per fragment
__global__ __launch_bounds__(256) void k_frag_tables(const int *starts, int Lx, int Ly, int Lz, int *fss,
                                                     const int *foff_dev) {
    __shared__ int lw[4];
    __shared__ int carry;
    const int f = blockIdx.x;
    const int dx = (f + 2) % 3 - 1, dy = (f / 3 + 2) % 3 - 1, dz = (f / 9 + 2) % 3 - 1;
    const int ox = dx == 1 ? Lx - 1 : 0, oy = dy == 1 ? Ly - 1 : 0, oz = dz == 1 ? Lz - 1 : 0;
    const int ex = dx ? 1 : Lx, ey = dy ? 1 : Ly, ez = dz ? 1 : Lz;
    const int nc = ex * ey * ez;
    int *out = fss + foff_dev[f];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int h0 = 0; h0 < nc; h0 += 256) {
        const int h = h0 + threadIdx.x;
        int v = 0;
        if (h < nc) {
            const int x = ox + h % ex, y = oy + (h / ex) % ey, z = oz + h / (ex * ey);
            const int cid = x + Lx * (y + Ly * z);
            v = starts[cid + 1] - starts[cid];
        }
        int inc = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        if (lane == 63) lw[wave] = inc;
        __syncthreads();
        int add = carry;
        for (int w = 0; w < wave; w++) add += lw[w];
        if (h < nc) out[h] = add + inc - v;
        __syncthreads();
        if (threadIdx.x == 255) carry = add + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[nc] = carry;
}

template <int WR, bool EXACT>
int launch_dpd(hipStream_t s, const DpdK &k, int nmax) {
    hipLaunchKernelGGL((k_dpd<WR, EXACT>), dim3(ceil_div(nmax, DPD_THREADS)), dim3(DPD_THREADS), 0, s, k);
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
    UDX_TRY(dev_alloc(&t->fss, off + 27));
    UDX_HIP(hipMemcpy(t->fss + off, t->off, 27 * sizeof(int), hipMemcpyHostToDevice));
    return UDX_OK;
}

void frag_free(FragTables *t) { dev_free(t->fss); }

int frag_build(hipStream_t s, const FragTables *t, const Box &b, const int *starts) {
    hipLaunchKernelGGL(k_frag_tables, dim3(26), dim3(256), 0, s, starts, b.L[0], b.L[1], b.L[2], t->fss,
                       t->fss + t->size);
    UDX_LAUNCHED();
    return UDX_OK;
}

// exact = 1 selects correctly rounded sqrt/division in the pair force
int dpd_forces_mode(hipStream_t s, const DpdLaunch &a, int exact) {
    UDX_CHECK(!a.periodic || (a.box.L[0] >= 3 && a.box.L[1] >= 3 && a.box.L[2] >= 3), UDX_EINVAL,
              "dpd: a periodic box needs at least 3 cells per side (got %d %d %d)", a.box.L[0], a.box.L[1],
              a.box.L[2]);
    UDX_CHECK(a.nmax < (1 << 27), UDX_ECAP, "dpd: %d particles exceed the 2^27 queue index range", a.nmax);
    if (a.nmax <= 0) return UDX_OK;
    DpdK k;
    k.pos = a.pos;
    k.vel = a.vel;
    k.starts = a.starts;
    k.fss = a.frag->fss;
    for (int f = 0; f < 27; f++) k.foff[f] = a.frag->off[f];
    k.Lx = a.box.L[0];
    k.Ly = a.box.L[1];
    k.Lz = a.box.L[2];
    k.nmax = a.nmax;
    k.periodic = a.periodic;
    k.par = DpdPar{a.par.a, a.par.g, a.par.s, a.par.spow};
    k.seeds = a.seeds;
    k.ff = a.ff;
    k.ff_stride = a.ff_stride;
    k.accumulate = a.accumulate;
    k.pairs = a.pairs;
    const float sp = a.par.spow;
    if (exact) {
        if (sp == 0.25f) return launch_dpd<WR_QUARTER, true>(s, k, a.nmax);
        if (sp == 0.5f) return launch_dpd<WR_HALF, true>(s, k, a.nmax);
        if (sp == 1.0f) return launch_dpd<WR_ONE, true>(s, k, a.nmax);
        return launch_dpd<WR_POW, true>(s, k, a.nmax);
    }
    if (sp == 0.25f) return launch_dpd<WR_QUARTER, false>(s, k, a.nmax);
    if (sp == 0.5f) return launch_dpd<WR_HALF, false>(s, k, a.nmax);
    if (sp == 1.0f) return launch_dpd<WR_ONE, false>(s, k, a.nmax);
    return launch_dpd<WR_POW, false>(s, k, a.nmax);
}

int dpd_forces(hipStream_t s, const DpdLaunch &a) { return dpd_forces_mode(s, a, 0); }

}  // namespace udx
