// Solvent-membrane DPD (FSI) on MI355X, one periodic domain.
//
// Reference: 3 threads per membrane vertex (one z plane each) walk the
// solvent cells of the vertex's 3x3x3 block clipped to the domain and add +f
// to the vertex and -f to each solvent particle by float atomics
// (src/fsi/dev/bulk.h, src/fsi/dev/map/bulk.h); vertices outside the domain
// are sent to the neighbours, which on one periodic domain means images
// shifted back by a period that interact the same way with their own seed
// and their halo-list position as noise index (src/exch/obj,
// src/exch/map/dev.h:8-82, src/fsi/dev/halo.h).
//
// Here both sides gather, so no atomics are needed and the sums have a
// fixed order:
//  - k_fsi_instance: one thread per instance (vertex or image) in cell
//    order, one sum over its block; k_fsi_vertex_sum adds to each vertex
//    its own sum, then its images' in halo-list order;
//  - k_fsi_solvent: one thread per solvent particle over the vertex/image
//    instances of the 27 cells around it, from a cell list of the instances
//    on the (L+2)^3 grid of cells -1..L (the only cells whose clipped block
//    reaches into the domain).
// Each pair is evaluated on both sides with the same arguments, so the two
// values are identical and the momentum exchange is exact per pair.
#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {

namespace {

constexpr int FSI_THREADS = 128;

// per-axis -1 / 0 / +1 of a vertex below / inside / above the domain
// (src/exch/map/dev.h:8-18)
__device__ __forceinline__ int emap_axis(float r, int L) {
    return -1 + (r >= (float)(-L / 2)) + (r >= (float)(L / 2));
}

// fragment i of the halo list holds the vertices outside toward -d_i
// (every non-empty subset of their outside axes, src/exch/map/dev.h:43-82)
__device__ __forceinline__ bool in_fragment(int i, int ex, int ey, int ez) {
    const int sx = -((i + 2) % 3 - 1), sy = -((i / 3 + 2) % 3 - 1), sz = -((i / 9 + 2) % 3 - 1);
    if ((sx | sy | sz) == 0) return false;
    return (sx == 0 || sx == ex) && (sy == 0 || sy == ey) && (sz == 0 || sz == ez);
}

__global__ __launch_bounds__(256) void k_fsi_flags(int no, const float *opp, int Lx, int Ly, int Lz, int *flag) {
    const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (g >= 26L * no) return;
    const int i = (int)(g / no), v = (int)(g % no);
    const float *p = opp + 6 * (long)v;
    const int ex = emap_axis(p[0], Lx), ey = emap_axis(p[1], Ly), ez = emap_axis(p[2], Lz);
    flag[g] = in_fragment(i, ex, ey, ez) ? 1 : 0;
}

// instance AoS: vertices [0, no) as they are, images at no + aid shifted by
// +d_i L; meta = (noise index, 0 vertex / 1 image).  The unused slots up to
// 8 no (at most 7 images per vertex) get a position outside every cell so
// the instance cell list drops them.  The force sum of every real instance
// starts at zero (instances the cell list drops have no solvent in reach).
__global__ __launch_bounds__(256) void k_fsi_instances(int no, const float *opp, int Lx, int Ly, int Lz,
                                                       const int *flag, const int *aid, float *inst, int2 *meta,
                                                       float *sinst) {
    const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (g < no) {
        const float *p = opp + 6 * g;
        float *o = inst + 6 * g;
        for (int c = 0; c < 6; c++) o[c] = p[c];
        meta[g] = make_int2((int)g, 0);
        sinst[3 * g] = sinst[3 * g + 1] = sinst[3 * g + 2] = 0.f;
    } else if (g < 8L * no && g >= no + aid[26L * no]) {
        float *o = inst + 6 * g;
        o[0] = o[1] = o[2] = 1e30f;
        o[3] = o[4] = o[5] = 0.f;
        meta[g] = make_int2(0, 0);
    }
    if (g >= 26L * no || !flag[g]) return;
    const int i = (int)(g / no), v = (int)(g % no);
    const int dx = (i + 2) % 3 - 1, dy = (i / 3 + 2) % 3 - 1, dz = (i / 9 + 2) % 3 - 1;
    const float *p = opp + 6 * (long)v;
    const long k = no + aid[g];
    float *o = inst + 6 * k;
    o[0] = p[0] + (float)(dx * Lx);
    o[1] = p[1] + (float)(dy * Ly);
    o[2] = p[2] + (float)(dz * Lz);
    o[3] = p[3];
    o[4] = p[4];
    o[5] = p[5];
    meta[k] = make_int2(aid[g], 1);
    sinst[3 * k] = sinst[3 * k + 1] = sinst[3 * k + 2] = 0.f;
}

struct FsiK {
    int Lx, Ly, Lz;
    const float4 *spos, *svel;
    const int *starts;
    DpdPar par;
    const float *seeds;      // device: [bulk, halo]
};

// Both sides queue the candidates that pass the cutoff in a per-lane ring
// in LDS and drain the rings together, two pairs per lane per step, so the
// pair body (noise + force) runs with most lanes active: tested inline, about
// one candidate in six passes and every lane of the wave would wait for the
// body at nearly every candidate.  The queues keep candidate order, so the
// sums are the same as a sequential walk.
constexpr int FQ_V = 16, FQ_S = 16;   // ring entries per lane: instance side, solvent side
constexpr int FSI_DRAIN_STEPS = 2;
constexpr int FSI_SCAP = 160;         // staged instances per wave and row (solvent side)
constexpr int FSI_VCAP = 160;         // staged solvent particles per wave and row (instance side)
struct FsiIn {
    float dx, dy, dz, r2, dvx, dvy, dvz, x;
    bool pair;
};

// Vertex side, one thread per instance (the vertex or one of its images)
// in cell order of the instance grid, so the lanes of a wave read the same
// solvent cells: the instance's sum over its clipped block, z planes 0..2,
// rows y - 1..y + 1, solvent in cell order (src/fsi/dev/map/bulk.h), written
// to sinst at the instance's own index.  k_fsi_vertex_sum then adds, per
// vertex, its own sum and its images' sums in halo-list order.
template <int WR, bool EXACT>
__global__ __launch_bounds__(FSI_THREADS) void k_fsi_instance(FsiK K, const int *n_dev, const float *inst,
                                                              const int2 *meta, const int *order, float *sinst,
                                                              unsigned long long *pairs) {
    __shared__ unsigned q[FQ_V * FSI_THREADS];
    __shared__ float4 vstage[FSI_THREADS / 64][FSI_VCAP];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = t < *n_dev;
    float4 *const sv = vstage[threadIdx.x >> 6];
    const float *p = inst + 6 * (long)(active ? t : 0);
    const float rx = p[0], ry = p[1], rz = p[2], ux = p[3], uy = p[4], uz = p[5];
    const int2 m = active ? meta[t] : make_int2(0, 0);
    const unsigned lid = (unsigned)m.x;
    const float seed = m.y ? K.seeds[1] : K.seeds[0];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int fill = 0, head = 0, np = 0;
    unsigned *ql = q + threadIdx.x;
    auto ring = [&](int k) {
        const int r = head + k;
        return r >= FQ_V ? r - FQ_V : r;
    };
    auto in_of = [&](unsigned b) {
        FsiIn o;
        const float4 pb = K.spos[b], vb = K.svel[b];
        o.dx = rx - pb.x;
        o.dy = ry - pb.y;
        o.dz = rz - pb.z;
        o.r2 = __builtin_fmaf(o.dz, o.dz, __builtin_fmaf(o.dx, o.dx, o.dy * o.dy));
        o.dvx = ux - vb.x;
        o.dvy = uy - vb.y;
        o.dvz = uz - vb.z;
        // mean0var1uu(seed, lid, b) before the logistic map (src/math/rnd/dev.h:65-72)
        o.x = seed - low_discrepancy((float)(lid & 0x3FFu), (float)lid, (float)(b & 0x3FFu), (float)b);
        o.pair = true;
        return o;
    };
    auto drain_step = [&]() {
        const bool va = fill > 0, vb = fill > 1;
        const unsigned ea = va ? ql[ring(0) * FSI_THREADS] : 0u;     // solvent 0 exists when anything is queued
        const unsigned eb = vb ? ql[ring(1) * FSI_THREADS] : ea;
        const FsiIn a = in_of(ea), b = in_of(eb);
        float xa = a.x, xb = b.x;
        logistic18x2(&xa, &xb);
        const float3 fa = dpd_pair<WR, EXACT>(K.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
        const float3 fb = dpd_pair<WR, EXACT>(K.par, b.dx, b.dy, b.dz, b.r2, b.dvx, b.dvy, b.dvz, xb * R_SQRT2);
        if (va) {
            sx += fa.x;
            sy += fa.y;
            sz += fa.z;
        }
        if (vb) {
            sx += fb.x;
            sy += fb.y;
            sz += fb.z;
        }
        const int took = fill < 2 ? fill : 2;
        np += took;
        head = ring(took);
        fill -= took;
    };
    const int xc = K.Lx / 2 + (int)__builtin_floorf(rx);
    const int yc = K.Ly / 2 + (int)__builtin_floorf(ry);
    const int zc = K.Lz / 2 + (int)__builtin_floorf(rz);
    const bool xok = active && !(xc - 1 >= K.Lx || xc + 2 <= 0);
    const int xs = xc - 1 > 0 ? xc - 1 : 0, xe = xc + 2 < K.Lx ? xc + 2 : K.Lx;
    for (int row = 0; row < 9; row++) {
        const int z = zc - 1 + row / 3, y = yc - 1 + row % 3;
        int b0 = 0, len = 0;
        if (xok && z >= 0 && z < K.Lz && y >= 0 && y < K.Ly) {
            const int cid = xs + K.Lx * (y + K.Ly * z);
            b0 = K.starts[cid];
            len = K.starts[cid + (xe - xs)] - b0;
        }
        // the lanes are consecutive instances in cell order: the union of
        // their solvent runs of this row is usually one small range, staged
        // in LDS with coalesced loads (one round trip instead of one per four
        // candidates); the candidate order is the same either way
        int lo = len > 0 ? b0 : 0x7fffffff, hi = len > 0 ? b0 + len : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
        }
        if (hi <= lo) continue;
        if (hi - lo <= FSI_VCAP) {
            __builtin_amdgcn_wave_barrier();
            for (int k = threadIdx.x & 63; k < hi - lo; k += 64) sv[k] = K.spos[lo + k];
            __builtin_amdgcn_wave_barrier();
            const int o = len > 0 ? b0 - lo : 0;
            for (int c = 0; wave_any(c < len); c += 4) {
                if (wave_any(fill + 4 > FQ_V))
#pragma unroll 1
                    for (int k = 0; k < FSI_DRAIN_STEPS; k++) drain_step();
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = o + c + u < FSI_VCAP ? o + c + u : 0;
                    const float4 pb = c + u < len ? sv[k] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
                    const float dx = rx - pb.x, dy = ry - pb.y, dz = rz - pb.z;
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
                    if (c + u < len && r2 < 1.0f) {
                        ql[ring(fill) * FSI_THREADS] = (unsigned)(b0 + c + u);
                        fill++;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        // four candidates per iteration: four loads in flight per lane
        for (int c = 0; wave_any(c < len); c += 4) {
            if (wave_any(fill + 4 > FQ_V))
#pragma unroll 1
                for (int k = 0; k < FSI_DRAIN_STEPS; k++) drain_step();
            float4 pb[4];
#pragma unroll
            for (int u = 0; u < 4; u++) pb[u] = c + u < len ? K.spos[b0 + c + u] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float dx = rx - pb[u].x, dy = ry - pb[u].y, dz = rz - pb[u].z;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
                if (c + u < len && r2 < 1.0f) {
                    ql[ring(fill) * FSI_THREADS] = (unsigned)(b0 + c + u);
                    fill++;
                }
            }
        }
    }
    while (wave_any(fill > 0)) drain_step();
    if (active) {
        float *o = sinst + 3 * (long)order[t];
        o[0] = sx;
        o[1] = sy;
        o[2] = sz;
    }
    if (pairs) {
        unsigned long long c = (unsigned long long)np;
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if ((threadIdx.x & 63) == 0 && c) atomicAdd(pairs, c);
    }
}

// vertex force += its own instance sum, then its images' in halo-list order
__global__ __launch_bounds__(256) void k_fsi_vertex_sum(int no, const int *flag, const int *aid, const float *sinst,
                                                        float *fo) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= no) return;
    float *o = fo + 3 * (long)v;
    float ox = o[0] + sinst[3 * (long)v], oy = o[1] + sinst[3 * (long)v + 1], oz = o[2] + sinst[3 * (long)v + 2];
    for (int i = 0; i < 26; i++) {
        if (!flag[(long)i * no + v]) continue;
        const float *s = sinst + 3 * ((long)no + aid[(long)i * no + v]);
        ox += s[0];
        oy += s[1];
        oz += s[2];
    }
    o[0] = ox;
    o[1] = oy;
    o[2] = oz;
}

// solvent side: -f of every instance within one cell, instances in cell
// order on the (L+2)^3 grid, stable within a cell
template <int WR, bool EXACT>
__global__ __launch_bounds__(FSI_THREADS) void k_fsi_solvent(FsiK K, int ns, const float *inst, const int2 *meta,
                                                             const int *istarts, float *fs, int fs_stride) {
    __shared__ unsigned q[FQ_S * FSI_THREADS];
    __shared__ float4 stage[FSI_THREADS / 64][FSI_SCAP];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = b < ns;
    float4 *sw = stage[threadIdx.x >> 6];
    float4 pb = make_float4(0.f, 0.f, 0.f, 0.f), vb = pb;
    int cx = 0, cy = 0, cz = 0;
    if (active) {
        pb = K.spos[b];
        vb = K.svel[b];
        cell_coord(pb.x, K.Lx, &cx);
        cell_coord(pb.y, K.Ly, &cy);
        cell_coord(pb.z, K.Lz, &cz);
    }
    const int Wx = K.Lx + 2, Wy = K.Ly + 2;
    const float sb = K.seeds[0], sh = K.seeds[1];
    float ax = 0.f, ay = 0.f, az = 0.f;
    int fill = 0, head = 0;
    unsigned *ql = q + threadIdx.x;
    auto ring = [&](int k) {
        const int r = head + k;
        return r >= FQ_S ? r - FQ_S : r;
    };
    auto in_of = [&](unsigned t) {
        FsiIn o;
        const float *qq = inst + 6 * (long)t;
        const int2 m = meta[t];
        o.dx = qq[0] - pb.x;
        o.dy = qq[1] - pb.y;
        o.dz = qq[2] - pb.z;
        o.r2 = __builtin_fmaf(o.dz, o.dz, __builtin_fmaf(o.dx, o.dx, o.dy * o.dy));
        o.dvx = qq[3] - vb.x;
        o.dvy = qq[4] - vb.y;
        o.dvz = qq[5] - vb.z;
        const unsigned lid = (unsigned)m.x, bu = (unsigned)b;
        o.x = (m.y ? sh : sb) - low_discrepancy((float)(lid & 0x3FFu), (float)lid, (float)(bu & 0x3FFu), (float)bu);
        o.pair = true;
        return o;
    };
    auto drain_step = [&]() {
        const bool va = fill > 0, vb2 = fill > 1;
        const unsigned ea = va ? ql[ring(0) * FSI_THREADS] : 0u;   // instance 0 exists (no > 0)
        const unsigned eb = vb2 ? ql[ring(1) * FSI_THREADS] : 0u;
        const FsiIn a = in_of(ea), c = in_of(eb);
        float xa = a.x, xc = c.x;
        logistic18x2(&xa, &xc);
        const float3 fa = dpd_pair<WR, EXACT>(K.par, a.dx, a.dy, a.dz, a.r2, a.dvx, a.dvy, a.dvz, xa * R_SQRT2);
        const float3 fc = dpd_pair<WR, EXACT>(K.par, c.dx, c.dy, c.dz, c.r2, c.dvx, c.dvy, c.dvz, xc * R_SQRT2);
        if (va) {
            ax -= fa.x;
            ay -= fa.y;
            az -= fa.z;
        }
        if (vb2) {
            ax -= fc.x;
            ay -= fc.y;
            az -= fc.z;
        }
        const int took = fill < 2 ? fill : 2;
        head = ring(took);
        fill -= took;
    };
    auto make_room = [&](int need) {
        if (wave_any(fill + need > FQ_S))
#pragma unroll 1
            for (int k = 0; k < FSI_DRAIN_STEPS; k++) drain_step();
    };
    for (int dz = 0; dz < 3; dz++)
        for (int dy = 0; dy < 3; dy++) {
            int t0 = 0, len = 0;
            if (active) {
                const int w0 = cx + Wx * ((cy + dy) + Wy * (cz + dz));
                t0 = istarts[w0];
                len = istarts[w0 + 3] - t0;
            }
            // the lanes are consecutive cell-ordered solvent particles: the
            // union of their instance runs usually fits one small range,
            // staged in LDS with coalesced loads
            int lo = active ? t0 : 0x7fffffff, hi = active ? t0 + len : 0;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                lo = min(lo, __shfl_xor(lo, d, 64));
                hi = max(hi, __shfl_xor(hi, d, 64));
            }
            if (hi > lo && hi - lo <= FSI_SCAP) {
                __builtin_amdgcn_wave_barrier();
                for (int k = threadIdx.x & 63; k < hi - lo; k += 64) {
                    const float *qq = inst + 6 * (long)(lo + k);
                    sw[k] = make_float4(qq[0], qq[1], qq[2], 0.f);
                }
                __builtin_amdgcn_wave_barrier();
                const int o = active ? t0 - lo : 0;
                for (int t = 0; wave_any(t < len); t += 4) {
                    make_room(4);
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const float4 c = sw[o + t + u < FSI_SCAP ? o + t + u : 0];
                        const float dx = c.x - pb.x, ddy = c.y - pb.y, ddz = c.z - pb.z;
                        const float r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(dx, dx, ddy * ddy));
                        if (t + u < len && r2 < 1.0f) {
                            ql[ring(fill) * FSI_THREADS] = (unsigned)(t0 + t + u);
                            fill++;
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                continue;
            }
            for (int t = 0; wave_any(t < len); t += 2) {
                if (wave_any(fill + 2 > FQ_S))
#pragma unroll 1
                    for (int k = 0; k < FSI_DRAIN_STEPS; k++) drain_step();
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const bool in = t + u < len;
                    const int tt = t0 + t + u;
                    float r2 = 2.0f;
                    if (in) {
                        const float *qq = inst + 6 * (long)tt;
                        const float dx = qq[0] - pb.x, ddy = qq[1] - pb.y, ddz = qq[2] - pb.z;
                        r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(dx, dx, ddy * ddy));
                    }
                    if (in && r2 < 1.0f) {
                        ql[ring(fill) * FSI_THREADS] = (unsigned)tt;
                        fill++;
                    }
                }
            }
        }
    while (wave_any(fill > 0)) drain_step();
    if (active) {
        float *o = fs + (long)fs_stride * b;
        o[0] += ax;
        o[1] += ay;
        o[2] += az;
    }
}

__global__ void k_gather_int2(int n, const int *n_dev, const int *order, const int2 *in, int2 *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || k >= *n_dev) return;
    out[k] = in[order[k]];
}

}  // namespace

int fsi_alloc(FsiBufs *f, const Box &b, int max_obj) {
    *f = FsiBufs{};
    f->box = b;
    f->max_obj = max_obj;
    const long ni = 8L * max_obj;    // vertices + at most 7 images each
    UDX_TRY(dev_alloc(&f->flag, 26L * max_obj));
    UDX_TRY(dev_alloc(&f->aid, 26L * max_obj + 1));
    UDX_TRY(dev_alloc(&f->scan_tmp, scan_scratch_size(26 * max_obj)));
    UDX_TRY(dev_alloc(&f->inst, 6 * ni));
    UDX_TRY(dev_alloc(&f->inst_sorted, 6 * ni));
    UDX_TRY(dev_alloc(&f->meta, ni));
    UDX_TRY(dev_alloc(&f->meta_sorted, ni));
    UDX_TRY(dev_alloc(&f->sinst, 3 * ni));
    Box m{{b.L[0] + 2, b.L[1] + 2, b.L[2] + 2}};
    UDX_TRY(clist_alloc(&f->cl, m, (int)ni));
    return UDX_OK;
}

void fsi_free(FsiBufs *f) {
    dev_free(f->flag);
    dev_free(f->aid);
    dev_free(f->scan_tmp);
    dev_free(f->inst);
    dev_free(f->inst_sorted);
    dev_free(f->meta);
    dev_free(f->meta_sorted);
    dev_free(f->sinst);
    clist_free(&f->cl);
}

int fsi_forces(hipStream_t s, FsiBufs *f, int ns, const float4 *spos, const float4 *svel, const int *starts,
               int no, const float *opp, const udx_pair_dpd &p, const float *seeds_dev, float *fo, float *fs,
               int fs_stride, unsigned long long *pairs, int exact) {
    const Box &b = f->box;
    UDX_CHECK(no >= 0 && no <= f->max_obj, UDX_ECAP, "fsi: %d membrane vertices above the capacity %d", no,
              f->max_obj);
    if (no == 0) return UDX_OK;
    const long g = 26L * no;
    hipLaunchKernelGGL(k_fsi_flags, dim3(ceil_div(g, 256)), dim3(256), 0, s, no, opp, b.L[0], b.L[1], b.L[2],
                       f->flag);
    UDX_LAUNCHED();
    UDX_TRY(scan_exclusive(s, f->flag, (int)g, f->aid, f->scan_tmp));
    hipLaunchKernelGGL(k_fsi_instances, dim3(ceil_div(g, 256)), dim3(256), 0, s, no, opp, b.L[0], b.L[1], b.L[2],
                       f->flag, f->aid, f->inst, f->meta, f->sinst);
    UDX_LAUNCHED();
    // at most 7 images per vertex (faces, edges, corner of its outside axes)
    const int nmax = 8 * no;
    UDX_TRY(clist_sort_aos(s, &f->cl, nmax, f->inst));
    UDX_TRY(gather_aos(s, &f->cl, nmax, f->inst, f->inst_sorted));
    const int nc = f->cl.box.ncells();
    hipLaunchKernelGGL(k_gather_int2, dim3(ceil_div(nmax, 256)), dim3(256), 0, s, nmax, f->cl.starts + nc,
                       f->cl.order, f->meta, f->meta_sorted);
    UDX_LAUNCHED();
    FsiK K;
    K.Lx = b.L[0];
    K.Ly = b.L[1];
    K.Lz = b.L[2];
    K.spos = spos;
    K.svel = svel;
    K.starts = starts;
    K.par = DpdPar{p.a, p.g, p.s, p.spow};
    K.seeds = seeds_dev;
#define UDX_FSI_LAUNCH(WR, EX)                                                                                  \
    do {                                                                                                        \
        hipLaunchKernelGGL((k_fsi_instance<WR, EX>), dim3(ceil_div(nmax, FSI_THREADS)), dim3(FSI_THREADS), 0,  \
                           s, K, f->cl.starts + nc, f->inst_sorted, f->meta_sorted, f->cl.order, f->sinst,      \
                           pairs);                                                                              \
        if (ns > 0)                                                                                             \
            hipLaunchKernelGGL((k_fsi_solvent<WR, EX>), dim3(ceil_div(ns, FSI_THREADS)), dim3(FSI_THREADS), 0,  \
                               s, K, ns, f->inst_sorted, f->meta_sorted, f->cl.starts, fs, fs_stride);          \
    } while (0)
    const float sp = p.spow;
    if (exact) {
        if (sp == 0.25f) UDX_FSI_LAUNCH(WR_QUARTER, true);
        else if (sp == 0.5f) UDX_FSI_LAUNCH(WR_HALF, true);
        else if (sp == 1.0f) UDX_FSI_LAUNCH(WR_ONE, true);
        else UDX_FSI_LAUNCH(WR_POW, true);
    } else {
        if (sp == 0.25f) UDX_FSI_LAUNCH(WR_QUARTER, false);
        else if (sp == 0.5f) UDX_FSI_LAUNCH(WR_HALF, false);
        else if (sp == 1.0f) UDX_FSI_LAUNCH(WR_ONE, false);
        else UDX_FSI_LAUNCH(WR_POW, false);
    }
#undef UDX_FSI_LAUNCH
    UDX_LAUNCHED();
    hipLaunchKernelGGL(k_fsi_vertex_sum, dim3(ceil_div(no, 256)), dim3(256), 0, s, no, f->flag, f->aid, f->sinst, fo);
    UDX_LAUNCHED();
    return UDX_OK;
}

}  // namespace udx

using namespace udx;

struct udx_fsi {
    udx_ctx *ctx;
    FsiBufs f;
    udx_rnd trunk;
    int mode, cap_s;
    float4 *spos, *svel;         // the bound solvent in float4 form
    const float *spp;            // bound solvent (AoS, cell ordered)
    float *sff;
    const int *starts;
    int ns;
    float *seeds_dev;
    unsigned long long *pairs_dev;
};

namespace udx {
FsiBufs *fsi_bufs(udx_fsi *q) { return &q->f; }
udx_ctx *fsi_ctx(const udx_fsi *q) { return q->ctx; }
float fsi_draw(udx_fsi *q) { return udx_rnd_get(&q->trunk); }
unsigned long long *fsi_pairs_dev(udx_fsi *q) { return q->pairs_dev; }
}  // namespace udx

extern "C" {

int udx_fsi_fin(udx_fsi *q) {
    if (!q) return UDX_OK;
    fsi_free(&q->f);
    dev_free(q->spos);
    dev_free(q->svel);
    dev_free(q->seeds_dev);
    dev_free(q->pairs_dev);
    delete q;
    return UDX_OK;
}

int udx_fsi_ini(udx_ctx *ctx, int Lx, int Ly, int Lz, int max_obj, udx_fsi **out) {
    UDX_CHECK(ctx && out && max_obj > 0, UDX_EINVAL, "udx_fsi_ini: bad arguments");
    UDX_CHECK(Lx >= 3 && Ly >= 3 && Lz >= 3 && Lx % 2 == 0 && Ly % 2 == 0 && Lz % 2 == 0, UDX_EINVAL,
              "udx_fsi_ini: the box sides must be even and at least 4 (got %d %d %d)", Lx, Ly, Lz);
    udx_fsi *q = new udx_fsi();
    q->ctx = ctx;
    udx_rnd_ini(1908, 1409, 290, 12968, &q->trunk);
    Box b{{Lx, Ly, Lz}};
    int s = fsi_alloc(&q->f, b, max_obj);
    if (!s) s = dev_alloc(&q->seeds_dev, 2);
    if (!s) s = dev_alloc(&q->pairs_dev, 1);
    if (!s && hipMemset(q->pairs_dev, 0, sizeof(unsigned long long)) != hipSuccess) s = UDX_EHIP;
    if (s) {
        udx_fsi_fin(q);
        return s;
    }
    *out = q;
    return UDX_OK;
}

int udx_fsi_set_mode(udx_fsi *q, int mode) {
    UDX_CHECK(q && (mode == UDX_MODE_FAST || mode == UDX_MODE_EXACT), UDX_EINVAL, "udx_fsi_set_mode: bad mode");
    q->mode = mode;
    return UDX_OK;
}

int udx_fsi_bind_solvent(udx_fsi *q, const float *pp, float *ff, int n, const int *starts) {
    UDX_CHECK(q && starts && (n == 0 || (pp && ff)), UDX_EINVAL, "udx_fsi_bind_solvent: null argument");
    if (n > q->cap_s) {
        dev_free(q->spos);
        dev_free(q->svel);
        UDX_TRY(dev_alloc(&q->spos, (size_t)n));
        UDX_TRY(dev_alloc(&q->svel, (size_t)n));
        q->cap_s = n;
    }
    q->spp = pp;
    q->sff = ff;
    q->ns = n;
    q->starts = starts;
    return UDX_OK;
}

int udx_fsi_apply_seeds(udx_fsi *q, const udx_pair_dpd *p, int no, const float *opp, float *off,
                        const float seeds[2]) {
    UDX_CHECK(q && p && seeds && (no == 0 || (opp && off)), UDX_EINVAL, "udx_fsi_apply_seeds: null argument");
    UDX_CHECK(q->starts, UDX_ESTATE, "udx_fsi_apply: bind the solvent first");
    hipStream_t s = q->ctx->stream;
    if (no == 0) return UDX_OK;
    UDX_HIP(hipMemcpyAsync(q->seeds_dev, seeds, 2 * sizeof(float), hipMemcpyHostToDevice, s));
    if (q->ns > 0) UDX_TRY(aos_to_f4(s, q->ns, q->spp, q->spos, q->svel));
    UDX_TRY(fsi_forces(s, &q->f, q->ns, q->spos, q->svel, q->starts, no, opp, *p, q->seeds_dev, off, q->sff, 3,
                       q->pairs_dev, q->mode == UDX_MODE_EXACT));
    // the seeds live on the host stack: wait for the copy before returning
    UDX_HIP(hipStreamSynchronize(s));
    return UDX_OK;
}

int udx_fsi_apply(udx_fsi *q, const udx_pair_dpd *p, int no, const float *opp, float *off) {
    UDX_CHECK(q, UDX_EINVAL, "udx_fsi_apply: null handle");
    float seeds[2];
    seeds[0] = udx_rnd_get(&q->trunk);
    seeds[1] = udx_rnd_get(&q->trunk);
    return udx_fsi_apply_seeds(q, p, no, opp, off, seeds);
}

int udx_fsi_pairs(udx_fsi *q, unsigned long long *npairs, int reset) {
    UDX_CHECK(q && npairs, UDX_EINVAL, "udx_fsi_pairs: null argument");
    hipStream_t s = q->ctx->stream;
    UDX_HIP(hipMemcpyAsync(npairs, q->pairs_dev, sizeof *npairs, hipMemcpyDeviceToHost, s));
    if (reset) UDX_HIP(hipMemsetAsync(q->pairs_dev, 0, sizeof(unsigned long long), s));
    UDX_HIP(hipStreamSynchronize(s));
    return UDX_OK;
}

}  // extern "C"
