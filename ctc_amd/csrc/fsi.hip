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
//  - k_fsi_vertex: one thread per vertex, its three z-plane partial sums and
//    then its images in halo-list order;
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
// the instance cell list drops them.
__global__ __launch_bounds__(256) void k_fsi_instances(int no, const float *opp, int Lx, int Ly, int Lz,
                                                       const int *flag, const int *aid, float *inst, int2 *meta) {
    const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (g < no) {
        const float *p = opp + 6 * g;
        float *o = inst + 6 * g;
        for (int c = 0; c < 6; c++) o[c] = p[c];
        meta[g] = make_int2((int)g, 0);
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
}

struct FsiK {
    int Lx, Ly, Lz;
    const float4 *spos, *svel;
    const int *starts;
    DpdPar par;
    const float *seeds;      // device: [bulk, halo]
};

// sum over one z plane of the clipped block around r (src/fsi/dev/map/bulk.h)
template <int WR, bool EXACT>
__device__ int plane_sum(const FsiK &K, float rx, float ry, float rz, const float3 &v, int zplane, unsigned lid,
                         float seed, float *fx, float *fy, float *fz) {
    const int xc = K.Lx / 2 + (int)__builtin_floorf(rx);
    const int yc = K.Ly / 2 + (int)__builtin_floorf(ry);
    const int zc = K.Lz / 2 + (int)__builtin_floorf(rz);
    if (xc - 1 >= K.Lx || xc + 2 <= 0) return 0;
    const int xs = xc - 1 > 0 ? xc - 1 : 0, xe = xc + 2 < K.Lx ? xc + 2 : K.Lx;
    const int z = zc - 1 + zplane;
    if (z < 0 || z >= K.Lz) return 0;
    int np = 0;
    for (int k = 0; k < 3; k++) {
        const int y = yc - 1 + k;
        if (y < 0 || y >= K.Ly) continue;
        const int cid = xs + K.Lx * (y + K.Ly * z);
        const int b0 = K.starts[cid], b1 = K.starts[cid + (xe - xs)];
        for (int b = b0; b < b1; b++) {
            const float4 pb = K.spos[b], vb = K.svel[b];
            const float dx = rx - pb.x, dy = ry - pb.y, dz = rz - pb.z;
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
            if (r2 < 1.0f) {
                const float rnd = mean0var1uu(seed, lid, (unsigned)b);
                const float3 f = dpd_pair<WR, EXACT>(K.par, dx, dy, dz, r2, v.x - vb.x, v.y - vb.y, v.z - vb.z, rnd);
                *fx += f.x;
                *fy += f.y;
                *fz += f.z;
                np++;
            }
        }
    }
    return np;
}

template <int WR, bool EXACT>
__global__ __launch_bounds__(FSI_THREADS) void k_fsi_vertex(FsiK K, int no, const float *opp, const int *flag,
                                                            const int *aid, float *fo,
                                                            unsigned long long *pairs) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    int np = 0;
    if (v < no) {
        const float *p = opp + 6 * (long)v;
        const float3 u = make_float3(p[3], p[4], p[5]);
        const float sb = K.seeds[0], sh = K.seeds[1];
        float *o = fo + 3 * (long)v;
        float ox = o[0], oy = o[1], oz = o[2];
        for (int zp = 0; zp < 3; zp++) {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            np += plane_sum<WR, EXACT>(K, p[0], p[1], p[2], u, zp, (unsigned)v, sb, &sx, &sy, &sz);
            ox += sx;
            oy += sy;
            oz += sz;
        }
        const int ex = emap_axis(p[0], K.Lx), ey = emap_axis(p[1], K.Ly), ez = emap_axis(p[2], K.Lz);
        if (ex | ey | ez)
            for (int i = 0; i < 26; i++) {
                if (!flag[(long)i * no + v]) continue;
                const int dx = (i + 2) % 3 - 1, dy = (i / 3 + 2) % 3 - 1, dz = (i / 9 + 2) % 3 - 1;
                const float rx = p[0] + (float)(dx * K.Lx), ry = p[1] + (float)(dy * K.Ly),
                            rz = p[2] + (float)(dz * K.Lz);
                const unsigned lid = (unsigned)aid[(long)i * no + v];
                float sx = 0.f, sy = 0.f, sz = 0.f;
                for (int zp = 0; zp < 3; zp++)
                    np += plane_sum<WR, EXACT>(K, rx, ry, rz, u, zp, lid, sh, &sx, &sy, &sz);
                ox += sx;
                oy += sy;
                oz += sz;
            }
        o[0] = ox;
        o[1] = oy;
        o[2] = oz;
    }
    if (pairs) {
        unsigned long long c = (unsigned long long)np;
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if ((threadIdx.x & 63) == 0 && c) atomicAdd(pairs, c);
    }
}

// solvent side: -f of every instance within one cell, instances in cell
// order on the (L+2)^3 grid, stable within a cell
template <int WR, bool EXACT>
__global__ __launch_bounds__(FSI_THREADS) void k_fsi_solvent(FsiK K, int ns, const float *inst, const int2 *meta,
                                                             const int *istarts, float *fs, int fs_stride) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ns) return;
    const float4 pb = K.spos[b], vb = K.svel[b];
    int cx = 0, cy = 0, cz = 0;
    cell_coord(pb.x, K.Lx, &cx);
    cell_coord(pb.y, K.Ly, &cy);
    cell_coord(pb.z, K.Lz, &cz);
    const int Wx = K.Lx + 2, Wy = K.Ly + 2;
    const float sb = K.seeds[0], sh = K.seeds[1];
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int dz = 0; dz < 3; dz++)
        for (int dy = 0; dy < 3; dy++) {
            const int w0 = cx + Wx * ((cy + dy) + Wy * (cz + dz));
            const int t0 = istarts[w0], t1 = istarts[w0 + 3];
            for (int t = t0; t < t1; t++) {
                const float *q = inst + 6 * (long)t;
                const float dx = q[0] - pb.x, ddy = q[1] - pb.y, ddz = q[2] - pb.z;
                const float r2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(dx, dx, ddy * ddy));
                if (r2 < 1.0f) {
                    const int2 m = meta[t];
                    const float rnd = mean0var1uu(m.y ? sh : sb, (unsigned)m.x, (unsigned)b);
                    const float3 f = dpd_pair<WR, EXACT>(K.par, dx, ddy, ddz, r2, q[3] - vb.x, q[4] - vb.y,
                                                         q[5] - vb.z, rnd);
                    ax -= f.x;
                    ay -= f.y;
                    az -= f.z;
                }
            }
        }
    float *o = fs + (long)fs_stride * b;
    o[0] += ax;
    o[1] += ay;
    o[2] += az;
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
                       f->flag, f->aid, f->inst, f->meta);
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
        hipLaunchKernelGGL((k_fsi_vertex<WR, EX>), dim3(ceil_div(no, FSI_THREADS)), dim3(FSI_THREADS), 0, s, K, \
                           no, opp, f->flag, f->aid, fo, pairs);                                                \
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
