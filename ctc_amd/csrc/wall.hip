// SDF walls for MI355X: field resample, evaluation, wall-particle freeze,
// solvent-wall DPD, bounce-back and repulsion.
//
// Reference: the SDF is resampled once with a periodic cubic B-spline onto a
// texture grid with a 6-cell margin and read back with point-filtered,
// wrap-addressed fetches (src/wall/sdf/imp/{main,gen}.h,
// src/wall/sdf/field/imp/sample.h, src/wall/sdf/dev.h).  Here the "texture" is
// a plain array in HBM with the same wrap addressing; all setup runs on the
// GPU.  Frozen wall particles and their periodic images are cell-sorted on
// the margin box, as the reference's wall ticket (src/wall/imp/generate.h).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {

// ------------------------------------------------------------ transforms
// (src/math/tform/imp/main.h; host code, no contraction)
namespace {
void os1(float a0, float a1, float b0, float b1, float *o, float *s) {
    *o = (a1 * b0 - a0 * b1) / (b0 - a0);
    *s = (b1 - a1) / (b0 - a0);
}
void tvec(const float a0[3], const float a1[3], const float b0[3], const float b1[3], float o[3], float s[3]) {
    for (int c = 0; c < 3; c++) os1(a0[c], a1[c], b0[c], b1[c], &o[c], &s[c]);
}
void tconv(const float o[3], const float s[3], const float a[3], float b[3]) {
    for (int c = 0; c < 3; c++) b[c] = s[c] * a[c] + o[c];
}
void to_grid(const float lo[3], const float hi[3], const int n[3], float o[3], float s[3]) {
    float a1[3], b1[3];
    for (int c = 0; c < 3; c++) { a1[c] = -0.5f; b1[c] = n[c] - 0.5f; }
    tvec(lo, a1, hi, b1, o, s);
}
void from_grid(const float lo[3], const float hi[3], const int n[3], float o[3], float s[3]) {
    float a1[3], b1[3];
    for (int c = 0; c < 3; c++) { a1[c] = -0.5f; b1[c] = n[c] - 0.5f; }
    tvec(a1, lo, b1, hi, o, s);
}
void grid2grid(const float flo[3], const float fhi[3], const int fn[3], const float tlo[3], const float thi[3],
               const int tn[3], float o[3], float s[3]) {
    float o1[3], s1[3], o2[3], s2[3];
    float a0[3] = {0, 0, 0}, b0[3] = {1, 1, 1}, a1[3], b1[3], a2[3], b2[3];
    from_grid(flo, fhi, fn, o1, s1);
    to_grid(tlo, thi, tn, o2, s2);
    tconv(o1, s1, a0, a1);
    tconv(o1, s1, b0, b1);
    tconv(o2, s2, a1, a2);
    tconv(o2, s2, b1, b2);
    tvec(a0, a2, b0, b2, o, s);
}
int ceiln(int a, int b) { return (a + b - 1) / b; }
}  // namespace

void sdf_texture_size(const int L[3], int T[3]) {
    enum { PAD = 16 };
    const int M = WALL_MARGIN;
    const int x = L[0] + 2 * M, y = L[1] + 2 * M, z = L[2] + 2 * M;
    T[0] = PAD * PAD;
    T[1] = PAD * ceiln(ceiln(y * T[0], x), PAD);
    T[2] = PAD * ceiln(ceiln(z * T[0], x), PAD);
}

namespace {

// ------------------------------------------------------------ resample
// cubic B-spline weight (src/wall/sdf/field/imp/sample.h:28-36)
__device__ float spl(float x) {
    return x <= 0 ? 0.0f
         : x <= 1 ? x * x * x / 6
         : x <= 2 ? (x * ((12 - 3 * x) * x - 12) + 4) / 6
         : x <= 3 ? (x * (x * (3 * x - 24) + 60) - 44) / 6
         : x <= 4 ? (x * ((12 - x) * x - 48) + 64) / 6
         : 0.0f;
}

struct Grid {
    int N[3];
    float o[3], s[3];
};

// one texel per thread: the periodic B-spline sample of the scaled field at
// the texel's field-grid coordinate (src/wall/sdf/field/imp/sample.h:38-89)
__global__ __launch_bounds__(256) void k_sdf_resample(const float *D, Grid g, int Tx, int Ty, int Tz, float *tex) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)Tx * Ty * Tz) return;
    const int ix = (int)(t % Tx), iy = (int)((t / Tx) % Ty), iz = (int)(t / ((long)Tx * Ty));
    const float a[3] = {(float)ix, (float)iy, (float)iz};
    float r[3];
    for (int c = 0; c < 3; c++) r[c] = g.s[c] * a[c] + g.o[c];
    int anchor[3];
    float w[3][4];
    for (int c = 0; c < 3; c++) anchor[c] = (int)__builtin_floorf(r[c]);
    for (int c = 0; c < 3; c++)
        for (int i = 0; i < 4; i++) w[c][i] = spl(r[c] - (float)(anchor[c] - 1 + i) + 2);
    float tmp[4][4], partial[4];
    for (int sz = 0; sz < 4; sz++)
        for (int sy = 0; sy < 4; sy++) {
            float s = 0;
            const int gz = (sz - 1 + anchor[2] + g.N[2]) % g.N[2];
            const int gy = (sy - 1 + anchor[1] + g.N[1]) % g.N[1];
            for (int sx = 0; sx < 4; sx++) {
                const int gx = (sx - 1 + anchor[0] + g.N[0]) % g.N[0];
                s += w[0][sx] * D[gx + g.N[0] * (gy + (long)g.N[1] * gz)];
            }
            tmp[sz][sy] = s;
        }
    for (int sz = 0; sz < 4; sz++) {
        float s = 0;
        for (int sy = 0; sy < 4; sy++) s += w[1][sy] * tmp[sz][sy];
        partial[sz] = s;
    }
    float val = 0;
    for (int sz = 0; sz < 4; sz++) val += w[2][sz] * partial[sz];
    tex[t] = val;
}

__global__ void k_scale(float *D, long n, float sc) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) D[i] = D[i] * sc;
}

// ------------------------------------------------------------ freeze
// label (src/wall/sdf/label/dev/main.h:1-14): 0 bulk, 1 wall, 2 deep
__global__ __launch_bounds__(256) void k_wall_label(SdfView v, int n, const float *pp, int *bulk, int *wall) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = pp + 6 * (long)i;
    const float s = sdf_eval(v, p[0], p[1], p[2]);
    bulk[i] = s < 0 ? 1 : 0;
    wall[i] = (s >= 0 && s <= 2) ? 1 : 0;
}

// stable split by the exclusive scans of the flags
__global__ __launch_bounds__(256) void k_wall_split(int n, const float2 *pp, const int *bulk, const int *bs,
                                                    const int *wall, const int *ws, float2 *keep, float2 *frozen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 *dst = bulk[i] ? keep + 3 * (long)bs[i] : (wall[i] ? frozen + 3 * (long)ws[i] : nullptr);
    if (!dst) return;
    dst[0] = pp[3 * (long)i];
    dst[1] = pp[3 * (long)i + 1];
    dst[2] = pp[3 * (long)i + 2];
}

// periodic images in the margin box: flags for (bag j, particle i), bag j
// holding the images shifted by -d_j L (src/wall/exch/imp.cpp:14-120)
__global__ __launch_bounds__(256) void k_image_flags(int nw, const float *wall, int Lx, int Ly, int Lz, int *flag) {
    const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (g >= 26L * nw) return;
    const int j = (int)(g / nw), i = (int)(g % nw);
    const int dx = -((j + 2) % 3 - 1), dy = -((j / 3 + 2) % 3 - 1), dz = -((j / 9 + 2) % 3 - 1);
    const float *p = wall + 6 * (long)i;
    const float r[3] = {p[0] + (float)(dx * Lx), p[1] + (float)(dy * Ly), p[2] + (float)(dz * Lz)};
    const double lw[3] = {0.5 * (Lx + 2 * WALL_MARGIN), 0.5 * (Ly + 2 * WALL_MARGIN), 0.5 * (Lz + 2 * WALL_MARGIN)};
    bool in = true;
    for (int c = 0; c < 3; c++) in = in && (double)r[c] >= -lw[c] && (double)r[c] < lw[c];
    flag[g] = in ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_image_scatter(int nw, const float *wall, int Lx, int Ly, int Lz,
                                                       const int *flag, const int *pos, float *out) {
    const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (g >= 26L * nw || !flag[g]) return;
    const int j = (int)(g / nw), i = (int)(g % nw);
    const int dx = -((j + 2) % 3 - 1), dy = -((j / 3 + 2) % 3 - 1), dz = -((j / 9 + 2) % 3 - 1);
    const float *p = wall + 6 * (long)i;
    float *o = out + 6 * ((long)nw + pos[g]);
    o[0] = p[0] + (float)(dx * Lx);
    o[1] = p[1] + (float)(dy * Ly);
    o[2] = p[2] + (float)(dz * Lz);
    o[3] = p[3];
    o[4] = p[4];
    o[5] = p[5];
}

// ------------------------------------------------------------ wall forces
// solvent x frozen wall particles (src/wall/force/dev/main.h:1-50), one
// thread per cell-ordered solvent particle, 3x3x3 wall cells around it on
// the margin box; wall velocity 0; noise mean0var1ii(seed, solvent, wall).
// Hits are queued per lane in an LDS ring and drained two pairs per lane per
// step (in candidate order, so the sum is that of a sequential walk): tested
// inline, about one candidate in six passes and the whole wave would wait
// for the pair body at nearly every candidate.
constexpr int WALL_THREADS = 128, WQ = 16;
struct WallArgs {
    const float4 *wpos;
    const int *wstarts;
    int Wx, Wy, Wz;
    DpdPar par;
    const float *seedp;
    float *ff;
    int ff_stride;
    unsigned long long *pairs;
    int assign;
};

// particle a (valid or not: every lane of the wave calls it)
template <int WR, bool EXACT>
__device__ __forceinline__ void wall_one(const SdfView &v, const WallVel &wv, int a, bool valid, const float4 *pos,
                                         const float4 *vel, const WallArgs &W) {
    __shared__ unsigned q[WQ * WALL_THREADS];
    const float4 *const wpos = W.wpos;
    const int *const wstarts = W.wstarts;
    const int Wx = W.Wx, Wy = W.Wy, Wz = W.Wz;
    const DpdPar par = W.par;
    float *const ff = W.ff;
    int npair = 0;
    bool near = false;
    float4 pa = make_float4(0.f, 0.f, 0.f, 0.f), va = pa;
    if (valid) {
        pa = pos[a];
        near = !sdf_far(v, pa.x, pa.y, pa.z);
    }
    if (!wave_any(near)) return;
    if (near) va = vel[a];
    const float seed = *W.seedp;
    int b0[3] = {1, 1, 1};
    if (near) {
        const float pc[3] = {pa.x, pa.y, pa.z};
        const int W[3] = {Wx, Wy, Wz};
        for (int c = 0; c < 3; c++) {
            const int qq = (int)__builtin_floor((double)pc[c] + 0.5 * (double)W[c]);
            b0[c] = qq < 1 ? 1 : (qq > W[c] - 2 ? W[c] - 2 : qq);
        }
    }
    float fx = 0.f, fy = 0.f, fz = 0.f;
    int fill = 0, head = 0;
    unsigned *ql = q + threadIdx.x;
    auto ring = [&](int k) {
        const int r = head + k;
        return r >= WQ ? r - WQ : r;
    };
    const float fa_lo = (float)(a & 0x3FF), fa_hi = (float)a;
    auto drain_step = [&]() {
        const bool ok0 = fill > 0, ok1 = fill > 1;
        const unsigned e0 = ok0 ? ql[ring(0) * WALL_THREADS] : 0u;   // wall particle 0 exists
        const unsigned e1 = ok1 ? ql[ring(1) * WALL_THREADS] : e0;
        const float4 w0 = wpos[e0], w1 = wpos[e1];
        const float dx0 = pa.x - w0.x, dy0 = pa.y - w0.y, dz0 = pa.z - w0.z;
        const float dx1 = pa.x - w1.x, dy1 = pa.y - w1.y, dz1 = pa.z - w1.z;
        const float r20 = __builtin_fmaf(dz0, dz0, __builtin_fmaf(dx0, dx0, dy0 * dy0));
        const float r21 = __builtin_fmaf(dz1, dz1, __builtin_fmaf(dx1, dx1, dy1 * dy1));
        // mean0var1ii(seed, a, b) (src/math/rnd/dev.h:57-63), two chains together
        float x0 = seed - low_discrepancy(fa_lo, fa_hi, (float)((int)e0 & 0x3FF), (float)(int)e0);
        float x1 = seed - low_discrepancy(fa_lo, fa_hi, (float)((int)e1 & 0x3FF), (float)(int)e1);
        logistic18x2(&x0, &x1);
        // wall particle velocities at their positions (src/wall/force/dev/main.h:1-12)
        float u0[3], u1[3];
        wvel_at(wv, w0.x, w0.y, w0.z, u0);
        wvel_at(wv, w1.x, w1.y, w1.z, u1);
        const float3 f0 = dpd_pair<WR, EXACT>(par, dx0, dy0, dz0, r20, va.x - u0[0], va.y - u0[1], va.z - u0[2],
                                              x0 * R_SQRT2);
        const float3 f1 = dpd_pair<WR, EXACT>(par, dx1, dy1, dz1, r21, va.x - u1[0], va.y - u1[1], va.z - u1[2],
                                              x1 * R_SQRT2);
        if (ok0) {
            fx += f0.x;
            fy += f0.y;
            fz += f0.z;
        }
        if (ok1) {
            fx += f1.x;
            fy += f1.y;
            fz += f1.z;
        }
        const int took = fill < 2 ? fill : 2;
        npair += took;
        head = ring(took);
        fill -= took;
    };
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            int c0 = 0, len = 0;
            if (near) {
                c0 = (b0[0] - 1) + Wx * ((b0[1] + dy) + Wy * (b0[2] + dz));
                c0 = wstarts[c0];
                len = wstarts[(b0[0] - 1) + Wx * ((b0[1] + dy) + Wy * (b0[2] + dz)) + 3] - c0;
            }
            for (int t = 0; wave_any(t < len); t += 4) {
                if (wave_any(fill + 4 > WQ))
#pragma unroll 1
                    for (int k = 0; k < 2; k++) drain_step();
                float4 pw[4];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    pw[u] = t + u < len ? wpos[c0 + t + u] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float dx = pa.x - pw[u].x, dyy = pa.y - pw[u].y, dzz = pa.z - pw[u].z;
                    const float r2 = __builtin_fmaf(dzz, dzz, __builtin_fmaf(dx, dx, dyy * dyy));
                    if (t + u < len && r2 < 1.0f) {
                        ql[ring(fill) * WALL_THREADS] = (unsigned)(c0 + t + u);
                        fill++;
                    }
                }
            }
        }
    while (wave_any(fill > 0)) drain_step();
    if (near) {
        if (W.ff_stride == 4) {
            float4 *f4 = (float4 *)ff + a;
            if (W.assign) {
                *f4 = make_float4(fx, fy, fz, 0.f);
            } else {
                const float4 o = *f4;
                *f4 = make_float4(o.x + fx, o.y + fy, o.z + fz, o.w);
            }
        } else {
            float *f3 = ff + 3 * (long)a;
            f3[0] += fx;
            f3[1] += fy;
            f3[2] += fz;
        }
    }
    if (W.pairs) {
        unsigned long long c = (unsigned long long)npair;
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if ((threadIdx.x & 63) == 0 && c) atomicAdd(W.pairs, c);
    }
}

template <int WR, bool EXACT>
__global__ __launch_bounds__(WALL_THREADS) void k_wall_force(SdfView v, WallVel wv, int nmax, const int *n_dev,
                                                             const float4 *pos, const float4 *vel, WallArgs W) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    wall_one<WR, EXACT>(v, wv, a, a < nmax && a < *n_dev, pos, vel, W);
}

// the same over the cell-ordered particles of the runs of cells that can
// hold a particle near a wall (SdfField::segs): one block per run
template <int WR, bool EXACT>
__global__ __launch_bounds__(WALL_THREADS) void k_wall_force_cells(SdfView v, WallVel wv, const int2 *segs,
                                                                   const int *starts, const float4 *pos,
                                                                   const float4 *vel, WallArgs W) {
    const int2 sg = segs[blockIdx.x];
    const int a0 = starts[sg.x], a1 = starts[sg.y];
    for (int base = a0; base < a1; base += WALL_THREADS) {
        const int a = base + (int)threadIdx.x;
        wall_one<WR, EXACT>(v, wv, a, a < a1, pos, vel, W);
    }
}

// Cells of the box that can hold a particle that is not sdf_far: the texels
// sdf_cheap can pick for a position in cell c (the rounded texture coordinates
// of the cell's float bounds, monotone in the position) hold a value above
// cheap_threshold - 1.
__global__ __launch_bounds__(256) void k_sdf_near_cells(SdfView v, int Lx, int Ly, int Lz, int *flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Lx * Ly * Lz) return;
    const int cc[3] = {c % Lx, (c / Lx) % Ly, c / (Lx * Ly)};
    const int L[3] = {Lx, Ly, Lz};
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        const float xlo = (float)((double)cc[d] - 0.5 * (double)L[d]);
        const float xup = (float)((double)(cc[d] + 1) - 0.5 * (double)L[d]);   // exact (a half-integer)
        const unsigned ub = __float_as_uint(xup);
        // the float below xup
        const float xhi = xup > 0.f ? __uint_as_float(ub - 1u) : (xup < 0.f ? __uint_as_float(ub + 1u) : -1e-45f);
        lo[d] = sdf_iround(__builtin_fmaf(v.s[d], xlo, v.o[d]));
        hi[d] = sdf_iround(__builtin_fmaf(v.s[d], xhi, v.o[d]));
    }
    const float far = v.cheap_threshold - 1;
    int near = 0;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) near |= !(sdf_fetch(v, i, j, k) <= far);
    flag[c] = near;
}

__global__ __launch_bounds__(256) void k_wall_bounce_aos(SdfView v, WallVel wv, float dt, int n, float *pp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *p = pp + 6 * (long)i;
    float r[3] = {p[0], p[1], p[2]}, u[3] = {p[3], p[4], p[5]};
    if (sdf_bounce(v, wv, dt, r, u)) {
        p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
        p[3] = u[0]; p[4] = u[1]; p[5] = u[2];
    }
}

// src/wall/force/dev/main.h:52-91
__global__ __launch_bounds__(256) void k_wall_repulse(SdfView v, float lambda, int n, const float *pp, float *ff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = pp + 6 * (long)i;
    if (sdf_far(v, p[0], p[1], p[2])) return;
    const float s = sdf_eval(v, p[0], p[1], p[2]);
    float g[3];
    sdf_ugrad(v, p[0], p[1], p[2], g);
    const float val = expf(lambda * (s + 1.f)) - 1.f;
    const float f0 = -(val < 1e4f ? val : 1e4f);
    float *f = ff + 3 * (long)i;
    f[0] += f0 * g[0];
    f[1] += f0 * g[1];
    f[2] += f0 * g[2];
}

__global__ void k_sdf_probe(SdfView v, int n, const float *r3, int what, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = r3[3 * i], y = r3[3 * i + 1], z = r3[3 * i + 2];
    if (what == 0) out[i] = sdf_eval(v, x, y, z);
    else if (what == 1) out[i] = sdf_cheap(v, x, y, z);
    else {
        float g[3];
        sdf_grad(v, x, y, z, g);
        out[3 * i] = g[0];
        out[3 * i + 1] = g[1];
        out[3 * i + 2] = g[2];
    }
}

__global__ void k_aos_pos4(int n, const float *pp, float4 *pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = pp + 6 * (long)i;
    pos[i] = make_float4(p[0], p[1], p[2], 0.f);
}

}  // namespace

// ------------------------------------------------------------ host side
// the runs of consecutive cells (at most SEG_CELLS each) that can hold a
// particle near a wall, for the wall force of cell-ordered particles
constexpr int SEG_CELLS = 16;
static int sdf_near_segments(hipStream_t s, SdfField *q) {
    const int nc = q->L[0] * q->L[1] * q->L[2];
    dev_free(q->segs);
    q->segs = nullptr;
    q->nseg = 0;
    int *flag = nullptr;
    UDX_TRY(dev_alloc(&flag, nc));
    std::vector<int> h(nc);
    hipLaunchKernelGGL(k_sdf_near_cells, dim3(ceil_div(nc, 256)), dim3(256), 0, s, q->view, q->L[0], q->L[1], q->L[2],
                       flag);
    int st = hipGetLastError() == hipSuccess ? UDX_OK : UDX_EHIP;
    if (!st && (hipMemcpyAsync(h.data(), flag, sizeof(int) * (size_t)nc, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        st = UDX_EHIP;
    dev_free(flag);
    if (st) {
        set_error("sdf: near-wall cells failed");
        return st;
    }
    std::vector<int2> seg;
    for (int c = 0; c < nc;) {
        if (!h[c]) {
            c++;
            continue;
        }
        int e = c;
        while (e < nc && h[e] && e - c < SEG_CELLS) e++;
        seg.push_back(make_int2(c, e));
        c = e;
    }
    if (seg.empty()) return UDX_OK;
    UDX_TRY(dev_alloc(&q->segs, (long)seg.size()));
    UDX_HIP(hipMemcpyAsync(q->segs, seg.data(), sizeof(int2) * seg.size(), hipMemcpyHostToDevice, s));
    UDX_HIP(hipStreamSynchronize(s));
    q->nseg = (int)seg.size();
    return UDX_OK;
}

int sdf_build(hipStream_t s, SdfField *q, const float ext[3], const int N[3], const float *D_host) {
    const long n0 = (long)N[0] * N[1] * N[2];
    UDX_CHECK(n0 > 0, UDX_EINVAL, "sdf: empty field");
    float *D = nullptr;
    UDX_TRY(dev_alloc(&D, n0));
    int st = UDX_OK;
    if (hipMemcpyAsync(D, D_host, n0 * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess) st = UDX_EHIP;
    // field_scale by xdomain / ext[x] (src/wall/sdf/imp/gen.h:36-46)
    const float sc = (float)q->L[0] / ext[0];
    if (!st) hipLaunchKernelGGL(k_scale, dim3(ceil_div(n0, 256)), dim3(256), 0, s, D, n0, sc);
    // texture grid -> field grid (src/wall/sdf/tform/imp.cpp:74-88)
    const int M = WALL_MARGIN;
    float tlo[3], thi[3], slo[3] = {0, 0, 0}, shi[3];
    Grid g;
    for (int c = 0; c < 3; c++) {
        tlo[c] = (float)(0 - M);
        thi[c] = (float)(q->L[c] + M);
        shi[c] = (float)q->L[c];
        g.N[c] = N[c];
    }
    grid2grid(tlo, thi, q->T, slo, shi, N, g.o, g.s);
    const long nt = (long)q->T[0] * q->T[1] * q->T[2];
    if (!st) hipLaunchKernelGGL(k_sdf_resample, dim3(ceil_div(nt, 256)), dim3(256), 0, s, D, g, q->T[0], q->T[1],
                                q->T[2], q->tex);
    if (!st && hipGetLastError() != hipSuccess) st = UDX_EHIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = UDX_EHIP;
    dev_free(D);
    if (st) {
        set_error("sdf: resample failed");
        return st;
    }
    // local coordinates -> texture grid (sub2tex_ini, :101-117)
    float lo[3], hi[3];
    for (int c = 0; c < 3; c++) {
        lo[c] = (float)(-q->L[c] / 2 - M);
        hi[c] = (float)(q->L[c] / 2 + M);
    }
    to_grid(lo, hi, q->T, q->view.o, q->view.s);
    q->view.cheap_threshold = -sqrtf(3.f) * ((float)(q->L[0] + 2 * M) / (float)q->T[0]);
    UDX_TRY(sdf_near_segments(s, q));
    q->ready = 1;
    return UDX_OK;
}

int sdf_probe(hipStream_t s, const SdfField *q, int n, const float *r3, int what, float *out) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_sdf_probe, dim3(ceil_div(n, 256)), dim3(256), 0, s, q->view, n, r3, what, out);
    UDX_LAUNCHED();
    return UDX_OK;
}

// freeze: pp (n, AoS) keeps the bulk particles in order (n updated), the
// frozen ones and their images go to the wall particle set, which is then
// cell-sorted on the margin box
int wall_build(hipStream_t s, WallSet *w, const SdfField *q, int *n, float *pp) {
    const int n0 = *n;
    int *flags = nullptr, *scan = nullptr, *tmp = nullptr;
    float *keep = nullptr, *frozen = nullptr;
    int st = UDX_OK;
    const int nf = n0 > 0 ? n0 : 1;
    st = dev_alloc(&flags, 2L * nf);
    if (!st) st = dev_alloc(&scan, 2L * (nf + 1));
    if (!st) st = dev_alloc(&tmp, scan_scratch_size(nf));
    if (!st) st = dev_alloc(&keep, 6L * nf);
    if (!st) st = dev_alloc(&frozen, 6L * nf);
    int nb = 0, nw = 0;
    if (!st && n0 > 0) {
        hipLaunchKernelGGL(k_wall_label, dim3(ceil_div(n0, 256)), dim3(256), 0, s, q->view, n0, pp, flags, flags + nf);
        st = scan_exclusive(s, flags, n0, scan, tmp);
        if (!st) st = scan_exclusive(s, flags + nf, n0, scan + nf + 1, tmp);
        if (!st) {
            hipLaunchKernelGGL(k_wall_split, dim3(ceil_div(n0, 256)), dim3(256), 0, s, n0, (const float2 *)pp, flags,
                               scan, flags + nf, scan + nf + 1, (float2 *)keep, (float2 *)frozen);
            if (hipMemcpyAsync(&nb, scan + n0, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(&nw, scan + nf + 1 + n0, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                st = UDX_EHIP;
        }
        if (!st && hipMemcpyAsync(pp, keep, 6L * nb * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = UDX_EHIP;
    }
    dev_free(keep);
    dev_free(flags);
    dev_free(scan);
    dev_free(tmp);
    if (st) {
        dev_free(frozen);
        set_error("wall: freeze failed");
        return st;
    }
    *n = nb;
    return wall_build_frozen(s, w, q, nw, frozen);
}

// the frozen particles (nw, AoS, device; owned by w from here on) -> periodic
// images and the wall cell list on the margin box
int wall_build_frozen(hipStream_t s, WallSet *w, const SdfField *q, int nw, float *frozen) {
    int st = UDX_OK;
    w->frozen = frozen;
    w->nfrozen = nw;
    // images
    const long ni = 26L * nw;
    int *iflag = nullptr, *ipos = nullptr, *itmp = nullptr;
    float *all = nullptr;
    int nimg = 0;
    if (ni > 0) {
        st = dev_alloc(&iflag, ni);
        if (!st) st = dev_alloc(&ipos, ni + 1);
        if (!st) st = dev_alloc(&itmp, scan_scratch_size((int)ni));
        if (!st) {
            hipLaunchKernelGGL(k_image_flags, dim3(ceil_div(ni, 256)), dim3(256), 0, s, nw, frozen, q->L[0], q->L[1],
                               q->L[2], iflag);
            st = scan_exclusive(s, iflag, (int)ni, ipos, itmp);
        }
        if (!st && (hipMemcpyAsync(&nimg, ipos + ni, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess))
            st = UDX_EHIP;
    }
    const int ntot = nw + nimg;
    if (!st) st = dev_alloc(&all, 6L * (ntot > 0 ? ntot : 1));
    if (!st && nw > 0) {
        if (hipMemcpyAsync(all, frozen, 6L * nw * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = UDX_EHIP;
        if (!st && nimg > 0)
            hipLaunchKernelGGL(k_image_scatter, dim3(ceil_div(ni, 256)), dim3(256), 0, s, nw, frozen, q->L[0],
                               q->L[1], q->L[2], iflag, ipos, all);
    }
    dev_free(iflag);
    dev_free(ipos);
    dev_free(itmp);
    if (st) {
        dev_free(all);
        set_error("wall: images failed");
        return st;
    }
    // wall cell list on the margin box
    Box wb;
    for (int c = 0; c < 3; c++) wb.L[c] = q->L[c] + 2 * WALL_MARGIN;
    w->n = ntot;
    st = clist_alloc(&w->cl, wb, ntot > 0 ? ntot : 1);
    float *sorted = nullptr;
    if (!st) st = dev_alloc(&sorted, 6L * (ntot > 0 ? ntot : 1));
    if (!st) st = dev_alloc(&w->pos, ntot > 0 ? ntot : 1);
    if (!st) st = clist_sort_aos(s, &w->cl, ntot, all);
    if (!st) st = gather_aos(s, &w->cl, ntot, all, sorted);
    if (!st && ntot > 0) {
        hipLaunchKernelGGL(k_aos_pos4, dim3(ceil_div(ntot, 256)), dim3(256), 0, s, ntot, sorted, w->pos);
        if (hipGetLastError() != hipSuccess) st = UDX_EHIP;
    }
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = UDX_EHIP;
    w->pp_aos = sorted;
    dev_free(all);
    if (st) set_error("wall: cell list failed");
    return st;
}

void wall_free(WallSet *w) {
    dev_free(w->frozen);
    clist_free(&w->cl);
    dev_free(w->pos);
    dev_free(w->pp_aos);
}

int wall_forces(hipStream_t s, const SdfField *q, const WallSet *w, int nmax, const int *n_dev, const float4 *pos,
                const float4 *vel, const udx_pair_dpd &p, const float *seed_dev, float *ff, int ff_stride,
                unsigned long long *pairs, int exact, const WallVel &wv, int assign, const int *cell_starts) {
    if (nmax <= 0 || w->n <= 0) return UDX_OK;
    const Box &b = w->cl.box;
    const WallArgs W{w->pos, w->cl.starts, b.L[0], b.L[1], b.L[2], DpdPar{p.a, p.g, p.s, p.spow}, seed_dev, ff,
                     ff_stride, pairs, assign};
    // cell-ordered particles: only the runs of cells that can hold a particle
    // near a wall
    const bool cells = cell_starts && q->segs;
    if (cells && q->nseg == 0) return UDX_OK;
    const dim3 grid(cells ? q->nseg : ceil_div(nmax, WALL_THREADS)), blk(WALL_THREADS);
#define UDX_WALL_LAUNCH(WR, EX)                                                                                \
    do {                                                                                                       \
        if (cells)                                                                                             \
            hipLaunchKernelGGL((k_wall_force_cells<WR, EX>), grid, blk, 0, s, q->view, wv, q->segs, cell_starts, \
                               pos, vel, W);                                                                   \
        else                                                                                                   \
            hipLaunchKernelGGL((k_wall_force<WR, EX>), grid, blk, 0, s, q->view, wv, nmax, n_dev, pos, vel, W);   \
    } while (0)
    const float sp = p.spow;
    if (exact) {
        if (sp == 0.25f) UDX_WALL_LAUNCH(WR_QUARTER, true);
        else if (sp == 0.5f) UDX_WALL_LAUNCH(WR_HALF, true);
        else if (sp == 1.0f) UDX_WALL_LAUNCH(WR_ONE, true);
        else UDX_WALL_LAUNCH(WR_POW, true);
    } else {
        if (sp == 0.25f) UDX_WALL_LAUNCH(WR_QUARTER, false);
        else if (sp == 0.5f) UDX_WALL_LAUNCH(WR_HALF, false);
        else if (sp == 1.0f) UDX_WALL_LAUNCH(WR_ONE, false);
        else UDX_WALL_LAUNCH(WR_POW, false);
    }
#undef UDX_WALL_LAUNCH
    UDX_LAUNCHED();
    return UDX_OK;
}

WallVel wall_vel_view(const udx_wall_vel &v, float t) {
    WallVel w{};
    w.type = WVEL_NONE;
    switch (v.type) {
        case UDX_WVEL_CSTE:
            w.type = WVEL_CSTE;
            for (int c = 0; c < 3; c++) w.u[c] = v.u[c];
            break;
        case UDX_WVEL_SHEAR:
        case UDX_WVEL_SHEAR_SIN:
            w.type = WVEL_SHEAR;
            w.gdot = v.type == UDX_WVEL_SHEAR ? v.gdot : v.gdot * sinf(v.w * t);
            w.vdir = v.vdir;
            w.gdir = v.gdir;
            w.half = v.half;
            break;
        case UDX_WVEL_HS:
            w.type = WVEL_HS;
            w.hu = v.hs_u;
            w.hh = v.hs_h;
            break;
        default:
            break;
    }
    return w;
}

int wall_bounce_aos(hipStream_t s, const SdfField *q, float dt, int n, float *pp, const WallVel &wv) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_wall_bounce_aos, dim3(ceil_div(n, 256)), dim3(256), 0, s, q->view, wv, dt, n, pp);
    UDX_LAUNCHED();
    return UDX_OK;
}

int wall_repulse_aos(hipStream_t s, const SdfField *q, float lambda, int n, const float *pp, float *ff) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_wall_repulse, dim3(ceil_div(n, 256)), dim3(256), 0, s, q->view, lambda, n, pp, ff);
    UDX_LAUNCHED();
    return UDX_OK;
}

}  // namespace udx
