// Integrator, body force, periodic wrap, layout conversion and FP64
// diagnostics for MI355X.  All are one pass over the particle streams and
// bound by HBM bandwidth.
#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {

namespace {

// With sc.counts set (and the wrap) the kernel also counts the new
// positions into the next step's x slabs, as the cell list's counting pass
// would: the next sort starts from its scan.
__global__ __launch_bounds__(256) void k_move_f4(int nmax, const int *n_dev, float Lx, float Ly, float Lz,
                                                 float dt, float mass, float ax, float ay, float az,
                                                 const float4 *ff, float4 *pos, float4 *vel, int wrap, SlabCount sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nmax && i < *n_dev;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        p = pos[i];
        float4 v = vel[i];
        const float4 f = ff[i];
        kick_drift(dt, mass, ax, f.x, &p.x, &v.x);
        kick_drift(dt, mass, ay, f.y, &p.y, &v.y);
        kick_drift(dt, mass, az, f.z, &p.z, &v.z);
        if (wrap) {
            p.x = wrap_coord(p.x, Lx);
            p.y = wrap_coord(p.y, Ly);
            p.z = wrap_coord(p.z, Lz);
        }
        pos[i] = p;
        vel[i] = v;
    }
    if (sc.counts) clist_count_one(valid, i, make_float3(p.x, p.y, p.z), sc.L[0], sc.L[1], sc.L[2], sc.sub, sc.counts,
                                   sc.cellslot);
}

// the same with the wall bounce-back between the drift and the wrap (the
// reference bounces after update_solvent, before the next redistribution,
// src/sim/imp/run.h:43-54)
__global__ __launch_bounds__(256) void k_move_f4_wall(int nmax, const int *n_dev, float Lx, float Ly, float Lz,
                                                      float dt, float mass, float ax, float ay, float az,
                                                      const float4 *ff, float4 *pos, float4 *vel, SdfView sdf,
                                                      WallVel wv, int wrap, SlabCount sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nmax && i < *n_dev;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        p = pos[i];
        float4 v = vel[i];
        const float4 f = ff[i];
        kick_drift(dt, mass, ax, f.x, &p.x, &v.x);
        kick_drift(dt, mass, ay, f.y, &p.y, &v.y);
        kick_drift(dt, mass, az, f.z, &p.z, &v.z);
        float r[3] = {p.x, p.y, p.z}, u[3] = {v.x, v.y, v.z};
        if (sdf_bounce(sdf, wv, dt, r, u)) {
            p.x = r[0]; p.y = r[1]; p.z = r[2];
            v.x = u[0]; v.y = u[1]; v.z = u[2];
        }
        if (wrap) {
            p.x = wrap_coord(p.x, Lx);
            p.y = wrap_coord(p.y, Ly);
            p.z = wrap_coord(p.z, Lz);
        }
        pos[i] = p;
        vel[i] = v;
    }
    if (sc.counts) clist_count_one(valid, i, make_float3(p.x, p.y, p.z), sc.L[0], sc.L[1], sc.L[2], sc.sub, sc.counts,
                                   sc.cellslot);
}

// reference kick-drift on AoS particles, no wrap (src/scheme/move/dev/main.h:1-12)
__global__ __launch_bounds__(256) void k_move_aos(int n, float dt, float mass, const float *ff, float *pp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *r = pp + 6 * (long)i, *v = r + 3;
    const float *f = ff + 3 * (long)i;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = __builtin_fmaf(f[c] / mass, dt, v[c]);
        r[c] = __builtin_fmaf(v[c], dt, r[c]);
    }
}

__global__ __launch_bounds__(256) void k_bforce_aos(int n, float mass, float ax, float ay, float az, float *ff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *f = ff + 3 * (long)i;
    f[0] = __builtin_fmaf(mass, ax, f[0]);
    f[1] = __builtin_fmaf(mass, ay, f[1]);
    f[2] = __builtin_fmaf(mass, az, f[2]);
}

__global__ __launch_bounds__(256) void k_wrap_aos(int n, float Lx, float Ly, float Lz, float *pp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *r = pp + 6 * (long)i;
    r[0] = wrap_coord(r[0], Lx);
    r[1] = wrap_coord(r[1], Ly);
    r[2] = wrap_coord(r[2], Lz);
}

__global__ __launch_bounds__(256) void k_aos_to_f4(int n, const float2 *pp, float4 *pos, float4 *vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = pp[3 * (long)i], b = pp[3 * (long)i + 1], c = pp[3 * (long)i + 2];
    pos[i] = make_float4(a.x, a.y, b.x, 0.f);
    vel[i] = make_float4(b.y, c.x, c.y, 0.f);
}

__global__ __launch_bounds__(256) void k_f4_to_aos(int nmax, const int *n_dev, const float4 *pos,
                                                   const float4 *vel, float2 *pp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nmax || (n_dev && i >= *n_dev)) return;
    const float4 p = pos[i], v = vel[i];
    pp[3 * (long)i] = make_float2(p.x, p.y);
    pp[3 * (long)i + 1] = make_float2(p.z, v.x);
    pp[3 * (long)i + 2] = make_float2(v.y, v.z);
}

__global__ __launch_bounds__(256) void k_f4_to_force(int nmax, const int *n_dev, const float4 *ff, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nmax || (n_dev && i >= *n_dev)) return;
    const float4 f = ff[i];
    out[3 * (long)i] = f.x;
    out[3 * (long)i + 1] = f.y;
    out[3 * (long)i + 2] = f.z;
}

// FP64 per-block partial sums (momentum, 0.5|v|^2, max |v|^2), reduced in a
// fixed order by k_diag_final so the result does not depend on timing
constexpr int DIAG_THREADS = 256;
constexpr int DIAG_BLOCKS = 1024;

__global__ __launch_bounds__(DIAG_THREADS) void k_diag_part(int n, const float *pp, double *part) {
    __shared__ double s[5][DIAG_THREADS];
    double m0 = 0, m1 = 0, m2 = 0, ke = 0, mx = 0;
    for (long i = blockIdx.x * (long)DIAG_THREADS + threadIdx.x; i < n; i += (long)gridDim.x * DIAG_THREADS) {
        const float *v = pp + 6 * i + 3;
        const double vx = v[0], vy = v[1], vz = v[2];
        m0 += vx;
        m1 += vy;
        m2 += vz;
        const double v2 = vx * vx + vy * vy + vz * vz;
        ke += 0.5 * v2;
        mx = v2 > mx ? v2 : mx;
    }
    const int t = threadIdx.x;
    s[0][t] = m0; s[1][t] = m1; s[2][t] = m2; s[3][t] = ke; s[4][t] = mx;
    __syncthreads();
    for (int w = DIAG_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int c = 0; c < 4; c++) s[c][t] += s[c][t + w];
            s[4][t] = s[4][t + w] > s[4][t] ? s[4][t + w] : s[4][t];
        }
        __syncthreads();
    }
    if (t < 5) part[5 * blockIdx.x + t] = s[t][0];
}

__global__ void k_diag_final(int nb, const double *part, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < nb; b++) {
        for (int c = 0; c < 4; c++) a[c] += part[5 * b + c];
        a[4] = part[5 * b + 4] > a[4] ? part[5 * b + 4] : a[4];
    }
    for (int c = 0; c < 5; c++) out[c] = a[c];
}

__global__ __launch_bounds__(256) void k_rnd_branch(float seed, int n, const int *u, const int *v, int uu,
                                                    float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = uu ? mean0var1uu(seed, (uint32_t)u[i], (uint32_t)v[i]) : mean0var1ii(seed, u[i], v[i]);
}

}  // namespace

int rnd_branch(hipStream_t s, float seed, int n, const int *u, const int *v, int uu, float *out) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_rnd_branch, dim3(ceil_div(n, 256)), dim3(256), 0, s, seed, n, u, v, uu, out);
    UDX_LAUNCHED();
    return UDX_OK;
}

int aos_to_f4(hipStream_t s, int n, const float *pp, float4 *pos, float4 *vel) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_aos_to_f4, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, (const float2 *)pp, pos, vel);
    UDX_LAUNCHED();
    return UDX_OK;
}

int f4_to_aos(hipStream_t s, int n, const int *n_dev, const float4 *pos, const float4 *vel, float *pp) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_f4_to_aos, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, n_dev, pos, vel, (float2 *)pp);
    UDX_LAUNCHED();
    return UDX_OK;
}

int f4_to_force_aos(hipStream_t s, int n, const int *n_dev, const float4 *ff, float *out) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_f4_to_force, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, n_dev, ff, out);
    UDX_LAUNCHED();
    return UDX_OK;
}

static SlabCount slab_count(ClistBufs *next, int wrap) {
    SlabCount c{};
    if (next && wrap) {
        for (int k = 0; k < 3; k++) c.L[k] = next->box.L[k];
        c.sub = next->sub;
        c.counts = next->bcounts;
        c.cellslot = next->cellslot;
    }
    return c;
}

int move_f4(hipStream_t s, int nmax, const int *n_dev, const Box &b, float dt, float mass, const float accel[3],
            const float4 *ff, float4 *pos, float4 *vel, int wrap, ClistBufs *next) {
    if (nmax <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_move_f4, dim3(ceil_div(nmax, 256)), dim3(256), 0, s, nmax, n_dev, (float)b.L[0],
                       (float)b.L[1], (float)b.L[2], dt, mass, accel[0], accel[1], accel[2], ff, pos, vel, wrap,
                       slab_count(next, wrap));
    UDX_LAUNCHED();
    return UDX_OK;
}

int move_f4_wall(hipStream_t s, int nmax, const int *n_dev, const Box &b, float dt, float mass, const float accel[3],
                 const float4 *ff, float4 *pos, float4 *vel, const SdfField *q, const WallVel &wv, int wrap,
                 ClistBufs *next) {
    if (!q) return move_f4(s, nmax, n_dev, b, dt, mass, accel, ff, pos, vel, wrap, next);
    if (nmax <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_move_f4_wall, dim3(ceil_div(nmax, 256)), dim3(256), 0, s, nmax, n_dev, (float)b.L[0],
                       (float)b.L[1], (float)b.L[2], dt, mass, accel[0], accel[1], accel[2], ff, pos, vel, q->view,
                       wv, wrap, slab_count(next, wrap));
    UDX_LAUNCHED();
    return UDX_OK;
}

int move_aos(hipStream_t s, int n, float dt, float mass, const float *ff, float *pp) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_move_aos, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, dt, mass, ff, pp);
    UDX_LAUNCHED();
    return UDX_OK;
}

int bforce_aos(hipStream_t s, int n, float mass, const float accel[3], float *ff) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_bforce_aos, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, mass, accel[0], accel[1],
                       accel[2], ff);
    UDX_LAUNCHED();
    return UDX_OK;
}

int wrap_aos(hipStream_t s, const Box &b, int n, float *pp) {
    if (n <= 0) return UDX_OK;
    hipLaunchKernelGGL(k_wrap_aos, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, (float)b.L[0], (float)b.L[1],
                       (float)b.L[2], pp);
    UDX_LAUNCHED();
    return UDX_OK;
}

int diag_aos(hipStream_t s, int n, const float *pp, double *partials_dev, double out[5]) {
    const int nb = n > 0 ? (ceil_div(n, DIAG_THREADS) < DIAG_BLOCKS ? ceil_div(n, DIAG_THREADS) : DIAG_BLOCKS) : 1;
    hipLaunchKernelGGL(k_diag_part, dim3(nb), dim3(DIAG_THREADS), 0, s, n, pp, partials_dev);
    UDX_LAUNCHED();
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, nb, partials_dev, partials_dev + 5 * DIAG_BLOCKS);
    UDX_LAUNCHED();
    double sums[5];
    UDX_HIP(hipMemcpyAsync(sums, partials_dev + 5 * DIAG_BLOCKS, sizeof sums, hipMemcpyDeviceToHost, s));
    UDX_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < 3; c++) out[c] = sums[c];
    out[3] = n > 0 ? sums[3] / (1.5 * n) : 0.0;
    out[4] = sums[4];
    return UDX_OK;
}

}  // namespace udx
