// Device-side building blocks shared by the udx kernels (gfx950 only).
//
// Floating-point operation order is written out explicitly (the library is
// compiled with -ffp-contract=off), so that every correctly rounded result
// is the same bit pattern the CPU restatement computes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace udx {

// ---------------------------------------------------------------- branch RNG
// src/math/rnd/dev.h:8-12
__device__ constexpr float R_GOLD   = 0.6180339887498948482f;
__device__ constexpr float R_SILVER = 0.4142135623730950488f;
__device__ constexpr float R_BRONZE = 0.00008877875787352212838853023f;
__device__ constexpr float R_TIN    = 0.00004602357186447026756768986f;
__device__ constexpr float R_SQRT2  = 1.41421356237309514547f;

// One round of the Chebyshev map with both fma rounded toward zero
// (src/math/rnd/dev.h:33-47):  r = fma_rz(fma_rz(8, x2, -8), x2, 1).
// The single-precision rounding field of the MODE register (bits 1:0) is
// set to round-toward-zero around the two fma and restored to nearest-even;
// both mode writes sit in one asm block so no other floating-point
// instruction can be scheduled inside the round-toward-zero window.  The
// nops cover the VALU -> s_setreg and s_setreg -> s_setreg wait states.
__device__ __forceinline__ float t4_round_rz(float x2, float c8, float cm8) {
    float t, r;
    asm volatile(
        "s_nop 1\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 3\n\t"
        "s_nop 1\n\t"
        "v_fma_f32 %1, %3, %2, %4\n\t"
        "v_fma_f32 %0, %1, %2, 1.0\n\t"
        "s_nop 1\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 0\n\t"
        "s_nop 1"
        : "=v"(r), "=&v"(t)
        : "v"(x2), "v"(c8), "v"(cm8));
    return r;
}

// the same round for two independent chains sharing one round-toward-zero
// window (halves the mode switches and interleaves the dependent fma)
__device__ __forceinline__ void t4_round_rz2(float a2, float b2, float c8, float cm8, float *ra, float *rb) {
    float ta, tb, oa, ob;
    asm volatile(
        "s_nop 1\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 3\n\t"
        "s_nop 1\n\t"
        "v_fma_f32 %2, %6, %4, %7\n\t"
        "v_fma_f32 %3, %6, %5, %7\n\t"
        "v_fma_f32 %0, %2, %4, 1.0\n\t"
        "v_fma_f32 %1, %3, %5, 1.0\n\t"
        "s_nop 1\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 0\n\t"
        "s_nop 1"
        : "=&v"(oa), "=&v"(ob), "=&v"(ta), "=&v"(tb)
        : "v"(a2), "v"(b2), "v"(c8), "v"(cm8));
    *ra = oa;
    *rb = ob;
}

// nine rounds (N = 18), the square rounded to nearest
__device__ __forceinline__ float logistic18(float x) {
    const float c8 = 8.0f, cm8 = -8.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        float x2 = x * x;
        x = t4_round_rz(x2, c8, cm8);
    }
    return x;
}

__device__ __forceinline__ void logistic18x2(float *xa, float *xb) {
    const float c8 = 8.0f, cm8 = -8.0f;
    float a = *xa, b = *xb;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        float a2 = a * a, b2 = b * b;
        t4_round_rz2(a2, b2, c8, cm8, &a, &b);
    }
    *xa = a;
    *xb = b;
}

// p = frac((u&1023) gold + u bronze + (v&1023) silver + v tin), evaluated as
//   s = fma(u&1023, gold, u*bronze); s = fma(v&1023, silver, s); s = fma(v, tin, s)
__device__ __forceinline__ float low_discrepancy(float fu_lo, float fu, float fv_lo, float fv) {
    float s = __builtin_fmaf(fu_lo, R_GOLD, fu * R_BRONZE);
    s = __builtin_fmaf(fv_lo, R_SILVER, s);
    s = __builtin_fmaf(fv, R_TIN, s);
    return __builtin_amdgcn_fractf(s);   // s - floor(s), exact for s >= 0
}

// src/math/rnd/dev.h:57-63
__device__ __forceinline__ float mean0var1ii(float seed, int u, int v) {
    float p = low_discrepancy((float)(u & 0x3FF), (float)u, (float)(v & 0x3FF), (float)v);
    return logistic18(seed - p) * R_SQRT2;
}

// src/math/rnd/dev.h:65-72
__device__ __forceinline__ float mean0var1uu(float seed, uint32_t u, uint32_t v) {
    float p = low_discrepancy((float)(u & 0x3FFu), (float)u, (float)(v & 0x3FFu), (float)v);
    return logistic18(seed - p) * R_SQRT2;
}

// ---------------------------------------------------------------- cells
// floor(r + L/2), exact in double (src/clist/dev/main.h:21-24 for even L);
// returns false for positions outside [0, L)
__device__ __forceinline__ bool cell_coord(float r, int L, int *c) {
    double v = __builtin_floor((double)r + 0.5 * (double)L);
    if (!(v >= 0.0 && v < (double)L)) return false;
    *c = (int)v;
    return true;
}

// x slab floor((r + L/2) sub), exact in double (sub a power of two); false
// outside [0, L sub).  The slab's cell is slab / sub.
__device__ __forceinline__ bool slab_coord(float r, int L, int sub, int *c) {
    double v = __builtin_floor(((double)r + 0.5 * (double)L) * (double)sub);
    if (!(v >= 0.0 && v < (double)L * (double)sub)) return false;
    *c = (int)v;
    return true;
}

// where a move kernel counts the next step's x slabs (counts null: no count)
struct SlabCount {
    int L[3], sub;
    int *counts;
    int2 *cellslot;
    int slab0;   // the box's first slab (boxes end to end in one list)
};

// One particle of a counting pass (every lane of the wave calls it; valid
// selects the lanes with a particle): its x slab (cell * sub + slab) and a
// provisional slot in it, cellslot[i] = (slab or -1 outside the box, slot).
// Consecutive lanes with the same slab (the usual case: the particles are
// the previous step's cell-ordered array) form a segment that takes its slots
// with one atomic.
// slab0: the first slab of the particle's box (boxes laid end to end).
__device__ __forceinline__ void clist_count_one(bool valid, int i, float3 r, int Lx, int Ly, int Lz, int sub,
                                                int *counts, int2 *cellslot, int slab0 = 0) {
    const int lane = threadIdx.x & 63;
    int cid = -1;
    if (valid) {
        int bx, cy, cz;
        if (slab_coord(r.x, Lx, sub, &bx) && cell_coord(r.y, Ly, &cy) && cell_coord(r.z, Lz, &cz))
            cid = slab0 + bx + Lx * sub * (cy + Ly * cz);
    }
    const int prev = __shfl_up(cid, 1, 64);
    const bool head = lane == 0 || prev != cid;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int hlane = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int next = above ? __ffsll(above) - 1 : 64;
    int base = 0;
    if (head && cid >= 0) base = atomicAdd(counts + cid, next - lane);
    base = __shfl(base, hlane, 64);
    if (valid) cellslot[i] = make_int2(cid, base + (lane - hlane));
}

// ---------------------------------------------------------------- DPD pair
// Correctly rounded square root of a positive normal float (or 0): the
// hardware root (within 1 ulp) moved to a neighbour when an fma residual
// shows that neighbour is nearer.  The compiler's IEEE square root does the
// same after scaling denormal inputs, which the pair force never has.  Every
// x in [2^-24, 1] is checked on the host against each value a 1-ulp root
// can give (the oracle's orc_verify_sqrt_rn).
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1), su = __int_as_float(__float_as_int(s) + 1);
    const float rd = __builtin_fmaf(-sd, s, x), ru = __builtin_fmaf(-su, s, x);
    const float t = rd <= 0.f ? sd : s;
    return ru > 0.f ? su : t;
}

// Correctly rounded reciprocal of r in [2^-24, 1]: the hardware reciprocal
// (within 1 ulp) refined by one Newton step with fused multiply-adds.  The
// result equals the IEEE quotient 1/r for every float of that range (checked
// exhaustively on gfx950 against the double quotient rounded to float; the
// library's self-check entry point repeats the check), at 3 instructions
// instead of the 10 of a general division.
__device__ __forceinline__ float rcp_rn(float r) {
    const float y = __builtin_amdgcn_rcpf(r);
    const float e = __builtin_fmaf(-r, y, 1.0f);
    return __builtin_fmaf(y, e, y);
}

// w_R = (1 - r)^spow for the exponents the kernels specialise
enum WrKind { WR_QUARTER = 0, WR_HALF = 1, WR_ONE = 2, WR_POW = 3 };

template <int WR, bool EXACT>
__device__ __forceinline__ float wr_pow(float rm, float spow) {
    if (WR == WR_ONE) return rm;
    if (WR == WR_HALF) return sqrt_rn(rm);
    if (WR == WR_QUARTER) return sqrt_rn(sqrt_rn(rm));
    return __builtin_powf(rm, spow);
}

struct DpdPar { float a, g, s, spow; };

// DPD force on a from b given dr = ra - rb, dv = va - vb and r2 = |dr|^2 < 1
// (src/pair/dev.h:12-47,110-139).  r, 1/r and the w_R roots are correctly
// rounded, so the force equals the CPU restatement bit for bit (the
// reference's fast-math r2 rsqrtf(r2) is 1e-3 wrong in w_R near the cutoff,
// where w_R = (1 - r)^spow has an unbounded slope).  The template argument
// EXACT is kept for the generic exponent (powf).
template <int WR, bool EXACT>
__device__ __forceinline__ float3 dpd_pair(const DpdPar &p, float dx, float dy, float dz, float r2,
                                           float dvx, float dvy, float dvz, float rnd) {
    float ex, ey, ez, r;
    if (r2 < 1e-6f) {
        ex = ey = ez = r = 0.0f;
    } else {
        r = sqrt_rn(r2);
        const float invr = rcp_rn(r);   // r >= 1e-3
        ex = dx * invr; ey = dy * invr; ez = dz * invr;
    }
    float ev = __builtin_fmaf(ez, dvz, __builtin_fmaf(ex, dvx, ey * dvy));
    float rm = 1.0f - r;
    rm = rm > 0.0f ? rm : 0.0f;
    float wr = wr_pow<WR, EXACT>(rm, p.spow);
    float f0 = __builtin_fmaf(-p.g * wr, ev, p.s * rnd) * wr;
    f0 = __builtin_fmaf(p.a, rm, f0);
    return make_float3(f0 * ex, f0 * ey, f0 * ez);
}

// ---------------------------------------------------------------- integrator
// periodic wrap into [-L/2, L/2); a result rounded onto +L/2 maps to -L/2
__device__ __forceinline__ float wrap_coord(float x, float L) {
    const float h = 0.5f * L;
    if (x < -h) x = x + L;
    if (x >= h) x = x - L;
    return x;
}

// kick-drift with the constant body force folded in
// (src/scheme/move/dev/vv.h:1-10, src/scheme/force/dev/main.h:1-4)
__device__ __forceinline__ void kick_drift(float dt, float mass, float a, float f, float *r, float *v) {
    const float ft = __builtin_fmaf(mass, a, f);
    *v = __builtin_fmaf(ft / mass, dt, *v);
    *r = __builtin_fmaf(*v, dt, *r);
}

// ---------------------------------------------------------------- fragments
// src/frag/dev.h:19-36
__device__ __forceinline__ int frag_d2i(int x, int y, int z) {
    return ((x + 2) % 3) + 3 * ((y + 2) % 3) + 9 * ((z + 2) % 3);
}

// index of cell c inside the box of fragment direction d (x fastest;
// src/exch/flu/dev/map.h:1-20)
__device__ __forceinline__ int frag_cell(int dx, int dy, int dz, int cx, int cy, int cz, int Lx, int Ly) {
    int ex = dx ? 1 : Lx, ey = dy ? 1 : Ly;
    int hx = dx ? 0 : cx, hy = dy ? 0 : cy, hz = dz ? 0 : cz;
    return hx + ex * (hy + ey * hz);
}

// ---------------------------------------------------------------- waves
__device__ __forceinline__ bool wave_any(bool p) { return __any(p); }

// ---------------------------------------------------------------- SDF
// wall margin in cells (XWM/YWM/ZWM, src/inc/def.h:15-18)
constexpr int WALL_MARGIN = 6;

// texture of the signed distance (positive inside the walls) with the map
// from local position to texture coordinate t = s r + o
struct SdfView {
    const float *tex;
    int T[3];
    float o[3], s[3];
    float cheap_threshold;
};

// point fetch with wrap addressing (src/wall/sdf/tex3d/imp/main.h:12-29)
__device__ __forceinline__ float sdf_fetch(const SdfView &v, int i, int j, int k) {
    i %= v.T[0]; if (i < 0) i += v.T[0];
    j %= v.T[1]; if (j < 0) j += v.T[1];
    k %= v.T[2]; if (k < 0) k += v.T[2];
    return v.tex[i + v.T[0] * ((long)j + (long)v.T[1] * k)];
}

__device__ __forceinline__ void sdf_tex(const SdfView &v, float x, float y, float z, float t[3]) {
    t[0] = __builtin_fmaf(v.s[0], x, v.o[0]);
    t[1] = __builtin_fmaf(v.s[1], y, v.o[1]);
    t[2] = __builtin_fmaf(v.s[2], z, v.o[2]);
}

// trilinear interpolation between texel centres (src/wall/sdf/dev.h:79-112)
__device__ __forceinline__ float sdf_eval(const SdfView &v, float x, float y, float z) {
    float t[3], l[3];
    int c0[3];
    sdf_tex(v, x, y, z, t);
    for (int c = 0; c < 3; c++) {
        c0[c] = (int)t[c];
        l[c] = t[c] - (float)c0[c];
    }
    const float s000 = sdf_fetch(v, c0[0], c0[1], c0[2]), s001 = sdf_fetch(v, c0[0] + 1, c0[1], c0[2]);
    const float s010 = sdf_fetch(v, c0[0], c0[1] + 1, c0[2]), s011 = sdf_fetch(v, c0[0] + 1, c0[1] + 1, c0[2]);
    const float s100 = sdf_fetch(v, c0[0], c0[1], c0[2] + 1), s101 = sdf_fetch(v, c0[0] + 1, c0[1], c0[2] + 1);
    const float s110 = sdf_fetch(v, c0[0], c0[1] + 1, c0[2] + 1), s111 = sdf_fetch(v, c0[0] + 1, c0[1] + 1, c0[2] + 1);
    // A (1 - p) + p B with the first product fused into the sum
    auto wav = [](float A, float B, float p) { return __builtin_fmaf(A, 1 - p, p * B); };
    const float s00x = wav(s000, s001, l[0]), s01x = wav(s010, s011, l[0]);
    const float s10x = wav(s100, s101, l[0]), s11x = wav(s110, s111, l[0]);
    const float s0yx = wav(s00x, s01x, l[1]), s1yx = wav(s10x, s11x, l[1]);
    return wav(s0yx, s1yx, l[2]);
}

// forward difference (src/wall/sdf/dev.h:32-46)
__device__ __forceinline__ void sdf_grad(const SdfView &v, float x, float y, float z, float g[3]) {
    float t[3];
    sdf_tex(v, x, y, z, t);
    const int i = (int)t[0], j = (int)t[1], k = (int)t[2];
    const float s0 = sdf_fetch(v, i, j, k);
    g[0] = v.s[0] * (sdf_fetch(v, i + 1, j, k) - s0);
    g[1] = v.s[1] * (sdf_fetch(v, i, j + 1, k) - s0);
    g[2] = v.s[2] * (sdf_fetch(v, i, j, k + 1) - s0);
}

__device__ __forceinline__ void sdf_ugrad(const SdfView &v, float x, float y, float z, float g[3]) {
    sdf_grad(v, x, y, z, g);
    const float mag = __builtin_sqrtf(__builtin_fmaf(g[2], g[2], __builtin_fmaf(g[0], g[0], g[1] * g[1])));
    if (mag > 1e-6f) {
        g[0] = g[0] / mag;
        g[1] = g[1] / mag;
        g[2] = g[2] / mag;
    }
}

__device__ __forceinline__ int sdf_iround(float x) { return (x > 0.5f) ? (int)(x + 0.5f) : (int)(x - 0.5f); }

// nearest texel (src/wall/sdf/dev.h:61-75)
__device__ __forceinline__ float sdf_cheap(const SdfView &v, float x, float y, float z) {
    float t[3];
    sdf_tex(v, x, y, z, t);
    return sdf_fetch(v, sdf_iround(t[0]), sdf_iround(t[1]), sdf_iround(t[2]));
}

__device__ __forceinline__ bool sdf_far(const SdfView &v, float x, float y, float z) {
    return sdf_cheap(v, x, y, z) <= v.cheap_threshold - 1;
}

// wall velocity at r (src/wall/wvel/dev.h:1-56): resting, constant, shear
// (a time-dependent rate resolved on the host) or Hele-Shaw.  On one domain
// the centre coordinates are the local ones (src/coords/dev.h:17-33).
enum { WVEL_NONE = 0, WVEL_CSTE = 1, WVEL_SHEAR = 2, WVEL_HS = 3 };
struct WallVel {
    int type;
    float u[3];
    float gdot;
    int vdir, gdir, half;
    float hu, hh;
};

__device__ __forceinline__ void wvel_at(const WallVel &w, float x, float y, float z, float v[3]) {
    v[0] = v[1] = v[2] = 0.f;
    if (w.type == WVEL_CSTE) {
        v[0] = w.u[0];
        v[1] = w.u[1];
        v[2] = w.u[2];
    } else if (w.type == WVEL_SHEAR) {
        float d = w.gdir == 0 ? x : (w.gdir == 1 ? y : z);
        if (w.half) d = d < 0.f ? d : 0.f;
        const float s = d * w.gdot;
        if (w.vdir == 0) v[0] = s;
        else if (w.vdir == 1) v[1] = s;
        else v[2] = s;
    } else if (w.type == WVEL_HS) {
        const float r2inv = 1.f / (x * x + y * y);
        float hfac = w.hu * (1.f - z * z * 4 / (w.hh * w.hh));
        hfac = hfac > 0.f ? hfac : 0.f;
        v[0] = r2inv * hfac * x;
        v[1] = r2inv * hfac * y;
    }
}

// v = 2 u_wall(rw) - v (src/wall/wvel/dev.h:57-66)
__device__ __forceinline__ void bounce_vel(const WallVel &w, const float rw[3], float u[3]) {
    float vw[3];
    wvel_at(w, rw[0], rw[1], rw[2], vw);
    for (int c = 0; c < 3; c++) u[c] = 2.0f * vw[c] - u[c];
}

// bounce-back (src/wall/sdf/bounce/dev/main.h:1-114); returns true when the
// particle was inside and has been moved
__device__ __forceinline__ bool sdf_bounce(const SdfView &v, const WallVel &wv, float dt, float r[3], float u[3]) {
    if (sdf_cheap(v, r[0], r[1], r[2]) < v.cheap_threshold) return false;
    const float currsdf = sdf_eval(v, r[0], r[1], r[2]);
    if (!(currsdf >= 0)) return false;
    auto axpy = [](float a, const float *x, float *y) {
        for (int c = 0; c < 3; c++) y[c] = __builtin_fmaf(a, x[c], y[c]);
    };
    float r0[3] = {r[0], r[1], r[2]};
    axpy(-dt, u, r0);
    if (sdf_eval(v, r0[0], r0[1], r0[2]) >= 0) {
        // rescue: walk back along the gradient
        float g[3];
        sdf_ugrad(v, r[0], r[1], r[2], g);
        const float sdf0 = currsdf;
        axpy(-sdf0, g, r);
        for (int l = 8; l >= 1; --l) {
            if (sdf_eval(v, r[0], r[1], r[2]) < 0) {
                bounce_vel(wv, r, u);
                return true;
            }
            const float jump = 1.1f * sdf0 / (float)(1 << l);
            axpy(-jump, g, r);
        }
        return true;
    }
    // Newton iterations on sdf(r + t u) = 0, t in [-dt, 0]
    float t = 0;
    for (int l = 0; l < 2; l++) {
        float rc[3] = {r[0], r[1], r[2]}, g[3];
        axpy(t, u, rc);
        const float phi = sdf_eval(v, rc[0], rc[1], rc[2]);
        sdf_grad(v, rc[0], rc[1], rc[2], g);
        const float dphi = __builtin_fmaf(u[2], g[2], __builtin_fmaf(u[0], g[0], u[1] * g[1]));
        if (__builtin_fabsf(dphi) < 1e-6f) break;
        t -= phi / dphi;
        if (t < -dt) t = -dt;
        if (t > 0) t = 0;
    }
    float rw[3] = {r[0], r[1], r[2]};
    axpy(t, u, rw);
    bounce_vel(wv, rw, u);
    r[0] = rw[0]; r[1] = rw[1]; r[2] = rw[2];
    axpy(-t, u, r);
    if (sdf_eval(v, r[0], r[1], r[2]) >= 0) {
        r[0] = r0[0]; r[1] = r0[1]; r[2] = r0[2];
    }
    return true;
}

}  // namespace udx
