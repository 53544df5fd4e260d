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
    return s - __builtin_floorf(s);
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

// ---------------------------------------------------------------- DPD pair
// w_R = (1 - r)^spow for the exponents the kernels specialise
enum WrKind { WR_QUARTER = 0, WR_HALF = 1, WR_ONE = 2, WR_POW = 3 };

template <int WR, bool EXACT>
__device__ __forceinline__ float wr_pow(float rm, float spow) {
    if (WR == WR_ONE) return rm;
    if (WR == WR_HALF) return EXACT ? __builtin_sqrtf(rm) : __builtin_amdgcn_sqrtf(rm);
    if (WR == WR_QUARTER)
        return EXACT ? __builtin_sqrtf(__builtin_sqrtf(rm))
                     : __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(rm));
    return __builtin_powf(rm, spow);
}

struct DpdPar { float a, g, s, spow; };

// DPD force on a from b given dr = ra - rb, dv = va - vb and r2 = |dr|^2 < 1
// (src/pair/dev.h:12-47,110-139).  r is always the correctly rounded square
// root: w_R = (1 - r)^spow has an unbounded slope at the cutoff, so 1 - r
// must not carry approximation error.  EXACT also rounds 1/r and the
// w_R roots correctly (bitwise equal to the CPU restatement); otherwise the
// hardware reciprocal and square root (1 ulp, relative error only) are used.
template <int WR, bool EXACT>
__device__ __forceinline__ float3 dpd_pair(const DpdPar &p, float dx, float dy, float dz, float r2,
                                           float dvx, float dvy, float dvz, float rnd) {
    float ex, ey, ez, r;
    if (r2 < 1e-6f) {
        ex = ey = ez = r = 0.0f;
    } else {
        r = __builtin_sqrtf(r2);
        float invr = EXACT ? 1.0f / r : __builtin_amdgcn_rcpf(r);
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

}  // namespace udx
