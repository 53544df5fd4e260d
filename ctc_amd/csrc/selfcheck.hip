// Exhaustive checks of the device arithmetic the kernels rely on for bit
// equality with the CPU restatement: over every float of a range, the
// library's routine against the correctly rounded result, computed in double
// and rounded once to float (double rounding of a quotient or a square root
// from 53 to 24 bits is innocuous).  Counts the floats that differ.
#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {
namespace {

__global__ __launch_bounds__(256) void k_selfcheck(int what, uint32_t lo, uint32_t n, unsigned long long *bad) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    bool differ = false;
    if (k < n) {
        const float x = __uint_as_float(lo + k);
        float got, want;
        if (what == UDX_CHECK_RCP || what == UDX_CHECK_RCP_ESTIMATE) {
            got = what == UDX_CHECK_RCP ? rcp_rn(x) : __builtin_amdgcn_rcpf(x);
            want = (float)(1.0 / (double)x);
        } else if (what == UDX_CHECK_SQRT) {
            got = sqrt_rn(x);
            want = (float)__builtin_sqrt((double)x);
        } else {
            got = __builtin_amdgcn_fractf(x);
            want = x - __builtin_floorf(x);
        }
        differ = __float_as_uint(got) != __float_as_uint(want);
    }
    const unsigned long long c = __popcll(__ballot(differ));
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(bad, c);
}

}  // namespace
}  // namespace udx

using namespace udx;

extern "C" int udx_selfcheck(udx_ctx *ctx, int what, float lo, float hi, unsigned long long *mismatches) {
    UDX_CHECK(ctx && mismatches, UDX_EINVAL, "udx_selfcheck: null argument");
    UDX_CHECK(what == UDX_CHECK_RCP || what == UDX_CHECK_SQRT || what == UDX_CHECK_FRACT ||
                  what == UDX_CHECK_RCP_ESTIMATE,
              UDX_EINVAL,
              "udx_selfcheck: unknown check %d", what);
    UDX_CHECK(lo >= 0.0f && hi >= lo && hi < 3.0e38f, UDX_EINVAL, "udx_selfcheck: range [%g, %g]", (double)lo,
              (double)hi);
    const uint32_t a = __builtin_bit_cast(uint32_t, lo), b = __builtin_bit_cast(uint32_t, hi);
    const uint64_t n = (uint64_t)b - a + 1;
    UDX_CHECK(n < (1ull << 31), UDX_EINVAL, "udx_selfcheck: range of %llu floats is too large",
              (unsigned long long)n);
    unsigned long long *bad = nullptr;
    UDX_TRY(dev_alloc(&bad, 1));
    hipStream_t s = ctx->stream;
    int st = UDX_OK;
    if (hipMemsetAsync(bad, 0, sizeof *bad, s) != hipSuccess) st = UDX_EHIP;
    if (!st) {
        hipLaunchKernelGGL(k_selfcheck, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, what, a, (uint32_t)n,
                           bad);
        if (hipGetLastError() != hipSuccess) st = UDX_EHIP;
    }
    if (!st && hipMemcpyAsync(mismatches, bad, sizeof *bad, hipMemcpyDeviceToHost, s) != hipSuccess) st = UDX_EHIP;
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = UDX_EHIP;
    dev_free(bad);
    UDX_CHECK(!st, UDX_EHIP, "udx_selfcheck: HIP error");
    return UDX_OK;
}
