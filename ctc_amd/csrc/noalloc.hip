// udx_clist_build_noalloc: the solvent cell list with no device allocation,
// for callers that cannot allocate (code under HIP graph capture, serving
// loops).  The caller passes every output; the working arrays (slab counts
// and starts, provisional slots, scan partials) live in one file-scope
// device array, `scratch`, that each GPU holds once for the whole process.
//
// The array's name is a host-side stand-in in host code; its device address
// is looked up once per device with hipGetSymbolAddress and only that
// address is handed to kernels.  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.  The slab counts are cleared by
// a kernel at the start of each call (no memset node if the caller captures
// the call into a graph).
#include <mutex>

#include "udx_internal.h"

namespace udx {
namespace {

constexpr int NA_SLABS = UDX_NOALLOC_MAXCELLS * SOLVENT_SUB;
// layout in ints, each part 256-byte aligned
constexpr long NA_ALIGN = 64;
constexpr long na_up(long x) { return (x + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
constexpr long NA_BCOUNTS = 0;
constexpr long NA_BSTARTS = NA_BCOUNTS + na_up(NA_SLABS);
constexpr long NA_CELLSLOT = NA_BSTARTS + na_up(NA_SLABS + 1);
constexpr long NA_SCAN = NA_CELLSLOT + na_up(2L * UDX_NOALLOC_MAXP);
constexpr long NA_SCAN_INTS = 2L * ((NA_SLABS + 1023) / 1024) + 2;   // scan_scratch_size(NA_SLABS)
constexpr long NA_TOTAL = NA_SCAN + na_up(NA_SCAN_INTS);

static __device__ __align__(256) int scratch[NA_TOTAL];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_lock;
int *g_scratch[NA_MAX_DEVICES] = {};

// the device address of `scratch` on `dev` (looked up on first use, with
// that device current)
int scratch_of(int dev, int **out) {
    UDX_CHECK(dev >= 0 && dev < NA_MAX_DEVICES, UDX_EINVAL, "udx_clist_build_noalloc: device %d", dev);
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch[dev]) {
        int cur = 0;
        UDX_HIP(hipGetDevice(&cur));
        if (cur != dev) UDX_HIP(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) UDX_HIP(hipSetDevice(cur));
        UDX_CHECK(e == hipSuccess && p, UDX_EHIP, "udx_clist_build_noalloc: hipGetSymbolAddress: %s",
                  hipGetErrorString(e));
        g_scratch[dev] = (int *)p;
    }
    *out = g_scratch[dev];
    return UDX_OK;
}

}  // namespace
}  // namespace udx

using namespace udx;

extern "C" int udx_clist_build_noalloc(udx_ctx *ctx, int Lx, int Ly, int Lz, int n, const float *pp, float *ppout,
                                       int *starts, int *counts, int *order) {
    UDX_CHECK(ctx && (n == 0 || pp) && starts && counts && (n == 0 || order), UDX_EINVAL,
              "udx_clist_build_noalloc: null argument");
    UDX_CHECK(Lx > 0 && Ly > 0 && Lz > 0 && n >= 0, UDX_EINVAL, "udx_clist_build_noalloc: bad box %d %d %d or n %d",
              Lx, Ly, Lz, n);
    const long nc = (long)Lx * Ly * Lz;
    UDX_CHECK(nc <= UDX_NOALLOC_MAXCELLS, UDX_ECAP,
              "udx_clist_build_noalloc: %ld cells exceed the static working space (%d cells)", nc,
              UDX_NOALLOC_MAXCELLS);
    UDX_CHECK(n <= UDX_NOALLOC_MAXP, UDX_ECAP,
              "udx_clist_build_noalloc: %d particles exceed the static working space (%d particles)", n,
              UDX_NOALLOC_MAXP);
    UDX_CHECK(scan_scratch_size(NA_SLABS) <= NA_SCAN_INTS, UDX_ECAP,
              "udx_clist_build_noalloc: the scan needs %d ints of working space, %ld reserved",
              scan_scratch_size(NA_SLABS), NA_SCAN_INTS);
    int *base = nullptr;
    UDX_TRY(scratch_of(ctx->device, &base));
    ClistBufs c{};
    c.box = Box{{Lx, Ly, Lz}};
    c.maxp = n;
    c.sub = SOLVENT_SUB;
    c.counts = counts;
    c.starts = starts;
    c.bcounts = base + NA_BCOUNTS;
    c.bstarts = base + NA_BSTARTS;
    c.cellslot = (int2 *)(base + NA_CELLSLOT);
    c.order = order;
    c.scan_tmp = base + NA_SCAN;
    hipStream_t s = ctx->stream;
    UDX_TRY(clist_sort_aos(s, &c, n, pp));
    if (ppout) UDX_TRY(gather_aos(s, &c, n, pp, ppout));
    return UDX_OK;
}
