// Cell list for MI355X: counting sort of particles into unit cells.
//
// The reference assigns each particle a slot with atomicAdd (so its in-cell
// order depends on thread timing, src/clist/dev/main.h:44-68) and builds the
// start table with a three-kernel uchar scan (src/algo/scan/dev.h).  Here the
// in-cell order is defined: by x slab (each cell cut into `sub` slabs along
// x), ties keeping the input order (a stable counting sort over slabs).  Atomics still hand out provisional slots; a per-cell insertion
// sort of the (few) slot owners makes the result independent of timing.
// Start/count tables come from a wavefront scan (64-lane shuffles, one LDS
// exchange per block).
#include "udx_dev.h"
#include "udx_internal.h"

namespace udx {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// block-wide inclusive scan of one value per thread (lds_waves: one entry
// per wave of the block)
__device__ __forceinline__ int block_inclusive_scan(int v, int *lds_waves) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_inclusive_scan(v);
    if (lane == 63) lds_waves[wave] = v;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wave; w++) add += lds_waves[w];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int *in, int n, int *part) {
    __shared__ int lw[4];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += (base + k < n) ? in[base + k] : 0;
    s = block_inclusive_scan(s, lw);
    if (threadIdx.x == SCAN_THREADS - 1) part[blockIdx.x] = s;
}

// Each tile's offset is the sum of the tile totals before it, added up by
// the tile's own block (integers, so the order of the additions does not
// matter): no launch between the two passes.  zero: the input is cleared
// after it is read (counts consumed by the scan are ready for the next
// counting pass).  total: out[n] = the total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(int *in, int n, const int *part, int nb, int *out,
                                                             int zero, int total) {
    __shared__ int lw[4];
    __shared__ int tile0;
    int pre = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += SCAN_THREADS) pre += part[k];
    pre = block_inclusive_scan(pre, lw);
    if (threadIdx.x == SCAN_THREADS - 1) tile0 = pre;
    __syncthreads();
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    if (zero)
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++)
            if (base + k < n) in[base + k] = 0;
    const int inc = block_inclusive_scan(s, lw);
    int run = tile0 + inc - s;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (total && blockIdx.x == nb - 1 && threadIdx.x == SCAN_THREADS - 1) out[n] = tile0 + inc;
}

// The whole exclusive scan in one block when n <= SCAN1_MAX (the two-kernel
// path costs a second launch, which dominates for the small arrays of a small
// box: slab counts of 16^3 cells, fragment counts, membrane and contact
// tables).  Wave w owns the entries [1024 w, 1024 (w + 1)) in 16 rounds of 64
// consecutive entries (lane l takes entry 64 k + l of round k), so every load
// and store is coalesced; the 16 rounds are scanned side by side across the
// wave, chained by their totals, and the waves by one pass through LDS.
constexpr int SCAN1_THREADS = 1024;
constexpr int SCAN1_ITEMS = 16;
constexpr int SCAN1_MAX = SCAN1_THREADS * SCAN1_ITEMS;

__global__ __launch_bounds__(SCAN1_THREADS) void k_scan_one(int *in, int n, int *out, int zero, int total) {
    __shared__ int lw[SCAN1_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = wave * 64 * SCAN1_ITEMS + lane;
    int x[SCAN1_ITEMS], v[SCAN1_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; k++) {
        const int i = base + 64 * k;
        x[k] = i < n ? in[i] : 0;
        v[k] = x[k];
    }
    if (zero)
#pragma unroll
        for (int k = 0; k < SCAN1_ITEMS; k++)
            if (base + 64 * k < n) in[base + 64 * k] = 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
#pragma unroll
        for (int k = 0; k < SCAN1_ITEMS; k++) {
            const int t = __shfl_up(v[k], d, 64);
            if (lane >= d) v[k] += t;
        }
    int carry = 0;   // the wave's entries before round k
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; k++) {
        const int tot = __builtin_amdgcn_readlane(v[k], 63);
        v[k] += carry;
        carry += tot;
    }
    if (lane == 0) lw[wave] = carry;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wave; w++) add += lw[w];
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; k++) {
        const int i = base + 64 * k;
        if (i < n) out[i] = add + v[k] - x[k];
    }
    if (total && threadIdx.x == SCAN1_THREADS - 1) out[n] = add + carry;
}

__global__ void k_fill_int(int *p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- counting sort --------------------------------------------------------
struct PosF4 {
    const float4 *p;
    __device__ float3 operator()(int i) const { float4 q = p[i]; return make_float3(q.x, q.y, q.z); }
};
struct PosAos {
    const float *p;
    __device__ float3 operator()(int i) const {
        const float *q = p + 6 * (long)i;
        return make_float3(q[0], q[1], q[2]);
    }
};

template <typename Pos>
__global__ __launch_bounds__(256) void k_clist_count(Pos pos, int n, int Lx, int Ly, int Lz, int sub, int *counts,
                                                     int2 *cellslot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float3 r = make_float3(0.f, 0.f, 0.f);
    if (i < n) r = pos(i);
    clist_count_one(i < n, i, r, Lx, Ly, Lz, sub, counts, cellslot);
}

// slot -> order; with x slabs (sub > 1) threads i <= ncells also write the
// cell tables from the slab starts, cstarts[c] = starts[c sub] and the counts
// (they need only the scan, so they take no launch of their own)
__global__ __launch_bounds__(256) void k_clist_scatter(int n, const int2 *cellslot, const int *starts,
                                                       int *order, int ncells, int sub, int *cstarts,
                                                       int *ccounts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (sub > 1 && i <= ncells) {
        const int s0 = starts[i * sub];
        cstarts[i] = s0;
        if (i < ncells) ccounts[i] = starts[(i + 1) * sub] - s0;
    }
    if (i >= n) return;
    int2 cs = cellslot[i];
    if (cs.x >= 0) order[starts[cs.x] + cs.y] = i;
}

// ascending bitonic network over the entries [s, e) of order (e - s <= N),
// in registers: one load and one store per entry
template <int N>
__device__ __forceinline__ void sort_net(int s, int e, int *order) {
    int v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = s + k < e ? order[s + k] : 0x7fffffff;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int k = 0; k < N; k++) {
                const int j = k ^ stride;
                if (j > k) {
                    const int lo = min(v[k], v[j]), hi = max(v[k], v[j]);
                    const bool up = (k & size) == 0;
                    v[k] = up ? lo : hi;
                    v[j] = up ? hi : lo;
                }
            }
#pragma unroll
    for (int k = 0; k < N; k++)
        if (s + k < e) order[s + k] = v[k];
}

// sort each cell's (slab's) input indices ascending: the stable order.  A
// wave sorts its cells with the network sized for its largest cell (2, 4, 8
// or 16 entries: the same for every lane; slabs hold a few particles, so the
// 16-entry network would mostly sort padding); cells above 16 by insertion
// in memory.
__global__ __launch_bounds__(256) void k_clist_cellsort(int ncells, const int *starts, int *order) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int s = 0, e = 0;
    if (c < ncells) {
        s = starts[c];
        e = starts[c + 1];
    }
    int m = e - s;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if (m <= 1) return;
    if (m <= 2) {
        sort_net<2>(s, e, order);
        return;
    }
    if (m <= 4) {
        sort_net<4>(s, e, order);
        return;
    }
    if (m <= 8) {
        sort_net<8>(s, e, order);
        return;
    }
    if (e - s <= 1) return;
    if (e - s <= 16) {
        sort_net<16>(s, e, order);
        return;
    }
    for (int k = s + 1; k < e; k++) {
        int v = order[k], w = k - 1;
        while (w >= s && order[w] > v) {
            order[w + 1] = order[w];
            w--;
        }
        order[w + 1] = v;
    }
}

__global__ __launch_bounds__(256) void k_gather_f4(int nmax, const int *nkept, const int *order,
                                                   const float4 *pin, const float4 *vin, float4 *pout,
                                                   float4 *vout) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmax || k >= *nkept) return;
    int i = order[k];
    pout[k] = pin[i];
    vout[k] = vin[i];
}

__global__ __launch_bounds__(256) void k_gather_aos(int nmax, const int *nkept, const int *order,
                                                    const float2 *pin, float2 *pout) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmax || k >= *nkept) return;
    long i = order[k];
    pout[3 * (long)k + 0] = pin[3 * i + 0];
    pout[3 * (long)k + 1] = pin[3 * i + 1];
    pout[3 * (long)k + 2] = pin[3 * i + 2];
}

}  // namespace

int scan_scratch_size(int n) {
    int nb = ceil_div(n, SCAN_TILE);
    return 2 * nb + 2;
}

static int scan_impl(hipStream_t s, int *in, int n, int *out, int *scratch, int zero, int total = 1) {
    if (n <= SCAN1_MAX) {
        hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(SCAN1_THREADS), 0, s, in, n, out, zero, total);
        UDX_LAUNCHED();
        return UDX_OK;
    }
    int nb = ceil_div(n, SCAN_TILE);
    if (nb == 0) nb = 1;
    int *part = scratch;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, part);
    UDX_LAUNCHED();
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, part, nb, out, zero, total);
    UDX_LAUNCHED();
    return UDX_OK;
}

int scan_exclusive(hipStream_t s, const int *in, int n, int *out, int *scratch) {
    return scan_impl(s, (int *)in, n, out, scratch, 0);
}

int scan_exclusive_consume(hipStream_t s, int *in, int n, int *out, int *scratch) {
    return scan_impl(s, in, n, out, scratch, 1);
}

int scan_exclusive_n(hipStream_t s, const int *in, int n, int *out, int *scratch) {
    return scan_impl(s, (int *)in, n, out, scratch, 0, 0);
}

int clist_alloc(ClistBufs *c, const Box &b, int maxp, int sub) {
    *c = ClistBufs{};
    c->box = b;
    c->maxp = maxp;
    c->sub = sub;
    UDX_CHECK(sub == 1 || sub == 2 || sub == 4 || sub == 8, UDX_EINVAL, "clist: %d slabs per cell (1, 2, 4 or 8)", sub);
    const int nc = b.ncells();
    UDX_CHECK((long)nc * sub < (1L << 30), UDX_ECAP, "clist: %d cells x %d slabs are too many", nc, sub);
    UDX_TRY(dev_alloc(&c->counts, nc));
    UDX_TRY(dev_alloc(&c->starts, nc + 1));
    if (sub > 1) {
        UDX_TRY(dev_alloc(&c->bcounts, (size_t)nc * sub));
        UDX_TRY(dev_alloc(&c->bstarts, (size_t)nc * sub + 1));
    } else {
        c->bcounts = c->counts;
        c->bstarts = c->starts;
    }
    UDX_TRY(dev_alloc(&c->cellslot, maxp));
    UDX_TRY(dev_alloc(&c->order, maxp));
    UDX_TRY(dev_alloc(&c->scan_tmp, scan_scratch_size(nc * sub)));
    return UDX_OK;
}

void clist_free(ClistBufs *c) {
    if (c->sub > 1) {
        dev_free(c->bcounts);
        dev_free(c->bstarts);
    }
    c->bcounts = c->bstarts = nullptr;
    dev_free(c->counts);
    dev_free(c->starts);
    dev_free(c->cellslot);
    dev_free(c->order);
    dev_free(c->scan_tmp);
}

template <typename Pos>
static int clist_sort_impl(hipStream_t s, ClistBufs *c, int n, Pos pos) {
    UDX_CHECK(n >= 0 && n <= c->maxp, UDX_ECAP, "clist: %d particles exceed capacity %d", n, c->maxp);
    const Box &b = c->box;
    UDX_TRY(clist_clear(s, c));
    if (n > 0) {
        hipLaunchKernelGGL(k_clist_count<Pos>, dim3(ceil_div(n, 256)), dim3(256), 0, s, pos, n, b.L[0],
                           b.L[1], b.L[2], c->sub, c->bcounts, c->cellslot);
        UDX_LAUNCHED();
    }
    return clist_finish(s, c, n);
}

int clist_clear(hipStream_t s, ClistBufs *c) {
    const int nb = c->box.ncells() * c->sub;
    hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(nb, 256)), dim3(256), 0, s, c->bcounts, nb, 0);
    UDX_LAUNCHED();
    return UDX_OK;
}

int clist_finish(hipStream_t s, ClistBufs *c, int n) {
    const int nc = c->box.ncells(), nb = nc * c->sub;
    // with x slabs the slab counts are read by the scan only: it clears them,
    // so a move kernel can count the next step's slabs (counts[] of the
    // cells come from the slab starts); without slabs counts[] is the output
    if (c->sub > 1) UDX_TRY(scan_exclusive_consume(s, c->bcounts, nb, c->bstarts, c->scan_tmp));
    else UDX_TRY(scan_exclusive(s, c->bcounts, nb, c->bstarts, c->scan_tmp));
    // one launch covers the particles and, with x slabs, the nc + 1 cell entries
    const int nt = c->sub > 1 ? (n > nc + 1 ? n : nc + 1) : n;
    if (nt > 0) {
        hipLaunchKernelGGL(k_clist_scatter, dim3(ceil_div(nt, 256)), dim3(256), 0, s, n, c->cellslot,
                           c->bstarts, c->order, nc, c->sub, c->starts, c->counts);
        UDX_LAUNCHED();
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_clist_cellsort, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, c->bstarts,
                           c->order);
        UDX_LAUNCHED();
    }
    return UDX_OK;
}

int clist_sort(hipStream_t s, ClistBufs *c, int n, const float4 *pos4) {
    return clist_sort_impl(s, c, n, PosF4{pos4});
}

int clist_count_only(hipStream_t s, ClistBufs *c, int n, const float4 *pos4) {
    UDX_CHECK(n >= 0 && n <= c->maxp, UDX_ECAP, "clist: %d particles exceed capacity %d", n, c->maxp);
    UDX_TRY(clist_clear(s, c));
    if (n > 0) {
        const Box &b = c->box;
        hipLaunchKernelGGL(k_clist_count<PosF4>, dim3(ceil_div(n, 256)), dim3(256), 0, s, PosF4{pos4}, n, b.L[0],
                           b.L[1], b.L[2], c->sub, c->bcounts, c->cellslot);
        UDX_LAUNCHED();
    }
    return UDX_OK;
}

int clist_sort_aos(hipStream_t s, ClistBufs *c, int n, const float *pp) {
    return clist_sort_impl(s, c, n, PosAos{pp});
}

int gather_f4(hipStream_t s, const ClistBufs *c, int n, const float4 *pin, const float4 *vin, float4 *pout,
              float4 *vout) {
    if (n <= 0) return UDX_OK;
    const int *nkept = c->starts + c->box.ncells();
    hipLaunchKernelGGL(k_gather_f4, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, nkept, c->order, pin, vin,
                       pout, vout);
    UDX_LAUNCHED();
    return UDX_OK;
}

int gather_aos(hipStream_t s, const ClistBufs *c, int n, const float *pin, float *pout) {
    if (n <= 0) return UDX_OK;
    const int *nkept = c->starts + c->box.ncells();
    hipLaunchKernelGGL(k_gather_aos, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, nkept, c->order,
                       (const float2 *)pin, (float2 *)pout);
    UDX_LAUNCHED();
    return UDX_OK;
}

}  // namespace udx
