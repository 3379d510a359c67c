// Device-wide primitives used inside the detection pipeline: fill, exclusive scan and a stable LSD
// radix sort of (uint64 key, int32 value) pairs.  They are plain kernel launches -- no memset nodes, no
// host synchronisation, no allocation -- so the whole pipeline can be captured in a HIP graph.
//
// Radix sort: 8-bit digits, one pass per digit.  A pass is three launches:
//   1. per-tile digit histograms (LDS atomics), stored digit-major: hist[digit * ntiles + tile];
//   2. an exclusive scan of that array, giving each (digit, tile) its first output slot;
//   3. a scatter that re-reads the tile in rounds of 256 keys and ranks each key among the keys of its
//      round with the same digit: within a wavefront by matching digits with 8 ballots (the lanes
//      whose digit bits all agree), across the block's 4 wavefronts through LDS counters.  Rounds are
//      visited in order and lanes within a round in index order, so equal digits keep their input
//      order and the sort is stable.
#include <algorithm>
#include <utility>

#include "ccd_internal.h"

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 8;                          // keys per thread per tile
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;       // 2048 keys per tile
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 2048 values per scan block

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <class T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ p, int64_t n, T v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- exclusive scan (int64 running sums of int32 or int64 inputs) ----------------------------
// Block-level exclusive scan of SCAN_TILE values in LDS; returns the block total.
template <class T>
__device__ __forceinline__ int64_t block_scan_tile(const T *__restrict__ in, T *__restrict__ out, int64_t n,
                                                   int64_t base, int64_t *lds)
{
    int64_t v[SCAN_ITEMS];
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
        v[k] = i < n ? (int64_t)in[i] : 0;
        sum += v[k];
    }
    lds[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele over the 256 thread sums.
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        int64_t a = (int)threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += a;
        __syncthreads();
    }
    int64_t total = lds[SCAN_THREADS - 1];
    int64_t run = lds[threadIdx.x] - sum;   // exclusive prefix of this thread's items
    if (out) {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
            if (i < n) out[i] = (T)run;
            run += v[k];
        }
    }
    __syncthreads();
    return total;
}

template <class T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const T *__restrict__ in, T *__restrict__ out, int64_t n,
                                                             int64_t *__restrict__ tile_sums)
{
    __shared__ int64_t lds[SCAN_THREADS];
    int64_t total = block_scan_tile<T>(in, out, n, (int64_t)blockIdx.x * SCAN_TILE, lds);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block scans the tile sums (any count) sequentially in chunks of SCAN_TILE.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(int64_t *__restrict__ sums, int64_t ntiles,
                                                            int64_t *__restrict__ grand_total)
{
    __shared__ int64_t lds[SCAN_THREADS];
    int64_t carry = 0;
    for (int64_t base = 0; base < ntiles; base += SCAN_TILE) {
        int64_t t = block_scan_tile<int64_t>(sums, sums, ntiles, base, lds);
        // add the carry of earlier chunks to this chunk
        for (int64_t i = base + threadIdx.x; i < base + SCAN_TILE && i < ntiles; i += SCAN_THREADS) sums[i] += carry;
        __syncthreads();
        carry += t;
    }
    if (threadIdx.x == 0 && grand_total) *grand_total = carry;
}

template <class T>
__global__ __launch_bounds__(256) void k_scan_add(T *__restrict__ out, int64_t n, const int64_t *__restrict__ sums)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)((int64_t)out[i] + sums[i / SCAN_TILE]);
}

// ---- radix sort ------------------------------------------------------------------------------
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                        int ntiles, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin,
                                                           uint64_t *__restrict__ kout, int32_t *__restrict__ vout,
                                                           int64_t n, int shift, int ntiles,
                                                           const uint32_t *__restrict__ offs)
{
    constexpr int NW = RS_THREADS / 64;
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcount[NW][256];
    const int lane = __lane_id();
    const int w = threadIdx.x / 64;
    base[threadIdx.x] = offs[(int64_t)threadIdx.x * ntiles + blockIdx.x];
    for (int k = 0; k < NW; k++) wcount[k][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int64_t tile = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t i = tile + r * RS_THREADS + threadIdx.x;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0;
        int32_t val = valid ? vin[i] : 0;
        uint32_t d = (uint32_t)((key >> shift) & 255u);
        // Lanes of this wavefront holding the same digit.
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(match & lt);
        bool leader = valid && (match & lt) == 0;
        if (leader) wcount[w][d] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (int k = 0; k < w; k++) before += wcount[k][d];
            uint32_t pos = base[d] + before + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int k = 0; k < NW; k++) {
            add += wcount[k][threadIdx.x];
            wcount[k][threadIdx.x] = 0;
        }
        base[threadIdx.x] += add;
        __syncthreads();
    }
}

}   // namespace

namespace ccdi {

int fill_u32(ccd_ctx *ctx, uint32_t *p, int64_t n, uint32_t v)
{
    if (n <= 0) return CCD_OK;
    unsigned g = (unsigned)std::min<int64_t>(blocks_for(n, 256), 2048);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(g), dim3(256), 0, ctx->stream, p, n, v);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int fill_u64(ccd_ctx *ctx, uint64_t *p, int64_t n, uint64_t v)
{
    if (n <= 0) return CCD_OK;
    unsigned g = (unsigned)std::min<int64_t>(blocks_for(n, 256), 2048);
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(g), dim3(256), 0, ctx->stream, p, n, v);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

size_t scan_temp_bytes(int64_t n) { return (size_t)(blocks_for(n, SCAN_TILE) + 2) * sizeof(int64_t); }

template <class T>
static int scan_impl(ccd_ctx *ctx, const T *in, T *out, int64_t n, void *temp, int64_t *total)
{
    if (n <= 0) return CCD_OK;
    int64_t nt = blocks_for(n, SCAN_TILE);
    int64_t *sums = (int64_t *)temp;
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)nt), dim3(SCAN_THREADS), 0, ctx->stream, in, out, n, sums);
    CCD_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, sums, nt, total);
    CCD_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_scan_add<T>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, out, n, sums);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int exclusive_scan_i32(ccd_ctx *ctx, const int32_t *in, int32_t *out, int64_t n, void *temp, int64_t *total)
{
    return scan_impl<int32_t>(ctx, in, out, n, temp, total);
}

int exclusive_scan_i64(ccd_ctx *ctx, const int64_t *in, int64_t *out, int64_t n, void *temp, int64_t *total)
{
    return scan_impl<int64_t>(ctx, in, out, n, temp, total);
}

size_t radix_sort_temp_bytes(int64_t n)
{
    int64_t nt = blocks_for(n, RS_TILE);
    return (size_t)(256 * nt) * 4 + scan_temp_bytes(256 * nt) + 64;
}

int radix_sort_pairs(ccd_ctx *ctx, uint64_t *keys, int32_t *vals, uint64_t *keys_alt, int32_t *vals_alt, int64_t n,
                     int bits, void *temp, bool &result_in_alt)
{
    result_in_alt = false;
    if (n <= 1) return CCD_OK;
    int ntiles = (int)blocks_for(n, RS_TILE);
    uint32_t *hist = (uint32_t *)temp;
    void *scan_tmp = (char *)temp + (size_t)256 * ntiles * 4;
    uint64_t *ka = keys, *kb = keys_alt;
    int32_t *va = vals, *vb = vals_alt;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, ctx->stream, ka, n, shift, ntiles, hist);
        CCD_LAUNCH_CHECK(ctx);
        int rc = scan_impl<uint32_t>(ctx, hist, hist, (int64_t)256 * ntiles, scan_tmp, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, ctx->stream, ka, va, kb, vb, n, shift,
                           ntiles, hist);
        CCD_LAUNCH_CHECK(ctx);
        std::swap(ka, kb);
        std::swap(va, vb);
        result_in_alt = !result_in_alt;
    }
    return CCD_OK;
}

}   // namespace ccdi
