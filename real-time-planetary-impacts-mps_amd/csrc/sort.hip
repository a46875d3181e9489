/*
 * sort.hip -- Morton hash, stable LSD radix sort of (cell key, index) pairs,
 * and cell start/end, for gfx950.
 *
 * Reference: Compute::sort (Compute.cpp:265-326) running hash / hist / scan /
 * sum / sort x4 / initialise / findCellPositions (sort.metal:15-148).
 *
 * The result is bit-identical to the reference: its LSD passes are stable
 * and start from the identity permutation, so the sorted array is ordered by
 * (key, index).  Any stable LSD sort over the significant digits gives the
 * same array; keys hold 3*cell_power bits (24 at CELL_POWER 8), so
 * ceil(3*cell_power/8) passes suffice and the reference's 4th pass (the
 * identity) is skipped.
 *
 * Per pass, three kernels:
 *   k_hash_hist / k_hist  per-tile digit histogram (tile = 2048 pairs), LDS atomics
 *   k_scan_rows           one workgroup per digit: exclusive scan of the
 *                         digit's row of the digit-major [256][ntiles]
 *                         table and the digit's total
 *   k_scatter             per tile: destination base of each digit = prefix
 *                         of the digit totals + the tile's row offset; then
 *                         8 rounds of 256 consecutive pairs; a
 *                         round ranks its pairs stably with 8 wave ballots per
 *                         digit bit (lanes with equal digit = one peer mask) and
 *                         a per-wave LDS count, then scatters
 * Cell start/end (sort.metal:131-148) come from run boundaries of the sorted
 * keys instead of 2 atomics per particle, and instead of re-initialising all
 * 2^24 cells (128 MiB) every step, only the cells occupied after the previous
 * sort are reset -- the table content is the same.
 */
#include <cstdio>
#include <mutex>

#include "sph_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 8;
constexpr int kTile = kThreads * kRounds;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ uint32_t hash_key(float4 p, float cellSize, int cellsPerDim) {
    /* sort.metal:20-43 */
    const float nx = p.x / cellSize, ny = p.y / cellSize, nz = p.z / cellSize;
    int px = (int)floorf(nx) % cellsPerDim;
    int py = (int)floorf(ny) % cellsPerDim;
    int pz = (int)floorf(nz) % cellsPerDim;
    if (px < 0) px += cellsPerDim;
    if (py < 0) py += cellsPerDim;
    if (pz < 0) pz += cellsPerDim;
    return dev_expand_bits((uint32_t)px) | (dev_expand_bits((uint32_t)py) << 1) | (dev_expand_bits((uint32_t)pz) << 2);
}

__global__ __launch_bounds__(kThreads) void k_hash_hist(const float4* __restrict__ pos, int n, float cellSize,
                                                        int cellsPerDim, uint2* __restrict__ pairs,
                                                        uint32_t* __restrict__ tile_hist, int ntiles) {
    __shared__ uint32_t hist[256];
    const int t = blockIdx.x;
    hist[threadIdx.x] = 0;
    __syncthreads();
    for (int r = 0; r < kRounds; ++r) {
        const int i = t * kTile + r * kThreads + threadIdx.x;
        if (i < n) {
            const uint32_t key = hash_key(pos[i], cellSize, cellsPerDim);
            pairs[i] = make_uint2(key, (uint32_t)i);
            atomicAdd(&hist[key & 255u], 1u);
        }
    }
    __syncthreads();
    tile_hist[threadIdx.x * ntiles + t] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_hist(const uint2* __restrict__ pairs, int n, int shift,
                                                   uint32_t* __restrict__ tile_hist, int ntiles) {
    __shared__ uint32_t hist[256];
    const int t = blockIdx.x;
    hist[threadIdx.x] = 0;
    __syncthreads();
    for (int r = 0; r < kRounds; ++r) {
        const int i = t * kTile + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[(pairs[i].x >> shift) & 255u], 1u);
    }
    __syncthreads();
    tile_hist[threadIdx.x * ntiles + t] = hist[threadIdx.x];
}

/* Exclusive scan of one digit's row of the digit-major tile table (one
 * workgroup per digit), in place, and the digit's total.  The workgroup
 * walks the row in coalesced chunks of 1024 (4 consecutive entries per
 * lane): lane-local scan, wave scan of the lane sums with shuffles, a scan
 * of the 4 wave sums, and a carry from chunk to chunk. */
__global__ __launch_bounds__(kThreads) void k_scan_rows(uint32_t* __restrict__ table, int ntiles,
                                                        uint32_t* __restrict__ digit_total) {
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t carry_s;
    uint32_t* __restrict__ a = table + (int64_t)blockIdx.x * ntiles;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < ntiles; base += 4 * kThreads) {
        const int i0 = base + 4 * tid;
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = i0 + k < ntiles ? a[i0 + k] : 0u;
        const uint32_t t = v[0] + v[1] + v[2] + v[3];
        uint32_t incl = t;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        uint32_t run = carry_s + before + (incl - t);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < ntiles) a[i0 + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (tid == kThreads - 1) carry_s = run;
        __syncthreads();
    }
    if (tid == 0) digit_total[blockIdx.x] = carry_s;
}

/* Stable scatter of one tile: the destination of digit d's pairs is the
 * exclusive prefix of the digit totals (formed here from the 256 totals)
 * plus the tile's offset within digit d's row. */
__global__ __launch_bounds__(kThreads) void k_scatter(const uint2* __restrict__ src, uint2* __restrict__ dst, int n,
                                                      int shift, const uint32_t* __restrict__ tile_off, int ntiles,
                                                      const uint32_t* __restrict__ digit_total) {
    __shared__ uint32_t goff[256];
    __shared__ uint32_t running[256];
    __shared__ uint32_t wave_cnt[kWaves][256];
    __shared__ uint32_t wave_off[kWaves][256];
    __shared__ uint32_t wtot[kWaves];
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    {
        const uint32_t dt = digit_total[tid];
        uint32_t incl = dt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < wave; ++w) before += wtot[w];
        goff[tid] = before + (incl - dt) + tile_off[tid * ntiles + t];
    }
    running[tid] = 0;
    for (int w = 0; w < kWaves; ++w) wave_cnt[w][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kRounds; ++r) {
        const int i = t * kTile + r * kThreads + tid;
        const bool valid = i < n;
        uint2 pr = valid ? src[i] : make_uint2(0u, 0u);
        const uint32_t d = (pr.x >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t m = __ballot(set);
            peers &= set ? m : ~m;
        }
        const uint32_t lane_rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && lane_rank == 0) wave_cnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t base = running[tid];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                wave_off[w][tid] = base;
                base += wave_cnt[w][tid];
                wave_cnt[w][tid] = 0;
            }
            running[tid] = base;
        }
        __syncthreads();
        if (valid) dst[goff[d] + wave_off[wave][d] + lane_rank] = pr;
    }
}

/* Resets the cells that the previous sort occupied (sparse form of
 * `initialise`, sort.metal:131-138). */
__global__ void k_cells_reset(const uint2* __restrict__ prev, int n, uint2* __restrict__ cell_se) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = prev[i].x;
    if (i == 0 || prev[i - 1].x != key) cell_se[key] = make_uint2(0xFFFFFFFFu, 0u);
}

/* findCellPositions (sort.metal:140-148) from run boundaries, fused with the
 * gather of the Morton-ordered copies used by the neighbour passes. */
__global__ void k_cells_gather(const uint2* __restrict__ sorted, int n, uint2* __restrict__ cell_se,
                               const float4* __restrict__ pos, const float4* __restrict__ vel,
                               const float* __restrict__ m, float4* __restrict__ s_posm, float4* __restrict__ s_vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint2 pr = sorted[i];
    const uint32_t key = pr.x;
    if (i == 0 || sorted[i - 1].x != key) cell_se[key].x = (uint32_t)i;
    if (i == n - 1 || sorted[i + 1].x != key) cell_se[key].y = (uint32_t)i;
    const uint32_t j = pr.y;
    const float4 p = pos[j];
    const float4 v = vel[j];
    s_posm[i] = p; /* w: the mass */
    s_vel[i] = make_float4(v.x, v.y, v.z, 0.f);
}

__global__ void k_fill_cells(uint2* __restrict__ cell_se, int ncells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncells) cell_se[i] = make_uint2(0xFFFFFFFFu, 0u);
}

}  // namespace

int launch_fill_cells(sph_ctx* ctx) {
    const int threads = 256;
    k_fill_cells<<<(ctx->ncells + threads - 1) / threads, threads, 0, ctx->stream>>>(ctx->cell_se, ctx->ncells);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

/* LSD radix sort of (key, value) pairs by the low `passes` bytes of the key,
 * stable (ties keep their order), ping-ponging between a and b, with the
 * caller's tile table (256 x tile_cap) and digit totals (256).  Returns the
 * buffer that holds the result through `out`. */
int launch_sort_pairs(sph_ctx* ctx, uint2* a, uint2* b, int n, int passes, uint32_t* tile_hist, int tile_cap,
                      uint32_t* digit_total, uint2** out) {
    const int ntiles = (n + kTile - 1) / kTile;
    if (ntiles > tile_cap) return ctx_fail(ctx, SPH_E_STATE, "sort tile table too small");
    hipStream_t s = ctx->stream;
    uint2* bufs[2] = {a, b};
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        k_hist<<<ntiles, kThreads, 0, s>>>(bufs[cur], n, shift, tile_hist, ntiles);
        SPH_HIP_TRY(hipGetLastError());
        k_scan_rows<<<256, kThreads, 0, s>>>(tile_hist, ntiles, digit_total);
        SPH_HIP_TRY(hipGetLastError());
        k_scatter<<<ntiles, kThreads, 0, s>>>(bufs[cur], bufs[cur ^ 1], n, shift, tile_hist, ntiles, digit_total);
        SPH_HIP_TRY(hipGetLastError());
        cur ^= 1;
    }
    *out = bufs[cur];
    return SPH_OK;
}

int sort_tile_size() { return kTile; }

int launch_sort(sph_ctx* ctx) {
    const int n = ctx->n;
    const int ntiles = (n + kTile - 1) / kTile;
    if (ntiles > ctx->ntiles) return ctx_fail(ctx, SPH_E_STATE, "sort tile table too small");
    const int bits = 3 * ctx->prm.cell_power;
    const int passes = (bits + 7) / 8;
    const int threads = 256;
    const int blocks = (n + threads - 1) / threads;
    hipStream_t s = ctx->stream;
    if (ctx->have_prev) {
        k_cells_reset<<<blocks, threads, 0, s>>>(ctx->prev_keys_src, n, ctx->cell_se);
        SPH_HIP_TRY(hipGetLastError());
    }
    uint2* bufs[2] = {ctx->pairs_a, ctx->pairs_b};
    k_hash_hist<<<ntiles, kThreads, 0, s>>>(ctx->pos, n, ctx->prm.cell_width, 1 << ctx->prm.cell_power, bufs[0],
                                            ctx->tile_hist, ntiles);
    SPH_HIP_TRY(hipGetLastError());
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (p > 0) {
            k_hist<<<ntiles, kThreads, 0, s>>>(bufs[cur], n, shift, ctx->tile_hist, ntiles);
            SPH_HIP_TRY(hipGetLastError());
        }
        k_scan_rows<<<256, kThreads, 0, s>>>(ctx->tile_hist, ntiles, ctx->digit_total);
        SPH_HIP_TRY(hipGetLastError());
        k_scatter<<<ntiles, kThreads, 0, s>>>(bufs[cur], bufs[cur ^ 1], n, shift, ctx->tile_hist, ntiles,
                                              ctx->digit_total);
        SPH_HIP_TRY(hipGetLastError());
        cur ^= 1;
    }
    uint2* sorted = bufs[cur];
    k_cells_gather<<<blocks, threads, 0, s>>>(sorted, n, ctx->cell_se, ctx->pos, ctx->vel, ctx->m, ctx->s_posm,
                                              ctx->s_vel);
    SPH_HIP_TRY(hipGetLastError());
    /* keep the sorted array where the next sort will not overwrite it before
     * the reset: swap so that `pairs_a` is always the current result */
    if (cur == 1) std::swap(ctx->pairs_a, ctx->pairs_b);
    ctx->prev_keys_src = ctx->pairs_a;
    ctx->have_prev = true;
    return SPH_OK;
}

/* ------------------------------------------------------------------------
 * sph_sort_noalloc: the hash, radix sort and cell start/end of sph_sort
 * without a context and without any allocation.  The sort's working memory
 * (the second pair buffer of the ping-pong, the per-tile digit table and the
 * digit totals) lives in `scratch`, a file-scope device array of fixed size
 * that every GPU the library runs on holds until the process ends.  Host
 * code never passes `scratch` itself to a kernel (on the host it is only a
 * stand-in for the symbol): each device's address of it comes from
 * hipGetSymbolAddress, once per device on its first call, and is kept in
 * noalloc_scratch_ptr[].  Every call on a GPU uses that GPU's one array, so
 * calls that overlap on one GPU (two streams, two threads) race.  Nothing in
 * the array needs clearing: each call writes every entry it reads.
 * ---------------------------------------------------------------------- */
namespace {

constexpr int64_t kNoAllocMaxN = SPH_SORT_NOALLOC_MAX_N;
constexpr int64_t kNoAllocMaxTiles = (kNoAllocMaxN + kTile - 1) / kTile;
constexpr size_t kNoAllocPairBytes = (size_t)kNoAllocMaxN * sizeof(uint2);
constexpr size_t kNoAllocHistBytes = (size_t)256 * kNoAllocMaxTiles * sizeof(uint32_t);
constexpr size_t kNoAllocBytes = kNoAllocPairBytes + kNoAllocHistBytes + 256 * sizeof(uint32_t);
static_assert(kNoAllocPairBytes % 256 == 0 && kNoAllocHistBytes % 256 == 0, "scratch partitions stay 256-B aligned");

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex noalloc_mutex;
void* noalloc_scratch_ptr[kMaxDevices] = {};

thread_local char noalloc_err[256] = "null context";

int noalloc_fail(int code, const char* fmt, long long a = 0, long long b = 0) {
    snprintf(noalloc_err, sizeof(noalloc_err), fmt, a, b);
    return code;
}

/* The run boundaries of the sorted keys give each occupied cell's first and
 * last slot (sort.metal:140-148); the other cells keep the fill value. */
__global__ void k_cells_bounds(const uint2* __restrict__ sorted, int n, uint2* __restrict__ cell_se) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = sorted[i].x;
    if (i == 0 || sorted[i - 1].x != key) cell_se[key].x = (uint32_t)i;
    if (i == n - 1 || sorted[i + 1].x != key) cell_se[key].y = (uint32_t)i;
}

}  // namespace

const char* noalloc_last_error() { return noalloc_err; }

extern "C" int sph_sort_noalloc(const float* pos4, int32_t n, float cell_width, int32_t cell_power,
                                uint32_t* sorted_pairs, uint32_t* cell_se, void* stream) {
    if (!pos4 || !sorted_pairs || !cell_se) return noalloc_fail(SPH_E_ARG, "sph_sort_noalloc: null pointer");
    if (n < 1) return noalloc_fail(SPH_E_ARG, "sph_sort_noalloc: n = %lld, needs at least 1 particle", n);
    if (n > kNoAllocMaxN)
        return noalloc_fail(SPH_E_CAPACITY, "sph_sort_noalloc: n = %lld exceeds the scratch maximum of %lld particles",
                            n, kNoAllocMaxN);
    if (cell_power < 1 || cell_power > 10)
        return noalloc_fail(SPH_E_ARG, "sph_sort_noalloc: cell_power = %lld, must be 1..10", cell_power);
    if (!(cell_width > 0.f) || !(cell_width < 3.0e38f))
        return noalloc_fail(SPH_E_ARG, "sph_sort_noalloc: cell_width must be positive and finite");
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail((int)e, "sph_sort_noalloc: hipGetDevice failed (%lld)", (long long)e);
    if (dev < 0 || dev >= kMaxDevices)
        return noalloc_fail(SPH_E_ARG, "sph_sort_noalloc: device %lld beyond the %lld supported", dev, kMaxDevices);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(noalloc_mutex);
        if (!noalloc_scratch_ptr[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return noalloc_fail((int)(e != hipSuccess ? e : hipErrorInvalidSymbol),
                                    "sph_sort_noalloc: hipGetSymbolAddress(scratch) failed (%lld)", (long long)e);
            noalloc_scratch_ptr[dev] = p;
        }
        base = noalloc_scratch_ptr[dev];
    }
    uint2* spare = reinterpret_cast<uint2*>(base);
    uint32_t* tile_hist = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(base) + kNoAllocPairBytes);
    uint32_t* digit_total = tile_hist + 256 * kNoAllocMaxTiles;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int ntiles = (n + kTile - 1) / kTile;
    const int passes = (3 * cell_power + 7) / 8;
    /* the hash writes into the buffer that makes the last pass land in the
     * caller's array: an odd number of passes starts in the spare one */
    uint2* out = reinterpret_cast<uint2*>(sorted_pairs);
    uint2* bufs[2] = {(passes & 1) ? spare : out, (passes & 1) ? out : spare};
    const int cells_per_dim = 1 << cell_power;
    const int ncells = 1 << (3 * cell_power);
    k_fill_cells<<<(ncells + 255) / 256, 256, 0, s>>>(reinterpret_cast<uint2*>(cell_se), ncells);
    k_hash_hist<<<ntiles, kThreads, 0, s>>>(reinterpret_cast<const float4*>(pos4), n, cell_width, cells_per_dim,
                                            bufs[0], tile_hist, ntiles);
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (p > 0) k_hist<<<ntiles, kThreads, 0, s>>>(bufs[cur], n, shift, tile_hist, ntiles);
        k_scan_rows<<<256, kThreads, 0, s>>>(tile_hist, ntiles, digit_total);
        k_scatter<<<ntiles, kThreads, 0, s>>>(bufs[cur], bufs[cur ^ 1], n, shift, tile_hist, ntiles, digit_total);
        cur ^= 1;
    }
    k_cells_bounds<<<(n + 255) / 256, 256, 0, s>>>(bufs[cur], n, reinterpret_cast<uint2*>(cell_se));
    e = hipGetLastError();
    if (e != hipSuccess) return noalloc_fail((int)e, "sph_sort_noalloc: launch failed (%lld)", (long long)e);
    return SPH_OK;
}
