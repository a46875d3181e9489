/*
 * treebuild.hip -- octree built on the GPU from Morton keys, for gfx950.
 * No counterpart in the reference: octree.cpp:14-141 builds a mean-split
 * tree on a CPU thread, which this replaces when the caller asks for it
 * (sph_build_tree).  The result uses octree.cpp's flat layout -- branch
 * [0, id, c0..c7], leaf [n, id, p0..], ids and offsets in pre-order, one
 * list of node offsets per level -- so the gravity passes take it unchanged.
 *
 *   keys      the particles' bounding cube is quantised to 1024 cells per
 *             axis (FP64 arithmetic), key = x | y << 1 | z << 2 interleaved
 *             (30 bits = 10 octal digits, digit t selects the octant at level
 *             t), pairs (key, particle) sorted by the LSD radix sort
 *   nodes     a node at level L holds the particles whose keys share the
 *             first L digits; it is a leaf when it holds at most max_leaf
 *             particles or L == max_depth
 *   chains    the nodes that start at one sorted position i are a chain of
 *             first children, from level Lmin(i) down to a leaf, and each is
 *             found from neighbouring keys alone (common(a, b) = the number
 *             of leading digits a and b share):
 *               d      = common(key[i-1], key[i]): i starts a segment at
 *                        every level above d; the level-d segment holding
 *                        i - 1 and i is the parent of the top node
 *               exists iff i == 0, or d < max_depth and the parent holds
 *                        more than max_leaf particles (some window
 *                        [j, j + max_leaf] around i lies inside it)
 *               Lmin   = 0 for i == 0, else d + 1
 *               Lleaf  = min(max(Lmin, D + 1), max_depth) with
 *                        D = common(key[i], key[i + max_leaf])
 *             so one thread per position produces its chain, and the whole
 *             tree takes a fixed number of launches whatever its depth
 *   layout    in pre-order a node precedes another iff it starts at an
 *             earlier sorted position, or at the same one and a lower level:
 *             exclusive scans over the positions of the chains' sizes (10 per
 *             branch, 2 + n per leaf) and node counts give every node's
 *             offset and id.  A chain's first children are its own next
 *             nodes; every other child links itself into its parent (one
 *             binary search for the parent's start).  The level lists are the
 *             (level, offset) pairs in pre-order, stably sorted by level.
 */
#include <algorithm>
#include <climits>
#include <cstring>

#include "sph_internal.h"

namespace {

constexpr int kMaxDepth = 10; /* 30-bit keys */
/* device scalars: [0] non-finite flag, [1] flat-array ints, [2] nodes,
 * [3] positions (scan length), [kLevelCounts + L] nodes in level L, then
 * the flat-array ints as a 64-bit sum (two ints, 8-byte aligned) */
constexpr int kLevelCounts = 4;
constexpr int kTotal64 = kLevelCounts + kMaxDepth + 2;
constexpr int kScalars = kTotal64 + 2;
constexpr int kGridCap = 2048; /* grid-stride launches: at most 2048 workgroups */

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

/* leading octal digits two 30-bit keys share (10 if equal) */
__device__ __forceinline__ int common_digits(uint32_t a, uint32_t b) {
    const uint32_t x = a ^ b;
    return x == 0u ? kMaxDepth : (__clz(x) - 2) / 3;
}

/* octant of a key at level L + 1 (digit L + 1) */
__device__ __forceinline__ int digit_below(uint32_t key, int L) { return (int)((key >> (30 - 3 * (L + 1))) & 7u); }

__global__ void k_tb_init(int32_t* scalars, int n) {
    const int t = threadIdx.x;
    if (t < kScalars) scalars[t] = t == 3 ? n : 0;
}

static_assert(kTotal64 % 2 == 0, "the 64-bit total needs 8-byte alignment");

/* Per-workgroup bounds of the positions; any non-finite coordinate sets
 * flags[0]. */
__global__ __launch_bounds__(256) void k_tb_bbox_part(const float4* __restrict__ pos, int n, float* __restrict__ part,
                                                      int32_t* __restrict__ flags) {
    __shared__ float sh[6][256];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    bool bad = false;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 p = pos[i];
        bad = bad || !isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z);
        v[0] = fminf(v[0], p.x);
        v[1] = fminf(v[1], p.y);
        v[2] = fminf(v[2], p.z);
        v[3] = fmaxf(v[3], p.x);
        v[4] = fmaxf(v[4], p.y);
        v[5] = fmaxf(v[5], p.z);
    }
    if (bad) atomicOr(flags, 1);
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
            for (int k = 3; k < 6; ++k) sh[k][threadIdx.x] = fmaxf(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[blockIdx.x * 6 + k] = sh[k][0];
}

/* Bounds of all workgroups; the cube's side ext = max extent (1 if zero) and
 * scale = 1024 / ext in FP64. */
__global__ __launch_bounds__(256) void k_tb_bbox_final(const float* __restrict__ part, int nb, float* __restrict__ box,
                                                       double* __restrict__ scale) {
    __shared__ float sh[6][256];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += 256) {
        for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], part[b * 6 + k]);
        for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], part[b * 6 + k]);
    }
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
            for (int k = 3; k < 6; ++k) sh[k][threadIdx.x] = fmaxf(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; ++k) box[k] = sh[k][0];
        float ext = fmaxf(sh[3][0] - sh[0][0], fmaxf(sh[4][0] - sh[1][0], sh[5][0] - sh[2][0]));
        if (!(ext > 0.f)) ext = 1.f;
        scale[0] = 1024.0 / (double)ext;
    }
}

__global__ __launch_bounds__(256) void k_tb_keys(const float4* __restrict__ pos, int n, const float* __restrict__ box,
                                                 const double* __restrict__ scale, uint2* __restrict__ pairs) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    const double sc = scale[0];
    const double t[3] = {((double)p.x - (double)box[0]) * sc, ((double)p.y - (double)box[1]) * sc,
                         ((double)p.z - (double)box[2]) * sc};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = (uint32_t)min(max((int)floor(t[a]), 0), 1023);
    pairs[i] = make_uint2(spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2), (uint32_t)i);
}

/* The chain of nodes starting at each sorted position: (Lmin, Lleaf, leaf
 * size, 0), or Lmin = -1 when no node starts there; its flat-array ints and
 * node count (scanned in place into offsets and ids afterwards), and the
 * nodes per level. */
__global__ __launch_bounds__(256) void k_tb_chain(const uint2* __restrict__ sorted, int n, int max_leaf, int max_depth,
                                                  int4* __restrict__ chain, int32_t* __restrict__ psize,
                                                  int32_t* __restrict__ pcount, int32_t* __restrict__ level_counts,
                                                  unsigned long long* __restrict__ total64) {
    __shared__ int hist[kMaxDepth + 1];
    __shared__ unsigned long long bsize;
    if (threadIdx.x <= kMaxDepth) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) bsize = 0ull;
    __syncthreads();
    unsigned long long mysize = 0ull;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint32_t key = sorted[i].x;
        int lmin = 0;
        bool has = true;
        if (i > 0) {
            const int d = common_digits(sorted[i - 1].x, key);
            has = false;
            if (d < max_depth) {
                /* the level-d segment holding i holds more than max_leaf
                 * particles iff some window of max_leaf + 1 around i lies in it */
                for (int j = max(0, i - max_leaf); j <= i && j + max_leaf < n; ++j)
                    if (common_digits(sorted[j].x, sorted[j + max_leaf].x) >= d) {
                        has = true;
                        break;
                    }
            }
            lmin = d + 1;
        }
        if (!has) {
            chain[i] = make_int4(-1, -1, 0, 0);
            psize[i] = 0;
            pcount[i] = 0;
            continue;
        }
        const int D = i + max_leaf < n ? common_digits(key, sorted[i + max_leaf].x) : -1;
        const int lleaf = min(max(lmin, D + 1), max_depth);
        int cnt;
        if (D < lleaf) {
            /* at most max_leaf particles share the leaf's prefix */
            cnt = 1;
            while (cnt < max_leaf && i + cnt < n && common_digits(key, sorted[i + cnt].x) >= lleaf) ++cnt;
        } else {
            /* a depth-limited leaf: its end by binary search */
            int lo = i + max_leaf + 1, hi = n;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (common_digits(key, sorted[mid].x) >= lleaf) lo = mid + 1;
                else hi = mid;
            }
            cnt = lo - i;
        }
        chain[i] = make_int4(lmin, lleaf, cnt, 0);
        psize[i] = 10 * (lleaf - lmin) + 2 + cnt;
        pcount[i] = lleaf - lmin + 1;
        mysize += (unsigned long long)psize[i];
        for (int L = lmin; L <= lleaf; ++L) atomicAdd(&hist[L], 1);
    }
    if (mysize) atomicAdd(&bsize, mysize);
    __syncthreads();
    if (threadIdx.x <= kMaxDepth && hist[threadIdx.x]) atomicAdd(&level_counts[threadIdx.x], hist[threadIdx.x]);
    if (threadIdx.x == 0 && bsize) atomicAdd(total64, bsize);
}

/* The records of every chain, and its (level, offset) pairs at the nodes'
 * ids.  A branch gets its first child (the chain's next node) and -1 in the
 * other slots; k_tb_link fills those.  A record that would not fit in n_ints
 * is skipped (the tree check of the next gravity pass then reports the
 * tree). */
__global__ __launch_bounds__(256) void k_tb_fill(const uint2* __restrict__ sorted, int n,
                                                 const int4* __restrict__ chain, const int32_t* __restrict__ pbase,
                                                 const int32_t* __restrict__ pnbase, int32_t* __restrict__ octree,
                                                 int64_t n_ints, uint2* __restrict__ lv, int32_t node_count) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int4 c = chain[i];
        if (c.x < 0) continue;
        const uint32_t key = sorted[i].x;
        const int base = pbase[i], nbase = pnbase[i];
        for (int L = c.x; L <= c.y; ++L) {
            const int off = base + 10 * (L - c.x), id = nbase + (L - c.x);
            const bool leaf = L == c.y;
            if (off < 0 || off + (int64_t)(leaf ? 2 + c.z : 10) > n_ints || id >= node_count) break;
            lv[id] = make_uint2((uint32_t)L, (uint32_t)off);
            octree[off + 1] = id;
            if (leaf) {
                octree[off] = c.z;
                for (int q = 0; q < c.z; ++q) octree[off + 2 + q] = (int32_t)sorted[i + q].y;
            } else {
                octree[off] = 0;
                const int first = digit_below(key, L);
                for (int k = 0; k < 8; ++k) octree[off + 2 + k] = k == first ? off + 10 : -1;
            }
        }
    }
}

/* Every chain but the root's links its top node into its parent: the node
 * at level Lmin - 1 that starts where that level's segment holding i
 * starts. */
__global__ __launch_bounds__(256) void k_tb_link(const uint2* __restrict__ sorted, int n,
                                                 const int4* __restrict__ chain, const int32_t* __restrict__ pbase,
                                                 int32_t* __restrict__ octree, int64_t n_ints) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int lmin = chain[i].x;
        if (i == 0 || lmin < 1) continue;
        const uint32_t key = sorted[i].x;
        const int d = lmin - 1;
        int lo = 0, hi = i; /* first position sharing d digits with i */
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (common_digits(sorted[mid].x, key) >= d) hi = mid;
            else lo = mid + 1;
        }
        const int4 pc = chain[lo];
        if (pc.x < 0 || pc.x > d || pc.y <= d) continue;
        const int64_t poff = pbase[lo] + 10 * (d - pc.x);
        if (poff < 0 || poff + 10 > n_ints) continue;
        octree[poff + 2 + digit_below(key, d)] = pbase[i];
    }
}

__global__ __launch_bounds__(256) void k_tb_levels(const uint2* __restrict__ lv, int count,
                                                   int32_t* __restrict__ level_nodes) {
    for (int j = blockIdx.x * 256 + threadIdx.x; j < count; j += gridDim.x * 256) level_nodes[j] = (int32_t)lv[j].y;
}

/* Exclusive scan of int32 in three launches: 1024-element blocks (4 per
 * thread), a scan of the block sums by one workgroup in chunks, and the
 * block offsets added back.  In place is allowed.  The element count is read
 * from the device; the grid covers the most it can be.  total[0] = the sum. */
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& sum) {
    /* 256 threads: Hillis-Steele in LDS (deterministic) */
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int t = (int)threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const int incl = sh[threadIdx.x];
    sum = sh[255];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_local(const int32_t* in, int32_t* out, const int32_t* __restrict__ count,
                                                    int32_t* __restrict__ bsum) {
    __shared__ int sh[256];
    const int n = *count;
    if ((int)blockIdx.x * kScanBlock >= n) return;
    const int base = blockIdx.x * kScanBlock + 4 * threadIdx.x;
    int v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        t += v[k];
    }
    int sum = 0;
    int run = block_exclusive_scan(t, sh, sum);
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sum;
}

__global__ __launch_bounds__(256) void k_scan_sums(int32_t* __restrict__ bsum, const int32_t* __restrict__ count,
                                                   int32_t* __restrict__ total) {
    __shared__ int sh[256];
    const int nb = (*count + kScanBlock - 1) / kScanBlock;
    int carry = 0;
    for (int c = 0; c < nb; c += 256) {
        const int i = c + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int sum = 0;
        const int ex = block_exclusive_scan(v, sh, sum);
        if (i < nb) bsum[i] = carry + ex;
        carry += sum;
    }
    if (threadIdx.x == 0) total[0] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int32_t* __restrict__ out, const int32_t* __restrict__ count,
                                                  const int32_t* __restrict__ bsum) {
    const int n = *count;
    if ((int)blockIdx.x * kScanBlock >= n) return;
    const int base = blockIdx.x * kScanBlock + 4 * threadIdx.x;
    const int add = bsum[blockIdx.x];
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

int grid_for(int64_t items) { return (int)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, kGridCap)); }

/* in-place exclusive scan of a[0, *count); the grid covers max_count */
int exclusive_scan(sph_ctx* ctx, int32_t* a, const int32_t* count, int max_count, int32_t* total) {
    sph_treebuild& B = ctx->tb;
    const int nb = std::max((max_count + kScanBlock - 1) / kScanBlock, 1);
    if (nb > B.bsum_cap) return ctx_fail(ctx, SPH_E_STATE, "tree build scan buffer too small");
    hipStream_t s = ctx->stream;
    k_scan_local<<<nb, 256, 0, s>>>(a, a, count, B.bsum);
    k_scan_sums<<<1, 256, 0, s>>>(B.bsum, count, total);
    k_scan_add<<<nb, 256, 0, s>>>(a, count, B.bsum);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

/* buffers sized by the particle count */
int ensure_buffers(sph_ctx* ctx) {
    sph_treebuild& B = ctx->tb;
    const int n = ctx->n;
    if (B.n_cap >= n) return SPH_OK;
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    sph_treebuild_free(B);
    B.bbox_blocks = std::min((n + 255) / 256, 1024);
    SPH_HIP_TRY(dalloc(&B.pairs[0], (size_t)n));
    SPH_HIP_TRY(dalloc(&B.pairs[1], (size_t)n));
    SPH_HIP_TRY(dalloc(&B.bbox_part, (size_t)B.bbox_blocks * 6));
    SPH_HIP_TRY(dalloc(&B.box, 8));
    SPH_HIP_TRY(dalloc(&B.scale, 1));
    SPH_HIP_TRY(dalloc(&B.chain, (size_t)n));
    SPH_HIP_TRY(dalloc(&B.pbase, (size_t)n));
    SPH_HIP_TRY(dalloc(&B.pnbase, (size_t)n));
    B.bsum_cap = (n + kScanBlock - 1) / kScanBlock;
    SPH_HIP_TRY(dalloc(&B.bsum, (size_t)B.bsum_cap));
    SPH_HIP_TRY(dalloc(&B.scalars, kScalars));
    SPH_HIP_TRY(dalloc(&B.digit_total, 256));
    B.n_cap = n;
    return SPH_OK;
}

/* buffers sized by the node count (grown by 1.2x) */
int ensure_node_buffers(sph_ctx* ctx, int64_t nodes) {
    sph_treebuild& B = ctx->tb;
    const int64_t items = std::max<int64_t>(nodes, ctx->n);
    const int tiles = (int)((items + sort_tile_size() - 1) / sort_tile_size());
    if (nodes > B.lv_cap || tiles > B.tile_cap) SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (nodes > B.lv_cap) {
        const int64_t cap = (int64_t)(nodes * 1.2) + 64;
        dfree(B.lv[0]);
        dfree(B.lv[1]);
        SPH_HIP_TRY(dalloc(&B.lv[0], (size_t)cap));
        SPH_HIP_TRY(dalloc(&B.lv[1], (size_t)cap));
        B.lv_cap = cap;
    }
    if (tiles > B.tile_cap) {
        const int cap = (int)(tiles * 1.2) + 4;
        dfree(B.tile_hist);
        SPH_HIP_TRY(dalloc(&B.tile_hist, (size_t)256 * cap));
        B.tile_cap = cap;
    }
    return SPH_OK;
}

}  // namespace

void sph_treebuild_free(sph_treebuild& B) {
    dfree(B.pairs[0]);
    dfree(B.pairs[1]);
    dfree(B.bbox_part);
    dfree(B.box);
    dfree(B.scale);
    dfree(B.chain);
    dfree(B.pbase);
    dfree(B.pnbase);
    dfree(B.bsum);
    dfree(B.scalars);
    dfree(B.lv[0]);
    dfree(B.lv[1]);
    dfree(B.tile_hist);
    dfree(B.digit_total);
    B.n_cap = 0;
    B.bsum_cap = 0;
    B.tile_cap = 0;
    B.lv_cap = 0;
}

extern "C" {

int sph_build_tree(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    const int max_depth = ctx->prm.max_tree_depth, max_leaf = ctx->prm.max_children_in_leaf;
    if (max_depth < 0 || max_depth > kMaxDepth)
        return ctx_fail(ctx, SPH_E_ARG, "the GPU tree needs max_tree_depth 0..10");
    SPH_TRY_RC(join_gravity(ctx));
    SPH_TRY_RC(join_tree(ctx)); /* this build's conversion rewrites the node records */
    SPH_TRY_RC(ensure_buffers(ctx));
    SPH_TRY_RC(ensure_node_buffers(ctx, ctx->tb.lv_cap)); /* tile table for the Morton sort */
    SPH_TRY_RC(ctx_begin_pass(ctx, SPH_PASS_TREE_BUILD));
    sph_treebuild& B = ctx->tb;
    const int n = ctx->n;
    hipStream_t s = ctx->stream;
    const int g = grid_for(n);
    /* bounds (and the non-finite check), keys, stable sort by key */
    k_tb_init<<<1, 64, 0, s>>>(B.scalars, n);
    k_tb_bbox_part<<<B.bbox_blocks, 256, 0, s>>>(ctx->pos, n, B.bbox_part, B.scalars);
    k_tb_bbox_final<<<1, 256, 0, s>>>(B.bbox_part, B.bbox_blocks, B.box, B.scale);
    k_tb_keys<<<(n + 255) / 256, 256, 0, s>>>(ctx->pos, n, B.box, B.scale, B.pairs[0]);
    SPH_HIP_TRY(hipGetLastError());
    uint2* sorted = nullptr;
    SPH_TRY_RC(launch_sort_pairs(ctx, B.pairs[0], B.pairs[1], n, 4, B.tile_hist, B.tile_cap, B.digit_total, &sorted));
    /* chains, then offsets and ids */
    k_tb_chain<<<g, 256, 0, s>>>(sorted, n, max_leaf, max_depth, B.chain, B.pbase, B.pnbase,
                                 B.scalars + kLevelCounts, (unsigned long long*)(B.scalars + kTotal64));
    SPH_HIP_TRY(hipGetLastError());
    SPH_TRY_RC(exclusive_scan(ctx, B.pbase, B.scalars + 3, n, B.scalars + 1));
    SPH_TRY_RC(exclusive_scan(ctx, B.pnbase, B.scalars + 3, n, B.scalars + 2));
    /* the one host synchronisation: flag, totals and level sizes */
    int32_t h[kScalars] = {};
    SPH_HIP_TRY(hipMemcpyAsync(h, B.scalars, sizeof(h), hipMemcpyDeviceToHost, s));
    SPH_HIP_TRY(hipStreamSynchronize(s));
    if (h[0]) return ctx_fail(ctx, SPH_E_ARG, "non-finite particle position");
    uint64_t total64 = 0;
    std::memcpy(&total64, h + kTotal64, sizeof(total64));
    if (total64 > (uint64_t)INT32_MAX)
        return ctx_fail(ctx, SPH_E_CAPACITY, "octree larger than the int32 flat layout allows");
    int n_levels = 0;
    while (n_levels <= max_depth && h[kLevelCounts + n_levels] > 0) ++n_levels;
    const int32_t* counts = h + kLevelCounts;
    const int64_t n_ints = h[1];
    const int32_t node_count = h[2];
    SPH_TRY_RC(ensure_node_buffers(ctx, node_count));
    int k = 0;
    SPH_TRY_RC(tree_reserve(ctx, n_ints, counts, n_levels, node_count, &k));
    /* buffer k: the compute stream is past its release (recorded on it) */
    int32_t* octree = ctx->tree.octree_up[k];
    k_tb_fill<<<g, 256, 0, s>>>(sorted, n, B.chain, B.pbase, B.pnbase, octree, n_ints, B.lv[0], node_count);
    k_tb_link<<<g, 256, 0, s>>>(sorted, n, B.chain, B.pbase, octree, n_ints);
    SPH_HIP_TRY(hipGetLastError());
    /* level lists: (level, offset) in pre-order, stably sorted by level */
    uint2* by_level = nullptr;
    SPH_TRY_RC(launch_sort_pairs(ctx, B.lv[0], B.lv[1], node_count, 1, B.tile_hist, B.tile_cap, B.digit_total,
                                 &by_level));
    k_tb_levels<<<grid_for(node_count), 256, 0, s>>>(by_level, node_count, ctx->tree.level_tp_up[k]);
    SPH_HIP_TRY(hipGetLastError());
    SPH_TRY_RC(ctx_end_pass(ctx, SPH_PASS_TREE_BUILD));
    return tree_commit(ctx, k, n_ints, counts, n_levels, node_count, ctx->stream);
}

int sph_tree_info(sph_ctx* ctx, int64_t* n_ints, int32_t* node_count, int32_t* n_levels, int32_t* level_counts,
                  int32_t level_cap) {
    if (!ctx || !n_ints || !node_count || !n_levels) return SPH_E_ARG;
    if (!ctx->have_tree) return ctx_fail(ctx, SPH_E_NOTREE, "no tree");
    const sph_tree_dev& T = ctx->tree;
    *n_ints = T.n_ints;
    *node_count = T.node_count;
    *n_levels = (int32_t)T.level_counts.size();
    if (level_counts)
        for (size_t l = 0; l < T.level_counts.size() && (int64_t)l < level_cap; ++l) level_counts[l] = T.level_counts[l];
    return SPH_OK;
}

int sph_download_tree(sph_ctx* ctx, int32_t* octree, int32_t* level_nodes) {
    if (!ctx || !octree || !level_nodes) return SPH_E_ARG;
    if (!ctx->have_tree) return ctx_fail(ctx, SPH_E_NOTREE, "no tree");
    SPH_TRY_RC(join_gravity(ctx));
    SPH_TRY_RC(join_tree(ctx));
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const sph_tree_dev& T = ctx->tree;
    SPH_HIP_TRY(hipMemcpy(octree, T.octree, 4 * (size_t)T.n_ints, hipMemcpyDeviceToHost));
    SPH_HIP_TRY(hipMemcpy(level_nodes, T.level_tp, 4 * (size_t)T.node_count, hipMemcpyDeviceToHost));
    return SPH_OK;
}

}  // extern "C"
