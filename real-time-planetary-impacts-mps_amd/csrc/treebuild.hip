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
 *             (30 bits), pairs (key, index) sorted by the LSD radix sort
 *   levels    a node at level L holds the particles sharing the key's top
 *             3 L bits; it is a leaf when it holds at most max_leaf particles
 *             or L == max_depth; level L + 1 is emitted from level L with one
 *             thread per node (binary searches for the octant boundaries) and
 *             an exclusive scan of the child counts, so every level lists its
 *             nodes in key order
 *   layout    in pre-order a node precedes another iff it starts at an
 *             earlier sorted position, or at the same one and a lower level
 *             (the nodes starting at one position are a chain of first
 *             children ending in a leaf).  Scans over the positions of the
 *             chains' sizes (10 per branch, 2 + n per leaf) and node counts
 *             give every node's offset and id.
 */
#include <algorithm>
#include <climits>

#include "sph_internal.h"

namespace {

constexpr int kMaxDepth = 10; /* 30-bit keys */
/* device scalars: [0] non-finite flag, [1] flat-array ints, [2] nodes,
 * [kLevelCounts + L] nodes in level L */
constexpr int kLevelCounts = 4;
constexpr int kScalars = kLevelCounts + kMaxDepth + 2;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

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

__global__ void k_tb_init(int2* nodes, int n, int32_t* scalars) {
    nodes[0] = make_int2(0, n);
    scalars[0] = 0;
    scalars[kLevelCounts] = 1;
}

/* First index in [lo, hi) whose octant at `shift` is >= k (the octants are
 * non-decreasing there: the keys share every higher bit). */
__device__ __forceinline__ int octant_bound(const uint2* __restrict__ sorted, int lo, int hi, int shift, uint32_t k) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (((sorted[mid].x >> shift) & 7u) < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* Number of children of each node of level `level` (0 for a leaf); the
 * level's size is read from the device. */
__global__ __launch_bounds__(256) void k_tb_split(const int2* __restrict__ nodes, const int32_t* __restrict__ count,
                                                  int level, int max_leaf, int max_depth,
                                                  const uint2* __restrict__ sorted, int32_t* __restrict__ cnt) {
    const int c = *count;
    const int shift = 30 - 3 * (level + 1);
    for (int j = blockIdx.x * 256 + threadIdx.x; j < c; j += gridDim.x * 256) {
        const int2 nd = nodes[j];
        if (nd.y - nd.x <= max_leaf || level >= max_depth) {
            cnt[j] = 0;
            continue;
        }
        int b = nd.x, k_c = 0;
        for (uint32_t k = 1; k <= 8; ++k) {
            const int e = k < 8 ? octant_bound(sorted, b, nd.y, shift, k) : nd.y;
            k_c += e > b ? 1 : 0;
            b = e;
        }
        cnt[j] = k_c;
    }
}

/* The children of each node, in octant order, at the scanned offsets. */
__global__ __launch_bounds__(256) void k_tb_emit(const int2* __restrict__ nodes, const int32_t* __restrict__ cnt,
                                                 const int32_t* __restrict__ scn, const int32_t* __restrict__ count,
                                                 int level, const uint2* __restrict__ sorted,
                                                 int2* __restrict__ next) {
    const int c = *count;
    const int shift = 30 - 3 * (level + 1);
    for (int j = blockIdx.x * 256 + threadIdx.x; j < c; j += gridDim.x * 256) {
        if (cnt[j] == 0) continue;
        const int2 nd = nodes[j];
        int o = scn[j], b = nd.x;
        for (uint32_t k = 1; k <= 8; ++k) {
            const int e = k < 8 ? octant_bound(sorted, b, nd.y, shift, k) : nd.y;
            if (e > b) next[o++] = make_int2(b, e);
            b = e;
        }
    }
}

__global__ __launch_bounds__(256) void k_tb_fill_i32(int32_t* p, int n, int v) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = v;
}

/* Per sorted position: the top level of the chain of nodes starting there,
 * and the leaf that ends the chain. */
__global__ __launch_bounds__(256) void k_tb_mark(const int2* __restrict__ nodes, const int32_t* __restrict__ cnt,
                                                 const int32_t* __restrict__ count, int level,
                                                 int32_t* __restrict__ lmin, int32_t* __restrict__ lleaf,
                                                 int32_t* __restrict__ lcnt) {
    const int c = *count;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < c; j += gridDim.x * 256) {
        const int2 nd = nodes[j];
        atomicMin(&lmin[nd.x], level);
        if (cnt[j] == 0) {
            lleaf[nd.x] = level;
            lcnt[nd.x] = nd.y - nd.x;
        }
    }
}

/* Flat-array ints and node count of the chain starting at each position. */
__global__ __launch_bounds__(256) void k_tb_chain(int n, const int32_t* __restrict__ lmin,
                                                  const int32_t* __restrict__ lleaf, const int32_t* __restrict__ lcnt,
                                                  int32_t* __restrict__ psize, int32_t* __restrict__ pcount) {
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const int lm = lmin[p];
        if (lm == INT_MAX) {
            psize[p] = 0;
            pcount[p] = 0;
            continue;
        }
        const int d = lleaf[p] - lm;
        psize[p] = 10 * d + 2 + lcnt[p];
        pcount[p] = d + 1;
    }
}

/* The flat records of one level, and its list of node offsets.  A record
 * that would not fit in n_ints is skipped (the tree check of the next
 * gravity pass then reports the tree). */
__global__ __launch_bounds__(256) void k_tb_fill(const int2* __restrict__ nodes, const int32_t* __restrict__ cnt,
                                                 int count, int level, const int32_t* __restrict__ lmin,
                                                 const int32_t* __restrict__ pbase, const int32_t* __restrict__ pnbase,
                                                 const uint2* __restrict__ sorted, int32_t* __restrict__ octree,
                                                 int64_t n_ints, int32_t* __restrict__ level_nodes) {
    const int shift = 30 - 3 * (level + 1);
    for (int j = blockIdx.x * 256 + threadIdx.x; j < count; j += gridDim.x * 256) {
        const int2 nd = nodes[j];
        const int d = level - lmin[nd.x];
        const int off = pbase[nd.x] + 10 * d;
        level_nodes[j] = off;
        const int64_t size = cnt[j] == 0 ? 2 + (nd.y - nd.x) : 10;
        if (off < 0 || off + size > n_ints) continue;
        octree[off + 1] = pnbase[nd.x] + d;
        if (cnt[j] == 0) {
            octree[off] = nd.y - nd.x;
            for (int q = nd.x; q < nd.y; ++q) octree[off + 2 + (q - nd.x)] = (int32_t)sorted[q].y;
            continue;
        }
        octree[off] = 0;
        int b = nd.x;
        for (uint32_t k = 1; k <= 8; ++k) {
            const int e = k < 8 ? octant_bound(sorted, b, nd.y, shift, k) : nd.y;
            octree[off + 2 + (int)(k - 1)] = e > b ? pbase[b] + 10 * (level + 1 - lmin[b]) : -1;
            b = e;
        }
    }
}

/* Exclusive scan of int32 in three launches: 1024-element blocks (4 per
 * thread), a scan of the block sums by one workgroup in chunks, and the
 * block offsets added back.  The element count is read from the device;
 * the grid covers the most it can be.  total[0] = the sum. */
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

__global__ __launch_bounds__(256) void k_scan_local(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                    const int32_t* __restrict__ count, int32_t* __restrict__ bsum) {
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

}  // namespace

namespace {

constexpr int kGridCap = 2048; /* grid-stride launches: at most 2048 workgroups */

int grid_for(int64_t items) { return (int)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, kGridCap)); }

/* exclusive scan of in[0, *count) into out; the grid covers max_count */
int exclusive_scan(sph_ctx* ctx, const int32_t* in, int32_t* out, const int32_t* count, int max_count,
                   int32_t* total) {
    sph_treebuild& B = ctx->tb;
    const int nb = (max_count + kScanBlock - 1) / kScanBlock;
    if (nb > B.bsum_cap) return ctx_fail(ctx, SPH_E_STATE, "tree build scan buffer too small");
    hipStream_t s = ctx->stream;
    k_scan_local<<<std::max(nb, 1), 256, 0, s>>>(in, out, count, B.bsum);
    k_scan_sums<<<1, 256, 0, s>>>(B.bsum, count, total);
    k_scan_add<<<std::max(nb, 1), 256, 0, s>>>(out, count, B.bsum);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

int ensure_buffers(sph_ctx* ctx) {
    sph_treebuild& B = ctx->tb;
    const int n = ctx->n;
    const int64_t node_cap = (int64_t)(kMaxDepth + 1) * n;
    if (B.n_cap >= n && B.node_cap >= node_cap) return SPH_OK;
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    sph_treebuild_free(B);
    B.bbox_blocks = std::min((n + 255) / 256, 1024);
    SPH_HIP_TRY(dalloc(&B.pairs[0], (size_t)n));
    SPH_HIP_TRY(dalloc(&B.pairs[1], (size_t)n));
    SPH_HIP_TRY(dalloc(&B.bbox_part, (size_t)B.bbox_blocks * 6));
    SPH_HIP_TRY(dalloc(&B.box, 8));
    SPH_HIP_TRY(dalloc(&B.scale, 1));
    SPH_HIP_TRY(dalloc(&B.nodes, (size_t)node_cap));
    SPH_HIP_TRY(dalloc(&B.cnt, (size_t)node_cap));
    SPH_HIP_TRY(dalloc(&B.scn, (size_t)node_cap));
    int32_t** pn[] = {&B.lmin, &B.lleaf, &B.lcnt, &B.psize, &B.pcount, &B.pbase, &B.pnbase};
    for (auto p : pn) SPH_HIP_TRY(dalloc(p, (size_t)n));
    B.bsum_cap = (n + kScanBlock - 1) / kScanBlock;
    SPH_HIP_TRY(dalloc(&B.bsum, (size_t)B.bsum_cap));
    SPH_HIP_TRY(dalloc(&B.scalars, kScalars));
    B.n_cap = n;
    B.node_cap = node_cap;
    return SPH_OK;
}

}  // namespace

void sph_treebuild_free(sph_treebuild& B) {
    dfree(B.pairs[0]);
    dfree(B.pairs[1]);
    dfree(B.bbox_part);
    dfree(B.box);
    dfree(B.scale);
    dfree(B.nodes);
    dfree(B.cnt);
    dfree(B.scn);
    int32_t** pn[] = {&B.lmin, &B.lleaf, &B.lcnt, &B.psize, &B.pcount, &B.pbase, &B.pnbase, &B.bsum, &B.scalars};
    for (auto p : pn) dfree(*p);
    B.n_cap = 0;
    B.node_cap = 0;
    B.bsum_cap = 0;
}

extern "C" {

int sph_build_tree(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    const int max_depth = ctx->prm.max_tree_depth, max_leaf = ctx->prm.max_children_in_leaf;
    if (max_depth < 0 || max_depth > kMaxDepth)
        return ctx_fail(ctx, SPH_E_ARG, "the GPU tree needs max_tree_depth 0..10");
    SPH_TRY_RC(join_gravity(ctx));
    SPH_TRY_RC(ensure_buffers(ctx));
    sph_treebuild& B = ctx->tb;
    const int n = ctx->n;
    const int64_t stride = n; /* level L's nodes at nodes + L * n (a level holds at most n) */
    hipStream_t s = ctx->stream;
    int32_t* lvc = B.scalars + kLevelCounts;
    /* bounds (and the non-finite check), keys, stable sort by key */
    k_tb_init<<<1, 1, 0, s>>>(B.nodes, n, B.scalars);
    k_tb_bbox_part<<<B.bbox_blocks, 256, 0, s>>>(ctx->pos, n, B.bbox_part, B.scalars);
    k_tb_bbox_final<<<1, 256, 0, s>>>(B.bbox_part, B.bbox_blocks, B.box, B.scale);
    k_tb_keys<<<(n + 255) / 256, 256, 0, s>>>(ctx->pos, n, B.box, B.scale, B.pairs[0]);
    SPH_HIP_TRY(hipGetLastError());
    uint2* sorted = nullptr;
    SPH_TRY_RC(launch_sort_pairs(ctx, B.pairs[0], B.pairs[1], n, 4, &sorted));
    /* levels top down; every level's size stays on the device */
    const int g = grid_for(n);
    for (int L = 0; L <= max_depth; ++L) {
        int2* lv = B.nodes + L * stride;
        int32_t* cnt = B.cnt + L * stride;
        k_tb_split<<<g, 256, 0, s>>>(lv, lvc + L, L, max_leaf, max_depth, sorted, cnt);
        if (L == max_depth) break;
        SPH_TRY_RC(exclusive_scan(ctx, cnt, B.scn + L * stride, lvc + L, n, lvc + L + 1));
        k_tb_emit<<<g, 256, 0, s>>>(lv, cnt, B.scn + L * stride, lvc + L, L, sorted, lv + stride);
    }
    /* chains of nodes per sorted position, then offsets and ids */
    k_tb_fill_i32<<<g, 256, 0, s>>>(B.lmin, n, INT_MAX);
    for (int L = 0; L <= max_depth; ++L)
        k_tb_mark<<<g, 256, 0, s>>>(B.nodes + L * stride, B.cnt + L * stride, lvc + L, L, B.lmin, B.lleaf, B.lcnt);
    k_tb_chain<<<g, 256, 0, s>>>(n, B.lmin, B.lleaf, B.lcnt, B.psize, B.pcount);
    SPH_HIP_TRY(hipGetLastError());
    k_tb_fill_i32<<<1, 256, 0, s>>>(B.scalars + 3, 1, n); /* scan length */
    SPH_TRY_RC(exclusive_scan(ctx, B.psize, B.pbase, B.scalars + 3, n, B.scalars + 1));
    SPH_TRY_RC(exclusive_scan(ctx, B.pcount, B.pnbase, B.scalars + 3, n, B.scalars + 2));
    /* the one host synchronisation: flag, totals and level sizes */
    int32_t h[kScalars] = {};
    SPH_HIP_TRY(hipMemcpyAsync(h, B.scalars, sizeof(h), hipMemcpyDeviceToHost, s));
    SPH_HIP_TRY(hipStreamSynchronize(s));
    if (h[0]) return ctx_fail(ctx, SPH_E_ARG, "non-finite particle position");
    int n_levels = 0;
    while (n_levels <= max_depth && h[kLevelCounts + n_levels] > 0) ++n_levels;
    const int32_t* counts = h + kLevelCounts;
    const int64_t n_ints = h[1];
    const int32_t node_count = h[2];
    int k = 0;
    SPH_TRY_RC(tree_reserve(ctx, n_ints, counts, n_levels, node_count, &k));
    /* buffer k: the compute stream is past its release (recorded on it) */
    int64_t lo = 0;
    for (int L = 0; L < n_levels; ++L) {
        k_tb_fill<<<grid_for(counts[L]), 256, 0, s>>>(B.nodes + L * stride, B.cnt + L * stride, counts[L], L, B.lmin,
                                                      B.pbase, B.pnbase, sorted, ctx->tree.octree_up[k], n_ints,
                                                      ctx->tree.level_tp_up[k] + lo);
        lo += counts[L];
    }
    SPH_HIP_TRY(hipGetLastError());
    return tree_commit(ctx, k, n_ints, counts, n_levels, node_count);
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
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const sph_tree_dev& T = ctx->tree;
    SPH_HIP_TRY(hipMemcpy(octree, T.octree, 4 * (size_t)T.n_ints, hipMemcpyDeviceToHost));
    SPH_HIP_TRY(hipMemcpy(level_nodes, T.level_tp, 4 * (size_t)T.node_count, hipMemcpyDeviceToHost));
    return SPH_OK;
}

}  // extern "C"
