/*
 * sph_internal.h -- device memory layout of one simulation context and the
 * launchers shared between the translation units of libsphhip.so.
 *
 * Layout in HBM (n particles, C = 2^(3*cell_power) cells, K nodes):
 *   particle state, original-ID order (the order the caller sees):
 *     pos, vel, acc            float4 x n   (xyz + unused w, 16-B stride)
 *     rho u m P h f c du B dhdt gravnorm   float x n
 *     mat, next_active, nr_iters            int32 x n;  active uint8 x n
 *   cells: pairs_a / pairs_b  uint2 x n (sorted (key, index), ping-pong)
 *          cell_se            uint2 x C (first slot, last slot), 128 MiB at C = 2^24
 *   Morton-ordered copies written once per pass and read by every
 *   neighbour candidate (slot k = k-th sorted particle):
 *          s_posm  float4 (x, y, z, m)        -- written after the sort
 *          s_vel   float4 (vx, vy, vz, 0)     -- written after the sort
 *   and at the start of the acceleration pass, so that a rejected candidate
 *   costs one 16-byte load:
 *          s_posh  float4 (x, y, z, h)
 *          s_velb  float4 (vx, vy, vz, Balsara factor)
 *          s_hyd   float4 (m, f P / rho^2, rho, c)
 *   tree (node id = pre-order node number):
 *          mp_far  3 x float4 x K  (pos, size) (M, XX, XY, XZ) (YY, YZ, ZZ, power2)
 *          mp_aux  3 x float4 x K  (min, minGrav) (max, active) (X, Y, Z, -) dipoles
 *          loc     4 x float4 x K  (pos, L_M) (L_X, L_Y, L_Z, L_XX) (L_XY, L_XZ, L_YY, L_YZ) (L_ZZ, -, -, -)
 *          mp_oct, loc3  3 x float4 x K each: octupoles + power3, third-order
 *                  local terms (XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ), P = 3
 *          node_info int4 x K (nParticles, treePointer, level, -), parent_slot int32 x K
 *          subtree_end int32 x K (first node id after the node's pre-order subtree)
 *          node_child int32 x 8K (child node ids, -1)
 *          lists a/b int32 x (max nodes of a non-last level) x max_unchecked_pointers
 */
#ifndef SPH_INTERNAL_H
#define SPH_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sph_capi.h"

/* propagates an SPH_* return code */
#define SPH_TRY_RC(expr)                   \
    do {                                   \
        const int _rc = (expr);            \
        if (_rc != SPH_OK) return _rc;     \
    } while (0)
#define SPH_HIP_TRY(expr)                                              \
    do {                                                               \
        hipError_t _e = (expr);                                        \
        if (_e != hipSuccess) return ctx_fail_hip(ctx, _e, #expr);     \
    } while (0)

template <typename T>
inline hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
}

template <typename T>
inline void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

/* work buffers of the GPU octree build (treebuild.hip) */
struct sph_treebuild {
    uint2* pairs[2] = {nullptr, nullptr}; /* (Morton key, particle) */
    float* bbox_part = nullptr;           /* 6 per workgroup */
    float* box = nullptr;                 /* lo xyz, hi xyz */
    double* scale = nullptr;              /* 1024 / cube side */
    int4* chain = nullptr;                /* per sorted position: top level, leaf level, leaf size */
    int32_t *pbase = nullptr, *pnbase = nullptr; /* per position: flat-array offset, node id */
    int32_t* bsum = nullptr;              /* scan block sums */
    int32_t* scalars = nullptr;           /* flag, totals, scan length, level sizes */
    uint2* lv[2] = {nullptr, nullptr};    /* (level, offset) per node, sorted by level */
    uint32_t* tile_hist = nullptr;        /* radix sort tile table, 256 x tile_cap */
    uint32_t* digit_total = nullptr;      /* 256 */
    int bbox_blocks = 0, bsum_cap = 0, n_cap = 0, tile_cap = 0;
    int64_t lv_cap = 0;
};
void sph_treebuild_free(sph_treebuild& B);

struct sph_tree_dev {
    int32_t* octree = nullptr;      /* raw octreeData (current) */
    int32_t* octree_up[2] = {nullptr, nullptr}; /* upload double buffer */
    int64_t octree_cap = 0;         /* ints */
    int32_t* level_tp = nullptr;    /* concatenated treeLevels (tree offsets, current) */
    int32_t* level_tp_up[2] = {nullptr, nullptr};
    int up_cur = 0;                 /* buffer holding the current tree */
    hipEvent_t released[2] = {nullptr, nullptr}; /* compute stream done with buffer k */
    int32_t* level_id = nullptr;    /* same slots, node ids */
    int32_t* act_slots = nullptr;   /* same slots: per level, the slots of its active nodes (README rule) */
    int32_t* act_count = nullptr;   /* 2 kMaxLevels counters of act_slots: leaves, then branches */
    int64_t level_cap = 0;
    int4* node_info = nullptr;      /* (nParticles, treePointer, level, 0) */
    int32_t* parent_slot = nullptr; /* slot of the parent in its level list (parentIndexes) */
    int32_t* parent_level = nullptr;/* level of the parent (validation) */
    int32_t* seen = nullptr;        /* times each node id occurs in the level lists (validation) */
    int32_t* subtree_end = nullptr; /* node id + pre-order subtree size (written by the up pass) */
    int64_t n_ints = 0;             /* size of the current octreeData */
    int32_t* node_child = nullptr;  /* 8 per node */
    float4* mp_far = nullptr;       /* 3 per node */
    float4* mp_aux = nullptr;       /* 3 per node */
    float4* mp_oct = nullptr;       /* 3 per node: octupoles + power[3] (expansion order 3) */
    float4* mp_clo = nullptr;       /* 1 per node (precision form): the centre's low part, centre = pos + clo */
    float4* loc3 = nullptr;         /* 3 per node: third-order local terms (expansion order 3) */
    float4* loc = nullptr;          /* 4 per node */
    float4* loc_tot = nullptr;      /* 7 per node: a near-field leaf's complete local expansion */
    int32_t* near_rec = nullptr;    /* near_list_cap per node: a near-field leaf's rejected source leaves */
    int32_t* near_cnt = nullptr;    /* per node: their count, -1 when not recorded */
    int64_t node_cap = 0;
    int32_t node_count = 0;
    std::vector<int32_t> level_counts, level_offsets;
    int32_t max_nonlast = 0;        /* most nodes in any level but the last */
};

struct sph_ctx {
    sph_params prm{};
    int n = 0;
    int ncells = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;
    /* the gravity down pass runs on its own stream, concurrently with the
     * density pass (which neither reads nor writes what it does); every other
     * call first makes `stream` wait for it (join_gravity) */
    hipStream_t gstream = nullptr;
    hipEvent_t ev_up = nullptr, ev_down = nullptr;
    bool grav_pending = false;
    /* an uploaded tree is converted on copy_stream, overlapping the sort;
     * join_tree makes `stream` wait for it (convert_done) */
    hipEvent_t tree_free = nullptr, convert_done = nullptr;
    bool convert_pending = false;
    std::string err;

    /* particle state, original order */
    float4 *pos = nullptr, *vel = nullptr, *acc = nullptr;
    /* the acceleration pass's pair sums: the step (or any call that reads acc
     * first) adds them to acc, so that the pass need not wait for the gravity
     * down pass, which writes acc on its own stream */
    float4* dv = nullptr;
    bool dv_pending = false;
    float *rho = nullptr, *u = nullptr, *m = nullptr, *P = nullptr, *h = nullptr;
    float *f = nullptr, *c = nullptr, *du = nullptr, *B = nullptr, *dhdt = nullptr, *gravnorm = nullptr;
    float* dlnh = nullptr; /* (dh/dt)/h of the last active step (SPH_FIX_INACTIVE_DRIFT) */
    float* phi = nullptr; /* potential (precision form) */
    int32_t *mat = nullptr, *next_active = nullptr, *nr_iters = nullptr;
    int4* density_work = nullptr; /* NR (candidates, inside 2h), final pass (candidates, inside) */
    uint8_t* active = nullptr;
    int32_t* time = nullptr; /* [0] globalTime, [1] dt */

    /* cells and sort */
    uint2 *pairs_a = nullptr, *pairs_b = nullptr; /* sorted result always ends in pairs_a */
    uint2* prev_keys_src = nullptr;               /* sorted pairs of the previous sort (for the sparse reset) */
    bool have_prev = false;
    uint2* cell_se = nullptr;
    uint32_t* tile_hist = nullptr; /* 256 x ntiles, digit-major */
    uint32_t* digit_total = nullptr; /* 256: pairs per digit of the current pass */
    int ntiles = 0;

    /* Morton-ordered copies; morton_fresh: s_posm / s_vel still equal the
     * particles' positions, masses and velocities (set by a sort, cleared by a
     * step or an upload of those fields) */
    bool morton_fresh = false;
    /* the density pass fills density_work (and runs without its neighbour
     * lists) only while this is set: sph_set_work_counters */
    bool count_work = false;
    float4 *s_posm = nullptr, *s_vel = nullptr;
    float4 *s_posh = nullptr, *s_velb = nullptr, *s_hyd = nullptr;

    /* EOS tables: slot 0 = material 400, slot 1 = every other material */
    float2* eos[2] = {nullptr, nullptr};
    int eos_res = 0;

    /* tree */
    sph_tree_dev tree;
    bool have_tree = false;
    int32_t *list_a = nullptr, *list_b = nullptr;
    int64_t list_cap = 0; /* entries per list buffer */
    uint32_t* grav_flags = nullptr; /* [0] != 0: the last uploaded tree failed validation */
    bool tree_check_pending = false;
    /* gravnorm holds a field (a gravity pass ran or it was uploaded) */
    bool gravnorm_known = false;

    sph_treebuild tb; /* GPU octree build */

    /* snapshot raster (render.hip): material id per pixel, res x res */
    int32_t* snapshot_img = nullptr;
    int64_t snapshot_cap = 0;
    int snapshot_res = 0;

    /* octree handoff */
    float4* snap = nullptr;
    hipEvent_t snap_done = nullptr;
    hipEvent_t upload_done = nullptr;
    bool upload_pending = false;

    /* timing: event pairs recorded around each pass, folded on query */
    struct TimedPass {
        int pass;
        hipEvent_t start, stop;
    };
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<TimedPass> ev_pending;
    hipEvent_t ev_open[SPH_PASS_COUNT] = {};
    double pass_ms[SPH_PASS_COUNT] = {};
    int64_t pass_calls[SPH_PASS_COUNT] = {};

    /* learned smoothing-length guess (learned.hip) */
    float4* feats = nullptr;        /* 3 float4 (12 features) per sorted slot */
    float* guess = nullptr;         /* starting h per sorted slot */
    int32_t* nr_order = nullptr;
    /* per cell: the epoch of the last acceleration pass over the active
     * particles whose cell boxes reach it (1..255, cycling) */
    uint8_t* cell_mark = nullptr;
    int mark_epoch = 0;    /* density pass's slot order by previous iteration count (density_order) */
    int32_t* slot_list = nullptr;   /* sorted slots of the active particles (README rule) */
    int32_t* slot_count = nullptr;  /* their number (device) */
    bool slot_list_valid = false;   /* cleared by a sort and by a gravity pass */
    sph_mlp* learned = nullptr;     /* not owned */
    sph_samples* dump = nullptr;    /* not owned */
};

int ctx_fail(sph_ctx* ctx, int code, const std::string& msg);
int ctx_fail_hip(sph_ctx* ctx, hipError_t e, const char* what);
int ctx_begin_pass(sph_ctx* ctx, int pass);
int ctx_end_pass(sph_ctx* ctx, int pass);
int ctx_fold_times(sph_ctx* ctx);

/* sort.hip */
int launch_sort(sph_ctx* ctx);
/* hydro.hip */
/* regathers s_posm / s_vel from original order unless morton_fresh */
int refresh_morton(sph_ctx* ctx);
int launch_density(sph_ctx* ctx);
int launch_acceleration(sph_ctx* ctx);
/* acc += dv for the last acceleration pass's particles, if not yet done */
int fold_dv(sph_ctx* ctx);
int launch_step(sph_ctx* ctx);
/* learned.hip */
int launch_density_prologue(sph_ctx* ctx);
int launch_density_epilogue(sph_ctx* ctx);
/* gravity.hip */
/* node records, links and checks of the current tree's upload, on stream st */
int launch_tree_convert(sph_ctx* ctx, hipStream_t st);
int launch_gravity_up(sph_ctx* ctx);
/* makes ctx->stream wait for a gravity down pass still running on gstream */
int join_gravity(sph_ctx* ctx);
/* makes ctx->stream wait for a tree upload's conversion still running on
 * copy_stream (every call that reads the tree's node records or the
 * malformed-tree flag) */
int join_tree(sph_ctx* ctx);
/* stable LSD radix sort of (key, index) pairs by the low `passes` key bytes */
int launch_sort_pairs(sph_ctx* ctx, uint2* a, uint2* b, int n, int passes, uint32_t* tile_hist, int tile_cap,
                      uint32_t* digit_total, uint2** out);
int sort_tile_size(); /* pairs per tile of the radix sort */
/* Message of the last failed context-free call on this thread. */
const char* noalloc_last_error();
/* device tree buffers for a tree of the given size: checks, grows, and picks
 * the upload slot; tree_commit makes that slot current and converts it */
int tree_reserve(sph_ctx* ctx, int64_t n_ints, const int32_t* level_counts, int n_levels, int32_t node_count,
                 int* slot);
int tree_commit(sph_ctx* ctx, int slot, int64_t n_ints, const int32_t* level_counts, int n_levels,
                int32_t node_count, hipStream_t convert_stream);
/* sorted slots of the active particles into slot_list / slot_count (README rule) */
int launch_active_list(sph_ctx* ctx);
/* bootstrap: the as-written P = 2 down pass (leaves force-accept), used by
 * the near-field form to estimate |a| before its first evaluation */
int launch_gravity_down(sph_ctx* ctx, bool bootstrap = false);
int launch_fill_cells(sph_ctx* ctx);

/* ---------------------------------------------------------- device helpers */
/* MI355X has 8 XCDs, each with its own L2, and hands consecutive workgroup
 * ids to the XCDs in turn.  Launches over Morton-ordered particles or level
 * slots remap the hardware workgroup id so that XCD x runs one contiguous
 * 1/8 of the logical blocks: neighbouring particles then share an L2.
 * Launch with xcd_grid(blocks) workgroups; -1 marks a padding workgroup. */
constexpr int kXcds = 8;
/* levels the active-slot lists of the down pass cover, and the largest grid
 * of its persistent launches over them */
constexpr int kMaxLevels = 64;
constexpr int kDownBlocks = 2048;
/* the largest grid of the hydro passes' persistent launches over the active
 * particles' slots (README rule) */
constexpr int kListBlocks = 2048;
inline unsigned xcd_grid(unsigned blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ int xcd_block(unsigned blocks) {
    const unsigned per = (blocks + kXcds - 1) / kXcds;
    const unsigned b = blockIdx.x;
    const unsigned logical = (b % kXcds) * per + b / kXcds;
    return logical < blocks ? (int)logical : -1;
}

/* Morton interleave of three <= 10-bit values (morton.h:16-34). */
__device__ __forceinline__ uint32_t dev_expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

#endif
