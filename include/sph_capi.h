/*
 * sph_capi.h -- C ABI of the MI355X (gfx950) SPH + FMM step core, exported by
 * libsphhip.so.
 *
 * This is the drop-in boundary for the reference's `class Compute`
 * (SPH/Compute/Compute.hpp:40-148).  The reference encodes each pass into a
 * Metal command buffer; here each pass is enqueued on the context's HIP stream
 * and nothing synchronises until sph_sync().  Plain pointers and sizes only.
 *
 * Conventions
 *  - Return codes: 0 = OK, negative = SPH_E_* error, positive = hipError_t.
 *    sph_last_error() gives the message of the last failure on a context
 *    (with NULL: of the last failed context-free call on this thread).
 *  - Particle arrays are in original-ID order (the order of the initial
 *    conditions), as in the reference.  3-vectors use a 16-byte stride
 *    (x, y, z, unused), matching simd::float3 (Compute.cpp:95-97).
 *  - A context owns all device memory and one compute stream plus one copy
 *    stream.  Calls on one context must come from one thread at a time.
 *    Host buffers stay owned by the caller.
 */
#ifndef SPH_CAPI_H
#define SPH_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
#define SPH_OK 0
#define SPH_E_ARG (-1)       /* invalid argument / size                     */
#define SPH_E_STATE (-2)     /* call not valid in the current state         */
#define SPH_E_NOTREE (-3)    /* gravity requested before a tree was uploaded */
#define SPH_E_CAPACITY (-4)  /* a fixed-size resource would overflow        */
#define SPH_E_NODEVICE (-5)  /* no HIP device visible                       */

/* ------------------------------------------------------------------ params */
/* Reference quirks that have a corrected form (bit flags for sph_params.fixes).
 * Zero = the reference as written. */
#define SPH_FIX_M2M_ZZ        (1u << 0) /* M2M.metal:29 uses expansion[Z] in the ZZ shift */
#define SPH_FIX_SYNC_RANGE    (1u << 1) /* M2M.metal:150 childEnd from the parent's start  */
#define SPH_FIX_SOFTENING     (1u << 2) /* kernels.h:93,97 integer-literal divisions       */
#define SPH_FIX_DENSITY_CACHE (1u << 3) /* density.metal:173-199 density not reset on rescan */
/* Near field (helpers.h:83-147 force-accept every MAC-rejected leaf): with
 * this flag a branch carries a rejected leaf down its unchecked list, and a
 * leaf opens rejected nodes recursively and sums rejected source leaves
 * particle by particle (P2P), as a standard FMM near field does. */
#define SPH_FIX_NEAR_FIELD    (1u << 4)
/* Expansion order 3 (helpers.metal:76-88, accept.metal:47): power[3] with
 * its 1/3 and 1/6 weights summed (as written they are integer divisions and
 * three lines assign instead of add) and r^3 in the MAC (as written r^2). */
#define SPH_FIX_ORDER3        (1u << 5)
/* step.metal:34-45 drifts every particle with h <- h exp(dh/dt dt / h) and
 * rho <- rho exp(-3 dh/dt dt / h), dh/dt from the particle's last active step
 * and h the current one.  Under the README activity rule an inactive particle
 * with dh/dt < 0 then shrinks faster each step (the rate is divided by an h
 * that keeps falling) until h reaches 0 and rho infinity, within a few steps
 * at 1e6 particles.  With this flag the step keeps the log-rate (dh/dt)/h of
 * the last active step (SPH_F_DLNH) and drifts inactive particles with it,
 * and the timestep goal also bounds |dh/dt| dt / h (SPH_DLNH_LIMIT); active
 * particles are drifted exactly as written. */
#define SPH_FIX_INACTIVE_DRIFT (1u << 6)
/* With SPH_FIX_INACTIVE_DRIFT the timestep goal (accelerations.metal:130-136)
 * is also bounded by SPH_DLNH_LIMIT h / |dh/dt|: the kept log-rate then moves
 * h by at most exp(SPH_DLNH_LIMIT) before the particle is recomputed. */
#define SPH_DLNH_LIMIT 0.25f
#define SPH_FIX_ALL           (0x7Fu)

/* Particle activity rule used by the gravity up pass (P2M.metal:31). */
#define SPH_ACTIVE_AS_WRITTEN 0 /* active iff globalTime <= nextActiveTime (every particle, every step) */
#define SPH_ACTIVE_README     1 /* active iff nextActiveTime <= globalTime (README.md:62)            */

typedef struct sph_params {
    /* Hydro (Parameters.h:11-17) */
    float viscosity_alpha;      /* VISCOSITY_ALPHA */
    float density_eta;          /* DENSITY_ETA: NR successive-iterate tolerance */
    float max_smoothing_length; /* MAX_SMOOTHING_LENGTH */
    int32_t max_nr_iterations;  /* MAX_DENSITY_NR_ITTERATIONS */
    int32_t n_neighbours_estim; /* N_NEIGHBOURS_ESTIM: density neighbour-cache size */
    float resolution_eta;       /* RESOLUTION_ETA */
    /* EOS axes (Parameters.h:21-24) */
    float eos_min_rho, eos_max_rho, eos_min_u, eos_max_u;
    /* Time stepping (Parameters.h:27-33) */
    float cfl, dt0, gamma;
    int32_t r_max;              /* R_MAX */
    int32_t activity_rule;      /* SPH_ACTIVE_* */
    int32_t max_rungs;          /* >0: clamp each particle's tick step to dt_min << (max_rungs-1); 0 = no cap */
    /* Gravity (Parameters.h:36-43, poles.h:16) */
    float gravity_eta;          /* GRAVITY_ETA (MAC tolerance) */
    float gravity_g;            /* G in code units */
    float gravity_softening;    /* GRAVITY_SMOOTHING_LENGTH */
    int32_t gravity_max_recursion;   /* GRAVITY_MAX_RECURSION (1..4) */
    int32_t max_unchecked_pointers;  /* MAX_UNCHECKED_POINTERS: per-node list length */
    int32_t max_children_in_leaf;    /* MAX_CHILDREN_IN_LEAF */
    int32_t max_tree_depth;          /* MAX_TREE_DEPTH */
    /* Cells (Parameters.h:47-48) */
    float cell_width;           /* CELL_WIDTH */
    int32_t cell_power;         /* CELL_POWER: 2^(3*cell_power) cells, 1..8 */
    /* Corrections and forms */
    uint32_t fixes;             /* SPH_FIX_* bit set */
    /* Precision form (no counterpart in the reference; needs
     * SPH_FIX_NEAR_FIELD): compensated (Neumaier) sums in the M2L, L2P and
     * P2P accumulators, centres of mass summed relative to a member in
     * P2M/M2M and kept in two parts (FP32 rounding + residual) so that every
     * offset from a node centre (moments, M2M and L2L shifts, M2L separation,
     * L2P) is taken relative to it, and the gravitational potential per
     * particle (SPH_F_POTENTIAL) for the FP64 conservation sums of
     * sph_conservation. */
    int32_t precision;
    /* Lanes per particle in the density pass: 0 = one thread per particle
     * (the reference's summation order), 2/4/8/16/32/64 = the particle's cell
     * box split over that many lanes (same terms, tree-ordered sums). */
    int32_t density_lanes;
    /* The same for the acceleration pass (0 = one thread per particle). */
    int32_t accel_lanes;
    /* With lanes > 0: 0 = each lane takes whole cells of the box, 1 = the
     * lanes walk the box together and split each cell's candidates. */
    int32_t lane_split;
    /* Multipole / local expansion order P (Parameters.h `#define P 2`): 2 or 3.
     * At 3 the reference's `#if P > 2` terms are added: octupoles in P2M/M2M,
     * third derivatives in M2L, third-order locals in L2L/L2P. */
    int32_t expansion_order;
    /* Order of the error estimate in the acceptance test (accept.metal:14-55):
     * 0 = the expansion order (the reference), 2 = the P = 2 estimate with
     * P = 3 expansions.  The P = 3 estimate uses the octupole power alone at
     * order 3, which vanishes for a source symmetric about its centre of mass
     * (two equal masses, say); such a node is then accepted at any distance
     * beyond the geometric limit although its neglected order-4 moment is
     * not small. */
    int32_t mac_order;
    /* Density pass with density_lanes > 0 and lane_split 0: 1 = each lane
     * walks the candidates of its cells as one flat loop (the same
     * candidates in the same order as 0, so the same sums bit for bit);
     * 0 = one loop per cell. */
    int32_t density_flat;
    /* Stopping rule of the density pass's Newton-Raphson iteration
     * (density.metal:73,131-136).  0 = the reference's: stop once successive
     * iterates differ by at most density_eta.  1 (no counterpart in the
     * reference) = stop once the estimated distance of h from the fixed point
     * is at most density_eta: with eps the relative change of the last scan
     * and k = eps / (previous scan's eps) the observed contraction, stop when
     * eps / (1 - k) <= density_eta (two scans at least, never while k >=
     * SPH_NR_MAX_CONTRACTION, and with one FP32 ulp, SPH_NR_EPS_ULP, added
     * to eps and taken from the previous one), or when eps is at the FP32
     * rounding floor (SPH_NR_NOISE_FLOOR).
     * Two iterations started from different guesses then end within about
     * 2 density_eta of each other; under rule 0 they can end density_eta /
     * (1 - k) apart, 20x that for weakly contracting particles. */
    int32_t density_stop;
    /* Density pass with lanes per particle, every particle active: 0 = the
     * waves take the particles in slot order; W (a power of two, 16..512) =
     * within each window of W consecutive slots, in ascending order of the
     * particles' previous Newton-Raphson counts, so that waves of like counts
     * waste fewer scans on their slowest particle (the same results). */
    int32_t density_order;
    /* Lanes per particle of the density and acceleration passes when they
     * walk the active particles only (activity_rule README), as
     * density_lanes and accel_lanes (0 = one thread per particle).  Such a
     * pass has few particles, so its time is the longest particle's chain
     * of dependent loads: more lanes per particle shorten it. */
    int32_t list_density_lanes;
    int32_t list_accel_lanes;
    /* With the active particles only: 1 = the acceleration pass's
     * Morton-ordered records are gathered for the cells some active
     * particle's cell box reaches (the only records the pass reads); 0 =
     * for every particle.  The same results either way. */
    int32_t gather_reach;
    /* Near-field form, leaves in two launches (precision form, P = 3): the
     * M2L launch records the particles of each leaf's rejected source
     * leaves, up to this many (0..4096), and the pair-sum launch sums over
     * that record, a source particle per lane, instead of walking the tree
     * again (0, or more particles: it walks).  The same pairs; only which
     * lane sums which changes. */
    int32_t near_list_cap;
    /* Every particle active (no slot list) and fewer than this many
     * particles (0 = off): the density pass takes 32 lanes per particle and
     * the acceleration pass 16 below 65,536 particles, else 8, in place of
     * density_lanes and accel_lanes.  With one thread (or 8 lanes) per
     * particle such a pass leaves most of the GPU's wave slots empty and
     * lasts as long as its longest particle's chain of dependent loads.  The
     * variants are those of density_lanes and accel_lanes: the same terms,
     * tree-ordered sums. */
    int32_t small_n_lanes;
} sph_params;
/* Relative change of h (2 ulp) at which density_stop = 1 stops whatever k. */
#define SPH_NR_NOISE_FLOOR 2.5e-7f
#define SPH_NR_EPS_ULP 1.2e-7f
#define SPH_NR_MAX_CONTRACTION 0.995f

/* One EOS table: res x res float2 (pressure, sound speed) in code units.
 * Row index = density sample, column index = internal-energy sample, sample i
 * at exponent i/res between the axis limits (ANEOS.cpp:123-136). */
typedef struct sph_eos_table {
    int32_t material_id;        /* 400 = forsterite slot, anything else = Fe85Si15 slot (density.metal:45-50) */
    int32_t res;
    const float* pc;            /* 2*res*res floats, row-major [rho][u] */
} sph_eos_table;

/* Initial conditions in the reference's DataStruct layout (hdfHandler.hpp:15-26). */
typedef struct sph_ic_view {
    int32_t n;
    const float* pos4;          /* n x 4 floats; the fourth is ignored (the library keeps the mass there) */
    const float* vel4;          /* n x 4 floats */
    const float* rho;
    const float* u;
    const float* m;
    const int32_t* mat;
    const float* p;
    const float* h;
} sph_ic_view;

typedef struct sph_ctx sph_ctx;

/* Fills the reference's Parameters.h defaults. */
void sph_params_default(sph_params* out);

/* Compute::Compute + buildBuffers + loadInitialConditions (Compute.cpp:22-84,
 * 92-132): allocates every device buffer, uploads the initial conditions and
 * the two EOS tables, sets active = true, nextActiveTime = 0, globalTime = 0. */
int sph_create(const sph_params* params, const sph_ic_view* ic,
               const sph_eos_table* tables, int ntables, sph_ctx** out);
void sph_destroy(sph_ctx* ctx);
const char* sph_last_error(const sph_ctx* ctx);
int sph_num_particles(const sph_ctx* ctx);

/* ------------------------------------------------------------ octree handoff */
/* Pinned host memory for the octree thread's staging buffers. */
void* sph_host_alloc(size_t bytes);
void sph_host_free(void* p);

/* Copies the current positions into a device snapshot buffer on the compute
 * stream (ordered after every pass enqueued so far), then into `pinned_dst4`
 * (n x 4 floats) on the copy stream.  Replaces the CPU read of the shared
 * position buffer in Compute::updateOctreeData (Compute.cpp:511). */
int sph_snapshot_positions(sph_ctx* ctx, float* pinned_dst4);
/* Blocks the calling thread until the last snapshot has landed in host memory.
 * The one call that may come from another thread (the octree builder) while
 * the owning thread keeps enqueuing passes. */
int sph_wait_snapshot(sph_ctx* ctx);

/* Compute::updateOctreeBuffer (Compute.cpp:456-508): takes a tree in the
 * reference's flat layout (octree.cpp:14-141) -- `octree` = octreeData,
 * `level_nodes` = the concatenated treeLevels lists (tree offsets),
 * `level_counts[l]` = size of treeLevels[l], `node_count` = nodeValues.
 * The upload runs on the copy stream; passes enqueued after this call use the
 * new tree.  When `octree`/`level_nodes` are pinned the call does not block. */
int sph_upload_tree(sph_ctx* ctx, const int32_t* octree, int64_t n_ints,
                    const int32_t* level_nodes, const int32_t* level_counts,
                    int n_levels, int32_t node_count);
/* 1 once the last uploaded tree's host buffers may be reused, else 0. */
int sph_tree_upload_done(sph_ctx* ctx);

/* Builds the octree on the GPU from the particles' current positions, in
 * place of a CPU build (octree.cpp:14-141) and its upload: no counterpart in
 * the reference.  The tree splits the positions' bounding cube at its
 * midpoints (Morton keys of 10 bits per axis) rather than octree.cpp's
 * particle means, with the same leaf size (max_children_in_leaf) and depth
 * limit (max_tree_depth, at most 10 here), and is laid out as octree.cpp lays
 * out its own: flat records, pre-order node ids, one offset list per level.
 * Passes enqueued after this call use it.  SPH_E_ARG for a non-finite
 * position or max_tree_depth > 10.  Blocks until the tree's size is known. */
int sph_build_tree(sph_ctx* ctx);
/* Size of the current tree: octreeData ints, nodes, levels, and the first
 * `level_cap` level sizes into `level_counts` (may be NULL). */
int sph_tree_info(sph_ctx* ctx, int64_t* n_ints, int32_t* node_count, int32_t* n_levels,
                  int32_t* level_counts, int32_t level_cap);
/* Copies the current tree to the host: `octree` (n_ints ints) and the
 * concatenated level lists `level_nodes` (node_count ints). */
int sph_download_tree(sph_ctx* ctx, int32_t* octree, int32_t* level_nodes);

/* ---------------------------------------------------------- snapshots */
/* drawToSnapshot (render.metal:49-66) as Compute::drawSnapshot enqueues it
 * (Compute.cpp:396-414): clears a res x res image of int32 material ids and
 * marks each particle whose (x, y) lies in [250, 250 + 20 R_E)^2 with the
 * largest material id landing on its pixel (row 0 = largest y).  Enqueued on
 * the context's stream like a pass.  The reference's res is
 * SNAPSHOT_RESOLUTION. */
int sph_draw_snapshot(sph_ctx* ctx, int res);
/* Copies the last drawn image (res * res int32) to the host; synchronous. */
int sph_download_snapshot(sph_ctx* ctx, int32_t* out);
/* res of the last drawn image (0 if none). */
int sph_snapshot_resolution(const sph_ctx* ctx);

/* ----------------------------------------------------------------- passes */
int sph_sort(sph_ctx* ctx);          /* Compute::sort               Compute.cpp:265-326 */
int sph_gravity(sph_ctx* ctx);       /* Compute::gravitationalPass  Compute.cpp:328-377 */
int sph_density(sph_ctx* ctx);       /* Compute::densityPass        Compute.cpp:179-208 */
int sph_acceleration(sph_ctx* ctx);  /* Compute::accelerationPass   Compute.cpp:210-240 */
int sph_step(sph_ctx* ctx);          /* Compute::stepPass           Compute.cpp:242-263 */
/* Headless::step's commit + waitUntilCompleted (Headless.cpp:53-54). */
int sph_sync(sph_ctx* ctx);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the hash, radix sort and cell start/end of sph_sort,
 * without a context, calling no hipMalloc / hipFree and no allocating
 * library routine.  Inputs and outputs are the caller's device buffers:
 *   pos4          float x 4n, positions with a 16-byte stride (w unused)
 *   sorted_pairs  uint32 x 2n, out: (cell key, particle index) pairs in
 *                 (key, index) order, as SPH_F_CELL_ARRAY
 *   cell_se       uint32 x 2 x 2^(3 cell_power), out: per cell (first slot,
 *                 last slot), (UINT_MAX, 0) for an empty cell, as
 *                 SPH_F_CELL_START / SPH_F_CELL_END interleaved
 * cell_width and cell_power are sph_params' (CELL_WIDTH, CELL_POWER).  The
 * sort's working memory is a fixed device array of the library (about
 * 34 MiB, held by each GPU it runs on until the process ends), which limits
 * n to SPH_SORT_NOALLOC_MAX_N: larger n returns SPH_E_CAPACITY.  Work is
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) on the
 * current device and is not waited for.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (other streams or threads) race.
 * Failures return a code; sph_last_error(NULL) gives the message. */
#define SPH_SORT_NOALLOC_MAX_N 4194304
int sph_sort_noalloc(const float* pos4, int32_t n, float cell_width, int32_t cell_power,
                     uint32_t* sorted_pairs, uint32_t* cell_se, void* stream);

/* ------------------------------------------------------------- state access */
enum sph_field {
    SPH_F_POS = 0,        /* float4 x n: x, y, z and the particle's mass (kept by the library) */
    SPH_F_VEL,            /* float4 x n */
    SPH_F_ACC,            /* float4 x n */
    SPH_F_RHO,            /* float x n  */
    SPH_F_U,
    SPH_F_MASS,
    SPH_F_PRESSURE,
    SPH_F_H,
    SPH_F_MAT,            /* int32 x n  */
    SPH_F_GRADTERM,
    SPH_F_SOUNDSPEED,
    SPH_F_DU,
    SPH_F_BALSARA,
    SPH_F_DHDT,
    SPH_F_ACTIVE,         /* uint8 x n  */
    SPH_F_NEXT_ACTIVE,    /* int32 x n  */
    SPH_F_GRAVNORM,       /* float x n  */
    SPH_F_CELL_ARRAY,     /* uint32 x 2n: sorted (cell key, particle index) pairs */
    SPH_F_CELL_START,     /* uint32 x 2^(3*cell_power): first sorted slot, UINT_MAX if empty */
    SPH_F_CELL_END,       /* uint32 x 2^(3*cell_power): last sorted slot (inclusive), 0 if empty */
    SPH_F_GLOBAL_TIME,    /* int32 x 1 (ticks of DT_MIN1) */
    SPH_F_DT,             /* int32 x 1 (global tick step of the last acceleration pass) */
    SPH_F_NR_ITERS,       /* int32 x n: NR iterations of the last density pass (diagnostic) */
    SPH_F_MULTIPOLES,     /* float x 24 x node_count, see below */
    SPH_F_LOCALS,         /* float x 16 x node_count, see below */
    SPH_F_DENSITY_WORK,   /* int32 x 4n: neighbour candidates examined and candidates inside 2h
                             over the NR scans, then the same two counts for the final pass,
                             of the last density pass (diagnostic) */
    SPH_F_POTENTIAL,      /* float x n: gravitational potential (G included) of the last
                             gravity pass in the precision form, else 0 */
    SPH_F_DLNH,           /* float x n: (dh/dt)/h of each particle's last active step, kept by
                             the step pass under SPH_FIX_INACTIVE_DRIFT (0 until then) */
    SPH_F_COUNT
};
/* Multipole record (24 floats, poles.h:46-55 reordered):
 *   0-2 pos, 3 size, 4-6 min, 7 minGrav, 8-10 max, 11 active (0/1),
 *   12 M, 13-15 X Y Z (always 0: never consumed at P = 2),
 *   16-21 XX XY XZ YY YZ ZZ, 22 power[2], 23 power[1] (= 0); power[0] = M.
 * Local record (16 floats, poles.h:58-61): 0-2 pos, 3 unused,
 *   4-13 expansion M X Y Z XX XY XZ YY YZ ZZ, 14-15 unused. */

int64_t sph_field_bytes(const sph_ctx* ctx, int field);
/* Synchronous copies (they drain the compute stream first). */
int sph_download(sph_ctx* ctx, int field, void* dst);
int sph_upload(sph_ctx* ctx, int field, const void* src);
/* Device address and size of a per-particle field, for zero-copy readers
 * such as a renderer: replaces the reference's public positionBuffer,
 * materialIdBuffer and densityBuffer (Compute.hpp:55-57, bound by
 * Render.cpp:26), and the other per-particle buffers alike.  The layout is
 * sph_download's (original-ID order, float4 stride for 3-vectors).  The
 * address stays valid for the context's life; the contents are those of
 * the passes enqueued so far once sph_sync has returned (the reference reads
 * its buffers after waitUntilCompleted).  Read-only for the caller: writes
 * go through sph_upload, which keeps the context's derived copies in step.
 * SPH_E_ARG for SPH_F_ACC (until the step pass adds them, the acceleration
 * pass's pair sums sit in a buffer of their own: use sph_download) and for
 * fields that are not per-particle arrays. */
int sph_device_field(sph_ctx* ctx, int field, void** dev_ptr, int64_t* bytes);

/* ------------------------------------------------------------------ timing */
enum sph_pass {
    SPH_PASS_SORT = 0, SPH_PASS_GRAVITY_UP, SPH_PASS_GRAVITY_DOWN, SPH_PASS_DENSITY,
    SPH_PASS_ACCELERATION, SPH_PASS_STEP, SPH_PASS_TREE_BUILD, SPH_PASS_COUNT
};
/* Enables HIP-event timing of every pass (0 = off, the default). */
int sph_set_timing(sph_ctx* ctx, int enabled);
/* Accumulated milliseconds per pass since the last reset (syncs the stream).
 * `ms` has SPH_PASS_COUNT entries; `calls` (optional) counts the passes. */
int sph_pass_times(sph_ctx* ctx, double* ms, int64_t* calls);
int sph_reset_pass_times(sph_ctx* ctx);
/* Measurement aid (no counterpart in the reference): while enabled, density
 * passes record per particle the reference algorithm's work in
 * SPH_F_DENSITY_WORK (candidates of the cell boxes and those inside 2h, Newton-
 * Raphson scans and final pass).  Off by default: SPH_F_DENSITY_WORK then
 * keeps its last content. */
int sph_set_work_counters(sph_ctx* ctx, int enabled);

/* ------------------------------------------------------------ diagnostics */
/* FP64 sums over all particles, reduced on the device (no counterpart in
 * the reference): out[0] kinetic energy sum m v^2 / 2, out[1] internal
 * energy sum m u, out[2] potential energy sum m phi / 2 (precision form;
 * 0 otherwise), out[3..5] momentum sum m v, out[6..8] angular momentum
 * sum m x cross v about the origin, out[9] total mass. */
#define SPH_CONSERVATION_COUNT 10
int sph_conservation(sph_ctx* ctx, double* out);
/* Simulation time in seconds: the integer tick counter (ticks of
 * DT_MIN1 = DT0 / 2^R_MAX, exact) times DT_MIN1, in FP64. */
int sph_time(sph_ctx* ctx, double* seconds);

/* ------------------------------------------------ learned smoothing length */
/* None of this has a counterpart in the reference.  A small network
 * (12 -> 64 -> 64 -> 1, SiLU, BF16 matrix-core products with FP32
 * accumulation) predicts d = log(h_converged) - log(h_start) from 12
 * features of a particle at the start of a density pass; the density pass
 * then starts its Newton-Raphson iteration (density.metal:73-143) from
 * h_start exp(d) instead of h_start, and falls back to h_start when the
 * residual grows.  Convergence is still decided by sph_params.density_eta.
 * Features: 0 log h, 1 (dh/dt) dt / h (dt = DT_MIN1 x the last tick step),
 * 2 log rho, 3-7 log(1 + count) over the 27 cells around the particle's cell
 * (all, own, 6 faces, 12 edges, 8 corners), 8 |sum count x offset| /
 * (n27 + 1), 9 (max - min count) / (n27 + 1), 10-11 material one-hot
 * (400, other).  Parameters (5057 floats): W1[64][12], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  Normalisation (26 floats): feature means[12],
 * feature scales[12], target mean, target scale.  Sample record (16 floats):
 * 12 features, d, weight, 2 unused. */
#define SPH_MLP_INPUTS 12
#define SPH_MLP_HIDDEN 64
#define SPH_MLP_PARAMS 5057
#define SPH_MLP_NORM 26
#define SPH_SAMPLE_FLOATS 16

typedef struct sph_mlp sph_mlp;
typedef struct sph_samples sph_samples;

/* Network on the current device, weights drawn from
 * Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) with a splitmix64 stream. */
int sph_mlp_create(uint64_t seed, sph_mlp** out);
void sph_mlp_destroy(sph_mlp* m);
const char* sph_mlp_last_error(const sph_mlp* m);
/* Parameters / normalisation in and out (either pointer may be NULL).
 * Setting parameters resets the Adam state. */
int sph_mlp_get(sph_mlp* m, float* params, float* norm);
int sph_mlp_set(sph_mlp* m, const float* params, const float* norm);
/* Prediction d for n feature rows (host arrays, n x 12 -> n). */
int sph_mlp_forward(sph_mlp* m, const float* x, int64_t n, float* d_out);
/* Loss mean_s w_s (y_s - (d_s - mean) / scale)^2 over the n samples and its
 * gradient (5057 floats), computed by the training kernels. */
int sph_mlp_loss_grad(sph_mlp* m, const float* x, const float* d, const float* w, int64_t n,
                      float* loss, float* grad);
/* Normalisation from the weighted samples (means and standard deviations). */
int sph_mlp_fit_norm(sph_mlp* m, const sph_samples* s);
/* `steps` Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on batches of `batch`
 * samples (multiple of 32) drawn with a counter hash of `seed`; `losses`
 * (optional, `steps` floats) receives the loss of every step. */
int sph_mlp_train(sph_mlp* m, const sph_samples* s, int steps, int batch, float lr, uint64_t seed,
                  float* losses);

/* Device buffer of sample records. */
int sph_samples_create(int64_t capacity, sph_samples** out);
void sph_samples_destroy(sph_samples* s);
int64_t sph_samples_count(const sph_samples* s);
int sph_samples_read(const sph_samples* s, int64_t first, int64_t count, float* dst);
int sph_samples_append(sph_samples* s, const float* src, int64_t count);
int sph_samples_clear(sph_samples* s);

/* With a sample buffer set, every density pass appends one record per
 * particle (weight 1 if active, else 0); NULL switches dumping off. */
int sph_set_sample_dump(sph_ctx* ctx, sph_samples* s);
/* With a network set, density passes start from its guess; NULL = plain. */
int sph_set_learned_guess(sph_ctx* ctx, sph_mlp* m);
/* Features of the current state (n x 12, original order); needs a sort. */
int sph_features(sph_ctx* ctx, float* dst);

#ifdef __cplusplus
}
#endif
#endif /* SPH_CAPI_H */
