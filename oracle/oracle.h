/*
 * oracle.h -- CPU restatement of the reference's per-step SPH + FMM core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline: only the tests, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product (libsphhip.so and the host library)
 * never links or calls it.
 *
 * Every function restates one reference kernel or helper in plain C++ with
 * FP32 arithmetic (all literals float, as in Metal Shading Language, which has
 * no double), OpenMP over the reference's thread index where the reference
 * runs threads independently.  Each function cites the reference file:line it
 * follows.  Built with -ffp-contract=off so that no multiply-add is fused.
 *
 * Pins (which test checks which function) are listed in oracle/README.md.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>
#include "../include/sph_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device buffers of Compute (Compute.cpp:92-132) as host arrays.  3-vectors use
 * a 16-byte stride.  Arrays are owned by the caller. */
typedef struct orc_state {
    int32_t n;
    float* pos;      /* 4n */
    float* vel;      /* 4n */
    float* acc;      /* 4n */
    float* rho;
    float* u;
    float* m;
    float* pressure;
    float* h;
    int32_t* mat;
    float* gradterm;
    float* soundspeed;
    float* du;
    float* balsara;
    float* dhdt;
    uint8_t* active;
    int32_t* next_active;
    float* gravnorm;
    uint32_t* cell_array;   /* 2n, sorted (key, index) */
    uint32_t* cell_start;   /* 2^(3*cell_power) */
    uint32_t* cell_end;     /* 2^(3*cell_power) */
    int32_t* global_time;   /* 1 */
    int32_t* dt;            /* 1 */
    int32_t* nr_iters;      /* n, optional (may be NULL) */
    const float* eos_forsterite; /* 2*res*res */
    const float* eos_fe;         /* 2*res*res */
    int32_t eos_res;
    float* dlnh;                 /* n: (dh/dt)/h of the last active step (SPH_FIX_INACTIVE_DRIFT) */
} orc_state;

/* ---- utils/morton.h, sort.metal ---- */
uint32_t orc_morton3d(uint32_t x, uint32_t y, uint32_t z);
void orc_hash(const sph_params* prm, const orc_state* s);
/* hist/scan/sum/sort x SORTING_ITTERATIONS, ping-pong as Compute.cpp:276-312. */
void orc_radix_sort(const sph_params* prm, const orc_state* s);
void orc_find_cell_positions(const sph_params* prm, const orc_state* s);
void orc_sort(const sph_params* prm, const orc_state* s);

/* ---- utils/kernels.h (scalars exposed for the kernel tests) ---- */
float orc_W(float r, float h1);
void orc_gradW(const float* x_ij3, float r, float r1, float h1, float* out3);
float orc_dW_dh(float r, float h1);
float orc_dphi_dr(const float* x_ij3, float h, int fixed);
void orc_cell_range(const sph_params* prm, const float* pos3, float h, int32_t* out6);
void orc_eos_sample(const sph_params* prm, const float* table, int32_t res,
                    float u, float rho, float* out2);

/* ---- density.metal, accelerations.metal, step.metal ---- */
void orc_density(const sph_params* prm, const orc_state* s);
void orc_acceleration(const sph_params* prm, const orc_state* s);
void orc_step(const sph_params* prm, const orc_state* s);

/* ---- octree/octree.cpp ---- */
typedef struct orc_tree {
    int32_t* octree;        /* octreeData */
    int64_t n_ints;
    int32_t* level_nodes;   /* concatenated treeLevels */
    int32_t* level_counts;  /* per level */
    int32_t n_levels;
    int32_t node_count;     /* nodeValues */
} orc_tree;
/* Builds into library-owned storage; free with orc_free_tree. */
int orc_build_octree(const float* pos4, int32_t n, int32_t max_leaf, int32_t max_depth, orc_tree* out);
void orc_free_tree(orc_tree* t);
/* The build's GPU tree restated: an octree over 30-bit Morton keys of the
 * bounding cube (max_depth <= 10), in octree.cpp's flat layout.  -1 for a
 * non-finite coordinate, -2 for bad limits. */
int orc_build_octree_morton(const float* pos4, int32_t n, int32_t max_leaf, int32_t max_depth, orc_tree* out);

/* ---- gravity/** ---- */
/* Records: 40 floats per multipole and 32 per local; the first 24 / 16 are
 * the sph_capi.h SPH_F_MULTIPOLES / SPH_F_LOCALS layouts, the rest hold the
 * third-order terms when sph_params.expansion_order is 3. */
typedef struct orc_gravity_buffers {
    float* multipoles;       /* 40 x node_count */
    float* locals;           /* 32 x node_count */
    uint32_t* parent_index;  /* node_count */
    int32_t* list_a;         /* max_level_nodes * max_unchecked_pointers */
    int32_t* list_b;
    int64_t list_len;        /* entries per list buffer */
} orc_gravity_buffers;
void orc_gravity_up(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g);
void orc_gravity_down(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g);
/* The down pass (after orc_gravity_up) with an interaction account, for the
 * pins of the list logic (gravity.metal:64-143, helpers.h:15-147), which has
 * no closed form of its own: per active particle, the number of particles and
 * the FP64 mass of every source node M2L'd into the expansions on its path,
 * plus its leaf's other members and, in the near-field form, the members of
 * its P2P source leaves.  Each particle should receive every other particle
 * exactly once: count = n - 1 and mass = M_total - m_p, except for accepts of
 * a node that contains the receiving node (accept.metal:17-25 can accept an
 * ancestor), whose double-counted part is returned apart as the excess.
 * Returns the number of containing accepts that the list structure does not
 * explain (any source other than the receiving node's parent; 0 expected).
 * `plant` (test use) 0 = as restated, 1 = drop the last entry of every list,
 * 2 = a leaf scans its first list entry twice.  Inactive particles get count
 * -1.  Also writes acc and gravnorm as orc_gravity_down does. */
int orc_gravity_account(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g,
                        int plant, int64_t* count_out, double* mass_out, int64_t* excess_count_out,
                        double* excess_mass_out);

/* Single FMM operators on 24-float multipole / 16-float local records (for
 * unit pins): M2L (M2L.metal:12-108) added into `local`, transformLocal
 * (L2L.metal:12-88), L2P (L2P.metal:12-32), the acceptance test
 * (accept.metal:14-55), transformMultipole (M2M.metal:12-73) and the
 * derivative tensor of the softened 1/r (derivatives.metal:59-156). */
void orc_pole_m2l(float* local16, const float* mp24, float eps);
void orc_pole_l2l(const float* local16, const float* r3, float* out16);
void orc_pole_l2p(const float* local16, const float* x3, float* acc3);
int orc_pole_mac(const float* A24, const float* B24, float eta);
void orc_pole_shift_multipole(const float* mp24, const float* r3, int fix_zz, float* out24);
void orc_pole_derivatives(const float* r3, float eps, float* d10);
/* The same at P = 3 (the reference's `#if P > 2` branches): 40-float
 * multipole records (24-float layout + octupole XXX XXY XXZ XYY XYZ XZZ YYY
 * YYZ YZZ ZZZ at 24-33, power[3] at 34) and 32-float local records (16-float
 * layout + third-order terms at 16-25, same order); derivatives to third
 * order (20 floats, poles.h:18-39 index map). */
void orc_pole3_m2l(float* local32, const float* mp40, float eps);
void orc_pole3_l2l(const float* local32, const float* r3, float* out32);
void orc_pole3_l2p(const float* local32, const float* x3, float* acc3);
int orc_pole3_mac(const float* A40, const float* B40, float eta, int fix_p3);
void orc_pole3_shift_multipole(const float* mp40, const float* r3, int fix_zz, int fix_p3, float* out40);
void orc_pole3_derivatives(const float* r3, float eps, float* d20);

/* FP64 direct sum over all n sources at the sampled targets: acceleration
 * (G included) and potential.  `softening` > 0 uses the cubic-spline softened
 * force (the corrected dphi_dr form) with that length (support 2 softening);
 * < 0 uses the potential whose derivatives M2L uses (derivatives.metal:12-57)
 * with support |softening|; 0 = pure Newton. */
void orc_direct_sum_fp64(const float* pos4, const float* m, int32_t n,
                         const int32_t* targets, int32_t n_targets, double G,
                         double softening, double* acc3_out, double* phi_out);

#ifdef __cplusplus
}
#endif
#endif
