/*
 * Parameters.h -- compile-time defaults of the SPH + FMM step core.
 *
 * Macro names and values follow the reference's SPH/Parameters.h:11-53 so that
 * host code written against it compiles unchanged.  Every value here is only a
 * default: the runtime copy lives in `sph_params` (sph_capi.h), filled by
 * sph_params_default(), so tests can shrink caps (unchecked-list length,
 * neighbour-cache size, NR iteration cap) without recompiling.
 *
 * Differences from the reference header:
 *  - DT_MIN is parenthesised (Parameters.h:31 relies on left-to-right
 *    evaluation at both of its uses; the value is the same).
 *  - FILEPATH and the rendering block (Parameters.h:55-69) are dropped: the
 *    initial conditions come from sph_ic_view and rendering is out of scope.
 *  - New knobs (rung cap, activity rule, quirk fixes, forms) are runtime-only
 *    and documented in sph_capi.h.
 */
#ifndef SPH_PARAMETERS_H
#define SPH_PARAMETERS_H

/* Hydro (Parameters.h:11-17) */
#define VISCOSITY_ALPHA 1.5f
#define DENSITY_ETA .005f
#define MAX_SMOOTHING_LENGTH 1.f
#define MAX_DENSITY_NR_ITTERATIONS 50
#define N_NEIGHBOURS_ESTIM 128
#define RESOLUTION_ETA 1.2348f

/* EOS (Parameters.h:19-24).  The table resolution is a runtime property of
 * each sph_eos_table; 2048 is the reference's texture size. */
#define ANEOS_TEXTURE_RESOLUTION 2048
#define ANEOS_MIN_RHO 1e-6f
#define ANEOS_MAX_RHO 1.f
#define ANEOS_MIN_U 1e-10f
#define ANEOS_MAX_U 1.f

/* Time stepping (Parameters.h:26-33) */
#define CFL .2f
#define DT0 32.f
#define R_MAX 16
#define DT_MIN1 (DT0 / (1 << R_MAX))
#define DT_MIN (1 / DT_MIN1)
#define STEPS_PER_FRAME 1
#define GAMMA 1.825742f

/* Gravity (Parameters.h:35-44) */
#define GRAVITY_ETA .005f
#define P 2
#define N_EXPANSION_TERMS ((P + 1) * (P + 2) * (P + 3) / 6)
#define MAX_TREE_DEPTH 7
#define GRAVITY_SMOOTHING_LENGTH 0.04f
#define GRAVITY_MAX_RECURSION 1
#define MAX_UNCHECKED_POINTERS 1024
#define MAX_CHILDREN_IN_LEAF 4
#define EXTRA_MEMORY_MULTIPLIER 1.2

/* Gravitational constant in code units (poles.h:16): L = 1 Mm, M = 1e24 kg,
 * t = 1 s. */
#define GRAVITY_G 6.67e-5f

/* Cells and sorting (Parameters.h:46-53) */
#define CELL_WIDTH 0.18f
#define CELL_POWER 8
#define SORTING_BLOCK_SIZE 256
#define SORTING_MASK_LENGTH 8
#define SORTING_BUCKET_NUMBER (1 << SORTING_MASK_LENGTH)
#define SORTING_BIT_MASK (SORTING_BUCKET_NUMBER - 1)
#define SORTING_ITTERATIONS (32 / SORTING_MASK_LENGTH)

/* Snapshot raster (Parameters.h:67) */
#define SNAPSHOT_RESOLUTION 1024

/* Headless timing loop (Parameters.h:69) */
#define HEADLESS_ITTERATION_RATE_REFRESH 32

#endif /* SPH_PARAMETERS_H */
