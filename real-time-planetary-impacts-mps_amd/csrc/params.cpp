/*
 * params.cpp -- sph_params_default(): the reference's Parameters.h values
 * (Parameters.h:11-53) as the runtime defaults.
 */
#include <cstring>

#include "sph_capi.h"
#include "Parameters.h"

extern "C" void sph_params_default(sph_params* o) {
    std::memset(o, 0, sizeof(*o));
    o->viscosity_alpha = VISCOSITY_ALPHA;
    o->density_eta = DENSITY_ETA;
    o->max_smoothing_length = MAX_SMOOTHING_LENGTH;
    o->max_nr_iterations = MAX_DENSITY_NR_ITTERATIONS;
    o->n_neighbours_estim = N_NEIGHBOURS_ESTIM;
    o->resolution_eta = RESOLUTION_ETA;
    o->eos_min_rho = ANEOS_MIN_RHO;
    o->eos_max_rho = ANEOS_MAX_RHO;
    o->eos_min_u = ANEOS_MIN_U;
    o->eos_max_u = ANEOS_MAX_U;
    o->cfl = CFL;
    o->dt0 = DT0;
    o->gamma = GAMMA;
    o->r_max = R_MAX;
    o->activity_rule = SPH_ACTIVE_AS_WRITTEN;
    o->max_rungs = 0;
    o->gravity_eta = GRAVITY_ETA;
    o->gravity_g = GRAVITY_G;
    o->gravity_softening = GRAVITY_SMOOTHING_LENGTH;
    o->gravity_max_recursion = GRAVITY_MAX_RECURSION;
    o->max_unchecked_pointers = MAX_UNCHECKED_POINTERS;
    o->max_children_in_leaf = MAX_CHILDREN_IN_LEAF;
    o->max_tree_depth = MAX_TREE_DEPTH;
    o->cell_width = CELL_WIDTH;
    o->cell_power = CELL_POWER;
    o->fixes = 0;
    /* measured at 1e6 on MI355X: density 5.67 ms (thread per particle) ->
     * 4.61 ms (8 lanes, whole cells per lane); the acceleration pass is
     * fastest with one thread per particle */
    o->density_lanes = 8;
    o->accel_lanes = 0;
    o->lane_split = 0;
    o->expansion_order = P; /* Parameters.h */
    /* measured at 1e6 on MI355X (100 steps): density 3.95 ms in slot order,
     * 3.85 / 3.78 / 3.95 ms in windows of 16 / 32 / 64 slots ranked by the
     * previous iteration counts */
    o->density_order = 32;
    /* measured at 4e6 (head-on, README rule, 8 rungs, about 1.5% of the
     * particles active; 100 steps): density + acceleration passes 0.81 +
     * 1.15 ms at 8 / 0 lanes, 0.70 + 0.70 ms at 16 / 16; 290 -> 346 steps/s */
    o->list_density_lanes = 16;
    o->list_accel_lanes = 16;
    o->gather_reach = 1;
    /* measured at 4e6 (head-on, learned-h + precision, P = 3, 100 steps):
     * gravity down pass 3.90 ms walking twice, 3.49 / 3.43 / 3.49 / 3.44 ms
     * with records of 256 / 512 / 1024 / 2048 source particles */
    o->near_list_cap = 512;
    /* measured on MI355X, 45-degree impact, 100 steps (steps/s at density /
     * acceleration lanes 8 / 0 -> the lanes taken here): 2e4 181.5 -> 630.2
     * (32 / 16), 1e5 292.6 -> 430.1 (32 / 8); at 1e6, 8 / 0 is fastest */
    o->small_n_lanes = 262144;
}
