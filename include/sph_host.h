/*
 * sph_host.h -- C API of the host layer (libsphhost.so): synthetic inputs,
 * the CPU octree builder, and the headless driver built on sph_capi.h.
 *
 * The C++ classes behind it keep the reference's names (Compute, Headless,
 * buildOctree); this flat API exists so that Python tests and bench.py can
 * drive them through ctypes.
 */
#ifndef SPH_HOST_H
#define SPH_HOST_H

#include <stdint.h>
#include "sph_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- EOS tables */
/* Tillotson parameter set (SI units) for one material slot. */
typedef struct sph_tillotson {
    double rho0, A, B, E0, Eiv, Ecv, a, b, alpha, beta;
} sph_tillotson;

/* Published parameter sets recorded with the table: granite (used for the
 * forsterite slot, material 400) and iron (Fe85Si15 slot, material 402). */
void sph_tillotson_granite(sph_tillotson* out);
void sph_tillotson_iron(sph_tillotson* out);

/* Fills `pc_out` (2*res*res floats, row-major [rho][u]) with (P, c) in code
 * units on the reference's log-spaced axes (ANEOS.cpp:123-136): sample i at
 * min * (max/min)^(i/res).  `seed` is recorded for reproducibility; the table
 * is deterministic and adds no noise. */
int sph_make_eos_table(const sph_params* prm, const sph_tillotson* t, int res, uint64_t seed, float* pc_out);

/* loadANEOSDataFromFile (ANEOS.cpp:21-175): reads an ANEOS text table (13
 * header lines, "numRho numT", the densities, the temperatures, then
 * numRho * numT lines "u P c" in SI with the density index fastest) and
 * resamples it into `pc_out` (2*res*res floats, the layout of
 * sph_make_eos_table) on the axes of `prm` (eos_min/max_rho/u).  `fixed` = 0
 * reproduces the reference's resampler, quirks included (Q13: interpolation
 * weights inverted; the upper row's search bounded by the lower row's index);
 * 1 interpolates bilinearly.  SPH_E_ARG with sph_aneos_last_error() for a
 * missing or malformed file. */
int sph_load_aneos(const char* path, const sph_params* prm, int res, int fixed, float* pc_out);
const char* sph_aneos_last_error(void);

/* ------------------------------------------------------ initial conditions */
typedef struct sph_impact_config {
    int32_t n_total;
    float target_fraction;        /* share of particles in the target (0.89) */
    float target_radius;          /* code units (Mm) */
    float target_mass;            /* code units (1e24 kg) */
    float impactor_radius;
    float impactor_mass;
    float core_mass_fraction;     /* 0.3 */
    float core_radius_fraction;   /* 0.55 */
    float u0;                     /* initial specific internal energy, code units (1e12 J/kg) */
    float separation_factor;      /* centre distance / (R_t + R_i) */
    float impact_parameter;       /* b / (R_t + R_i) */
    float closing_speed;          /* code units (Mm/s) */
    uint64_t seed;
} sph_impact_config;

/* Oblique 45-degree impact of the configs (R_t 6.4, M_t 5.3, R_i 3.6, M_i 0.8,
 * u0 2e-6, separation 1.5, b = sin 45deg, v = 0.009). */
void sph_impact_config_default(sph_impact_config* out, int32_t n_total, uint64_t seed);
/* Head-on variant (b = 0.3, v = 0.012, u0 = 5e-6). */
void sph_impact_config_headon(sph_impact_config* out, int32_t n_total, uint64_t seed);

/* Writes the two-body initial conditions in DataStruct layout
 * (hdfHandler.hpp:15-26): arrays of length n_total (pos4/vel4: 4*n_total).
 * Pressure comes from the EOS tables (pc_400 for material 400, pc_402 for 402). */
int sph_make_impact(const sph_params* prm, const sph_impact_config* cfg, const float* pc_400,
                    const float* pc_402, int res, float* pos4, float* vel4, float* rho, float* u,
                    float* m, int32_t* mat, float* p, float* h);

/* ---------------------------------------------------------------- octree */
/* buildOctree (octree.cpp:111-141) output: octreeData, treeLevels flattened,
 * nodeValues.  Library-owned; free with sph_octree_free. */
typedef struct sph_octree {
    int32_t* octree;
    int64_t n_ints;
    int32_t* level_nodes;
    int32_t* level_counts;
    int32_t n_levels;
    int32_t node_count;
} sph_octree;
/* `nthreads` <= 1 builds serially; otherwise subtrees are built in parallel.
 * The result is identical for every thread count.  A non-finite coordinate
 * returns SPH_E_ARG (octree.cpp:68-86 would put every particle into one
 * depth-limited leaf). */
int sph_octree_build(const float* pos4, int32_t n, int32_t max_leaf, int32_t max_depth,
                     int32_t nthreads, sph_octree* out);
void sph_octree_free(sph_octree* t);

/* ------------------------------------------------------------ HDF5 files */
/* DataStruct (hdfHandler.hpp:15-26) as library-owned arrays; pos4 / vel4 hold
 * 4 floats per particle (w = 0).  Free with sph_particles_free. */
typedef struct sph_particles {
    int32_t n;
    float* pos4;
    float* vel4;
    float* rho;
    float* u;
    float* m;
    int32_t* mat;
    float* p;
    float* h;
} sph_particles;

/* readHDFFile (hdfHandler.cpp:12-136): PartType0/{Coordinates, Densities,
 * InternalEnergies, Masses, MaterialIDs, Pressures, SmoothingLengths,
 * Velocities} of an HDF5 file, converted as H5Dread converts to
 * H5T_NATIVE_FLOAT / H5T_NATIVE_INT32.  Reads superblock 0/1 files with
 * symbol-table groups and contiguous or compact, unfiltered storage (the
 * layout of the reference's files) without libhdf5.  A missing, malformed or
 * unsupported file returns SPH_E_ARG and sph_hdf_last_error() says why (the
 * reference exits). */
int sph_read_hdf(const char* path, sph_particles* out);
void sph_particles_free(sph_particles* p);
/* writeHDFFile (hdfHandler.cpp:139-254): PartType0 with the DataStruct
 * datasets, Entropies (empty), Potentials (`potentials`, or empty when NULL)
 * and ParticleIDs 0..n-1. */
int sph_write_hdf(const char* path, const sph_ic_view* ic, const float* potentials);
/* Checkpoint and resume (no counterpart: the reference never saves its
 * state).  The file is a valid initial-conditions file (PartType0 as
 * sph_write_hdf writes it) plus a Checkpoint group with what a step carries
 * over: accelerations, du/dt, dh/dt and its log-rate, grad-h terms, sound
 * speeds, Balsara switches, gravity norms, activity, next active times, the
 * global time and step.  sph_checkpoint_read restores it into a context of
 * the same particle count; with a tree rebuilt before every step the run then
 * continues bit-identically. */
int sph_checkpoint_write(sph_ctx* ctx, const char* path);
int sph_checkpoint_read(sph_ctx* ctx, const char* path);
/* Message of the last failing sph_read_hdf / sph_write_hdf / checkpoint call
 * on this thread. */
const char* sph_hdf_last_error(void);

/* -------------------------------------------------------------- headless */
#define SPH_TREE_ASYNC 0   /* rebuild on a CPU thread, swap when ready (Compute.cpp:456-508) */
#define SPH_TREE_EVERY 1   /* rebuild from the current positions before every step */
#define SPH_TREE_FIXED 2   /* keep the first tree */
#define SPH_TREE_GPU 3     /* rebuild on the GPU (sph_build_tree) before every step; no counterpart in the reference */

typedef struct sph_headless sph_headless;

/* Headless::Headless (Headless.cpp:19-23) with synthetic inputs instead of
 * the HDF5 file: builds the EOS tables and the impact, creates the context
 * and the initial tree. */
int sph_headless_create(const sph_params* prm, const sph_impact_config* cfg, int eos_res,
                        int tree_mode, int tree_threads, sph_headless** out);
/* Headless::Headless(filepath) (Headless.cpp:19-23): initial conditions read
 * from an HDF5 file as sph_read_hdf reads them; Tillotson EOS tables (the
 * reference's ANEOS files are not available). */
int sph_headless_create_from_file(const sph_params* prm, const char* path, int eos_res, int tree_mode,
                                  int tree_threads, sph_headless** out);
/* The same with the EOS tables of materials 400 and 402 read from ANEOS
 * text files (sph_load_aneos at eos_res, `fixed_resampler` as there), as the
 * reference's constructor reads ANEOS_forsterite_S19.txt and
 * ANEOS_Fe85Si15_S20.txt. */
int sph_headless_create_with_aneos(const sph_params* prm, const char* path, const char* aneos_400,
                                   const char* aneos_402, int eos_res, int fixed_resampler, int tree_mode,
                                   int tree_threads, sph_headless** out);
/* Compute::drawSnapshot (Compute.cpp:396-448) in every step, before the
 * passes, as Headless::step calls it (Headless.cpp:43): with `dir` non-NULL
 * the positions are rasterised each step (sph_draw_snapshot at `res`) and,
 * once the simulated time passes the next multiple of `interval` code time
 * units (the reference: 1000), the image drawn at the previous step is
 * written to dir/snapshot_<k>.png.  NULL turns snapshots off (the default). */
int sph_headless_set_snapshots(sph_headless* h, const char* dir, int res, int interval);
/* The snapshot PNG of a material-id image (Compute.cpp:419-446): 402 grey
 * (160,160,160), 400 orange (255,100,0), anything else white; 8-bit RGBA. */
int sph_write_snapshot_png(const int32_t* materials, int res, const char* path);
void sph_headless_destroy(sph_headless* h);
/* Headless::step (Headless.cpp:37-55): updateOctreeBuffer, then sort,
 * gravity, density, acceleration, step, then wait. */
int sph_headless_step(sph_headless* h);
/* Same sequence without the final wait (the caller syncs). */
int sph_headless_enqueue_step(sph_headless* h);
sph_ctx* sph_headless_ctx(sph_headless* h);
/* Tree statistics: number of trees swapped in so far and the last build ms. */
int sph_headless_tree_stats(sph_headless* h, int64_t* trees_swapped, double* last_build_ms);
/* Compute's public positionBuffer, materialIdBuffer and densityBuffer
 * (Compute.hpp:55-57, bound by Render.cpp:26) as device addresses; any
 * pointer may be NULL. */
int sph_headless_render_buffers(sph_headless* h, void** positions, void** material_ids, void** densities);
const char* sph_headless_last_error(const sph_headless* h);

#ifdef __cplusplus
}
#endif
#endif
