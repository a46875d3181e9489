/*
 * Compute.hpp -- host-side owner of one simulation, with the reference's
 * method names and call contract (SPH/Compute/Compute.hpp:40-148).
 *
 * The reference's Compute owns Metal buffers and encodes each pass into a
 * command buffer.  This one owns an sph_ctx (device memory + HIP stream) and
 * enqueues each pass through the C ABI; nothing waits until Headless::step
 * calls sph_sync, as the reference waits once per frame.
 *
 * The octree is still built on a CPU thread (Compute.cpp:456-516), with two
 * differences: positions reach the thread through an explicit pinned
 * snapshot (the reference reads the shared Metal buffer while the GPU may be
 * writing it), and the hand-off uses atomics instead of a plain bool.  As in
 * the reference, the call that swaps in a finished tree also starts the next
 * build: the thread writes into one of two pinned staging sets while the
 * other one's upload may still be in flight.
 */
#ifndef SPH_HOST_COMPUTE_HPP
#define SPH_HOST_COMPUTE_HPP

#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "Parameters.h"
#include "sph_capi.h"
#include "sph_host.h"

/* Initial conditions (hdfHandler.hpp:15-26); 3-vectors with a 4-float stride. */
struct DataStruct {
    std::vector<float> positions;
    std::vector<float> velocities;
    std::vector<float> densities;
    std::vector<float> entropies;
    std::vector<float> internalEnergy;
    std::vector<float> masses;
    std::vector<int32_t> materialIDs;
    std::vector<float> potentials;
    std::vector<float> pressures;
    std::vector<float> smoothingLengths;
};

/* One EOS table per material slot (ANEOSTable, ANEOS.hpp:16-29). */
struct EOSTable {
    int32_t materialId;
    int32_t resolution;
    std::vector<float> data; /* (P, c) pairs, row = density sample */
};

class Compute {
public:
    /* Compute::Compute + loadInitialConditions (Compute.cpp:22-84): uploads
     * the state and the tables and builds the first tree synchronously. */
    Compute(const sph_params& params, const DataStruct& data, const std::vector<EOSTable>& tables, int treeMode,
            int treeThreads);
    ~Compute();
    Compute(const Compute&) = delete;
    Compute& operator=(const Compute&) = delete;

    /* Compute.cpp:456-508: swap in a finished tree, start the next build. */
    void updateOctreeBuffer();
    void sort();               /* Compute.cpp:265-326 */
    void gravitationalPass();  /* Compute.cpp:328-377 */
    void densityPass();        /* Compute.cpp:179-208 */
    void accelerationPass();   /* Compute.cpp:210-240 */
    void stepPass();           /* Compute.cpp:242-263 */
    /* Compute.cpp:396-448: with snapshots on, writes snapshot_<k>.png from the
     * image drawn at the previous call once the simulated time passes the
     * next multiple of the interval, then draws the current positions. */
    void drawSnapshot();
    void setSnapshots(const std::string& dir, int resolution, int interval);

    int nParticles = 0;
    sph_ctx* ctx = nullptr;
    /* Compute.hpp:55-57: the buffers the renderer binds (Render.cpp:26), as
     * device addresses (sph_device_field): positions (4 floats per
     * particle), material ids (int32), densities (float); original-ID order,
     * read-only, current after each Headless::step. */
    void* positionBuffer = nullptr;
    void* materialIdBuffer = nullptr;
    void* densityBuffer = nullptr;

    /* first failing call (0 = none) and its message */
    int status() const { return status_; }
    const std::string& error() const { return error_; }
    int64_t treesSwapped() const { return treesSwapped_; }
    double lastBuildMs() const { return lastBuildMs_.load(); }

private:
    void check(int rc, const char* what);
    void buildAndUpload(const float* pos4);
    void buildOnDevice(); /* sph_build_tree (SPH_TREE_GPU) */
    void updateOctreeData(int set); /* body of the octree thread (Compute.cpp:510-516) */
    void startBuild();               /* snapshot + thread writing into the free staging set */
    void ensureStaging(int set, int64_t nInts, int64_t nNodes);

    sph_params params_{};
    int treeMode_ = SPH_TREE_ASYNC;
    int treeThreads_ = 1;
    int status_ = 0;
    std::string error_;

    /* octree hand-off */
    float* snapshot_ = nullptr;       /* pinned, 4 floats per particle */
    /* two pinned staging sets of octreeData and the level lists: the thread
     * fills set `buildSet_` while the other one may still be uploading */
    int32_t* stageOctree_[2] = {nullptr, nullptr};
    int32_t* stageLevels_[2] = {nullptr, nullptr};
    int64_t stageOctreeCap_[2] = {0, 0}, stageLevelsCap_[2] = {0, 0};
    std::vector<int32_t> stageCounts_[2];
    int64_t stageInts_[2] = {0, 0};
    int32_t stageNodes_[2] = {0, 0};
    int buildSet_ = 0;
    int stageBuildRc_ = 0;
    std::thread builder_;
    std::atomic<bool> updating_{false};
    std::atomic<bool> treeReady_{false};
    std::atomic<double> lastBuildMs_{0.0};
    bool haveTree_ = false;
    int64_t treesSwapped_ = 0;
    std::string snapshotDir_;   /* empty: snapshots off */
    int snapshotRes_ = SNAPSHOT_RESOLUTION;
    int snapshotInterval_ = 1000;
    int nextSnapshot_ = 0;
};

/* Headless driver (Headless.cpp:19-55). */
class Headless {
public:
    Headless(const sph_params& params, const sph_impact_config& config, int eosResolution, int treeMode,
             int treeThreads);
    /* Headless::Headless(filepath) (Headless.cpp:19-23): initial conditions
     * from an HDF5 file (readHDFFile); the EOS tables are the Tillotson ones
     * (the ANEOS files are not available). */
    Headless(const sph_params& params, const std::string& path, int eosResolution, int treeMode, int treeThreads);
    /* The same with the two ANEOS tables read from their files (material 400
     * and 402, loadANEOSDataFromFile at eosResolution), as the reference's
     * constructor loads them. */
    Headless(const sph_params& params, const std::string& path, const std::string& aneos400,
             const std::string& aneos402, int eosResolution, bool fixedResampler, int treeMode, int treeThreads);
    ~Headless();
    /* updateOctreeBuffer, sort, gravity, density, acceleration, step, wait */
    void step();
    /* the same sequence without the final wait */
    void enqueueStep();

    Compute* compute = nullptr;
    int status() const;
    std::string error() const;

private:
    int status_ = 0;
    std::string error_;
};

/* readHDFFile (hdfHandler.cpp:12-136): PartType0/{Coordinates, Densities,
 * InternalEnergies, Masses, MaterialIDs, Pressures, SmoothingLengths,
 * Velocities}, converted to float / int32 as H5Dread converts them.  Errors
 * are returned (the reference exits). */
bool readHDFFile(const std::string& path, DataStruct& out, std::string& err);
/* writeHDFFile (hdfHandler.cpp:139-254): the same datasets plus Entropies,
 * Potentials and ParticleIDs = 0..n-1; empty vectors give empty datasets. */
bool writeHDFFile(const std::string& path, const DataStruct& data, std::string& err);
/* Checkpoint of a context: writeHDFFile's datasets from the device state plus
 * a Checkpoint group with the rest of the state a step carries over (no
 * counterpart: the reference never writes its state). */
int writeCheckpoint(sph_ctx* ctx, const std::string& path, std::string& err);
/* Restores a checkpoint into a context of the same particle count. */
int readCheckpoint(sph_ctx* ctx, const std::string& path, std::string& err);
/* The snapshot PNG (Compute.cpp:416-446): 402 grey, 400 orange, else white. */
bool writeSnapshotPNG(const std::string& path, const int32_t* materials, int res, std::string& err);
bool writePNG(const std::string& path, const uint8_t* rgba, int w, int h, std::string& err);
/* sets what sph_hdf_last_error() reports on this thread */
void setHdfError(const std::string& msg);

/* loadANEOSDataFromFile (ANEOS.cpp:21-175): an ANEOS text table resampled
 * into `out` (res x res (P, c), code units) on the axes of `prm`; `fixed`
 * selects the corrected resampler instead of the reference's (Q13). */
bool loadANEOSDataFromFile(const std::string& path, const sph_params& prm, int resolution, bool fixed,
                           EOSTable& out, std::string& err);

/* Synthetic inputs of the configs. */
std::vector<EOSTable> makeEOSTables(const sph_params& params, int resolution);
DataStruct makeImpact(const sph_params& params, const sph_impact_config& config, const std::vector<EOSTable>& tables);

#endif
