/*
 * Compute.cpp -- host side of one simulation over the C ABI.  See Compute.hpp.
 */
#include "Compute.hpp"

#include <chrono>
#include <cstring>

#include "Parameters.h"

void Compute::check(int rc, const char* what) {
    if (rc != SPH_OK && status_ == 0) {
        status_ = rc;
        error_ = std::string(what) + ": " + (ctx ? sph_last_error(ctx) : "no context");
    }
}

Compute::Compute(const sph_params& params, const DataStruct& data, const std::vector<EOSTable>& tables, int treeMode,
                 int treeThreads)
    : params_(params), treeMode_(treeMode), treeThreads_(treeThreads < 1 ? 1 : treeThreads) {
    nParticles = (int)data.masses.size();
    sph_ic_view ic;
    ic.n = nParticles;
    ic.pos4 = data.positions.data();
    ic.vel4 = data.velocities.data();
    ic.rho = data.densities.data();
    ic.u = data.internalEnergy.data();
    ic.m = data.masses.data();
    ic.mat = data.materialIDs.data();
    ic.p = data.pressures.empty() ? nullptr : data.pressures.data();
    ic.h = data.smoothingLengths.data();
    std::vector<sph_eos_table> tv;
    for (const auto& t : tables) tv.push_back(sph_eos_table{t.materialId, t.resolution, t.data.data()});
    int rc = sph_create(&params_, &ic, tv.data(), (int)tv.size(), &ctx);
    if (rc != SPH_OK) {
        status_ = rc;
        error_ = "sph_create failed (" + std::to_string(rc) + ")";
        ctx = nullptr;
        return;
    }
    check(sph_device_field(ctx, SPH_F_POS, &positionBuffer, nullptr), "sph_device_field");
    check(sph_device_field(ctx, SPH_F_MAT, &materialIdBuffer, nullptr), "sph_device_field");
    check(sph_device_field(ctx, SPH_F_RHO, &densityBuffer, nullptr), "sph_device_field");
    snapshot_ = (float*)sph_host_alloc(16 * (size_t)nParticles);
    if (!snapshot_) {
        status_ = SPH_E_CAPACITY;
        error_ = "pinned snapshot allocation failed";
        return;
    }
    /* Compute.cpp:82-83: first tree built synchronously from the ICs */
    if (treeMode_ == SPH_TREE_GPU)
        buildOnDevice();
    else
        buildAndUpload(data.positions.data());
}

void Compute::buildOnDevice() {
    const auto t0 = std::chrono::steady_clock::now();
    check(sph_build_tree(ctx), "sph_build_tree");
    lastBuildMs_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    haveTree_ = true;
    treesSwapped_++;
}

Compute::~Compute() {
    if (builder_.joinable()) builder_.join();
    if (ctx) (void)sph_sync(ctx);
    sph_host_free(snapshot_);
    for (int k = 0; k < 2; ++k) {
        sph_host_free(stageOctree_[k]);
        sph_host_free(stageLevels_[k]);
    }
    sph_destroy(ctx);
}

void Compute::ensureStaging(int set, int64_t nInts, int64_t nNodes) {
    if (nInts > stageOctreeCap_[set]) {
        sph_host_free(stageOctree_[set]);
        stageOctreeCap_[set] = (int64_t)(nInts * 1.2) + 64;
        stageOctree_[set] = (int32_t*)sph_host_alloc(4 * (size_t)stageOctreeCap_[set]);
    }
    if (nNodes > stageLevelsCap_[set]) {
        sph_host_free(stageLevels_[set]);
        stageLevelsCap_[set] = (int64_t)(nNodes * 1.2) + 64;
        stageLevels_[set] = (int32_t*)sph_host_alloc(4 * (size_t)stageLevelsCap_[set]);
    }
}

void Compute::buildAndUpload(const float* pos4) {
    if (!ctx) return;
    sph_octree t;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = sph_octree_build(pos4, nParticles, params_.max_children_in_leaf, params_.max_tree_depth, treeThreads_, &t);
    lastBuildMs_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    check(rc, "sph_octree_build");
    if (rc != SPH_OK) return;
    check(sph_upload_tree(ctx, t.octree, t.n_ints, t.level_nodes, t.level_counts, t.n_levels, t.node_count),
          "sph_upload_tree");
    /* pageable source: make sure the copy has landed before freeing it */
    check(sph_sync(ctx), "sph_sync");
    sph_octree_free(&t);
    haveTree_ = true;
    treesSwapped_++;
}

void Compute::updateOctreeData(int set) {
    /* runs on the builder thread: touches only host memory and waits for the
     * snapshot event.  The snapshot's device-to-host copy is queued on the
     * copy stream after every tree upload issued so far, so once it has
     * landed the last upload from staging set `set` has landed too and the
     * set may be overwritten. */
    int rc = sph_wait_snapshot(ctx);
    sph_octree t;
    const auto t0 = std::chrono::steady_clock::now();
    if (rc == SPH_OK)
        rc = sph_octree_build(snapshot_, nParticles, params_.max_children_in_leaf, params_.max_tree_depth,
                              treeThreads_, &t);
    lastBuildMs_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == SPH_OK) {
        ensureStaging(set, t.n_ints, t.node_count);
        if (!stageOctree_[set] || !stageLevels_[set]) {
            rc = SPH_E_CAPACITY;
        } else {
            std::memcpy(stageOctree_[set], t.octree, 4 * (size_t)t.n_ints);
            std::memcpy(stageLevels_[set], t.level_nodes, 4 * (size_t)t.node_count);
            stageCounts_[set].assign(t.level_counts, t.level_counts + t.n_levels);
            stageInts_[set] = t.n_ints;
            stageNodes_[set] = t.node_count;
        }
        sph_octree_free(&t);
    }
    stageBuildRc_ = rc;
    treeReady_.store(rc == SPH_OK, std::memory_order_release);
    updating_.store(false, std::memory_order_release);
}

void Compute::updateOctreeBuffer() {
    if (!ctx || status_ != 0) return;
    if (treeMode_ == SPH_TREE_FIXED) {
        return;
    }
    if (treeMode_ == SPH_TREE_GPU) {
        /* rebuilt on the GPU from the current positions before the step */
        buildOnDevice();
        return;
    }
    if (treeMode_ == SPH_TREE_EVERY) {
        /* deterministic: rebuild from the current positions before the step */
        check(sph_snapshot_positions(ctx, snapshot_), "sph_snapshot_positions");
        check(sph_wait_snapshot(ctx), "sph_wait_snapshot");
        buildAndUpload(snapshot_);
        return;
    }
    /* SPH_TREE_ASYNC (Compute.cpp:456-508): while a build runs, keep the
     * current tree; once it has finished, swap its tree in and start the next
     * build in the same call, from the positions at this point of the stream
     * (those of the last completed step). */
    if (updating_.load(std::memory_order_acquire)) return;
    if (builder_.joinable()) builder_.join();
    if (stageBuildRc_ != 0) {
        check(stageBuildRc_, "octree thread");
        return;
    }
    if (treeReady_.load(std::memory_order_acquire)) {
        treeReady_.store(false);
        const int set = buildSet_;
        check(sph_upload_tree(ctx, stageOctree_[set], stageInts_[set], stageLevels_[set], stageCounts_[set].data(),
                              (int)stageCounts_[set].size(), stageNodes_[set]),
              "sph_upload_tree");
        treesSwapped_++;
        buildSet_ = 1 - set;
    }
    if (status_ == 0) startBuild();
}

void Compute::startBuild() {
    check(sph_snapshot_positions(ctx, snapshot_), "sph_snapshot_positions");
    if (status_ != 0) return;
    updating_.store(true, std::memory_order_release);
    builder_ = std::thread(&Compute::updateOctreeData, this, buildSet_);
}

void Compute::setSnapshots(const std::string& dir, int resolution, int interval) {
    snapshotDir_ = dir;
    snapshotRes_ = resolution;
    snapshotInterval_ = interval;
}

void Compute::drawSnapshot() {
    if (!ctx || status_ != 0 || snapshotDir_.empty()) return;
    /* Compute.cpp:416-418: the time of the last completed step */
    int32_t ticks = 0;
    check(sph_download(ctx, SPH_F_GLOBAL_TIME, &ticks), "sph_download");
    const float floatTime = ticks * DT_MIN1;
    if (floatTime > nextSnapshot_ && sph_snapshot_resolution(ctx) > 0) {
        /* the image drawn at the previous call (Compute.cpp:419-446) */
        const int res = sph_snapshot_resolution(ctx);
        std::vector<int32_t> img((size_t)res * res);
        check(sph_download_snapshot(ctx, img.data()), "sph_download_snapshot");
        std::string err;
        const std::string path =
            snapshotDir_ + "/snapshot_" + std::to_string(nextSnapshot_ / snapshotInterval_) + ".png";
        if (status_ == 0 && !writeSnapshotPNG(path, img.data(), res, err)) check(SPH_E_ARG, err.c_str());
        nextSnapshot_ += snapshotInterval_;
    }
    check(sph_draw_snapshot(ctx, snapshotRes_), "sph_draw_snapshot");
}

void Compute::sort() { if (ctx) check(sph_sort(ctx), "sph_sort"); }
void Compute::gravitationalPass() { if (ctx) check(sph_gravity(ctx), "sph_gravity"); }
void Compute::densityPass() { if (ctx) check(sph_density(ctx), "sph_density"); }
void Compute::accelerationPass() { if (ctx) check(sph_acceleration(ctx), "sph_acceleration"); }
void Compute::stepPass() { if (ctx) check(sph_step(ctx), "sph_step"); }

/* ------------------------------------------------------------------ inputs */
std::vector<EOSTable> makeEOSTables(const sph_params& params, int resolution) {
    std::vector<EOSTable> out(2);
    sph_tillotson rock, iron;
    sph_tillotson_granite(&rock);
    sph_tillotson_iron(&iron);
    out[0].materialId = 400;
    out[1].materialId = 402;
    for (auto& t : out) {
        t.resolution = resolution;
        t.data.resize(2 * (size_t)resolution * resolution);
    }
    sph_make_eos_table(&params, &rock, resolution, 400, out[0].data.data());
    sph_make_eos_table(&params, &iron, resolution, 402, out[1].data.data());
    return out;
}

DataStruct makeImpact(const sph_params& params, const sph_impact_config& cfg, const std::vector<EOSTable>& tables) {
    DataStruct d;
    const size_t n = (size_t)cfg.n_total;
    d.positions.resize(4 * n);
    d.velocities.resize(4 * n);
    d.densities.resize(n);
    d.internalEnergy.resize(n);
    d.masses.resize(n);
    d.materialIDs.resize(n);
    d.pressures.resize(n);
    d.smoothingLengths.resize(n);
    sph_make_impact(&params, &cfg, tables[0].data.data(), tables[1].data.data(), tables[0].resolution,
                    d.positions.data(), d.velocities.data(), d.densities.data(), d.internalEnergy.data(),
                    d.masses.data(), d.materialIDs.data(), d.pressures.data(), d.smoothingLengths.data());
    return d;
}

/* ---------------------------------------------------------------- headless */
Headless::Headless(const sph_params& params, const sph_impact_config& config, int eosResolution, int treeMode,
                   int treeThreads) {
    std::vector<EOSTable> tables = makeEOSTables(params, eosResolution);
    DataStruct data = makeImpact(params, config, tables);
    compute = new Compute(params, data, tables, treeMode, treeThreads);
}

Headless::Headless(const sph_params& params, const std::string& path, int eosResolution, int treeMode,
                   int treeThreads) {
    DataStruct data;
    if (!readHDFFile(path, data, error_)) {
        status_ = SPH_E_ARG;
        return;
    }
    std::vector<EOSTable> tables = makeEOSTables(params, eosResolution);
    compute = new Compute(params, data, tables, treeMode, treeThreads);
}

Headless::Headless(const sph_params& params, const std::string& path, const std::string& aneos400,
                   const std::string& aneos402, int eosResolution, bool fixedResampler, int treeMode,
                   int treeThreads) {
    DataStruct data;
    if (!readHDFFile(path, data, error_)) {
        status_ = SPH_E_ARG;
        return;
    }
    std::vector<EOSTable> tables(2);
    tables[0].materialId = 400;
    tables[1].materialId = 402;
    if (!loadANEOSDataFromFile(aneos400, params, eosResolution, fixedResampler, tables[0], error_) ||
        !loadANEOSDataFromFile(aneos402, params, eosResolution, fixedResampler, tables[1], error_)) {
        status_ = SPH_E_ARG;
        return;
    }
    compute = new Compute(params, data, tables, treeMode, treeThreads);
}

Headless::~Headless() { delete compute; }

void Headless::enqueueStep() {
    /* Headless.cpp:43-51 (drawSnapshot is visualisation and is not run) */
    if (!compute) return;
    compute->updateOctreeBuffer();
    compute->drawSnapshot();
    for (int i = 0; i < STEPS_PER_FRAME; i++) {
        compute->sort();
        compute->gravitationalPass();
        compute->densityPass();
        compute->accelerationPass();
        compute->stepPass();
    }
}

void Headless::step() {
    enqueueStep();
    if (compute && compute->ctx && compute->status() == 0) {
        int rc = sph_sync(compute->ctx);
        if (rc != SPH_OK && status_ == 0) {
            status_ = rc;
            error_ = std::string("sph_sync: ") + sph_last_error(compute->ctx);
        }
    }
}

int Headless::status() const { return status_ != 0 ? status_ : (compute ? compute->status() : SPH_E_STATE); }

std::string Headless::error() const { return status_ != 0 ? error_ : (compute ? compute->error() : "no compute"); }
