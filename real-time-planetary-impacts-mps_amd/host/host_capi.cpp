/*
 * host_capi.cpp -- flat C entry points of the headless driver (sph_host.h)
 * so that Python tests and bench.py can drive Headless through ctypes.
 */
#include <exception>
#include <string>

#include "Compute.hpp"

struct sph_headless {
    Headless* h = nullptr;
    std::string err;
};

extern "C" {

int sph_headless_create(const sph_params* prm, const sph_impact_config* cfg, int eos_res, int tree_mode,
                        int tree_threads, sph_headless** out) {
    if (!prm || !cfg || !out || eos_res < 2 || cfg->n_total < 2) return SPH_E_ARG;
    if (tree_mode < SPH_TREE_ASYNC || tree_mode > SPH_TREE_GPU) return SPH_E_ARG;
    *out = nullptr;
    sph_headless* w = new sph_headless();
    w->h = new Headless(*prm, *cfg, eos_res, tree_mode, tree_threads);
    const int st = w->h->status();
    if (st != 0) {
        delete w->h;
        delete w;
        return st;
    }
    *out = w;
    return SPH_OK;
}

int sph_headless_create_from_file(const sph_params* prm, const char* path, int eos_res, int tree_mode,
                                  int tree_threads, sph_headless** out) {
    if (!prm || !path || !out || eos_res < 2) return SPH_E_ARG;
    if (tree_mode < SPH_TREE_ASYNC || tree_mode > SPH_TREE_GPU) return SPH_E_ARG;
    *out = nullptr;
    sph_headless* w = new sph_headless();
    try {
        w->h = new Headless(*prm, std::string(path), eos_res, tree_mode, tree_threads);
    } catch (const std::exception&) {
        delete w;
        return SPH_E_CAPACITY;
    }
    const int st = w->h->status();
    if (st != 0) {
        setHdfError(w->h->error());
        delete w->h;
        delete w;
        return st;
    }
    *out = w;
    return SPH_OK;
}

int sph_headless_create_with_aneos(const sph_params* prm, const char* path, const char* aneos_400,
                                   const char* aneos_402, int eos_res, int fixed_resampler, int tree_mode,
                                   int tree_threads, sph_headless** out) {
    if (!prm || !path || !aneos_400 || !aneos_402 || !out || eos_res < 2) return SPH_E_ARG;
    if (tree_mode < SPH_TREE_ASYNC || tree_mode > SPH_TREE_GPU) return SPH_E_ARG;
    *out = nullptr;
    sph_headless* w = new sph_headless();
    try {
        w->h = new Headless(*prm, std::string(path), std::string(aneos_400), std::string(aneos_402), eos_res,
                            fixed_resampler != 0, tree_mode, tree_threads);
    } catch (const std::exception&) {
        delete w;
        return SPH_E_CAPACITY;
    }
    const int st = w->h->status();
    if (st != 0) {
        setHdfError(w->h->error());
        delete w->h;
        delete w;
        return st;
    }
    *out = w;
    return SPH_OK;
}

int sph_headless_set_snapshots(sph_headless* w, const char* dir, int res, int interval) {
    if (!w || !w->h->compute) return SPH_E_ARG;
    if (dir && (res < 1 || res > 16384 || interval < 1)) return SPH_E_ARG;
    w->h->compute->setSnapshots(dir ? std::string(dir) : std::string(), res, interval);
    return SPH_OK;
}

void sph_headless_destroy(sph_headless* w) {
    if (!w) return;
    delete w->h;
    delete w;
}

int sph_headless_step(sph_headless* w) {
    if (!w) return SPH_E_ARG;
    w->h->step();
    return w->h->status();
}

int sph_headless_enqueue_step(sph_headless* w) {
    if (!w) return SPH_E_ARG;
    w->h->enqueueStep();
    return w->h->status();
}

sph_ctx* sph_headless_ctx(sph_headless* w) { return w ? w->h->compute->ctx : nullptr; }

int sph_headless_tree_stats(sph_headless* w, int64_t* trees_swapped, double* last_build_ms) {
    if (!w) return SPH_E_ARG;
    if (trees_swapped) *trees_swapped = w->h->compute->treesSwapped();
    if (last_build_ms) *last_build_ms = w->h->compute->lastBuildMs();
    return SPH_OK;
}

int sph_headless_render_buffers(sph_headless* w, void** positions, void** material_ids, void** densities) {
    if (!w || !w->h->compute) return SPH_E_ARG;
    const Compute* c = w->h->compute;
    if (positions) *positions = c->positionBuffer;
    if (material_ids) *material_ids = c->materialIdBuffer;
    if (densities) *densities = c->densityBuffer;
    return SPH_OK;
}

const char* sph_headless_last_error(const sph_headless* w) {
    if (!w) return "null handle";
    const_cast<sph_headless*>(w)->err = w->h->error();
    return w->err.c_str();
}

}  // extern "C"
