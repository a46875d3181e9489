/*
 * capi.hip -- the C ABI of libsphhip.so (sph_capi.h): context life cycle,
 * uploads/downloads, the octree handoff and pass dispatch.
 *
 * Replaces the reference's Compute class (Compute.cpp:22-516) below its
 * public methods: Compute owns Metal buffers and encodes passes into a command
 * buffer; a context here owns HIP device memory and enqueues the passes on
 * its own stream.
 */
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <utility>

#include "sph_internal.h"

namespace {

bool valid_params(const sph_params* p, std::string& why) {
    if (p->cell_power < 1 || p->cell_power > 8) { why = "cell_power must be 1..8"; return false; }
    if (!(p->cell_width > 0.f)) { why = "cell_width must be > 0"; return false; }
    if (p->max_unchecked_pointers < 2) { why = "max_unchecked_pointers must be >= 2"; return false; }
    if (p->gravity_max_recursion < 1 || p->gravity_max_recursion > 4) { why = "gravity_max_recursion must be 1..4"; return false; }
    if (p->max_children_in_leaf < 1) { why = "max_children_in_leaf must be >= 1"; return false; }
    if (p->r_max < 0 || p->r_max > 24) { why = "r_max must be 0..24"; return false; }
    if (p->max_rungs < 0 || p->max_rungs > 24) { why = "max_rungs must be 0..24"; return false; }
    if (p->n_neighbours_estim < 0) { why = "n_neighbours_estim must be >= 0"; return false; }
    if (p->max_nr_iterations < 0) { why = "max_nr_iterations must be >= 0"; return false; }
    if (p->precision && !(p->fixes & SPH_FIX_NEAR_FIELD)) { why = "the precision form needs SPH_FIX_NEAR_FIELD"; return false; }
    auto lanes_ok = [](int v) { return v == 0 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32 || v == 64; };
    if (!lanes_ok(p->density_lanes)) { why = "density_lanes must be 0, 2, 4, 8, 16, 32 or 64"; return false; }
    if (!lanes_ok(p->accel_lanes)) { why = "accel_lanes must be 0, 2, 4, 8, 16, 32 or 64"; return false; }
    if (!lanes_ok(p->list_density_lanes)) {
        why = "list_density_lanes must be 0, 2, 4, 8, 16, 32 or 64";
        return false;
    }
    if (!lanes_ok(p->list_accel_lanes)) { why = "list_accel_lanes must be 0, 2, 4, 8, 16, 32 or 64"; return false; }
    if ((p->list_density_lanes == 2 || p->list_density_lanes == 4 || p->list_accel_lanes == 2 ||
         p->list_accel_lanes == 4) &&
        p->lane_split) { why = "lane_split needs 8 or more lanes"; return false; }
    if (p->gather_reach != 0 && p->gather_reach != 1) { why = "gather_reach must be 0 or 1"; return false; }
    if (p->near_list_cap < 0 || p->near_list_cap > 4096) { why = "near_list_cap must be 0..4096"; return false; }
    if (p->small_n_lanes < 0) { why = "small_n_lanes must be >= 0"; return false; }
    if ((p->density_lanes == 2 || p->density_lanes == 4 || p->accel_lanes == 2 || p->accel_lanes == 4) &&
        p->lane_split) { why = "lane_split needs 8 or more lanes"; return false; }
    if (p->lane_split != 0 && p->lane_split != 1) { why = "lane_split must be 0 or 1"; return false; }
    if (p->density_flat != 0 && p->density_flat != 1) { why = "density_flat must be 0 or 1"; return false; }
    if (p->density_stop != 0 && p->density_stop != 1) { why = "density_stop must be 0 or 1"; return false; }
    if (p->density_order != 0 && (p->density_order < 16 || p->density_order > 512 ||
                                  (p->density_order & (p->density_order - 1)))) {
        why = "density_order must be 0 or a power of two from 16 to 512";
        return false;
    }
    if (p->expansion_order != 2 && p->expansion_order != 3) { why = "expansion_order must be 2 or 3"; return false; }
    if (p->mac_order != 0 && p->mac_order != 2) { why = "mac_order must be 0 or 2"; return false; }
    if (p->expansion_order == 3 && p->gravity_max_recursion != 1 && !(p->fixes & SPH_FIX_NEAR_FIELD)) {
        why = "expansion_order 3 needs gravity_max_recursion 1 or SPH_FIX_NEAR_FIELD";
        return false;
    }
    return true;
}

void free_tree(sph_tree_dev& T) {
    for (int k = 0; k < 2; ++k) {
        dfree(T.octree_up[k]);
        dfree(T.level_tp_up[k]);
        if (T.released[k]) (void)hipEventDestroy(T.released[k]);
        T.released[k] = nullptr;
    }
    T.octree = nullptr;
    T.level_tp = nullptr;
    dfree(T.level_id);
    dfree(T.act_slots);
    dfree(T.act_count);
    dfree(T.node_info);
    dfree(T.parent_slot);
    dfree(T.parent_level);
    dfree(T.seen);
    dfree(T.subtree_end);
    dfree(T.node_child);
    dfree(T.mp_far);
    dfree(T.mp_aux);
    dfree(T.mp_oct);
    dfree(T.mp_clo);
    dfree(T.loc3);
    dfree(T.loc);
    dfree(T.loc_tot);
    dfree(T.near_rec);
    dfree(T.near_cnt);
    T.octree_cap = T.level_cap = T.node_cap = 0;
}

void destroy_ctx(sph_ctx* ctx) {
    if (!ctx) return;
    if (ctx->gstream) (void)hipStreamSynchronize(ctx->gstream);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    float4** f4[] = {&ctx->pos,    &ctx->vel,    &ctx->acc,   &ctx->s_posm, &ctx->s_vel,
                     &ctx->s_posh, &ctx->s_velb, &ctx->s_hyd, &ctx->snap, &ctx->dv};
    for (auto p : f4) dfree(*p);
    float** f1[] = {&ctx->rho, &ctx->u, &ctx->m, &ctx->P, &ctx->h, &ctx->f, &ctx->c,
                    &ctx->du, &ctx->B, &ctx->dhdt, &ctx->gravnorm, &ctx->phi, &ctx->dlnh};
    for (auto p : f1) dfree(*p);
    int32_t** i1[] = {&ctx->mat, &ctx->next_active, &ctx->nr_iters, &ctx->nr_order, &ctx->time, &ctx->list_a, &ctx->list_b};
    for (auto p : i1) dfree(*p);
    dfree(ctx->active);
    dfree(ctx->density_work);
    dfree(ctx->pairs_a);
    dfree(ctx->pairs_b);
    dfree(ctx->cell_se);
    dfree(ctx->cell_mark);
    dfree(ctx->tile_hist);
    dfree(ctx->grav_flags);
    dfree(ctx->digit_total);
    dfree(ctx->feats);
    dfree(ctx->guess);
    dfree(ctx->slot_list);
    dfree(ctx->slot_count);
    dfree(ctx->eos[0]);
    dfree(ctx->eos[1]);
    free_tree(ctx->tree);
    sph_treebuild_free(ctx->tb);
    dfree(ctx->snapshot_img);
    if (ctx->snap_done) (void)hipEventDestroy(ctx->snap_done);
    if (ctx->upload_done) (void)hipEventDestroy(ctx->upload_done);
    if (ctx->tree_free) (void)hipEventDestroy(ctx->tree_free);
    if (ctx->convert_done) (void)hipEventDestroy(ctx->convert_done);
    for (auto& t : ctx->ev_pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto e : ctx->ev_open)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev_up) (void)hipEventDestroy(ctx->ev_up);
    if (ctx->ev_down) (void)hipEventDestroy(ctx->ev_down);
    if (ctx->gstream) (void)hipStreamDestroy(ctx->gstream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    delete ctx;
}

int64_t field_bytes(const sph_ctx* ctx, int field) {
    const int64_t n = ctx->n;
    switch (field) {
        case SPH_F_POS: case SPH_F_VEL: case SPH_F_ACC: return 16 * n;
        case SPH_F_RHO: case SPH_F_U: case SPH_F_MASS: case SPH_F_PRESSURE: case SPH_F_H:
        case SPH_F_MAT: case SPH_F_GRADTERM: case SPH_F_SOUNDSPEED: case SPH_F_DU:
        case SPH_F_BALSARA: case SPH_F_DHDT: case SPH_F_NEXT_ACTIVE: case SPH_F_GRAVNORM:
        case SPH_F_NR_ITERS: case SPH_F_POTENTIAL: case SPH_F_DLNH: return 4 * n;
        case SPH_F_ACTIVE: return n;
        case SPH_F_CELL_ARRAY: return 8 * n;
        case SPH_F_CELL_START: case SPH_F_CELL_END: return 4 * (int64_t)ctx->ncells;
        case SPH_F_GLOBAL_TIME: case SPH_F_DT: return 4;
        case SPH_F_MULTIPOLES:
            return ctx->have_tree ? (ctx->prm.expansion_order == 3 ? 40 : 24) * 4 * (int64_t)ctx->tree.node_count : 0;
        case SPH_F_LOCALS:
            return ctx->have_tree ? (ctx->prm.expansion_order == 3 ? 32 : 16) * 4 * (int64_t)ctx->tree.node_count : 0;
        case SPH_F_DENSITY_WORK: return 16 * n;
        default: return -1;
    }
}

void* field_ptr(sph_ctx* ctx, int field) {
    switch (field) {
        case SPH_F_POS: return ctx->pos;
        case SPH_F_VEL: return ctx->vel;
        case SPH_F_ACC: return ctx->acc;
        case SPH_F_RHO: return ctx->rho;
        case SPH_F_U: return ctx->u;
        case SPH_F_MASS: return ctx->m;
        case SPH_F_PRESSURE: return ctx->P;
        case SPH_F_H: return ctx->h;
        case SPH_F_MAT: return ctx->mat;
        case SPH_F_GRADTERM: return ctx->f;
        case SPH_F_SOUNDSPEED: return ctx->c;
        case SPH_F_DU: return ctx->du;
        case SPH_F_BALSARA: return ctx->B;
        case SPH_F_DHDT: return ctx->dhdt;
        case SPH_F_ACTIVE: return ctx->active;
        case SPH_F_NEXT_ACTIVE: return ctx->next_active;
        case SPH_F_GRAVNORM: return ctx->gravnorm;
        case SPH_F_NR_ITERS: return ctx->nr_iters;
        case SPH_F_DENSITY_WORK: return ctx->density_work;
        case SPH_F_POTENTIAL: return ctx->phi;
        case SPH_F_DLNH: return ctx->dlnh;
        case SPH_F_CELL_ARRAY: return ctx->pairs_a;
        case SPH_F_GLOBAL_TIME: return ctx->time;
        case SPH_F_DT: return ctx->time + 1;
        default: return nullptr;
    }
}

}  // namespace

int ctx_fail(sph_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int ctx_fail_hip(sph_ctx* ctx, hipError_t e, const char* what) {
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
}

namespace {
/* The position record's fourth component holds the particle's mass (the
 * reference's float3 positions leave it unused), so that the gathers from
 * original order (cell copies, P2M, pair sums) read one record, not two. */
__global__ void k_pos_mass(int n, float4* __restrict__ pos, const float* __restrict__ m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pos[i].w = m[i];
}

hipError_t take_event(sph_ctx* ctx, hipEvent_t* e) {
    if (!ctx->ev_pool.empty()) {
        *e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(e);
}
}  // namespace

int ctx_begin_pass(sph_ctx* ctx, int pass) {
    if (!ctx->timing) return SPH_OK;
    hipEvent_t e;
    SPH_HIP_TRY(take_event(ctx, &e));
    SPH_HIP_TRY(hipEventRecord(e, ctx->stream));
    ctx->ev_open[pass] = e;
    return SPH_OK;
}

int ctx_end_pass(sph_ctx* ctx, int pass) {
    if (!ctx->timing || !ctx->ev_open[pass]) return SPH_OK;
    hipEvent_t e;
    SPH_HIP_TRY(take_event(ctx, &e));
    SPH_HIP_TRY(hipEventRecord(e, ctx->stream));
    ctx->ev_pending.push_back({pass, ctx->ev_open[pass], e});
    ctx->ev_open[pass] = nullptr;
    if (ctx->ev_pending.size() >= 4096) return ctx_fold_times(ctx);
    return SPH_OK;
}

int join_gravity(sph_ctx* ctx) {
    if (!ctx->grav_pending) return SPH_OK;
    ctx->grav_pending = false;
    SPH_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_down, 0));
    return SPH_OK;
}

int join_tree(sph_ctx* ctx) {
    if (!ctx->convert_pending) return SPH_OK;
    ctx->convert_pending = false;
    SPH_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->convert_done, 0));
    return SPH_OK;
}

int ctx_fold_times(sph_ctx* ctx) {
    SPH_TRY_RC(join_gravity(ctx));
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto& t : ctx->ev_pending) {
        float ms = 0.f;
        SPH_HIP_TRY(hipEventElapsedTime(&ms, t.start, t.stop));
        ctx->pass_ms[t.pass] += ms;
        ctx->pass_calls[t.pass] += 1;
        ctx->ev_pool.push_back(t.start);
        ctx->ev_pool.push_back(t.stop);
    }
    ctx->ev_pending.clear();
    return SPH_OK;
}

extern "C" {

int sph_create(const sph_params* params, const sph_ic_view* ic, const sph_eos_table* tables, int ntables,
               sph_ctx** out) {
    if (!out) return SPH_E_ARG;
    *out = nullptr;
    if (!params || !ic || !tables || ntables < 1 || ic->n < 1) return SPH_E_ARG;
    if (!ic->pos4 || !ic->vel4 || !ic->rho || !ic->u || !ic->m || !ic->mat || !ic->h) return SPH_E_ARG;
    std::string why;
    if (!valid_params(params, why)) return SPH_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SPH_E_NODEVICE;

    sph_ctx* ctx = new sph_ctx();
    ctx->prm = *params;
    ctx->n = ic->n;
    ctx->ncells = 1 << (3 * params->cell_power);
    const size_t n = (size_t)ic->n;
    auto fail = [&](int code) {
        destroy_ctx(ctx);
        return code;
    };
#define TRY(expr)                                  \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return fail((int)_e); \
    } while (0)
    TRY(hipGetDevice(&ctx->device));
    TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    TRY(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    TRY(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
    TRY(hipEventCreateWithFlags(&ctx->ev_up, hipEventDisableTiming));
    TRY(hipEventCreateWithFlags(&ctx->ev_down, hipEventDisableTiming));
    TRY(hipEventCreateWithFlags(&ctx->snap_done, hipEventDisableTiming));
    TRY(hipEventCreateWithFlags(&ctx->upload_done, hipEventDisableTiming));
    TRY(hipEventCreateWithFlags(&ctx->tree_free, hipEventDisableTiming));
    TRY(hipEventCreateWithFlags(&ctx->convert_done, hipEventDisableTiming));
    for (auto& e : ctx->tree.released) TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    TRY(dalloc(&ctx->pos, n));
    TRY(dalloc(&ctx->vel, n));
    TRY(dalloc(&ctx->acc, n));
    TRY(dalloc(&ctx->dv, n));
    TRY(dalloc(&ctx->snap, n));
    TRY(dalloc(&ctx->s_posm, n));
    TRY(dalloc(&ctx->s_vel, n));
    TRY(dalloc(&ctx->s_posh, n));
    TRY(dalloc(&ctx->s_velb, n));
    TRY(dalloc(&ctx->s_hyd, n));
    float** f1[] = {&ctx->rho, &ctx->u, &ctx->m, &ctx->P, &ctx->h, &ctx->f, &ctx->c, &ctx->du, &ctx->B, &ctx->dhdt, &ctx->gravnorm, &ctx->phi, &ctx->dlnh};
    for (auto p : f1) TRY(dalloc(p, n));
    TRY(hipMemset(ctx->phi, 0, 4 * n));
    TRY(hipMemset(ctx->dlnh, 0, 4 * n));
    TRY(dalloc(&ctx->mat, n));
    TRY(dalloc(&ctx->next_active, n));
    TRY(dalloc(&ctx->nr_iters, n));
    TRY(dalloc(&ctx->nr_order, n));
    TRY(dalloc(&ctx->active, n));
    TRY(dalloc(&ctx->density_work, n));
    TRY(hipMemset(ctx->density_work, 0, 16 * n));
    TRY(dalloc(&ctx->time, 4));
    TRY(dalloc(&ctx->pairs_a, n));
    TRY(dalloc(&ctx->pairs_b, n));
    TRY(dalloc(&ctx->cell_se, (size_t)ctx->ncells));
    TRY(dalloc(&ctx->cell_mark, (size_t)ctx->ncells));
    TRY(hipMemset(ctx->cell_mark, 0, (size_t)ctx->ncells));
    ctx->ntiles = (int)((n + 2047) / 2048);
    TRY(dalloc(&ctx->tile_hist, (size_t)ctx->ntiles * 256));
    TRY(dalloc(&ctx->digit_total, 256));
    TRY(dalloc(&ctx->grav_flags, 4));
    TRY(dalloc(&ctx->tree.act_count, 2 * kMaxLevels));
    TRY(hipMemset(ctx->grav_flags, 0, 16));

    /* initial conditions (Compute.cpp:52-80) */
    TRY(hipMemcpy(ctx->pos, ic->pos4, 16 * n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(ctx->vel, ic->vel4, 16 * n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(ctx->rho, ic->rho, 4 * n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(ctx->u, ic->u, 4 * n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(ctx->m, ic->m, 4 * n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(ctx->mat, ic->mat, 4 * n, hipMemcpyHostToDevice));
    TRY(hipMemcpy(ctx->h, ic->h, 4 * n, hipMemcpyHostToDevice));
    if (ic->p) TRY(hipMemcpy(ctx->P, ic->p, 4 * n, hipMemcpyHostToDevice));
    else TRY(hipMemset(ctx->P, 0, 4 * n));
    TRY(hipMemset(ctx->acc, 0, 16 * n));
    TRY(hipMemset(ctx->f, 0, 4 * n));
    TRY(hipMemset(ctx->c, 0, 4 * n));
    TRY(hipMemset(ctx->du, 0, 4 * n));
    TRY(hipMemset(ctx->B, 0, 4 * n));
    TRY(hipMemset(ctx->dhdt, 0, 4 * n));
    TRY(hipMemset(ctx->gravnorm, 0, 4 * n)); /* never initialised in the reference (Compute.cpp:125) */
    TRY(hipMemset(ctx->next_active, 0, 4 * n));
    TRY(hipMemset(ctx->nr_iters, 0, 4 * n));
    TRY(hipMemset(ctx->active, 1, n));
    TRY(hipMemset(ctx->time, 0, 16));
    k_pos_mass<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, ctx->pos, ctx->m);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(ctx->stream));

    /* EOS tables (Compute.cpp:45-49): slot 0 = material 400 */
    int res = tables[0].res;
    const float* slot[2] = {nullptr, nullptr};
    for (int k = 0; k < ntables; ++k) {
        if (tables[k].res != res || !tables[k].pc || res < 2) return fail(SPH_E_ARG);
        slot[tables[k].material_id == 400 ? 0 : 1] = tables[k].pc;
    }
    if (!slot[0]) slot[0] = slot[1];
    if (!slot[1]) slot[1] = slot[0];
    ctx->eos_res = res;
    for (int k = 0; k < 2; ++k) {
        TRY(dalloc(&ctx->eos[k], (size_t)res * res));
        TRY(hipMemcpy(ctx->eos[k], slot[k], 8 * (size_t)res * res, hipMemcpyHostToDevice));
    }
    int rc = launch_fill_cells(ctx);
    if (rc != SPH_OK) return fail(rc);
    TRY(hipStreamSynchronize(ctx->stream));
#undef TRY
    *out = ctx;
    return SPH_OK;
}

void sph_destroy(sph_ctx* ctx) { destroy_ctx(ctx); }

const char* sph_last_error(const sph_ctx* ctx) { return ctx ? ctx->err.c_str() : noalloc_last_error(); }

int sph_num_particles(const sph_ctx* ctx) { return ctx ? ctx->n : 0; }

void* sph_host_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void sph_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int sph_snapshot_positions(sph_ctx* ctx, float* pinned_dst4) {
    if (!ctx || !pinned_dst4) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    const size_t bytes = 16 * (size_t)ctx->n;
    SPH_HIP_TRY(hipMemcpyAsync(ctx->snap, ctx->pos, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    SPH_HIP_TRY(hipEventRecord(ctx->snap_done, ctx->stream));
    SPH_HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->snap_done, 0));
    SPH_HIP_TRY(hipMemcpyAsync(pinned_dst4, ctx->snap, bytes, hipMemcpyDeviceToHost, ctx->copy_stream));
    SPH_HIP_TRY(hipEventRecord(ctx->snap_done, ctx->copy_stream));
    return SPH_OK;
}

int sph_wait_snapshot(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    SPH_HIP_TRY(hipEventSynchronize(ctx->snap_done));
    return SPH_OK;
}

}  // extern "C"

int tree_reserve(sph_ctx* ctx, int64_t n_ints, const int32_t* level_counts, int n_levels, int32_t node_count,
                 int* slot) {
    sph_tree_dev& T = ctx->tree;
    int64_t total = 0;
    int32_t max_nonlast = 0;
    for (int l = 0; l < n_levels; ++l) {
        if (level_counts[l] < 0) return ctx_fail(ctx, SPH_E_ARG, "negative level count");
        total += level_counts[l];
        if (l != n_levels - 1) max_nonlast = std::max(max_nonlast, level_counts[l]);
    }
    if (total != node_count) return ctx_fail(ctx, SPH_E_ARG, "level lists do not cover node_count nodes");
    if (level_counts[0] != 1) return ctx_fail(ctx, SPH_E_ARG, "level 0 must hold only the root");
    const double grow = 1.2; /* EXTRA_MEMORY_MULTIPLIER */
    const int64_t list_need = (int64_t)std::max<int32_t>(max_nonlast, 1) * ctx->prm.max_unchecked_pointers;
    const bool need_realloc = n_ints > T.octree_cap || total > T.level_cap || node_count > T.node_cap ||
                              list_need > ctx->list_cap;
    if (need_realloc) {
        /* growth: drain both streams, then reallocate (Compute.cpp:464-502) */
        SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
        SPH_HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
        if (n_ints > T.octree_cap) {
            const int64_t cap = (int64_t)(n_ints * grow) + 16;
            for (int k = 0; k < 2; ++k) {
                dfree(T.octree_up[k]);
                SPH_HIP_TRY(dalloc(&T.octree_up[k], (size_t)cap));
            }
            T.octree_cap = cap;
            ctx->have_tree = false;
        }
        if (total > T.level_cap) {
            const int64_t cap = (int64_t)(total * grow) + 16;
            for (int k = 0; k < 2; ++k) {
                dfree(T.level_tp_up[k]);
                SPH_HIP_TRY(dalloc(&T.level_tp_up[k], (size_t)cap));
            }
            dfree(T.level_id);
            SPH_HIP_TRY(dalloc(&T.level_id, (size_t)cap));
            dfree(T.act_slots);
            SPH_HIP_TRY(dalloc(&T.act_slots, (size_t)cap));
            T.level_cap = cap;
            ctx->have_tree = false;
        }
        if (node_count > T.node_cap) {
            const int64_t cap = (int64_t)(node_count * grow) + 16;
            dfree(T.node_info);
            dfree(T.parent_slot);
            dfree(T.parent_level);
            dfree(T.seen);
            dfree(T.subtree_end);
            dfree(T.node_child);
            dfree(T.mp_far);
            dfree(T.mp_aux);
            dfree(T.mp_oct);
            dfree(T.mp_clo);
            dfree(T.loc3);
            dfree(T.loc);
            SPH_HIP_TRY(dalloc(&T.node_info, (size_t)cap));
            SPH_HIP_TRY(dalloc(&T.parent_slot, (size_t)cap));
            SPH_HIP_TRY(dalloc(&T.parent_level, (size_t)cap));
            SPH_HIP_TRY(dalloc(&T.seen, (size_t)cap));
            SPH_HIP_TRY(dalloc(&T.subtree_end, (size_t)cap));
            SPH_HIP_TRY(dalloc(&T.node_child, (size_t)cap * 8));
            SPH_HIP_TRY(dalloc(&T.mp_far, (size_t)cap * 3));
            SPH_HIP_TRY(dalloc(&T.mp_aux, (size_t)cap * 3));
            if (ctx->prm.expansion_order == 3) {
                SPH_HIP_TRY(dalloc(&T.mp_oct, (size_t)cap * 3));
                SPH_HIP_TRY(dalloc(&T.loc3, (size_t)cap * 3));
            }
            if (ctx->prm.precision) SPH_HIP_TRY(dalloc(&T.mp_clo, (size_t)cap));
            SPH_HIP_TRY(dalloc(&T.loc, (size_t)cap * 4));
            dfree(T.loc_tot);
            dfree(T.near_rec);
            dfree(T.near_cnt);
            if (ctx->prm.fixes & SPH_FIX_NEAR_FIELD) {
                SPH_HIP_TRY(dalloc(&T.loc_tot, (size_t)cap * 7));
                if (ctx->prm.near_list_cap > 0) {
                    SPH_HIP_TRY(dalloc(&T.near_rec, (size_t)cap * (size_t)ctx->prm.near_list_cap));
                    SPH_HIP_TRY(dalloc(&T.near_cnt, (size_t)cap));
                }
            }
            T.node_cap = cap;
        }
        if (list_need > ctx->list_cap) {
            const int64_t cap = (int64_t)(list_need * grow) + 16;
            dfree(ctx->list_a);
            dfree(ctx->list_b);
            SPH_HIP_TRY(dalloc(&ctx->list_a, (size_t)cap));
            SPH_HIP_TRY(dalloc(&ctx->list_b, (size_t)cap));
            ctx->list_cap = cap;
        }
    }
    /* the buffer the current tree does not use */
    *slot = ctx->have_tree ? (T.up_cur ^ 1) : T.up_cur;
    return SPH_OK;
}

int tree_commit(sph_ctx* ctx, int k, int64_t n_ints, const int32_t* level_counts, int n_levels, int32_t node_count,
                hipStream_t convert_stream) {
    sph_tree_dev& T = ctx->tree;
    int32_t max_nonlast = 0;
    for (int l = 0; l + 1 < n_levels; ++l) max_nonlast = std::max(max_nonlast, level_counts[l]);
    /* every pass enqueued from here on uses buffer k */
    if (ctx->have_tree && k != T.up_cur) SPH_HIP_TRY(hipEventRecord(T.released[T.up_cur], ctx->stream));
    T.up_cur = k;
    T.octree = T.octree_up[k];
    T.level_tp = T.level_tp_up[k];
    T.node_count = node_count;
    T.n_ints = n_ints;
    T.max_nonlast = max_nonlast;
    T.level_counts.assign(level_counts, level_counts + n_levels);
    T.level_offsets.assign(n_levels, 0);
    for (int l = 1; l < n_levels; ++l) T.level_offsets[l] = T.level_offsets[l - 1] + level_counts[l - 1];
    int rc = launch_tree_convert(ctx, convert_stream);
    if (rc != SPH_OK) return rc;
    ctx->have_tree = true;
    ctx->tree_check_pending = true;
    return SPH_OK;
}

extern "C" {

int sph_upload_tree(sph_ctx* ctx, const int32_t* octree, int64_t n_ints, const int32_t* level_nodes,
                    const int32_t* level_counts, int n_levels, int32_t node_count) {
    if (!ctx || !octree || !level_nodes || !level_counts || n_levels < 1 || node_count < 1 || n_ints < 2)
        return SPH_E_ARG;
    /* the tree buffers' release events must cover a down pass still running */
    SPH_TRY_RC(join_gravity(ctx));
    sph_tree_dev& T = ctx->tree;
    int k = 0;
    SPH_TRY_RC(tree_reserve(ctx, n_ints, level_counts, n_levels, node_count, &k));
    /* upload on the copy stream, once the compute stream has released buffer k */
    SPH_HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, T.released[k], 0));
    SPH_HIP_TRY(hipMemcpyAsync(T.octree_up[k], octree, 4 * (size_t)n_ints, hipMemcpyHostToDevice, ctx->copy_stream));
    SPH_HIP_TRY(hipMemcpyAsync(T.level_tp_up[k], level_nodes, 4 * (size_t)node_count, hipMemcpyHostToDevice,
                               ctx->copy_stream));
    SPH_HIP_TRY(hipEventRecord(ctx->upload_done, ctx->copy_stream));
    /* The conversion rewrites the node records, which every pass enqueued so
     * far may read: it runs on the copy stream after them, and the compute
     * stream waits for it only where the tree is next read (join_tree), so
     * that work that does not read it (the sort) overlaps the upload. */
    SPH_HIP_TRY(hipEventRecord(ctx->tree_free, ctx->stream));
    SPH_HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->tree_free, 0));
    SPH_TRY_RC(tree_commit(ctx, k, n_ints, level_counts, n_levels, node_count, ctx->copy_stream));
    SPH_HIP_TRY(hipEventRecord(ctx->convert_done, ctx->copy_stream));
    ctx->convert_pending = true;
    ctx->upload_pending = true;
    return SPH_OK;
}

int sph_tree_upload_done(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    if (!ctx->upload_pending) return 1;
    hipError_t e = hipEventQuery(ctx->upload_done);
    if (e == hipSuccess) {
        ctx->upload_pending = false;
        return 1;
    }
    if (e == hipErrorNotReady) return 0;
    return ctx_fail_hip(ctx, e, "hipEventQuery(upload_done)");
}

int sph_sort(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    ctx->slot_list_valid = false;
    int rc = ctx_begin_pass(ctx, SPH_PASS_SORT);
    if (rc == SPH_OK) rc = launch_sort(ctx);
    if (rc == SPH_OK) ctx->morton_fresh = true;
    if (rc == SPH_OK) rc = ctx_end_pass(ctx, SPH_PASS_SORT);
    return rc;
}

int sph_gravity(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    if (!ctx->have_tree) return ctx_fail(ctx, SPH_E_NOTREE, "gravity before any tree upload");
    SPH_TRY_RC(join_gravity(ctx));
    SPH_TRY_RC(join_tree(ctx));
    SPH_TRY_RC(fold_dv(ctx));
    /* the near-field form checks the node numbering in the up pass */
    if (ctx->prm.fixes & SPH_FIX_NEAR_FIELD) ctx->tree_check_pending = true;
    ctx->slot_list_valid = false; /* the up pass sets the activity */
    int rc = ctx_begin_pass(ctx, SPH_PASS_GRAVITY_UP);
    /* With gravNorm = 0 every MAC test rejects (minGrav = 0, Q6).  As written
     * the leaves then force-accept; in the near-field form every pair would be
     * summed directly, O(n^2).  Before its first evaluation the near-field form
     * therefore estimates |a| with one as-written pass (P = 2, leaves
     * force-accept) and rebuilds the multipoles' minGrav from it. */
    if (rc == SPH_OK && (ctx->prm.fixes & SPH_FIX_NEAR_FIELD) && !ctx->gravnorm_known) {
        rc = launch_gravity_up(ctx);
        if (rc == SPH_OK) rc = launch_gravity_down(ctx, true);
    }
    if (rc == SPH_OK) rc = launch_gravity_up(ctx);
    if (rc == SPH_OK) rc = ctx_end_pass(ctx, SPH_PASS_GRAVITY_UP);
    /* the down pass on gstream, after the up pass; the density pass may run
     * meanwhile on the context stream */
    if (rc == SPH_OK) rc = (int)hipEventRecord(ctx->ev_up, ctx->stream);
    if (rc == SPH_OK) rc = (int)hipStreamWaitEvent(ctx->gstream, ctx->ev_up, 0);
    if (rc != SPH_OK) return rc;
    hipStream_t main_stream = ctx->stream;
    ctx->stream = ctx->gstream;
    rc = ctx_begin_pass(ctx, SPH_PASS_GRAVITY_DOWN);
    if (rc == SPH_OK) rc = launch_gravity_down(ctx);
    if (rc == SPH_OK) rc = ctx_end_pass(ctx, SPH_PASS_GRAVITY_DOWN);
    ctx->stream = main_stream;
    if (rc == SPH_OK) rc = (int)hipEventRecord(ctx->ev_down, ctx->gstream);
    if (rc == SPH_OK) {
        ctx->grav_pending = true;
        ctx->gravnorm_known = true;
    }
    return rc;
}

int sph_density(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    if (!ctx->have_prev) return ctx_fail(ctx, SPH_E_STATE, "density before any sort");
    int rc = ctx_begin_pass(ctx, SPH_PASS_DENSITY);
    if (rc == SPH_OK) rc = refresh_morton(ctx);
    if (rc == SPH_OK) rc = launch_density_prologue(ctx);
    if (rc == SPH_OK) rc = launch_density(ctx);
    if (rc == SPH_OK) rc = launch_density_epilogue(ctx);
    if (rc == SPH_OK) rc = ctx_end_pass(ctx, SPH_PASS_DENSITY);
    return rc;
}

int sph_acceleration(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    /* no join: the pass writes its pair sums to dv, not to the acc the
     * gravity down pass may still be writing */
    if (!ctx->have_prev) return ctx_fail(ctx, SPH_E_STATE, "acceleration before any sort");
    int rc = ctx_begin_pass(ctx, SPH_PASS_ACCELERATION);
    if (rc == SPH_OK) rc = launch_acceleration(ctx);
    if (rc == SPH_OK) rc = ctx_end_pass(ctx, SPH_PASS_ACCELERATION);
    return rc;
}

int sph_step(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    int rc = ctx_begin_pass(ctx, SPH_PASS_STEP);
    if (rc == SPH_OK) rc = launch_step(ctx);
    if (rc == SPH_OK) rc = ctx_end_pass(ctx, SPH_PASS_STEP);
    return rc;
}

int sph_sync(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    SPH_TRY_RC(join_tree(ctx)); /* the malformed-tree flag below */
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    SPH_HIP_TRY(hipGetLastError());
    if (ctx->tree_check_pending) {
        uint32_t flag = 0;
        SPH_HIP_TRY(hipMemcpy(&flag, ctx->grav_flags, 4, hipMemcpyDeviceToHost));
        ctx->tree_check_pending = false;
        if (flag != 0) {
            ctx->have_tree = false;
            return ctx_fail(ctx, SPH_E_ARG,
                            flag == 2u ? "uploaded octree's node ids are not in pre-order (required by the "
                                         "near-field form); gravity was skipped"
                                       : "uploaded octree is malformed; gravity was skipped");
        }
    }
    return SPH_OK;
}

int64_t sph_field_bytes(const sph_ctx* ctx, int field) {
    if (!ctx) return SPH_E_ARG;
    return field_bytes(ctx, field);
}

int sph_download(sph_ctx* ctx, int field, void* dst) {
    if (!ctx || !dst) return SPH_E_ARG;
    const int64_t bytes = field_bytes(ctx, field);
    if (bytes < 0) return ctx_fail(ctx, SPH_E_ARG, "unknown field");
    SPH_TRY_RC(join_gravity(ctx));
    if (field == SPH_F_ACC) SPH_TRY_RC(fold_dv(ctx));
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (field == SPH_F_CELL_START || field == SPH_F_CELL_END) {
        std::vector<uint2> se((size_t)ctx->ncells);
        SPH_HIP_TRY(hipMemcpy(se.data(), ctx->cell_se, 8 * (size_t)ctx->ncells, hipMemcpyDeviceToHost));
        uint32_t* o = (uint32_t*)dst;
        for (int i = 0; i < ctx->ncells; ++i) o[i] = field == SPH_F_CELL_START ? se[i].x : se[i].y;
        return SPH_OK;
    }
    if (field == SPH_F_MULTIPOLES || field == SPH_F_LOCALS) {
        if (!ctx->have_tree) return ctx_fail(ctx, SPH_E_NOTREE, "no tree");
        const size_t K = (size_t)ctx->tree.node_count;
        float* o = (float*)dst;
        const bool p3 = ctx->prm.expansion_order == 3;
        std::vector<float4> third(p3 ? 3 * K : 0);
        if (field == SPH_F_MULTIPOLES) {
            std::vector<float4> far(3 * K), aux(3 * K);
            SPH_HIP_TRY(hipMemcpy(far.data(), ctx->tree.mp_far, 48 * K, hipMemcpyDeviceToHost));
            SPH_HIP_TRY(hipMemcpy(aux.data(), ctx->tree.mp_aux, 48 * K, hipMemcpyDeviceToHost));
            if (p3) SPH_HIP_TRY(hipMemcpy(third.data(), ctx->tree.mp_oct, 48 * K, hipMemcpyDeviceToHost));
            const size_t stride = p3 ? 40 : 24;
            for (size_t i = 0; i < K; ++i) {
                const float4 f0 = far[3 * i], f1 = far[3 * i + 1], f2 = far[3 * i + 2];
                const float4 x0 = aux[3 * i], x1 = aux[3 * i + 1], x2 = aux[3 * i + 2];
                float* r = o + stride * i;
                const float rec[24] = {f0.x, f0.y, f0.z, f0.w, x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w,
                                       f1.x, x2.x, x2.y, x2.z, f1.y, f1.z, f1.w, f2.x, f2.y, f2.z, f2.w, 0.f};
                std::memcpy(r, rec, sizeof(rec));
                if (p3) {
                    const float4 q0 = third[3 * i], q1 = third[3 * i + 1], q2 = third[3 * i + 2];
                    const float ext[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                           q2.x, q2.y, q2.z, 0.f, 0.f, 0.f, 0.f, 0.f};
                    std::memcpy(r + 24, ext, sizeof(ext));
                }
            }
        } else {
            std::vector<float4> loc(4 * K);
            SPH_HIP_TRY(hipMemcpy(loc.data(), ctx->tree.loc, 64 * K, hipMemcpyDeviceToHost));
            if (p3) SPH_HIP_TRY(hipMemcpy(third.data(), ctx->tree.loc3, 48 * K, hipMemcpyDeviceToHost));
            const size_t stride = p3 ? 32 : 16;
            for (size_t i = 0; i < K; ++i) {
                const float4 l0 = loc[4 * i], l1 = loc[4 * i + 1], l2 = loc[4 * i + 2], l3 = loc[4 * i + 3];
                const float rec[16] = {l0.x, l0.y, l0.z, 0.f, l0.w, l1.x, l1.y, l1.z,
                                       l1.w, l2.x, l2.y, l2.z, l2.w, l3.x, 0.f, 0.f};
                std::memcpy(o + stride * i, rec, sizeof(rec));
                if (p3) {
                    const float4 m0 = third[3 * i], m1 = third[3 * i + 1], m2 = third[3 * i + 2];
                    const float ext[16] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w,
                                           m2.x, m2.y, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    std::memcpy(o + stride * i + 16, ext, sizeof(ext));
                }
            }
        }
        return SPH_OK;
    }
    void* src = field_ptr(ctx, field);
    if (!src) return ctx_fail(ctx, SPH_E_ARG, "field not downloadable");
    SPH_HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_device_field(sph_ctx* ctx, int field, void** dev_ptr, int64_t* bytes) {
    if (!ctx || !dev_ptr) return SPH_E_ARG;
    switch (field) {
        case SPH_F_POS: case SPH_F_VEL: case SPH_F_RHO: case SPH_F_U: case SPH_F_MASS: case SPH_F_PRESSURE:
        case SPH_F_H: case SPH_F_MAT: case SPH_F_GRADTERM: case SPH_F_SOUNDSPEED: case SPH_F_DU:
        case SPH_F_BALSARA: case SPH_F_DHDT: case SPH_F_ACTIVE: case SPH_F_NEXT_ACTIVE: case SPH_F_GRAVNORM:
        case SPH_F_NR_ITERS: case SPH_F_POTENTIAL: case SPH_F_DLNH:
            break;
        case SPH_F_ACC:
            return ctx_fail(ctx, SPH_E_ARG, "acc holds the gravity part until the step adds the pair sums: "
                                            "use sph_download");
        default:
            return ctx_fail(ctx, SPH_E_ARG, "not a per-particle field");
    }
    *dev_ptr = field_ptr(ctx, field);
    if (bytes) *bytes = field_bytes(ctx, field);
    return SPH_OK;
}

int sph_upload(sph_ctx* ctx, int field, const void* src) {
    if (!ctx || !src) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    const int64_t bytes = field_bytes(ctx, field);
    if (bytes < 0) return ctx_fail(ctx, SPH_E_ARG, "unknown field");
    if (field == SPH_F_CELL_ARRAY || field == SPH_F_CELL_START || field == SPH_F_CELL_END ||
        field == SPH_F_MULTIPOLES || field == SPH_F_LOCALS || field == SPH_F_NR_ITERS ||
        field == SPH_F_DENSITY_WORK || field == SPH_F_POTENTIAL)
        return ctx_fail(ctx, SPH_E_ARG, "derived field cannot be uploaded");
    void* dst = field_ptr(ctx, field);
    if (!dst) return ctx_fail(ctx, SPH_E_ARG, "field not uploadable");
    if (field == SPH_F_GRAVNORM) ctx->gravnorm_known = true;
    /* acc replaced: the pending pair sums belonged to the old values; the
     * activity replaced: fold with the flags the pass ran under first */
    if (field == SPH_F_ACC) ctx->dv_pending = false;
    if (field == SPH_F_POS || field == SPH_F_VEL || field == SPH_F_MASS) ctx->morton_fresh = false;
    if (field == SPH_F_ACTIVE) SPH_TRY_RC(fold_dv(ctx));
    ctx->slot_list_valid = false;
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    SPH_HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    if (field == SPH_F_POS || field == SPH_F_MASS) {
        k_pos_mass<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, ctx->pos, ctx->m);
        SPH_HIP_TRY(hipGetLastError());
    }
    return SPH_OK;
}

int sph_set_timing(sph_ctx* ctx, int enabled) {
    if (!ctx) return SPH_E_ARG;
    ctx->timing = enabled != 0;
    return SPH_OK;
}

int sph_pass_times(sph_ctx* ctx, double* ms, int64_t* calls) {
    if (!ctx || !ms) return SPH_E_ARG;
    int rc = ctx_fold_times(ctx);
    if (rc != SPH_OK) return rc;
    for (int p = 0; p < SPH_PASS_COUNT; ++p) {
        ms[p] = ctx->pass_ms[p];
        if (calls) calls[p] = ctx->pass_calls[p];
    }
    return SPH_OK;
}

int sph_set_work_counters(sph_ctx* ctx, int enabled) {
    if (!ctx) return SPH_E_ARG;
    ctx->count_work = enabled != 0;
    return SPH_OK;
}

int sph_reset_pass_times(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    int rc = ctx_fold_times(ctx);
    if (rc != SPH_OK) return rc;
    for (int p = 0; p < SPH_PASS_COUNT; ++p) {
        ctx->pass_ms[p] = 0;
        ctx->pass_calls[p] = 0;
    }
    return SPH_OK;
}

}  // extern "C"
