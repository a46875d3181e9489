/*
 * render.hip -- the headless snapshot raster of drawToSnapshot
 * (render.metal:49-66, Compute.cpp:396-448): every particle whose (x, y)
 * falls in the window [250, 250 + 20 R_E)^2 (R_E = 6.371 Mm) marks its pixel
 * with the largest material id that lands there.  Row 0 of the image is the
 * window's top (largest y).
 *
 * The reference accepts uv == res exactly, which addresses one pixel past the
 * row (past the buffer for the top row); such particles are skipped here.
 */
#include <algorithm>

#include "sph_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_snapshot_clear(int32_t* __restrict__ img, int count) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) img[i] = 0;
}

__global__ __launch_bounds__(256) void k_draw_snapshot(const float4* __restrict__ pos,
                                                       const int32_t* __restrict__ mat, int n, int res,
                                                       int32_t* __restrict__ img) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float R_E = 6.3710f;
    const float scale = (float)res / (20.f * R_E);
    const float4 p = pos[i];
    const float ux = (p.x - 250.f) * scale, uy = (p.y - 250.f) * scale;
    if (!(ux >= 0.f) || !(uy >= 0.f) || ux > (float)res || uy > (float)res) return;
    const int px = (int)ux, py = (int)uy;
    if (px >= res || py >= res) return;
    atomicMax(&img[px + (res - 1 - py) * res], mat[i]);
}

}  // namespace

extern "C" {

int sph_draw_snapshot(sph_ctx* ctx, int res) {
    if (!ctx || res < 1 || res > 16384) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    const int64_t count = (int64_t)res * res;
    if (count > ctx->snapshot_cap) {
        SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
        dfree(ctx->snapshot_img);
        SPH_HIP_TRY(dalloc(&ctx->snapshot_img, (size_t)count));
        ctx->snapshot_cap = count;
    }
    ctx->snapshot_res = res;
    hipStream_t s = ctx->stream;
    k_snapshot_clear<<<(int)std::min<int64_t>((count + 255) / 256, 2048), 256, 0, s>>>(ctx->snapshot_img, (int)count);
    k_draw_snapshot<<<(ctx->n + 255) / 256, 256, 0, s>>>(ctx->pos, ctx->mat, ctx->n, res, ctx->snapshot_img);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

int sph_download_snapshot(sph_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return SPH_E_ARG;
    if (!ctx->snapshot_img || ctx->snapshot_res < 1) return ctx_fail(ctx, SPH_E_STATE, "no snapshot drawn");
    SPH_TRY_RC(join_gravity(ctx));
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    SPH_HIP_TRY(hipMemcpy(out, ctx->snapshot_img, 4 * (size_t)ctx->snapshot_res * ctx->snapshot_res,
                          hipMemcpyDeviceToHost));
    return SPH_OK;
}

int sph_snapshot_resolution(const sph_ctx* ctx) { return ctx ? ctx->snapshot_res : SPH_E_ARG; }

}  // extern "C"
