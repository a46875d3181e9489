/*
 * diagnostics.hip -- FP64 conservation sums and the FP64 time view.  No
 * counterpart in the reference, which reports neither.
 *
 * One pass over the particles: each thread forms its particle's ten terms
 * in FP64, a workgroup reduces them in LDS (fixed tree order) and writes one
 * partial per workgroup; the host adds the partials in workgroup order, so
 * the result does not depend on scheduling.
 */
#include <vector>

#include "sph_internal.h"

namespace {

constexpr int kTerms = SPH_CONSERVATION_COUNT;
constexpr int kThreads = 256;
constexpr int kBlocks = 512;

__global__ __launch_bounds__(kThreads) void k_conserve(int n, const float4* __restrict__ pos,
                                                       const float4* __restrict__ vel, const float* __restrict__ m,
                                                       const float* __restrict__ u, const float* __restrict__ phi,
                                                       double* __restrict__ part) {
    __shared__ double red[kThreads];
    double t[kTerms];
#pragma unroll
    for (int k = 0; k < kTerms; ++k) t[k] = 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double mi = m[i];
        const float4 x = pos[i], v = vel[i];
        const double vx = v.x, vy = v.y, vz = v.z, px = x.x, py = x.y, pz = x.z;
        t[0] += 0.5 * mi * (vx * vx + vy * vy + vz * vz);
        t[1] += mi * (double)u[i];
        t[2] += 0.5 * mi * (double)phi[i];
        t[3] += mi * vx;
        t[4] += mi * vy;
        t[5] += mi * vz;
        t[6] += mi * (py * vz - pz * vy);
        t[7] += mi * (pz * vx - px * vz);
        t[8] += mi * (px * vy - py * vx);
        t[9] += mi;
    }
    for (int k = 0; k < kTerms; ++k) {
        red[threadIdx.x] = t[k];
        __syncthreads();
        for (int off = kThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x * kTerms + k] = red[0];
        __syncthreads();
    }
}

}  // namespace

extern "C" {

int sph_conservation(sph_ctx* ctx, double* out) {
    if (!ctx || !out) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    double* part = nullptr;
    SPH_HIP_TRY(hipMalloc((void**)&part, sizeof(double) * kBlocks * kTerms));
    k_conserve<<<kBlocks, kThreads, 0, ctx->stream>>>(ctx->n, ctx->pos, ctx->vel, ctx->m, ctx->u, ctx->phi, part);
    hipError_t e = hipGetLastError();
    std::vector<double> host((size_t)kBlocks * kTerms);
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), part, host.size() * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(part);
    SPH_HIP_TRY(e);
    for (int k = 0; k < kTerms; ++k) {
        double s = 0.0;
        for (int b = 0; b < kBlocks; ++b) s += host[(size_t)b * kTerms + k];
        out[k] = s;
    }
    return SPH_OK;
}

int sph_time(sph_ctx* ctx, double* seconds) {
    if (!ctx || !seconds) return SPH_E_ARG;
    SPH_TRY_RC(join_gravity(ctx));
    int32_t ticks = 0;
    SPH_HIP_TRY(hipStreamSynchronize(ctx->stream));
    SPH_HIP_TRY(hipMemcpy(&ticks, ctx->time, 4, hipMemcpyDeviceToHost));
    const double dtmin1 = (double)ctx->prm.dt0 / (double)(1 << ctx->prm.r_max);
    *seconds = (double)ticks * dtmin1;
    return SPH_OK;
}

}  // extern "C"
