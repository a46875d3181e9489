/*
 * hydro.hip -- SPH density (Newton-Raphson smoothing length, EOS table,
 * Balsara switch), acceleration (pressure + viscosity, CFL timestep) and the
 * drift/kick step, for gfx950.
 *
 * References: density.metal:12-239, accelerations.metal:12-141,
 * step.metal:12-46, kernels.h:35-102, cellsToScan.h:21-59.
 *
 * One thread per Morton-sorted slot.  Neighbour candidates are read from the
 * Morton-ordered copies written after the sort (s_posm, s_vel) and at the
 * start of the acceleration pass (s_posh, s_velb, s_hyd), so a cell's particles are
 * contiguous in HBM and neighbouring threads read overlapping cells.  Summation order per
 * particle is the reference's: cells x -> y -> z over the particle's own cell
 * box, particles of a cell in sorted order.
 *
 * The reference keeps a per-thread cache of up to N_NEIGHBOURS_ESTIM
 * (index, r) pairs from the last cell scan (density.metal:66-70,113-118).  It
 * is used (a) inside the NR loop only when h < max_h, where max_h is an int
 * (so only after a particle sat at h = 1), and (b) for the second pass when
 * it did not fill up.  Its content is exactly "the members (r^2 <= 4 h_s^2) of
 * the cell box of the last scanned h_s, in scan order", so instead of a
 * private array (scratch memory on a GPU) the kernel rescans that box with
 * that membership test and the cache's own range test (r > 2h), which yields
 * the same terms in the same order.
 */
#include "sph_internal.h"

namespace {

constexpr float kPi1 = 1.f / 3.14159265358979323846f;

struct Cells {
    const uint2* cell_se;
    float cell_size;
    int mask;
};

/* cellsToScan.h:33-59.  Guard: a non-finite bound gives an empty box and an
 * extent is capped at 2^cell_power cells per axis (the periodic grid), so a
 * corrupted particle cannot make the scan unbounded.  Neither guard changes a
 * box of a finite particle with h <= MAX_SMOOTHING_LENGTH. */
__device__ __forceinline__ void cell_box(float3 x, float cs, float h, int cells_per_dim, int mn[3], int mx[3]) {
    const float radius = 2.f * h;
    const float lo[3] = {(x.x - radius) / cs, (x.y - radius) / cs, (x.z - radius) / cs};
    const float hi[3] = {(x.x + radius) / cs, (x.y + radius) / cs, (x.z + radius) / cs};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (!(lo[k] <= hi[k]) || !(fabsf(lo[k]) < 1e9f) || !(fabsf(hi[k]) < 1e9f)) {
            mn[k] = 0;
            mx[k] = -1;
            continue;
        }
        mn[k] = (int)floorf(lo[k]);
        mx[k] = (int)floorf(hi[k]);
        if (mx[k] - mn[k] >= cells_per_dim) mx[k] = mn[k] + cells_per_dim - 1;
    }
}

/* kernels.h:35-50 */
__device__ __forceinline__ float kW(float r, float h1) {
    const float q = r * h1;
    if (q > 2.0f) return 0.f;
    const float fac = kPi1 * h1 * h1 * h1;
    if (q > 1.0f) {
        const float tmp = 2.f - q;
        return fac * 0.25f * tmp * tmp * tmp;
    }
    return fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
}
/* kW with its normalisation fac = h^-3 / pi given, as selects instead of
 * branches: the same operations in the same order, so the same value. */
__device__ __forceinline__ float kW_sel(float q, float fac) {
    const float tmp = 2.f - q;
    const float outer = fac * 0.25f * tmp * tmp * tmp;
    const float inner = fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
    return q > 2.0f ? 0.f : (q > 1.0f ? outer : inner);
}
/* kernels.h:52-69: scalar factor g with gradW = g * x_ij / r */
__device__ __forceinline__ float kGradWFactor(float r, float h1) {
    const float q = r * h1;
    if (q > 2.f || q < 1e-12f) return 0.f;
    const float fac = kPi1 * h1 * h1 * h1 * h1;
    if (q > 1.f) {
        const float tmp = 2.f - q;
        return -0.75f * fac * tmp * tmp;
    }
    return 3.f * fac * (0.75f * q - 1.f) * q;
}
/* kernels.h:71-85 */
__device__ __forceinline__ float kdWdh(float r, float h1) {
    const float q = r * h1;
    if (q > 2.0f) return 0.f;
    const float fac = -3.f * kPi1 * h1 * h1 * h1 * h1;
    if (q > 1.0f) return fac * (2.f - 4.f * q + 2.5f * q * q - 0.5f * q * q * q);
    return fac * (1.f + 0.5f * (3.f * q - 5.f) * q * q);
}

/* The cubic spline of kernels.h:35-85 through its truncated powers
 * a = (2 - q)+ and b = (1 - q)+:  w(q) = a^3/4 - b^3,  w'(q) = 3 b^2 -
 * 3 a^2/4  and  w + q w'/3 = a^2 (a - q)/4 - b^2 (b - q).  These are the
 * reference's piecewise polynomials (each is 0 beyond q = 2) written without
 * branches or selects: the neighbour passes are bound by VALU issue, and the
 * pieces share a and b.  They differ from the reference's nested forms in
 * rounding only. */
__device__ __forceinline__ float kW_tp(float q, float fac) {
    const float a = fmaxf(2.f - q, 0.f), b = fmaxf(1.f - q, 0.f);
    return fac * (0.25f * a * a * a - b * b * b);
}
/* kernels.h:52-69: gradW = g x_ij / r with g = h^-4 / pi * w'(q) */
__device__ __forceinline__ float kGradWFactor_tp(float r, float h1) {
    const float q = r * h1;
    const float a = fmaxf(2.f - q, 0.f), b = fmaxf(1.f - q, 0.f);
    const float fac = kPi1 * h1 * h1 * h1 * h1;
    return fac * (3.f * b * b - 0.75f * a * a);
}
/* kernels.h:71-85: dW/dh = -3 h^-4 / pi (w + q w'/3) */
__device__ __forceinline__ float kdWdh_tp(float r, float h1) {
    const float q = r * h1;
    const float a = fmaxf(2.f - q, 0.f), b = fmaxf(1.f - q, 0.f);
    const float fac = -3.f * kPi1 * h1 * h1 * h1 * h1;
    return fac * (0.25f * a * a * (a - q) - b * b * (b - q));
}

/* Bilinear sample with Metal's texture conventions (density.metal:45-50):
 * normalised coordinates, texel centres at (i + 1/2)/res, clamp to edge.
 * Column = internal energy axis, row = density axis. */
__device__ __forceinline__ float2 eos_sample(const float2* __restrict__ tab, int res, float u, float rho, float min_u,
                                             float max_u, float min_rho, float max_rho) {
    float uc = logf(u / min_u) / logf(max_u / min_u);
    float rc = logf(rho / min_rho) / logf(max_rho / min_rho);
    uc = fminf(fmaxf(uc, 0.f), 1.f);
    rc = fminf(fmaxf(rc, 0.f), 1.f);
    const float tx = uc * (float)res - 0.5f, ty = rc * (float)res - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float a = tx - fx, b = ty - fy;
    const int ix = (int)fx, iy = (int)fy;
    const int i0 = min(max(ix, 0), res - 1), i1 = min(max(ix + 1, 0), res - 1);
    const int j0 = min(max(iy, 0), res - 1), j1 = min(max(iy + 1, 0), res - 1);
    const float2 t00 = tab[j0 * res + i0], t01 = tab[j0 * res + i1];
    const float2 t10 = tab[j1 * res + i0], t11 = tab[j1 * res + i1];
    const float w00 = (1.f - a) * (1.f - b), w01 = a * (1.f - b), w10 = (1.f - a) * b, w11 = a * b;
    return make_float2(w00 * t00.x + w01 * t01.x + w10 * t10.x + w11 * t11.x,
                       w00 * t00.y + w01 * t01.y + w10 * t10.y + w11 * t11.y);
}

struct DensityArgs {
    int n;
    const uint2* sorted;
    Cells cells;
    const float4* s_posm;
    const float4* s_vel;
    const uint8_t* active;
    const int32_t* mat;
    const float* u;
    float *rho, *h, *P, *c, *f, *B;
    int32_t* nr_iters;
    int4* work;
    int32_t* time;
    const float2* eos0;
    const float2* eos1;
    int eos_res;
    float min_u, max_u, min_rho, max_rho;
    float eta, res_eta, max_h_len;
    int max_it, ne;
    int fix_cache;
    int dt_reset;
    const float* guess;   /* learned starting h per sorted slot, or null */
    /* README rule: the inactive particles' EOS runs in k_inactive_eos (in
     * original order, coalesced) and their threads here return at once */
    int skip_inactive;
    int stop_rule; /* sph_params.density_stop */
    /* README rule: the sorted slots of the active particles and their count
     * on the device (k_active_slots); null: every slot */
    const int32_t* slots;
    const int32_t* nslots;
    /* every slot: the order in which the groups take them (k_nr_order), or
     * null for slot order */
    const int32_t* order;
};

/* Whether the Newton-Raphson iteration goes on after a scan whose relative
 * change of h is eps; eps_prev is the previous scan's (3e38 before the
 * second scan).  Rule 0 is the reference's successive-iterate test
 * (density.metal:73,131-136).  Rule 1 goes on until the estimated distance
 * of h from the fixed point, eps / (1 - k) with the observed contraction
 * k = eps / eps_prev, is within eta, or eps is at the rounding floor. */
__device__ __forceinline__ bool nr_goes_on(int rule, float eps, float eps_prev, float eta) {
    if (rule == 0) return eps > eta;
    if (eps <= SPH_NR_NOISE_FLOOR) return false;
    if (eps_prev > 1.0e38f) return true;
    /* each eps is a difference of FP32 values: allow one ulp either way */
    const float e = eps + SPH_NR_EPS_ULP;
    const float k = e / (eps_prev - SPH_NR_EPS_ULP);
    if (k >= SPH_NR_MAX_CONTRACTION) return true;
    return e > eta * (1.f - k);
}

constexpr float kGuessTrust = 0.05f;
/* Successive changes below this are left to the plain iteration's stopping
 * rule: near the fixed point they are rounding noise of the density sum and
 * may grow from one scan to the next without the guess being wrong. */
constexpr float kGuessNoise = 1e-5f;

/* The learned starting guess is judged on its first two scans only: the
 * first residual must be below kGuessTrust and the second must not exceed
 * the first (unless both are at the noise level).  A guess that fails either
 * test is dropped and the iteration restarts from the step's h; one that
 * passes both is then iterated exactly as the plain method would iterate
 * from that point.  Returns true when the iteration must restart. */
__device__ __forceinline__ bool guess_rejected(bool& guided, float eps, float& eps_last) {
    if (!guided) return false;
    const bool first = eps_last > 1.0e38f;
    if (first ? eps > kGuessTrust : (eps > eps_last && eps > kGuessNoise)) {
        guided = false;
        return true;
    }
    if (first) eps_last = eps;
    else guided = false;
    return false;
}

/* density.metal:12-239 */
__global__ __launch_bounds__(256) void k_density(DensityArgs a) {
    const int blk = xcd_block((a.n + 255) / 256);
    if (blk < 0) return;
    const int index = blk * 256 + threadIdx.x;
    if (index >= a.n) return;
    if (index == 0) a.time[1] = a.dt_reset;
    const uint32_t ind = a.sorted[index].y;
    const float2* table = a.mat[ind] == 400 ? a.eos0 : a.eos1;
    if (!a.active[ind]) {
        if (a.skip_inactive) return;
        const float rho = a.rho[ind];
        const float2 pc = eos_sample(table, a.eos_res, a.u[ind], rho, a.min_u, a.max_u, a.min_rho, a.max_rho);
        a.P[ind] = pc.x;
        a.c[ind] = pc.y;
        a.nr_iters[ind] = 0;
        if (a.work) a.work[ind] = make_int4(0, 0, 0, 0);
        return;
    }
    const float4 pm = a.s_posm[index];
    const float3 position = make_float3(pm.x, pm.y, pm.z);
    const float mass_i = pm.w;
    const uint2* __restrict__ cse = a.cells.cell_se;
    const float4* __restrict__ spm = a.s_posm;
    const float cs = a.cells.cell_size;
    const int mask = a.cells.mask;

    /* With a learned starting guess the iteration starts there and restarts
     * from the step's h when guess_rejected() says so.  No counterpart in the
     * reference, whose iteration always starts from h. */
    const float h_start = a.h[ind];
    float h_i = a.guess ? a.guess[index] : h_start;
    bool guided = a.guess != nullptr;
    float eps_last = 3.0e38f;
    int max_h = 0;
    float h2x4 = 4.f * h_i * h_i;
    float h1 = 1.f / h_i;
    float density = 0.f;
    float eps = 1.f;
    int count = 0;
    int nN = 0;             /* cached-neighbour count of the last scan (capped) */
    float h_scan = h_i;     /* h of the last scan */
    float hs2x4 = h2x4;     /* its 4 h^2 */
    int w_cand = 0, w_in = 0, f_cand = 0, f_in = 0;
    float eps_prev = 3.0e38f;
    bool go_on = a.stop_rule ? true : eps > a.eta;
    while (go_on && count < a.max_it) {
        density = 0.f;
        const bool use_cache = (h_i < (float)max_h) && (nN < a.ne);
        const float hb = use_cache ? h_scan : h_i;
        const float hb2 = use_cache ? hs2x4 : h2x4;
        const float hsum = h_i + h_i;
        int mn[3], mx[3];
        cell_box(position, cs, hb, mask + 1, mn, mx);
        int cnt = 0;
        for (int x = mn[0]; x <= mx[0]; x++) {
            const uint32_t ex = dev_expand_bits((uint32_t)(x & mask));
            for (int y = mn[1]; y <= mx[1]; y++) {
                const uint32_t exy = ex | (dev_expand_bits((uint32_t)(y & mask)) << 1);
                for (int z = mn[2]; z <= mx[2]; z++) {
                    const uint32_t key = exy | (dev_expand_bits((uint32_t)(z & mask)) << 2);
                    const uint2 se = cse[key];
                    if (se.x == 0xFFFFFFFFu) continue;
                    w_cand += (int)(se.y - se.x + 1);
                    for (uint32_t k = se.x; k <= se.y; k++) {
                        const float4 q = spm[k];
                        const float dx = q.x - position.x, dy = q.y - position.y, dz = q.z - position.z;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 > hb2) continue;
                        w_in++;
                        const float r = sqrtf(r2);
                        if (use_cache) {
                            if (r > hsum) continue;
                        } else {
                            cnt++;
                        }
                        density += kW(r, h1) * q.w;
                    }
                }
            }
        }
        if (!use_cache) {
            nN = min(cnt, a.ne);
            h_scan = h_i;
            hs2x4 = h2x4;
        }
        float newH = fminf(a.res_eta * powf(mass_i / density, 1.f / 3.f), a.max_h_len);
        newH = 0.5f * (h_i + newH);
        eps = fabsf(newH - h_i) / h_i;
        if (guess_rejected(guided, eps, eps_last)) {
            h_i = h_start;
            h1 = 1.f / h_i;
            h2x4 = 4.f * h_i * h_i;
            max_h = 0;
            eps = 1.f;
            eps_prev = 3.0e38f;
            go_on = a.stop_rule ? true : eps > a.eta;
            count++;
            continue;
        }
        go_on = nr_goes_on(a.stop_rule, eps, eps_prev, a.eta);
        eps_prev = eps;
        if (go_on) {
            max_h = (int)fmaxf((float)max_h, h_i);
            h_i = newH;
            h1 = 1.f / h_i;
            h2x4 = 4.f * h_i * h_i;
            count++;
        }
    }
    a.nr_iters[ind] = count;

    /* second pass (density.metal:146-213) */
    float density_h = 0.f, velDiv = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    const float4 vi4 = a.s_vel[index];
    const bool cached = nN < a.ne;
    if (!cached && a.fix_cache) density = 0.f;
    {
        const float hb = cached ? h_scan : h_i;
        const float hb2 = cached ? hs2x4 : h2x4;
        const float hsum = h_i + h_i;
        int mn[3], mx[3];
        cell_box(position, cs, hb, mask + 1, mn, mx);
        for (int x = mn[0]; x <= mx[0]; x++) {
            const uint32_t ex = dev_expand_bits((uint32_t)(x & mask));
            for (int y = mn[1]; y <= mx[1]; y++) {
                const uint32_t exy = ex | (dev_expand_bits((uint32_t)(y & mask)) << 1);
                for (int z = mn[2]; z <= mx[2]; z++) {
                    const uint32_t key = exy | (dev_expand_bits((uint32_t)(z & mask)) << 2);
                    const uint2 se = cse[key];
                    if (se.x == 0xFFFFFFFFu) continue;
                    f_cand += (int)(se.y - se.x + 1);
                    for (uint32_t k = se.x; k <= se.y; k++) {
                        const float4 q = spm[k];
                        const float dx = q.x - position.x, dy = q.y - position.y, dz = q.z - position.z;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 > hb2) continue;
                        f_in++;
                        const float r = sqrtf(r2);
                        if (cached && r > hsum) continue;
                        const float mass = q.w;
                        if (!cached) density += kW(r, h1) * mass;
                        density_h += mass * kdWdh(r, h1);
                        if (r > 1e-12f) {
                            const float r1 = 1.f / r;
                            const float4 vj = a.s_vel[k];
                            const float vx = vj.x - vi4.x, vy = vj.y - vi4.y, vz = vj.z - vi4.z;
                            const float g = kGradWFactor(r, h1);
                            const float gx = g * (dx * r1), gy = g * (dy * r1), gz = g * (dz * r1);
                            velDiv += mass * (vx * gx + vy * gy + vz * gz);
                            cx += mass * (vy * gz - vz * gy);
                            cy += mass * (vz * gx - vx * gz);
                            cz += mass * (vx * gy - vy * gx);
                        }
                    }
                }
            }
        }
    }
    velDiv = fabsf(velDiv) / density;
    const float absCurl = sqrtf(cx * cx + cy * cy + cz * cz) / density;
    const float fgrad = 1.f / (1.f + (h_i / (3.f * density)) * density_h);
    const float2 pc = eos_sample(table, a.eos_res, a.u[ind], density, a.min_u, a.max_u, a.min_rho, a.max_rho);
    const float bal = velDiv / (velDiv + absCurl + 1e-4f * (pc.y / h_i));
    a.h[ind] = h_i;
    a.rho[ind] = density;
    a.f[ind] = fgrad;
    a.P[ind] = pc.x;
    a.c[ind] = pc.y;
    a.B[ind] = bal;
    if (a.work) a.work[ind] = make_int4(w_cand, w_in, f_cand, f_in);
}

/* Group sums over G lanes in a fixed butterfly order (reproducible). */
template <int G>
__device__ __forceinline__ float gsum(float v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
template <int G>
__device__ __forceinline__ int gsum_i(int v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

/* density.metal:12-239 with G lanes per particle.  SPLIT 0: the lanes of a
 * group split the particle's cell box (lane l takes cells l, l + G, ... in
 * the box's x -> y -> z order).  SPLIT 1: the group walks the box's cells
 * together and lane l takes candidates l, l + G, ... of each cell, so the
 * lanes read consecutive records.  Either way the group sums the partial
 * results with a butterfly, so every lane holds the particle's sums and the
 * Newton-Raphson control flow stays uniform over the group.
 * The terms are the reference's; only the order of the additions differs
 * (per-cell partial sums, then a tree), which changes results at the
 * rounding level, and the distance's square root and the second pass's 1/r
 * are the hardware's (about 1 ulp, as under the reference's Metal fast-math
 * defaults) rather than the correctly rounded library sequences. */
/* Calls body(key) for the cells first, first + STEP, ... of the box
 * [mn, mx] in x -> y -> z order (z fastest), with the Morton key of the
 * periodic cell.  The position in the box advances by carries instead of a
 * division per cell, and the x/y part of the key is recomputed only when x
 * or y changes. */
template <int STEP, class Body>
__device__ __forceinline__ void box_cells(const int mn[3], const int mx[3], int first, int mask, Body body) {
    const int ny = max(mx[1] - mn[1] + 1, 0), nz = max(mx[2] - mn[2] + 1, 0);
    const int ncell = max(mx[0] - mn[0] + 1, 0) * ny * nz;
    if (first >= ncell) return;
    int iz = first % nz;
    const int t = first / nz;
    int iy = t % ny, ix = t / ny;
    uint32_t exy = dev_expand_bits((uint32_t)((mn[0] + ix) & mask)) |
                   (dev_expand_bits((uint32_t)((mn[1] + iy) & mask)) << 1);
    for (int c = first; c < ncell; c += STEP) {
        body(exy | (dev_expand_bits((uint32_t)((mn[2] + iz) & mask)) << 2));
        iz += STEP;
        if (iz >= nz) {
            do {
                iz -= nz;
                if (++iy >= ny) {
                    iy = 0;
                    ++ix;
                }
            } while (iz >= nz);
            exy = dev_expand_bits((uint32_t)((mn[0] + ix) & mask)) |
                  (dev_expand_bits((uint32_t)((mn[1] + iy) & mask)) << 1);
        }
    }
}

/* Calls on_cell(count) for each non-empty cell and body(k) for every
 * candidate slot k of the cells first, first + STEP, ... of the box
 * [mn, mx] in x -> y -> z order, like box_cells with a loop over each
 * cell's slots, but as one flat loop: a lane moves on to its next cell as
 * soon as it has used up the current one, instead of every lane of the wave
 * waiting at each cell for the lane with the most candidates.  The slot range
 * of a lane's next cell is loaded one cell ahead.  The slots and their order
 * are those of box_cells, so the sums are bit for bit the same. */
template <int STEP, class CellFn, class Body>
__device__ __forceinline__ void box_slots_flat(const int mn[3], const int mx[3], int first, int mask,
                                               const uint2* __restrict__ cse, CellFn on_cell, Body body) {
    const int ny = max(mx[1] - mn[1] + 1, 0), nz = max(mx[2] - mn[2] + 1, 0);
    const int ncell = max(mx[0] - mn[0] + 1, 0) * ny * nz;
    int c = first;
    if (c >= ncell) return;
    int iz = c % nz;
    const int t = c / nz;
    int iy = t % ny, ix = t / ny;
    uint32_t exy = dev_expand_bits((uint32_t)((mn[0] + ix) & mask)) |
                   (dev_expand_bits((uint32_t)((mn[1] + iy) & mask)) << 1);
    /* the cell after the current one: its slot range, loaded ahead */
    uint2 nxt = cse[exy | (dev_expand_bits((uint32_t)((mn[2] + iz) & mask)) << 2)];
    bool more = true;
    uint32_t k = 1, end = 0; /* current range, empty */
    for (;;) {
        if (k > end) {
            if (!more) break;
            k = nxt.x;
            end = nxt.y;
            if (k == 0xFFFFFFFFu) {
                k = 1;
                end = 0;
            } else {
                on_cell(end - k + 1);
            }
            c += STEP;
            if (c < ncell) {
                iz += STEP;
                if (iz >= nz) {
                    do {
                        iz -= nz;
                        if (++iy >= ny) {
                            iy = 0;
                            ++ix;
                        }
                    } while (iz >= nz);
                    exy = dev_expand_bits((uint32_t)((mn[0] + ix) & mask)) |
                          (dev_expand_bits((uint32_t)((mn[1] + iy) & mask)) << 1);
                }
                nxt = cse[exy | (dev_expand_bits((uint32_t)((mn[2] + iz) & mask)) << 2)];
            } else {
                more = false;
            }
            if (k > end) continue;
        }
        body(k);
        ++k;
    }
}

/* The candidate slots of a particle's box for lane `lane` of its G lanes:
 * SPLIT 0 whole cells per lane (per-cell loops, or FLAT: one flat loop over
 * the lane's cells), SPLIT 1 the candidates of each cell split over the
 * lanes.  on_cell(count) per non-empty cell (on lane 0 only with SPLIT). */
template <int G, int SPLIT, int FLAT, class CellFn, class Body>
__device__ __forceinline__ void box_slots(const int mn[3], const int mx[3], int lane, int mask,
                                          const uint2* __restrict__ cse, CellFn on_cell, Body body) {
    if (FLAT && !SPLIT) {
        box_slots_flat<G>(mn, mx, lane, mask, cse, on_cell, body);
        return;
    }
    box_cells<SPLIT ? 1 : G>(mn, mx, SPLIT ? 0 : lane, mask, [&](uint32_t key) {
        const uint2 se = cse[key];
        if (se.x == 0xFFFFFFFFu) return;
        if (!SPLIT || lane == 0) on_cell(se.y - se.x + 1);
        for (uint32_t k = se.x + (SPLIT ? lane : 0); k <= se.y; k += SPLIT ? G : 1) body(k);
    });
}

/* One particle (sorted slot `index`) on its G lanes (`lane` = 0..G-1).
 * COUNT: also count the work (SPH_F_DENSITY_WORK); the pass is bound by
 * VALU issue, so the production kernel leaves those counts out. */
template <int G, int SPLIT, int FLAT, bool COUNT>
__device__ __forceinline__ void density_group_particle(const DensityArgs& a, int index, int lane) {
    if (index == 0 && lane == 0) a.time[1] = a.dt_reset;
    const uint32_t ind = a.sorted[index].y;
    const float2* table = a.mat[ind] == 400 ? a.eos0 : a.eos1;
    if (!a.active[ind]) {
        if (a.skip_inactive) return;
        if (lane == 0) {
            const float rho = a.rho[ind];
            const float2 pc = eos_sample(table, a.eos_res, a.u[ind], rho, a.min_u, a.max_u, a.min_rho, a.max_rho);
            a.P[ind] = pc.x;
            a.c[ind] = pc.y;
            a.nr_iters[ind] = 0;
            if (a.work) a.work[ind] = make_int4(0, 0, 0, 0);
        }
        return;
    }
    const float4 pm = a.s_posm[index];
    const float3 position = make_float3(pm.x, pm.y, pm.z);
    const float mass_i = pm.w;
    const uint2* __restrict__ cse = a.cells.cell_se;
    const float4* __restrict__ spm = a.s_posm;
    const float cs = a.cells.cell_size;
    const int mask = a.cells.mask;

    const float h_start = a.h[ind];
    float h_i = a.guess ? a.guess[index] : h_start;
    bool guided = a.guess != nullptr;
    float eps_last = 3.0e38f;
    int max_h = 0;
    float h2x4 = 4.f * h_i * h_i;
    float h1 = 1.f / h_i;
    float density = 0.f;
    float eps = 1.f;
    int count = 0;
    int nN = 0;
    float h_scan = h_i;
    float hs2x4 = h2x4;
    int w_cand = 0, w_in = 0, f_cand = 0, f_in = 0;
    float eps_prev = 3.0e38f;
    bool go_on = a.stop_rule ? true : eps > a.eta;
    while (go_on && count < a.max_it) {
        const bool use_cache = (h_i < (float)max_h) && (nN < a.ne);
        const float hb = use_cache ? h_scan : h_i;
        const float hb2 = use_cache ? hs2x4 : h2x4;
        const float hsum = h_i + h_i;
        int mn[3], mx[3];
        cell_box(position, cs, hb, mask + 1, mn, mx);
        float part = 0.f;
        int cnt = 0; /* candidates inside the box radius (the NR neighbour count without the cache) */
        /* the candidate body without branches: a wave of 64 lanes almost
         * always holds a lane whose candidate is inside 2h, so the kernel
         * term costs the same either way, and selects leave out the exec-mask
         * bookkeeping of the nested tests (same terms, same order) */
        const float fac = kPi1 * h1 * h1 * h1;
        auto on_cell = [&](uint32_t nc) { if constexpr (COUNT) w_cand += (int)nc; };
        if (!use_cache) {
            /* the box of this h: w(q) is 0 beyond 2h, so the kernel sum needs
             * no test (a candidate beyond 2h adds 0, or a few ulp of a term
             * where rounding puts q just under 2), and the normalisation
             * h^-3 / pi is applied once to the sum */
            box_slots<G, SPLIT, FLAT>(mn, mx, lane, mask, cse, on_cell, [&](uint32_t k) {
                const float4 q = spm[k];
                const float dx = q.x - position.x, dy = q.y - position.y, dz = q.z - position.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                cnt += r2 > hb2 ? 0 : 1;
                const float r = __builtin_amdgcn_sqrtf(r2);
                const float ta = fmaxf(fmaf(-h1, r, 2.f), 0.f), tb = fmaxf(fmaf(-h1, r, 1.f), 0.f);
                part = fmaf(0.25f * ta * ta * ta - tb * tb * tb, q.w, part);
            });
            part *= fac;
        } else {
            box_slots<G, SPLIT, FLAT>(mn, mx, lane, mask, cse, on_cell, [&](uint32_t k) {
                const float4 q = spm[k];
                const float dx = q.x - position.x, dy = q.y - position.y, dz = q.z - position.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                const bool in = !(r2 > hb2);
                const float r = __builtin_amdgcn_sqrtf(r2);
                const bool take = in && !(r > hsum);
                cnt += in ? 1 : 0;
                const float w = kW_tp(r * h1, fac);
                part += take ? w * q.w : 0.f;
            });
        }
        if constexpr (COUNT) w_in += cnt;
        density = gsum<G>(part);
        if (!use_cache) {
            nN = min(gsum_i<G>(cnt), a.ne);
            h_scan = h_i;
            hs2x4 = h2x4;
        }
        float newH = fminf(a.res_eta * powf(mass_i / density, 1.f / 3.f), a.max_h_len);
        newH = 0.5f * (h_i + newH);
        eps = fabsf(newH - h_i) / h_i;
        if (guess_rejected(guided, eps, eps_last)) {
            h_i = h_start;
            h1 = 1.f / h_i;
            h2x4 = 4.f * h_i * h_i;
            max_h = 0;
            eps = 1.f;
            eps_prev = 3.0e38f;
            go_on = a.stop_rule ? true : eps > a.eta;
            count++;
            continue;
        }
        go_on = nr_goes_on(a.stop_rule, eps, eps_prev, a.eta);
        eps_prev = eps;
        if (go_on) {
            max_h = (int)fmaxf((float)max_h, h_i);
            h_i = newH;
            h1 = 1.f / h_i;
            h2x4 = 4.f * h_i * h_i;
            count++;
        }
    }

    /* second pass (density.metal:146-213) */
    float density_h = 0.f, velDiv = 0.f, dpart = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    const float4 vi4 = a.s_vel[index];
    const bool cached = nN < a.ne;
    {
        const float hb = cached ? h_scan : h_i;
        const float hb2 = cached ? hs2x4 : h2x4;
        const float hsum = h_i + h_i;
        int mn[3], mx[3];
        cell_box(position, cs, hb, mask + 1, mn, mx);
        box_slots<G, SPLIT, FLAT>(mn, mx, lane, mask, cse, [&](uint32_t nc) { if constexpr (COUNT) f_cand += (int)nc; },
                                  [&](uint32_t k) {
            const float4 q = spm[k];
            const float dx = q.x - position.x, dy = q.y - position.y, dz = q.z - position.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > hb2) return;
            if constexpr (COUNT) f_in++;
            const float r = __builtin_amdgcn_sqrtf(r2);
            if (cached && r > hsum) return;
            const float mass = q.w;
            if (!cached) dpart += kW_tp(r * h1, kPi1 * h1 * h1 * h1) * mass;
            density_h += mass * kdWdh_tp(r, h1);
            if (r > 1e-12f) {
                const float r1 = __builtin_amdgcn_rcpf(r);
                const float4 vj = a.s_vel[k];
                const float vx = vj.x - vi4.x, vy = vj.y - vi4.y, vz = vj.z - vi4.z;
                const float g = kGradWFactor_tp(r, h1);
                const float gx = g * (dx * r1), gy = g * (dy * r1), gz = g * (dz * r1);
                velDiv += mass * (vx * gx + vy * gy + vz * gz);
                cx += mass * (vy * gz - vz * gy);
                cy += mass * (vz * gx - vx * gz);
                cz += mass * (vx * gy - vy * gx);
            }
        });
    }
    /* as written, a rescan adds to the last NR density (Q7) unless fixed */
    if (!cached) density = (a.fix_cache ? 0.f : density) + gsum<G>(dpart);
    density_h = gsum<G>(density_h);
    velDiv = gsum<G>(velDiv);
    cx = gsum<G>(cx);
    cy = gsum<G>(cy);
    cz = gsum<G>(cz);
    if constexpr (COUNT) {
        w_cand = gsum_i<G>(w_cand);
        w_in = gsum_i<G>(w_in);
        f_cand = gsum_i<G>(f_cand);
        f_in = gsum_i<G>(f_in);
    }
    if (lane != 0) return;
    a.nr_iters[ind] = count;
    velDiv = fabsf(velDiv) / density;
    const float absCurl = sqrtf(cx * cx + cy * cy + cz * cz) / density;
    const float fgrad = 1.f / (1.f + (h_i / (3.f * density)) * density_h);
    const float2 pc = eos_sample(table, a.eos_res, a.u[ind], density, a.min_u, a.max_u, a.min_rho, a.max_rho);
    const float bal = velDiv / (velDiv + absCurl + 1e-4f * (pc.y / h_i));
    a.h[ind] = h_i;
    a.rho[ind] = density;
    a.f[ind] = fgrad;
    a.P[ind] = pc.x;
    a.c[ind] = pc.y;
    a.B[ind] = bal;
    if (COUNT && a.work) a.work[ind] = make_int4(w_cand, w_in, f_cand, f_in);
}

/* Every slot: one particle per group of G lanes, workgroups XCD-remapped.
 * With a slot list (README rule, most particles inactive): a persistent grid
 * whose XCD x walks the x-th eighth of the list, so no workgroup is spent on
 * an inactive particle. */
/* The sorted slots of each window of W consecutive slots in ascending order
 * of their particles' previous Newton-Raphson counts (ties by slot), for the
 * density pass's `order`: a wave runs as many scans as its slowest particle,
 * and the counts change little from one step to the next, so waves of like
 * counts waste fewer scans.  One workgroup of W threads per window. */
template <int W>
__global__ __launch_bounds__(W) void k_nr_order(int n, const uint2* __restrict__ sorted,
                                               const int32_t* __restrict__ nr_iters, int32_t* __restrict__ order) {
    __shared__ int s_pred[W];
    const int t = (int)threadIdx.x;
    const int base = (int)blockIdx.x * W;
    const int pk = base + t < n ? nr_iters[sorted[base + t].y] : 0x7FFFFFFF;
    s_pred[t] = pk;
    __syncthreads();
    int r = 0;
    for (int j = 0; j < W; ++j) {
        const int pj = s_pred[j];
        r += (pj < pk || (pj == pk && j < t)) ? 1 : 0;
    }
    if (base + r < n) order[base + r] = base + t;
}

/* Every slot: one particle per group of G lanes, workgroups XCD-remapped;
 * with `order` (k_nr_order) the groups take the slots in that order instead
 * (only which lanes take a particle changes, not its result). */
/* The list form at 7 waves per SIMD (79 -> 72 VGPRs, no spills): density
 * pass 0.70 -> 0.69 ms at 4e6 with individual timesteps. */
template <int G, int SPLIT, int FLAT, bool LIST, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LIST ? 7 : 1))) void k_density_group(DensityArgs a) {
    constexpr int PPB = 256 / G;
    const int lane = (int)threadIdx.x % G, g = (int)threadIdx.x / G;
    if constexpr (!LIST) {
        const int blk = xcd_block((a.n + PPB - 1) / PPB);
        if (blk < 0) return;
        const int k = blk * PPB + g;
        if (k < a.n) density_group_particle<G, SPLIT, FLAT, COUNT>(a, a.order ? a.order[k] : k, lane);
        return;
    }
    const int total = *a.nslots;
    const int nlog = (total + PPB - 1) / PPB;
    const int per = (nlog + kXcds - 1) / kXcds;
    const int xcd = (int)(blockIdx.x % kXcds), nper = (int)(gridDim.x / kXcds);
    for (int k = (int)(blockIdx.x / kXcds); k < per; k += nper) {
        const int lb = xcd * per + k;
        if (lb >= nlog) break;
        const int e = lb * PPB + g;
        if (e < total) density_group_particle<G, SPLIT, FLAT, COUNT>(a, a.slots[e], lane);
    }
}

/* density.metal:40-55 for the inactive particles, in original order
 * (coalesced) instead of through the sorted slots. */
__global__ __launch_bounds__(256) void k_inactive_eos(DensityArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) a.time[1] = a.dt_reset; /* (the density pass's first thread does too) */
    if (i >= a.n || a.active[i]) return;
    const float2* table = a.mat[i] == 400 ? a.eos0 : a.eos1;
    const float2 pc = eos_sample(table, a.eos_res, a.u[i], a.rho[i], a.min_u, a.max_u, a.min_rho, a.max_rho);
    a.P[i] = pc.x;
    a.c[i] = pc.y;
    a.nr_iters[i] = 0;
    if (a.work) a.work[i] = make_int4(0, 0, 0, 0);
}

/* Sorted slots of the active particles (for the learned guess's features
 * and inference): one ballot and one counter update per wave (the list order
 * across waves is not fixed; each particle's result depends on its own slot
 * only). */
__global__ __launch_bounds__(256) void k_active_slots(int n, const uint2* __restrict__ sorted,
                                                      const uint8_t* __restrict__ active, int32_t* list,
                                                      int32_t* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n && active[sorted[i].y];
    const uint64_t b = __ballot(act);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0 && b) base = atomicAdd(count, (int)__popcll(b));
    base = __shfl(base, 0, 64);
    if (act) list[base + (int)__popcll(b & ((1ull << lane) - 1ull))] = i;
}

__global__ void k_zero_i32(int32_t* p) { p[0] = 0; }

/* Morton-ordered copies of what the acceleration loop reads for every
 * candidate, packed so that the range test needs one 16-byte load:
 * (x, y, z, h), (v, Balsara factor), (m, f P / rho^2, rho, c).  The product
 * f P / rho^2 is formed exactly as accelerations.metal:45,91 forms it. */
/* The acceleration pass's Morton-ordered records.  MORTON: position, mass
 * and velocity come from the sort's Morton-ordered copies (unchanged since
 * the sort), read in order; only the density pass's outputs are gathered
 * from original order.  Otherwise every field is gathered. */
/* With `mark` (the active particles only), a slot whose cell does not hold
 * `epoch` is skipped: no active particle's cell box reaches it, so the
 * acceleration pass reads none of its records. */
template <bool MORTON>
__global__ __launch_bounds__(256) void k_gather_hydro(int n, const uint2* __restrict__ sorted,
                                                      const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                      const float* __restrict__ m, const float4* __restrict__ s_posm,
                                                      const float4* __restrict__ s_vel, const float* __restrict__ h,
                                                      const float* __restrict__ f, const float* __restrict__ P,
                                                      const float* __restrict__ rho, const float* __restrict__ c,
                                                      const float* __restrict__ B, float4* __restrict__ s_posh,
                                                      float4* __restrict__ s_velb, float4* __restrict__ s_hyd,
                                                      const uint8_t* __restrict__ mark, uint8_t epoch) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint2 sk = sorted[k];
    if (mark && mark[sk.x] != epoch) return;
    const uint32_t j = sk.y;
    const float r = rho[j];
    float4 x, v;
    float mj;
    if (MORTON) {
        x = s_posm[k];
        v = s_vel[k];
        mj = x.w;
    } else {
        x = pos[j];
        v = vel[j];
        mj = x.w;
    }
    s_posh[k] = make_float4(x.x, x.y, x.z, h[j]);
    s_velb[k] = make_float4(v.x, v.y, v.z, B[j]);
    s_hyd[k] = make_float4(mj, f[j] * P[j] / (r * r), r, c[j]);
}

/* Marks with `epoch` every cell of the acceleration pass's cell box of each
 * listed (active) particle: the same box from the same position and h as
 * that pass forms (the sort's Morton copy of the position, or the position
 * itself when that copy is stale, and h).  G lanes per particle stride over
 * the box; a persistent grid walks the list, XCD x taking the x-th eighth.
 * Marks left from 255 calls before only make the gather copy a few records
 * no pass reads. */
template <int G>
__global__ __launch_bounds__(256) void k_mark_cells(const int32_t* __restrict__ slots, const int32_t* __restrict__ nslots,
                                                    const uint2* __restrict__ sorted,
                                                    const float4* __restrict__ s_posm, const float4* __restrict__ pos,
                                                    const float* __restrict__ h, float cell_size, int mask,
                                                    float max_h_len, uint8_t epoch, uint8_t* __restrict__ mark) {
    constexpr int PPB = 256 / G;
    const int lane = (int)threadIdx.x % G, g = (int)threadIdx.x / G;
    const int total = *nslots;
    const int nlog = (total + PPB - 1) / PPB;
    const int per = (nlog + kXcds - 1) / kXcds;
    const int xcd = (int)(blockIdx.x % kXcds), nper = (int)(gridDim.x / kXcds);
    for (int q = (int)(blockIdx.x / kXcds); q < per; q += nper) {
        const int lb = xcd * per + q;
        if (lb >= nlog) break;
        const int e = lb * PPB + g;
        if (e >= total) continue;
        const int k = slots[e];
        const uint32_t j = sorted[k].y;
        const float4 x = s_posm ? s_posm[k] : pos[j];
        int mn[3], mx[3];
        cell_box(make_float3(x.x, x.y, x.z), cell_size, fminf(h[j], max_h_len), mask + 1, mn, mx);
        const int ny = max(mx[1] - mn[1] + 1, 0), nz = max(mx[2] - mn[2] + 1, 0);
        const int ncell = max(mx[0] - mn[0] + 1, 0) * ny * nz;
        for (int c = lane; c < ncell; c += G) {
            const int iz = c % nz, t = c / nz;
            const int iy = t % ny, ix = t / ny;
            const uint32_t key = dev_expand_bits((uint32_t)((mn[0] + ix) & mask)) |
                                 (dev_expand_bits((uint32_t)((mn[1] + iy) & mask)) << 1) |
                                 (dev_expand_bits((uint32_t)((mn[2] + iz) & mask)) << 2);
            mark[key] = epoch;
        }
    }
}

/* The sort's Morton-ordered copies (position + mass, velocity) gathered
 * again from original order, for a density pass that follows a change of
 * positions, velocities or masses without a new sort (the reference's
 * density pass reads the position buffer itself). */
__global__ __launch_bounds__(256) void k_refresh_morton(int n, const uint2* __restrict__ sorted,
                                                        const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel, const float* __restrict__ m,
                                                        float4* __restrict__ s_posm, float4* __restrict__ s_vel) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t j = sorted[k].y;
    s_posm[k] = pos[j]; /* w: the mass */
    s_vel[k] = vel[j];
}

struct AccelArgs {
    int n;
    const uint2* sorted;
    Cells cells;
    const float4* s_posh;
    const float4* s_velb;
    const float4* s_hyd;
    const uint8_t* active;
    float4* dv;       /* the pair sum; the step adds it to the gravity acceleration */
    float* du;
    float* dhdt;
    int32_t* next_active;
    int32_t* time;
    float alpha, cfl, gamma, dt0, dtmin, max_h_len;
    int r_max;
    float dlnh_limit; /* > 0: SPH_FIX_INACTIVE_DRIFT's bound on |dh/dt| dt / h */
    /* README rule: the sorted slots of the active particles and their count
     * on the device; null: every slot */
    const int32_t* slots;
    const int32_t* nslots;
};

/* accelerations.metal:130-136: the step goal 2 CFL Gamma h / v_sig; with
 * SPH_FIX_INACTIVE_DRIFT also dlnh_limit h / |dh/dt|, so that no particle's
 * h or rho drifts by more than exp(dlnh_limit) (exp(3 dlnh_limit)) before it
 * is recomputed. */
__device__ __forceinline__ float step_goal(const AccelArgs& a, float h_i, float v_sigi, float dhdt) {
    float goaldt = 2.0f * a.cfl * a.gamma * h_i / v_sigi;
    if (a.dlnh_limit > 0.f) goaldt = fminf(goaldt, a.dlnh_limit * h_i / fabsf(dhdt));
    return goaldt;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

/* accelerations.metal:12-141 */
/* One particle (sorted slot `index`, -1 for none); returns its tick step
 * (INT_MAX for none or an inactive particle).  The pair term takes the
 * hardware square root and reciprocals (about 1 ulp, as under the
 * reference's Metal fast-math defaults) for r, 1/r, 1/h_j, 1/rho_ij and
 * 1/rho_j instead of correctly rounded sequences: the pass is bound by
 * VALU issue, and those sequences were a quarter of its pair term. */
__device__ __forceinline__ int accel_particle(const AccelArgs& a, int index) {
    int integerDt = 0x7FFFFFFF;
    if (index >= 0 && index < a.n) {
        const uint32_t ind = a.sorted[index].y;
        if (a.active[ind]) {
            const float4 xi4 = a.s_posh[index];
            const float4 vi4 = a.s_velb[index];
            const float4 hyd = a.s_hyd[index];
            const float h_i = xi4.w, fact_i = hyd.y, rho_i = hyd.z, c_i = hyd.w;
            const float B_i = vi4.w;
            const float hi1 = 1.f / h_i;
            const float hi2x4 = h_i * h_i * 4.f;
            float dvx = 0.f, dvy = 0.f, dvz = 0.f, du_dt = 0.f, dhdt = 0.f;
            float v_sigi = 1e-12f;
            const uint2* __restrict__ cse = a.cells.cell_se;
            const int mask = a.cells.mask;
            int mn[3], mx[3];
            cell_box(make_float3(xi4.x, xi4.y, xi4.z), a.cells.cell_size, fminf(h_i, a.max_h_len), mask + 1, mn, mx);
            for (int x = mn[0]; x <= mx[0]; x++) {
                const uint32_t ex = dev_expand_bits((uint32_t)(x & mask));
                for (int y = mn[1]; y <= mx[1]; y++) {
                    const uint32_t exy = ex | (dev_expand_bits((uint32_t)(y & mask)) << 1);
                    for (int z = mn[2]; z <= mx[2]; z++) {
                        const uint32_t key = exy | (dev_expand_bits((uint32_t)(z & mask)) << 2);
                        const uint2 se = cse[key];
                        if (se.x == 0xFFFFFFFFu) continue;
                        float4 nxt = a.s_posh[se.x];
                        for (uint32_t k = se.x; k <= se.y; k++) {
                            /* the next candidate's record is requested before this
                             * one is used, so its latency overlaps this step */
                            const float4 xj4 = nxt;
                            nxt = a.s_posh[min(k + 1u, se.y)];
                            const float dx = xi4.x - xj4.x, dy = xi4.y - xj4.y, dz = xi4.z - xj4.z;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            const float h_j = xj4.w;
                            const float hj2x4 = h_j * h_j * 4.f;
                            if ((r2 > hi2x4 && r2 > hj2x4) || r2 < 1e-12f) continue;
                            const float r = __builtin_amdgcn_sqrtf(r2);
                            const float r1 = __builtin_amdgcn_rcpf(r);
                            const float4 hj = a.s_hyd[k];
                            const float4 vj4 = a.s_velb[k];
                            const float m_j = hj.x;
                            const float fact_j = hj.y;
                            const float rho_j = hj.z;
                            const float c_j = hj.w;
                            const float hj1 = __builtin_amdgcn_rcpf(h_j);
                            const float vx = vi4.x - vj4.x, vy = vi4.y - vj4.y, vz = vi4.z - vj4.z;
                            const float rho_ij = 0.5f * (rho_i + rho_j);
                            const float v_dot_r = vx * dx + vy * dy + vz * dz;
                            const float mu_ij = v_dot_r < 0.f ? v_dot_r * r1 : 0.f;
                            const float v_sigij = c_i + c_j - 3.f * mu_ij;
                            v_sigi = fmaxf(v_sigij, v_sigi);
                            const float B_ij = 0.5f * (B_i + vj4.w);
                            const float nu_ij = -0.5f * a.alpha * B_ij * mu_ij * v_sigij * __builtin_amdgcn_rcpf(rho_ij);
                            const float rx = dx * r1, ry = dy * r1, rz = dz * r1;
                            const float gi = kGradWFactor_tp(r, hi1), gj = kGradWFactor_tp(r, hj1);
                            const float gix = gi * rx, giy = gi * ry, giz = gi * rz;
                            const float gjx = gj * rx, gjy = gj * ry, gjz = gj * rz;
                            const float gmx = 0.5f * (gix + gjx), gmy = 0.5f * (giy + gjy), gmz = 0.5f * (giz + gjz);
                            dvx -= m_j * (fact_i * gix + fact_j * gjx + nu_ij * gmx);
                            dvy -= m_j * (fact_i * giy + fact_j * gjy + nu_ij * gmy);
                            dvz -= m_j * (fact_i * giz + fact_j * gjz + nu_ij * gmz);
                            const float vgi = vx * gix + vy * giy + vz * giz;
                            const float vgm = vx * gmx + vy * gmy + vz * gmz;
                            du_dt += m_j * (fact_i * vgi + 0.5f * nu_ij * vgm);
                            dhdt += (m_j * __builtin_amdgcn_rcpf(rho_j)) * vgi;
                        }
                    }
                }
            }
            a.dv[ind] = make_float4(dvx, dvy, dvz, 0.f);
            a.du[ind] = du_dt;
            a.dhdt[ind] = dhdt;
            const float goaldt = step_goal(a, h_i, v_sigi, dhdt);
            int r = 0;
            float dtCandidate = a.dt0;
            while (dtCandidate > goaldt && r < a.r_max) {
                dtCandidate *= 0.5f;
                r++;
            }
            integerDt = (int)ceilf(a.dtmin * dtCandidate);
            a.next_active[ind] = a.time[0] + integerDt;
        }
    }
    return integerDt;
}

/* One thread per slot, workgroups XCD-remapped; with a slot list (README
 * rule) a persistent grid over the list, XCD x taking its x-th eighth.  The
 * global timestep is the minimum of the tick steps: wave minimum, then one
 * atomic per wave. */
/* 8 waves per SIMD (64 VGPRs, no spills) with the candidate prefetch */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_acceleration(AccelArgs a) {
    int integerDt = 0x7FFFFFFF;
    if (!a.slots) {
        const int blk = xcd_block((a.n + 255) / 256);
        if (blk < 0) return;
        integerDt = accel_particle(a, blk * 256 + (int)threadIdx.x);
    } else {
        const int total = *a.nslots;
        const int nlog = (total + 255) / 256;
        const int per = (nlog + kXcds - 1) / kXcds;
        const int xcd = (int)(blockIdx.x % kXcds), nper = (int)(gridDim.x / kXcds);
        for (int k = (int)(blockIdx.x / kXcds); k < per; k += nper) {
            const int lb = xcd * per + k;
            if (lb >= nlog) break;
            const int e = lb * 256 + (int)threadIdx.x;
            integerDt = min(integerDt, accel_particle(a, e < total ? a.slots[e] : -1));
        }
    }
    const int wmin = wave_min(integerDt);
    if ((threadIdx.x & 63) == 0 && wmin != 0x7FFFFFFF) atomicMin(&a.time[1], wmin);
}

/* accelerations.metal:12-141 with G lanes per particle, the box split as
 * in k_density_group (SPLIT 0: whole cells per lane; SPLIT 1: a cell's
 * candidates over the lanes); the group combines the partial sums with a
 * butterfly (sums) and a butterfly maximum (v_sig).  Same terms as the reference; only the
 * order of the additions differs. */
template <int G, int SPLIT>
__device__ __forceinline__ int accel_group_particle(const AccelArgs& a, int index, int lane) {
    int integerDt = 0x7FFFFFFF;
    const uint32_t ind = index >= 0 && index < a.n ? a.sorted[index].y : 0u;
    if (index >= 0 && index < a.n && a.active[ind]) {
        const float4 xi4 = a.s_posh[index];
        const float4 vi4 = a.s_velb[index];
        const float4 hyd = a.s_hyd[index];
        const float h_i = xi4.w, fact_i = hyd.y, rho_i = hyd.z, c_i = hyd.w;
        const float B_i = vi4.w;
        const float hi1 = 1.f / h_i;
        const float hi2x4 = h_i * h_i * 4.f;
        float dvx = 0.f, dvy = 0.f, dvz = 0.f, du_dt = 0.f, dhdt = 0.f;
        float v_sigi = 1e-12f;
        const uint2* __restrict__ cse = a.cells.cell_se;
        const int mask = a.cells.mask;
        int mn[3], mx[3];
        cell_box(make_float3(xi4.x, xi4.y, xi4.z), a.cells.cell_size, fminf(h_i, a.max_h_len), mask + 1, mn, mx);
        const int ny = max(mx[1] - mn[1] + 1, 0), nz = max(mx[2] - mn[2] + 1, 0);
        const int ncell = max(mx[0] - mn[0] + 1, 0) * ny * nz;
        for (int c = SPLIT ? 0 : lane; c < ncell; c += SPLIT ? 1 : G) {
            const int iz = c % nz, t = c / nz;
            const int iy = t % ny, ix = t / ny;
            const uint32_t key = dev_expand_bits((uint32_t)((mn[0] + ix) & mask)) |
                                 (dev_expand_bits((uint32_t)((mn[1] + iy) & mask)) << 1) |
                                 (dev_expand_bits((uint32_t)((mn[2] + iz) & mask)) << 2);
            const uint2 se = cse[key];
            if (se.x == 0xFFFFFFFFu) continue;
            constexpr uint32_t STEP = SPLIT ? G : 1;
            uint32_t k = se.x + (SPLIT ? lane : 0);
            if (k > se.y) continue;
            float4 nxt = a.s_posh[k];
            for (; k <= se.y; k += STEP) {
                /* the next candidate's record is requested before this one is used */
                const float4 xj4 = nxt;
                nxt = a.s_posh[min(k + STEP, se.y)];
                const float dx = xi4.x - xj4.x, dy = xi4.y - xj4.y, dz = xi4.z - xj4.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float h_j = xj4.w;
                const float hj2x4 = h_j * h_j * 4.f;
                if ((r2 > hi2x4 && r2 > hj2x4) || r2 < 1e-12f) continue;
                const float r = __builtin_amdgcn_sqrtf(r2);
                const float r1 = __builtin_amdgcn_rcpf(r);
                const float4 hj = a.s_hyd[k];
                const float4 vj4 = a.s_velb[k];
                const float m_j = hj.x;
                const float fact_j = hj.y;
                const float rho_j = hj.z;
                const float c_j = hj.w;
                const float hj1 = __builtin_amdgcn_rcpf(h_j);
                const float vx = vi4.x - vj4.x, vy = vi4.y - vj4.y, vz = vi4.z - vj4.z;
                const float rho_ij = 0.5f * (rho_i + rho_j);
                const float v_dot_r = vx * dx + vy * dy + vz * dz;
                const float mu_ij = v_dot_r < 0.f ? v_dot_r * r1 : 0.f;
                const float v_sigij = c_i + c_j - 3.f * mu_ij;
                v_sigi = fmaxf(v_sigij, v_sigi);
                const float B_ij = 0.5f * (B_i + vj4.w);
                const float nu_ij = -0.5f * a.alpha * B_ij * mu_ij * v_sigij * __builtin_amdgcn_rcpf(rho_ij);
                const float rx = dx * r1, ry = dy * r1, rz = dz * r1;
                const float gi = kGradWFactor_tp(r, hi1), gj = kGradWFactor_tp(r, hj1);
                const float gix = gi * rx, giy = gi * ry, giz = gi * rz;
                const float gjx = gj * rx, gjy = gj * ry, gjz = gj * rz;
                const float gmx = 0.5f * (gix + gjx), gmy = 0.5f * (giy + gjy), gmz = 0.5f * (giz + gjz);
                dvx -= m_j * (fact_i * gix + fact_j * gjx + nu_ij * gmx);
                dvy -= m_j * (fact_i * giy + fact_j * gjy + nu_ij * gmy);
                dvz -= m_j * (fact_i * giz + fact_j * gjz + nu_ij * gmz);
                const float vgi = vx * gix + vy * giy + vz * giz;
                const float vgm = vx * gmx + vy * gmy + vz * gmz;
                du_dt += m_j * (fact_i * vgi + 0.5f * nu_ij * vgm);
                dhdt += (m_j * __builtin_amdgcn_rcpf(rho_j)) * vgi;
            }
        }
        dvx = gsum<G>(dvx);
        dvy = gsum<G>(dvy);
        dvz = gsum<G>(dvz);
        du_dt = gsum<G>(du_dt);
        dhdt = gsum<G>(dhdt);
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) v_sigi = fmaxf(v_sigi, __shfl_xor(v_sigi, off, 64));
        if (lane == 0) {
            a.dv[ind] = make_float4(dvx, dvy, dvz, 0.f);
            a.du[ind] = du_dt;
            a.dhdt[ind] = dhdt;
            const float goaldt = step_goal(a, h_i, v_sigi, dhdt);
            int r = 0;
            float dtCandidate = a.dt0;
            while (dtCandidate > goaldt && r < a.r_max) {
                dtCandidate *= 0.5f;
                r++;
            }
            integerDt = (int)ceilf(a.dtmin * dtCandidate);
            a.next_active[ind] = a.time[0] + integerDt;
        }
    }
    return integerDt;
}

/* LIST false: every slot, workgroups XCD-remapped; LIST true: a persistent
 * grid over the active particles' slots, XCD x taking the x-th eighth of the
 * list.  Every group of a wave runs the same number of rounds, so the
 * butterflies see whole groups. */
template <int G, int SPLIT, bool LIST>
__global__ __launch_bounds__(256) void k_acceleration_group(AccelArgs a) {
    constexpr int PPB = 256 / G;
    const int lane = threadIdx.x % G, g = (int)threadIdx.x / G;
    int integerDt = 0x7FFFFFFF;
    if constexpr (!LIST) {
        const int blk = xcd_block((a.n + PPB - 1) / PPB);
        if (blk < 0) return;
        integerDt = accel_group_particle<G, SPLIT>(a, blk * PPB + g, lane);
    } else {
        const int total = *a.nslots;
        const int nlog = (total + PPB - 1) / PPB;
        const int per = (nlog + kXcds - 1) / kXcds;
        const int xcd = (int)(blockIdx.x % kXcds), nper = (int)(gridDim.x / kXcds);
        for (int k = (int)(blockIdx.x / kXcds); k < per; k += nper) {
            const int lb = xcd * per + k;
            if (lb >= nlog) break;
            const int e = lb * PPB + g;
            integerDt = min(integerDt, accel_group_particle<G, SPLIT>(a, e < total ? a.slots[e] : -1, lane));
        }
    }
    const int wmin = wave_min(integerDt);
    if ((threadIdx.x & 63) == 0 && wmin != 0x7FFFFFFF) atomicMin(&a.time[1], wmin);
}

struct StepArgs {
    int n;
    float4 *pos, *vel;
    float4* acc;
    const float4* dv; /* pair sums of this step's acceleration pass, or null */
    float *rho, *u, *h;
    const float *du, *dhdt;
    const uint8_t* active;
    int32_t* next_active;
    int32_t* time;
    float* dlnh;
    float dtmin1, max_h_len;
    int max_rungs;
    int fix_drift;
};

/* step.metal:12-46 */
__global__ __launch_bounds__(256) void k_step(StepArgs a) {
    const int index = blockIdx.x * blockDim.x + threadIdx.x;
    const int dtInt = a.time[1];
    if (index >= a.n) return;
    const float dt = a.dtmin1 * (float)dtInt;
    const float4 velocity = a.vel[index];
    float4 acceleration = a.acc[index];
    if (a.dv && a.active[index]) {
        /* accelerations.metal:127: acc += dv, for the particles the pass ran on */
        const float4 d = a.dv[index];
        acceleration.x += d.x;
        acceleration.y += d.y;
        acceleration.z += d.z;
        a.acc[index] = acceleration;
    }
    const float h_i = a.h[index];
    const float h1 = 1.f / h_i;
    const float dhdt = a.dhdt[index];
    float rate = h1 * dhdt;
    if (a.fix_drift) {
        /* inactive particles keep the log-rate of their last active step */
        if (a.active[index]) a.dlnh[index] = rate;
        else rate = a.dlnh[index];
    }
    a.h[index] = fminf(h_i * expf(rate * dt), a.max_h_len);
    a.rho[index] *= expf(-3.f * rate * dt);
    a.vel[index] = make_float4(velocity.x + dt * acceleration.x, velocity.y + dt * acceleration.y,
                               velocity.z + dt * acceleration.z, velocity.w);
    a.u[index] += dt * a.du[index];
    const float4 p = a.pos[index];
    const float hx = 0.5f * acceleration.x * dt * dt, hy = 0.5f * acceleration.y * dt * dt,
                hz = 0.5f * acceleration.z * dt * dt;
    a.pos[index] = make_float4(p.x + (dt * velocity.x + hx), p.y + (dt * velocity.y + hy),
                               p.z + (dt * velocity.z + hz), p.w);
    if (a.max_rungs > 0) {
        const int cap = dtInt << (a.max_rungs - 1);
        const int gt = a.time[0];
        if (a.active[index] && a.next_active[index] - gt > cap) a.next_active[index] = gt + cap;
    } else if (index == 0) {
        a.time[0] += dtInt;
    }
}

__global__ void k_advance_time(int32_t* time) { time[0] += time[1]; }

/* acc += dv for the particles the last acceleration pass ran on */
__global__ __launch_bounds__(256) void k_fold_dv(int n, const uint8_t* __restrict__ active,
                                                 const float4* __restrict__ dv, float4* __restrict__ acc) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !active[i]) return;
    float4 ac = acc[i];
    const float4 d = dv[i];
    ac.x += d.x;
    ac.y += d.y;
    ac.z += d.z;
    acc[i] = ac;
}

Cells make_cells(const sph_ctx* ctx) {
    Cells c;
    c.cell_se = ctx->cell_se;
    c.cell_size = ctx->prm.cell_width;
    c.mask = (1 << ctx->prm.cell_power) - 1;
    return c;
}

}  // namespace

/* The sorted slots of the active particles, rebuilt only after a sort or a
 * gravity pass (which set the order and the activity). */
int launch_active_list(sph_ctx* ctx) {
    if (ctx->slot_list_valid) return SPH_OK;
    if (!ctx->slot_list) SPH_HIP_TRY(hipMalloc((void**)&ctx->slot_list, sizeof(int32_t) * (size_t)ctx->n));
    if (!ctx->slot_count) SPH_HIP_TRY(hipMalloc((void**)&ctx->slot_count, sizeof(int32_t)));
    k_zero_i32<<<1, 1, 0, ctx->stream>>>(ctx->slot_count);
    k_active_slots<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, ctx->pairs_a, ctx->active, ctx->slot_list,
                                                                  ctx->slot_count);
    SPH_HIP_TRY(hipGetLastError());
    ctx->slot_list_valid = true;
    return SPH_OK;
}

int refresh_morton(sph_ctx* ctx) {
    if (ctx->morton_fresh) return SPH_OK;
    k_refresh_morton<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, ctx->pairs_a, ctx->pos, ctx->vel, ctx->m,
                                                                     ctx->s_posm, ctx->s_vel);
    SPH_HIP_TRY(hipGetLastError());
    ctx->morton_fresh = true;
    return SPH_OK;
}

int launch_density(sph_ctx* ctx) {
    const sph_params& p = ctx->prm;
    DensityArgs a;
    a.n = ctx->n;
    a.sorted = ctx->pairs_a;
    a.cells = make_cells(ctx);
    a.s_posm = ctx->s_posm;
    a.s_vel = ctx->s_vel;
    a.active = ctx->active;
    a.mat = ctx->mat;
    a.u = ctx->u;
    a.rho = ctx->rho;
    a.h = ctx->h;
    a.P = ctx->P;
    a.c = ctx->c;
    a.f = ctx->f;
    a.B = ctx->B;
    a.nr_iters = ctx->nr_iters;
    a.work = ctx->count_work ? ctx->density_work : nullptr;
    a.time = ctx->time;
    a.eos0 = ctx->eos[0];
    a.eos1 = ctx->eos[1];
    a.eos_res = ctx->eos_res;
    a.min_u = p.eos_min_u;
    a.max_u = p.eos_max_u;
    a.min_rho = p.eos_min_rho;
    a.max_rho = p.eos_max_rho;
    a.eta = p.density_eta;
    a.res_eta = p.resolution_eta;
    a.max_h_len = p.max_smoothing_length;
    a.max_it = p.max_nr_iterations;
    a.ne = p.n_neighbours_estim;
    a.fix_cache = (p.fixes & SPH_FIX_DENSITY_CACHE) ? 1 : 0;
    const float dtmin1 = p.dt0 / (float)(1 << p.r_max);
    a.dt_reset = (int)(p.dt0 * (1.f / dtmin1));
    a.guess = ctx->learned ? ctx->guess : nullptr;
    /* With individual timesteps most particles are inactive: their EOS is
     * cheaper in original order than through the sorted slots.  (A compacted
     * list of the active slots would leave too few waves to hide the
     * neighbour loops' latency.) */
    a.skip_inactive = p.activity_rule == SPH_ACTIVE_README ? 1 : 0;
    a.stop_rule = p.density_stop;
    a.slots = nullptr;
    a.nslots = nullptr;
    a.order = nullptr;
    if (!a.skip_inactive && p.density_order > 0) {
#define SPH_NR_ORDER(W)                                                                                   \
    case W:                                                                                               \
        k_nr_order<W><<<(ctx->n + (W) - 1) / (W), W, 0, ctx->stream>>>(ctx->n, ctx->pairs_a, ctx->nr_iters, \
                                                                      ctx->nr_order);                     \
        break
        switch (p.density_order) {
            SPH_NR_ORDER(16);
            SPH_NR_ORDER(32);
            SPH_NR_ORDER(64);
            SPH_NR_ORDER(128);
            SPH_NR_ORDER(256);
            SPH_NR_ORDER(512);
            default: return SPH_E_ARG;
        }
#undef SPH_NR_ORDER
        SPH_HIP_TRY(hipGetLastError());
        a.order = ctx->nr_order;
    }
    if (a.skip_inactive) {
        k_inactive_eos<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(a);
        SPH_HIP_TRY(hipGetLastError());
        /* the grouped passes walk the active particles' slots only */
        SPH_TRY_RC(launch_active_list(ctx));
        a.slots = ctx->slot_list;
        a.nslots = ctx->slot_count;
    }
    const int threads = 256;
    const int flat = p.density_flat ? 1 : 0;
    const int dlanes = a.slots ? p.list_density_lanes : (ctx->n < p.small_n_lanes ? 32 : p.density_lanes);
    const int dsel = (dlanes * 2 + (p.lane_split ? 1 : 0)) * 2 + flat;
    /* over a slot list the grid is persistent: at most kListBlocks workgroups */
#define SPH_DENS_C(G, S, F, PPB, C)                                                                      \
    if (a.slots)                                                                                         \
        k_density_group<G, S, F, true, C>                                                                \
            <<<std::min(xcd_grid((ctx->n + (PPB) - 1) / (PPB)), (unsigned)kListBlocks), threads, 0,     \
               ctx->stream>>>(a);                                                                        \
    else                                                                                                 \
        k_density_group<G, S, F, false, C><<<xcd_grid((ctx->n + (PPB) - 1) / (PPB)), threads, 0, ctx->stream>>>(a)
    /* the work-counting form only while sph_set_work_counters is on */
#define SPH_DENS(G, S, F, PPB)                                                                           \
    if (a.work) {                                                                                        \
        SPH_DENS_C(G, S, F, PPB, true);                                                                  \
    } else {                                                                                             \
        SPH_DENS_C(G, S, F, PPB, false);                                                                 \
    }
    switch (dsel) {
        case 8: SPH_DENS(2, 0, 0, 128); break;
        case 9: SPH_DENS(2, 0, 1, 128); break;
        case 16: SPH_DENS(4, 0, 0, 64); break;
        case 17: SPH_DENS(4, 0, 1, 64); break;
        case 32: SPH_DENS(8, 0, 0, 32); break;
        case 33: SPH_DENS(8, 0, 1, 32); break;
        case 34: case 35: SPH_DENS(8, 1, 0, 32); break;
        case 64: SPH_DENS(16, 0, 0, 16); break;
        case 65: SPH_DENS(16, 0, 1, 16); break;
        case 66: case 67: SPH_DENS(16, 1, 0, 16); break;
        case 128: SPH_DENS(32, 0, 0, 8); break;
        case 129: SPH_DENS(32, 0, 1, 8); break;
        case 130: case 131: SPH_DENS(32, 1, 0, 8); break;
        case 256: SPH_DENS(64, 0, 0, 4); break;
        case 257: SPH_DENS(64, 0, 1, 4); break;
        case 258: case 259: SPH_DENS(64, 1, 0, 4); break;
        default: k_density<<<xcd_grid((ctx->n + threads - 1) / threads), threads, 0, ctx->stream>>>(a); break;
    }
#undef SPH_DENS
#undef SPH_DENS_C
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

int launch_acceleration(sph_ctx* ctx) {
    /* a second pass before the step: the first one's sums go in first */
    SPH_TRY_RC(fold_dv(ctx));
    const sph_params& p = ctx->prm;
    AccelArgs a;
    a.n = ctx->n;
    a.sorted = ctx->pairs_a;
    a.cells = make_cells(ctx);
    a.s_posh = ctx->s_posh;
    a.s_velb = ctx->s_velb;
    a.s_hyd = ctx->s_hyd;
    a.active = ctx->active;
    a.dv = ctx->dv;
    a.du = ctx->du;
    a.dhdt = ctx->dhdt;
    a.next_active = ctx->next_active;
    a.time = ctx->time;
    a.alpha = p.viscosity_alpha;
    a.cfl = p.cfl;
    a.gamma = p.gamma;
    a.dlnh_limit = (p.fixes & SPH_FIX_INACTIVE_DRIFT) ? SPH_DLNH_LIMIT : 0.f;
    a.dt0 = p.dt0;
    const float dtmin1 = p.dt0 / (float)(1 << p.r_max);
    a.dtmin = 1.f / dtmin1;
    a.max_h_len = p.max_smoothing_length;
    a.r_max = p.r_max;
    a.slots = nullptr;
    a.nslots = nullptr;
    if (p.activity_rule == SPH_ACTIVE_README) {
        /* the thread-per-particle pass walks the active particles' slots only */
        SPH_TRY_RC(launch_active_list(ctx));
        a.slots = ctx->slot_list;
        a.nslots = ctx->slot_count;
    }
    const int threads = 256;
    /* with the active particles only, the records are gathered for the
     * cells their boxes reach */
    const uint8_t* mark = nullptr;
    uint8_t epoch = 0;
    if (a.slots && p.gather_reach) {
        ctx->mark_epoch = ctx->mark_epoch % 255 + 1;
        epoch = (uint8_t)ctx->mark_epoch;
        k_mark_cells<16><<<std::min(xcd_grid((ctx->n + 15) / 16), (unsigned)kListBlocks), threads, 0, ctx->stream>>>(
            a.slots, a.nslots, ctx->pairs_a, ctx->morton_fresh ? ctx->s_posm : nullptr, ctx->pos, ctx->h,
            a.cells.cell_size, a.cells.mask, p.max_smoothing_length, epoch, ctx->cell_mark);
        SPH_HIP_TRY(hipGetLastError());
        mark = ctx->cell_mark;
    }
    if (ctx->morton_fresh)
        k_gather_hydro<true><<<(ctx->n + threads - 1) / threads, threads, 0, ctx->stream>>>(
            ctx->n, ctx->pairs_a, ctx->pos, ctx->vel, ctx->m, ctx->s_posm, ctx->s_vel, ctx->h, ctx->f, ctx->P,
            ctx->rho, ctx->c, ctx->B, ctx->s_posh, ctx->s_velb, ctx->s_hyd, mark, epoch);
    else
        k_gather_hydro<false><<<(ctx->n + threads - 1) / threads, threads, 0, ctx->stream>>>(
            ctx->n, ctx->pairs_a, ctx->pos, ctx->vel, ctx->m, ctx->s_posm, ctx->s_vel, ctx->h, ctx->f, ctx->P,
            ctx->rho, ctx->c, ctx->B, ctx->s_posh, ctx->s_velb, ctx->s_hyd, mark, epoch);
    SPH_HIP_TRY(hipGetLastError());
    const int alanes = a.slots ? p.list_accel_lanes
                               : (ctx->n < p.small_n_lanes ? (ctx->n < 65536 ? 16 : 8) : p.accel_lanes);
    const int asel = alanes * 2 + (p.lane_split ? 1 : 0);
    /* over a slot list the grid is persistent: at most kListBlocks workgroups */
#define SPH_ACC(G, S)                                                                                    \
    if (a.slots)                                                                                         \
        k_acceleration_group<G, S, true>                                                                 \
            <<<std::min(xcd_grid((ctx->n + (256 / (G)) - 1) / (256 / (G))), (unsigned)kListBlocks), threads, 0, \
               ctx->stream>>>(a);                                                                        \
    else                                                                                                 \
        k_acceleration_group<G, S, false>                                                                \
            <<<xcd_grid((ctx->n + (256 / (G)) - 1) / (256 / (G))), threads, 0, ctx->stream>>>(a)
    switch (asel) {
        case 4: SPH_ACC(2, 0); break;
        case 8: SPH_ACC(4, 0); break;
        case 16: SPH_ACC(8, 0); break;
        case 17: SPH_ACC(8, 1); break;
        case 32: SPH_ACC(16, 0); break;
        case 33: SPH_ACC(16, 1); break;
        case 64: SPH_ACC(32, 0); break;
        case 65: SPH_ACC(32, 1); break;
        case 128: SPH_ACC(64, 0); break;
        case 129: SPH_ACC(64, 1); break;
        default: {
            const unsigned g = xcd_grid((ctx->n + threads - 1) / threads);
            k_acceleration<<<a.slots ? std::min(g, (unsigned)kListBlocks) : g, threads, 0, ctx->stream>>>(a);
            break;
        }
    }
#undef SPH_ACC
    SPH_HIP_TRY(hipGetLastError());
    ctx->dv_pending = true;
    return SPH_OK;
}

int fold_dv(sph_ctx* ctx) {
    if (!ctx->dv_pending) return SPH_OK;
    SPH_TRY_RC(join_gravity(ctx));
    k_fold_dv<<<(ctx->n + 255) / 256, 256, 0, ctx->stream>>>(ctx->n, ctx->active, ctx->dv, ctx->acc);
    SPH_HIP_TRY(hipGetLastError());
    ctx->dv_pending = false;
    return SPH_OK;
}

int launch_step(sph_ctx* ctx) {
    const sph_params& p = ctx->prm;
    StepArgs a;
    a.n = ctx->n;
    a.pos = ctx->pos;
    a.vel = ctx->vel;
    a.acc = ctx->acc;
    a.dv = ctx->dv_pending ? ctx->dv : nullptr;
    a.rho = ctx->rho;
    a.u = ctx->u;
    a.h = ctx->h;
    a.du = ctx->du;
    a.dhdt = ctx->dhdt;
    a.active = ctx->active;
    a.next_active = ctx->next_active;
    a.time = ctx->time;
    a.dtmin1 = p.dt0 / (float)(1 << p.r_max);
    a.max_h_len = p.max_smoothing_length;
    a.max_rungs = p.max_rungs;
    a.dlnh = ctx->dlnh;
    a.fix_drift = (p.fixes & SPH_FIX_INACTIVE_DRIFT) ? 1 : 0;
    const int threads = 256;
    k_step<<<(ctx->n + threads - 1) / threads, threads, 0, ctx->stream>>>(a);
    SPH_HIP_TRY(hipGetLastError());
    ctx->dv_pending = false;
    ctx->morton_fresh = false; /* positions and velocities moved */
    if (p.max_rungs > 0) {
        k_advance_time<<<1, 1, 0, ctx->stream>>>(ctx->time);
        SPH_HIP_TRY(hipGetLastError());
    }
    return SPH_OK;
}
