/*
 * InitialConditions.cpp -- seeded two-body impact in the reference's
 * DataStruct layout (hdfHandler.hpp:15-26).
 *
 * The reference reads its initial conditions from an HDF5 file
 * (hdfHandler.cpp:12-136) that is not part of the checkout; the impacts here
 * are generated instead.  Each body is a uniform-density iron core (material
 * 402) inside a uniform-density rock mantle (material 400), sampled uniformly
 * in volume with equal-mass particles.  The generator uses std::mt19937_64
 * (whose output sequence the C++ standard fixes) with its own uniform and
 * Box-Muller transforms, so every machine produces the same particles.
 */
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "sph_host.h"

namespace {

struct Rng {
    std::mt19937_64 g;
    bool have_spare = false;
    double spare = 0;
    explicit Rng(uint64_t seed) : g(seed) {}
    double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); } /* [0, 1) */
    double normal() {
        if (have_spare) {
            have_spare = false;
            return spare;
        }
        const double u1 = 1.0 - uniform(); /* (0, 1] */
        const double u2 = uniform();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const double ang = 2.0 * 3.14159265358979323846 * u2;
        spare = rad * std::sin(ang);
        have_spare = true;
        return rad * std::cos(ang);
    }
};

/* Bilinear lookup with the texture convention of the density pass
 * (texel centres at (i + 1/2) / res, clamp to edge). */
float table_pressure(const sph_params* prm, const float* pc, int res, float u, float rho) {
    float uc = std::log(u / prm->eos_min_u) / std::log(prm->eos_max_u / prm->eos_min_u);
    float rc = std::log(rho / prm->eos_min_rho) / std::log(prm->eos_max_rho / prm->eos_min_rho);
    uc = std::fmin(std::fmax(uc, 0.f), 1.f);
    rc = std::fmin(std::fmax(rc, 0.f), 1.f);
    const float tx = uc * res - 0.5f, ty = rc * res - 0.5f;
    const float fx = std::floor(tx), fy = std::floor(ty);
    const float a = tx - fx, b = ty - fy;
    auto cl = [res](int i) { return i < 0 ? 0 : (i > res - 1 ? res - 1 : i); };
    const int i0 = cl((int)fx), i1 = cl((int)fx + 1), j0 = cl((int)fy), j1 = cl((int)fy + 1);
    auto T = [&](int j, int i) { return pc[2 * ((int64_t)j * res + i)]; };
    return (1 - a) * (1 - b) * T(j0, i0) + a * (1 - b) * T(j0, i1) + (1 - a) * b * T(j1, i0) + a * b * T(j1, i1);
}

struct Body {
    int n;
    double R, M;
    double cx, cy, cz, vx, vy, vz;
};

void make_body(const sph_params* prm, const sph_impact_config* cfg, const Body& b, int offset, Rng& rng,
               const float* pc400, const float* pc402, int res, float* pos4, float* vel4, float* rho, float* u,
               float* m, int32_t* mat, float* p, float* h) {
    const double pi = 3.14159265358979323846;
    const int n_core = (int)std::llround(cfg->core_mass_fraction * b.n);
    const double mass = b.M / b.n;
    const double rc = cfg->core_radius_fraction * b.R;
    const double v_core = 4.0 / 3.0 * pi * rc * rc * rc;
    const double v_mantle = 4.0 / 3.0 * pi * (b.R * b.R * b.R - rc * rc * rc);
    const double rho_core = n_core > 0 ? n_core * mass / v_core : 0.0;
    const double rho_mantle = (b.n - n_core) * mass / v_mantle;
    for (int k = 0; k < b.n; ++k) {
        const bool core = k < n_core;
        const double r0 = core ? 0.0 : rc, r1 = core ? rc : b.R;
        const double U = rng.uniform();
        const double r = std::cbrt(r0 * r0 * r0 + U * (r1 * r1 * r1 - r0 * r0 * r0));
        double gx, gy, gz, gn;
        do {
            gx = rng.normal();
            gy = rng.normal();
            gz = rng.normal();
            gn = std::sqrt(gx * gx + gy * gy + gz * gz);
        } while (gn == 0.0);
        const int i = offset + k;
        pos4[4 * i] = (float)(b.cx + r * gx / gn);
        pos4[4 * i + 1] = (float)(b.cy + r * gy / gn);
        pos4[4 * i + 2] = (float)(b.cz + r * gz / gn);
        pos4[4 * i + 3] = 0.f;
        vel4[4 * i] = (float)b.vx;
        vel4[4 * i + 1] = (float)b.vy;
        vel4[4 * i + 2] = (float)b.vz;
        vel4[4 * i + 3] = 0.f;
        const double rho0 = core ? rho_core : rho_mantle;
        rho[i] = (float)rho0;
        u[i] = cfg->u0;
        m[i] = (float)mass;
        mat[i] = core ? 402 : 400;
        h[i] = (float)(1.2348 * std::cbrt(mass / rho0));
        p[i] = table_pressure(prm, core ? pc402 : pc400, res, u[i], rho[i]);
    }
}

}  // namespace

extern "C" {

void sph_impact_config_default(sph_impact_config* o, int32_t n_total, uint64_t seed) {
    o->n_total = n_total;
    o->target_fraction = 0.89f;
    o->target_radius = 6.4f;
    o->target_mass = 5.3f;
    o->impactor_radius = 3.6f;
    o->impactor_mass = 0.8f;
    o->core_mass_fraction = 0.3f;
    o->core_radius_fraction = 0.55f;
    o->u0 = 2e-6f;
    o->separation_factor = 1.5f;
    o->impact_parameter = 0.70710678f;
    o->closing_speed = 0.009f;
    o->seed = seed;
}

void sph_impact_config_headon(sph_impact_config* o, int32_t n_total, uint64_t seed) {
    sph_impact_config_default(o, n_total, seed);
    o->impact_parameter = 0.3f;
    o->closing_speed = 0.012f;
    o->u0 = 5e-6f;
}

int sph_make_impact(const sph_params* prm, const sph_impact_config* cfg, const float* pc400, const float* pc402,
                    int res, float* pos4, float* vel4, float* rho, float* u, float* m, int32_t* mat, float* p,
                    float* h) {
    if (!prm || !cfg || !pc400 || !pc402 || cfg->n_total < 2) return SPH_E_ARG;
    const int n_t = (int)std::llround((double)cfg->target_fraction * cfg->n_total);
    const int n_i = cfg->n_total - n_t;
    if (n_t < 1 || n_i < 1) return SPH_E_ARG;
    const double Rs = (double)cfg->target_radius + cfg->impactor_radius;
    const double d = cfg->separation_factor * Rs;
    const double bimp = cfg->impact_parameter * Rs;
    const double dx = -std::sqrt(d * d - bimp * bimp);
    /* target at rest at the origin, impactor at (dx, b, 0) moving +x; the
     * system is shifted to its centre-of-mass frame below */
    Body target{n_t, cfg->target_radius, cfg->target_mass, 0, 0, 0, 0, 0, 0};
    Body imp{n_i, cfg->impactor_radius, cfg->impactor_mass, dx, bimp, 0, cfg->closing_speed, 0, 0};
    Rng rng(cfg->seed);
    make_body(prm, cfg, target, 0, rng, pc400, pc402, res, pos4, vel4, rho, u, m, mat, p, h);
    make_body(prm, cfg, imp, n_t, rng, pc400, pc402, res, pos4, vel4, rho, u, m, mat, p, h);
    const double Mt = cfg->target_mass, Mi = cfg->impactor_mass;
    const double com[3] = {Mi * imp.cx / (Mt + Mi), Mi * imp.cy / (Mt + Mi), 0.0};
    const double vcom[3] = {Mi * imp.vx / (Mt + Mi), 0.0, 0.0};
    for (int i = 0; i < cfg->n_total; ++i)
        for (int k = 0; k < 3; ++k) {
            pos4[4 * i + k] = (float)((double)pos4[4 * i + k] - com[k]);
            vel4[4 * i + k] = (float)((double)vel4[4 * i + k] - vcom[k]);
        }
    return SPH_OK;
}

}  // extern "C"
