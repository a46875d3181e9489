/*
 * EOSTable.cpp -- synthetic Tillotson (P, c) tables on the reference's
 * log-spaced (rho, u) axes.
 *
 * The reference resamples ANEOS text tables into a res x res float2 texture
 * (ANEOS.cpp:21-175) whose row index is the density sample and column index
 * the internal-energy sample, sample i at min * (max/min)^(i/res)
 * (ANEOS.cpp:123-136), values in code units (P * 1e-18, c * 1e-6,
 * ANEOS.cpp:108).  The ANEOS files are not part of the reference checkout, so
 * the tables here are evaluated from the Tillotson form (Tillotson 1962;
 * Melosh 1989, appendix II) on the same axes and with the same layout.
 */
#include <cmath>
#include <cstdint>

#include "sph_host.h"

namespace {

/* Tillotson pressure in SI units (Pa) for density rho (kg/m^3) and specific
 * internal energy u (J/kg).  Compressed or cold states use the condensed
 * branch, hot expanded states the vapour branch, and the band between the
 * incipient and complete vaporisation energies blends the two linearly in u.
 * Tension is floored at zero. */
double tillotson_pressure(const sph_tillotson& t, double rho, double u) {
    const double eta = rho / t.rho0;
    const double mu = eta - 1.0;
    const double w = u / (t.E0 * eta * eta) + 1.0;
    const double Pc = (t.a + t.b / w) * rho * u + t.A * mu + t.B * mu * mu;
    double P;
    if (rho >= t.rho0 || u <= t.Eiv) {
        P = Pc;
    } else {
        const double z = 1.0 / eta - 1.0;
        const double Pe = t.a * rho * u + (t.b * rho * u / w + t.A * mu * std::exp(-t.beta * z)) * std::exp(-t.alpha * z * z);
        if (u >= t.Ecv) P = Pe;
        else P = ((u - t.Eiv) * Pe + (t.Ecv - u) * Pc) / (t.Ecv - t.Eiv);
    }
    return P;
}

/* c^2 = dP/drho|_u + (P / rho^2) dP/du|_rho by central differences of the
 * unfloored pressure (so a cold, slightly expanded state keeps its bulk sound
 * speed), floored at (1% of the bulk sound speed sqrt(A/rho0))^2. */
double tillotson_sound_speed(const sph_tillotson& t, double rho, double u) {
    const double dr = rho * 1e-6, du = u * 1e-6;
    const double P = tillotson_pressure(t, rho, u);
    const double dPdr = (tillotson_pressure(t, rho + dr, u) - tillotson_pressure(t, rho - dr, u)) / (2 * dr);
    const double dPdu = (tillotson_pressure(t, rho, u + du) - tillotson_pressure(t, rho, u - du)) / (2 * du);
    double c2 = dPdr + P / (rho * rho) * dPdu;
    const double cmin = 0.01 * std::sqrt(t.A / t.rho0);
    if (!(c2 > cmin * cmin)) c2 = cmin * cmin;
    return std::sqrt(c2);
}

}  // namespace

extern "C" {

void sph_tillotson_granite(sph_tillotson* o) {
    *o = sph_tillotson{2680.0, 1.8e10, 1.8e10, 1.6e7, 3.5e6, 1.8e7, 0.5, 1.3, 5.0, 5.0};
}

void sph_tillotson_iron(sph_tillotson* o) {
    *o = sph_tillotson{7800.0, 1.28e11, 1.05e11, 9.5e6, 2.4e6, 8.67e6, 0.5, 1.5, 5.0, 5.0};
}

int sph_make_eos_table(const sph_params* prm, const sph_tillotson* t, int res, uint64_t seed, float* pc_out) {
    (void)seed;
    if (!prm || !t || !pc_out || res < 2) return SPH_E_ARG;
    /* Axis limits in SI, as ANEOS.cpp:125-128 (code density unit 1e6 kg/m^3,
     * code energy unit 1e12 J/kg). */
    const double minRho = (double)prm->eos_min_rho * 1e6, maxRho = (double)prm->eos_max_rho * 1e6;
    const double minU = (double)prm->eos_min_u * 1e12, maxU = (double)prm->eos_max_u * 1e12;
    for (int i = 0; i < res; ++i) {
        const double rho = minRho * std::pow(maxRho / minRho, (double)i / res);
        for (int j = 0; j < res; ++j) {
            const double u = minU * std::pow(maxU / minU, (double)j / res);
            const double P = tillotson_pressure(*t, rho, u);
            const double c = tillotson_sound_speed(*t, rho, u);
            pc_out[2 * ((int64_t)i * res + j)] = (float)((P > 0.0 ? P : 0.0) * 1e-18);
            pc_out[2 * ((int64_t)i * res + j) + 1] = (float)(c * 1e-6);
        }
    }
    return SPH_OK;
}

}  // extern "C"
