/*
 * ANEOS.cpp -- loadANEOSDataFromFile (ANEOS.cpp:21-175 of the reference):
 * reads an ANEOS text table of (u, P, c) on a (rho, T) grid and resamples it
 * into the (P, c) table the density pass samples, on log-spaced (rho, u)
 * axes (sample i at min * (max/min)^(i/res), FP32).
 *
 * File layout the reference expects: 13 header lines; "numRho numT"; numRho
 * densities (kg/m^3), then numT temperatures, each list over as many lines as
 * it takes (values after the last one on a line are dropped); then
 * numRho * numT lines "u P c ..." (J/kg, Pa, m/s) with the density index
 * running fastest.  P and c are scaled to code units (1e-18, 1e-6) on read.
 *
 * Resampling, as written (Q13):
 *   - the bracketing indexes advance monotonically along each axis and stop
 *     at the second-to-last interval (values beyond the table extrapolate);
 *   - the interpolation weights are inverted: the lower node gets the
 *     fraction (x - x_lo) / (x_hi - x_lo) that belongs to the upper node;
 *   - the search along u in the upper density row tests the lower row's index
 *     against its bound, so it can run past the row; the reference then reads
 *     the next row of its contiguous array (or past the array's end, which is
 *     undefined there and clamped to the last element here).
 * The fixed form weights the nodes correctly and bounds both searches.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <exception>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "Compute.hpp"

namespace {

struct RawTable {
    int numRho = 0, numT = 0;
    std::vector<float> rho, T;
    std::vector<float> u;        /* [r * numT + t] */
    std::vector<float> p, c;     /* code units */
};

bool read_list(std::ifstream& in, int count, std::vector<float>& out) {
    out.assign(count, 0.f);
    int n = 0;
    std::string line;
    while (n < count && std::getline(in, line)) {
        std::istringstream iss(line);
        float v;
        while (n < count && iss >> v) out[n++] = v;
    }
    return n == count;
}

bool parse(const std::string& path, RawTable& t, std::string& err) {
    std::ifstream in(path);
    if (!in.is_open()) {
        err = path + ": cannot open the table file";
        return false;
    }
    std::string line;
    for (int i = 0; i < 13; ++i)
        if (!std::getline(in, line)) {
            err = path + ": truncated header";
            return false;
        }
    if (!std::getline(in, line)) {
        err = path + ": no table size line";
        return false;
    }
    {
        std::istringstream iss(line);
        if (!(iss >> t.numRho >> t.numT) || t.numRho < 2 || t.numT < 2 ||
            (int64_t)t.numRho * t.numT > (int64_t)1 << 26) {
            err = path + ": bad table size line '" + line + "'";
            return false;
        }
    }
    if (!read_list(in, t.numRho, t.rho)) {
        err = path + ": not enough density values";
        return false;
    }
    if (!read_list(in, t.numT, t.T)) {
        err = path + ": not enough temperature values";
        return false;
    }
    const size_t N = (size_t)t.numRho * t.numT;
    t.u.assign(N, 0.f);
    t.p.assign(N, 0.f);
    t.c.assign(N, 0.f);
    const float pscale = std::pow(10.f, -18.f), cscale = std::pow(10.f, -6.f);
    int ri = 0, ti = 0;
    while (ti < t.numT && std::getline(in, line)) {
        std::istringstream iss(line);
        float u, p, c;
        if (!(iss >> u >> p >> c)) {
            err = path + ": malformed data line '" + line + "'";
            return false;
        }
        const size_t k = (size_t)ri * t.numT + ti;
        t.u[k] = u;
        t.p[k] = p * pscale;
        t.c[k] = c * cscale;
        if (++ri == t.numRho) {
            ri = 0;
            ++ti;
        }
    }
    if (ti < t.numT) {
        err = path + ": the table ends after " + std::to_string(ti) + " of " + std::to_string(t.numT) +
              " temperature rows";
        return false;
    }
    return true;
}

void resample(const RawTable& t, const sph_params& prm, int res, bool fixed, float* pc) {
    const int nR = t.numRho, nT = t.numT;
    const int64_t N = (int64_t)nR * nT;
    /* flat reads of the row-major arrays, as the reference's 2-D arrays are
     * read past a row's end; past the array's end clamped */
    auto flat = [&](const std::vector<float>& a, int r, int c) {
        int64_t k = (int64_t)r * nT + c;
        if (k < 0) k = 0;
        if (k >= N) k = N - 1;
        return a[(size_t)k];
    };
    auto U = [&](int r, int c) { return flat(t.u, r, c); };
    const float minRho = (float)(prm.eos_min_rho * 1e+6), maxRho = (float)(prm.eos_max_rho * 1e+6);
    const float minU = (float)(prm.eos_min_u * 1e+12), maxU = (float)(prm.eos_max_u * 1e+12);
    std::vector<float> rhoAxis(res), uAxis(res);
    for (int i = 0; i < res; ++i) {
        rhoAxis[i] = minRho * std::pow(maxRho / minRho, (float)i / res);
        uAxis[i] = minU * std::pow(maxU / minU, (float)i / res);
    }
    int rT = 0;
    for (int ri = 0; ri < res; ++ri) {
        const float rho = rhoAxis[ri];
        while (rho > t.rho[rT + 1] && rT < nR - 2) ++rT;
        /* fr: position of rho inside its density bracket [rT, rT + 1] */
        const float fr = (rho - t.rho[rT]) / (t.rho[rT + 1] - t.rho[rT]);
        int lo = 0, hi = 0; /* energy brackets in density rows rT and rT + 1 */
        for (int ui = 0; ui < res; ++ui) {
            const float u = uAxis[ui];
            while (u > U(rT, lo + 1) && lo < nT - 2) ++lo;
            if (fixed) {
                while (u > U(rT + 1, hi + 1) && hi < nT - 2) ++hi;
            } else {
                /* as written the upper row's search is bounded by the lower
                 * row's index (it may run into the next row) */
                while (u > U(rT + 1, hi + 1) && lo < nT - 2 && (int64_t)(rT + 1) * nT + hi + 1 < N - 1) ++hi;
            }
            /* f0, f1: position of u inside the bracket of row rT, row rT + 1 */
            const float f0 = (u - U(rT, lo)) / (U(rT, lo + 1) - U(rT, lo));
            const float f1 = (u - U(rT + 1, hi)) / (U(rT + 1, hi + 1) - U(rT + 1, hi));
            for (int comp = 0; comp < 2; ++comp) {
                const std::vector<float>& D = comp == 0 ? t.p : t.c;
                const float v00 = flat(D, rT, lo), v01 = flat(D, rT, lo + 1);
                const float v10 = flat(D, rT + 1, hi), v11 = flat(D, rT + 1, hi + 1);
                float v;
                if (fixed) /* bilinear */
                    v = (1 - fr) * (v00 * (1 - f0) + v01 * f0) + fr * (v10 * (1 - f1) + v11 * f1);
                else /* as written: every node takes the weight of the opposite node */
                    v = fr * (v00 * f0 + v01 * (1 - f0)) + (1 - fr) * (v10 * f1 + v11 * (1 - f1));
                pc[2 * ((int64_t)ri * res + ui) + comp] = v;
            }
        }
    }
}

thread_local std::string g_aneos_err;

}  // namespace

bool loadANEOSDataFromFile(const std::string& path, const sph_params& prm, int resolution, bool fixed,
                           EOSTable& out, std::string& err) {
    RawTable t;
    if (!parse(path, t, err)) return false;
    out.resolution = resolution;
    out.data.assign(2 * (size_t)resolution * resolution, 0.f);
    resample(t, prm, resolution, fixed, out.data.data());
    return true;
}

extern "C" {

int sph_load_aneos(const char* path, const sph_params* prm, int res, int fixed, float* pc_out) {
    if (!path || !prm || !pc_out || res < 2) {
        g_aneos_err = "bad argument";
        return SPH_E_ARG;
    }
    try {
        EOSTable t;
        std::string err;
        if (!loadANEOSDataFromFile(path, *prm, res, fixed != 0, t, err)) {
            g_aneos_err = err;
            return SPH_E_ARG;
        }
        std::copy(t.data.begin(), t.data.end(), pc_out);
        return SPH_OK;
    } catch (const std::exception& e) {
        g_aneos_err = std::string("sph_load_aneos: ") + e.what();
        return SPH_E_CAPACITY;
    }
}

const char* sph_aneos_last_error(void) { return g_aneos_err.c_str(); }

}  // extern "C"
