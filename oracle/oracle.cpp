/*
 * oracle.cpp -- CPU restatement of the reference's per-step SPH + FMM core.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).  See oracle.h.
 *
 * Arithmetic is FP32 with float literals everywhere: Metal Shading Language
 * has no double, so every unsuffixed literal in the reference shaders is a
 * float.  The reference builds its shaders with MTL_FAST_MATH
 * (project.pbxproj:518,573); this restatement uses IEEE libm (sqrtf, powf,
 * expf, logf) instead, so transcendental results can differ from the Metal
 * build by a few ULP.  That is why GPU parity is stated to FP32 tolerance.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

struct f3 {
    float x, y, z;
};
inline f3 mk(float x, float y, float z) { return {x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
inline f3& operator-=(f3& a, f3 b) { a = a - b; return a; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len2(f3 a) { return dot(a, a); }
inline float len(f3 a) { return sqrtf(len2(a)); }
inline f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline f3 fmin3(f3 a, f3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline f3 fmax3(f3 a, f3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline f3 ld3(const float* p, int64_t i) { return {p[4 * i], p[4 * i + 1], p[4 * i + 2]}; }
inline void st3(float* p, int64_t i, f3 v) { p[4 * i] = v.x; p[4 * i + 1] = v.y; p[4 * i + 2] = v.z; }

const float kPi1 = 1.f / 3.14159265358979323846f; /* kernels.h:14 */
const float kMaxFloat = 3.40282347e+38f;          /* MAXFLOAT */

inline float dt_min1(const sph_params* p) { return p->dt0 / (float)(1 << p->r_max); } /* Parameters.h:30 */
inline int n_cells(const sph_params* p) { return 1 << (3 * p->cell_power); }

/* ---------------------------------------------------------------- morton.h */
/* morton.h:16-23 */
inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
/* morton.h:26-34 */
inline uint32_t morton(uint32_t x, uint32_t y, uint32_t z) {
    return expand_bits(x) | (expand_bits(y) << 1) | (expand_bits(z) << 2);
}

/* ------------------------------------------------------------ cellsToScan.h */
struct CellRange { int mn[3], mx[3]; };
/* cellsToScan.h:21-31 */
inline uint32_t cell_index(const sph_params* p, int x, int y, int z) {
    int mask = (1 << p->cell_power) - 1;
    return morton((uint32_t)(x & mask), (uint32_t)(y & mask), (uint32_t)(z & mask));
}
/* cellsToScan.h:33-59.  Guard shared with the GPU path: a non-finite bound
 * gives an empty range and an extent is capped at the periodic grid size;
 * neither changes the range of a finite particle with h <= 1. */
inline CellRange cell_range(f3 position, float cellSize, float h, int cellsPerDim) {
    float radius = 2.f * h;
    f3 minRange = position - mk(radius, radius, radius);
    f3 maxRange = position + mk(radius, radius, radius);
    f3 minNorm = minRange / cellSize;
    f3 maxNorm = maxRange / cellSize;
    const float lo[3] = {minNorm.x, minNorm.y, minNorm.z};
    const float hi[3] = {maxNorm.x, maxNorm.y, maxNorm.z};
    CellRange r;
    for (int k = 0; k < 3; ++k) {
        if (!(lo[k] <= hi[k]) || !(std::fabs(lo[k]) < 1e9f) || !(std::fabs(hi[k]) < 1e9f)) {
            r.mn[k] = 0;
            r.mx[k] = -1;
            continue;
        }
        r.mn[k] = (int)floorf(lo[k]);
        r.mx[k] = (int)floorf(hi[k]);
        if (r.mx[k] - r.mn[k] >= cellsPerDim) r.mx[k] = r.mn[k] + cellsPerDim - 1;
    }
    return r;
}

/* ---------------------------------------------------------------- kernels.h */
/* kernels.h:35-50 */
inline float W(float r, float h1) {
    float q = r * h1;
    if (q > 2.0f) return 0;
    float fac = kPi1 * h1 * h1 * h1;
    if (q > 1.0f) {
        float tmp = 2.f - q;
        return fac * 0.25f * tmp * tmp * tmp;
    }
    return fac * (1 - 1.5f * q * q * (1 - 0.5f * q));
}
/* kernels.h:52-69 */
inline f3 gradW(f3 x_ij, float r, float r1, float h1) {
    float q = r * h1;
    if (q > 2.f || q < 1e-12f) return mk(0, 0, 0);
    f3 r_hat = x_ij * r1;
    float fac = kPi1 * h1 * h1 * h1 * h1;
    if (q > 1.f) {
        float tmp = 2.f - q;
        return (-0.75f * fac * tmp * tmp) * r_hat;
    }
    return (3 * fac * (0.75f * q - 1) * q) * r_hat;
}
/* kernels.h:71-85 */
inline float dW_dh(float r, float h1) {
    float q = r * h1;
    if (q > 2.0f) return 0;
    float fac = -3 * kPi1 * h1 * h1 * h1 * h1;
    if (q > 1.0f) return fac * (2 - 4 * q + 2.5f * q * q - 0.5f * q * q * q);
    return fac * (1 + 0.5f * (3 * q - 5) * q * q);
}
/* kernels.h:87-102.  As written the literals (4/3), (6/5), (8/3), (1/6),
 * (1/15) are integer divisions (= 1, 1, 2, 0, 0); `fixed` restores them. */
inline float dphi_dr(f3 x_ij, float h, bool fixed) {
    float h1 = 1.f / std::min(h, 0.16f);
    float r = len(x_ij);
    float q = r * h1;
    const float c43 = fixed ? 4.f / 3.f : 1.f, c65 = fixed ? 6.f / 5.f : 1.f;
    const float c83 = fixed ? 8.f / 3.f : 2.f, c16 = fixed ? 1.f / 6.f : 0.f, c115 = fixed ? 1.f / 15.f : 0.f;
    if (q <= 1) return (h1 * h1) * (c43 * q - c65 * q * q * q + 0.5f * q * q * q * q);
    if (q <= 2) {
        float q1 = 1.f / q;
        return (h1 * h1) * (c83 * q - 3 * q * q + c65 * q * q * q - c16 * q * q * q * q - c115 * q1 * q1);
    }
    float r1 = 1.f / r;
    return r1 * r1;
}

/* density.metal:45-50 / 228-235: Metal sample() of an RG32Float texture with
 * linear filtering, clamp_to_edge and normalised coordinates.  Texel centres
 * sit at (i + 1/2)/res.  u is the x (column) axis, rho the y (row) axis. */
inline void eos_sample(const sph_params* p, const float* table, int res, float u, float density,
                       float* P_out, float* c_out) {
    float uCoord = logf(u / p->eos_min_u) / logf(p->eos_max_u / p->eos_min_u);
    float rhoCoord = logf(density / p->eos_min_rho) / logf(p->eos_max_rho / p->eos_min_rho);
    /* clamp(uv, 0, 1); a NaN coordinate (log of a non-positive value) maps to 0 */
    uCoord = std::fmin(std::fmax(uCoord, 0.f), 1.f);
    rhoCoord = std::fmin(std::fmax(rhoCoord, 0.f), 1.f);
    float tx = uCoord * (float)res - 0.5f;
    float ty = rhoCoord * (float)res - 0.5f;
    float fx = floorf(tx), fy = floorf(ty);
    float a = tx - fx, b = ty - fy;
    int i0 = (int)fx, j0 = (int)fy;
    int i1 = std::min(std::max(i0 + 1, 0), res - 1), j1 = std::min(std::max(j0 + 1, 0), res - 1);
    i0 = std::min(std::max(i0, 0), res - 1);
    j0 = std::min(std::max(j0, 0), res - 1);
    for (int c = 0; c < 2; ++c) {
        float t00 = table[2 * ((int64_t)j0 * res + i0) + c];
        float t01 = table[2 * ((int64_t)j0 * res + i1) + c];
        float t10 = table[2 * ((int64_t)j1 * res + i0) + c];
        float t11 = table[2 * ((int64_t)j1 * res + i1) + c];
        float v = (1.f - a) * (1.f - b) * t00 + a * (1.f - b) * t01 + (1.f - a) * b * t10 + a * b * t11;
        if (c == 0) *P_out = v; else *c_out = v;
    }
}

inline const float* eos_table_for(const orc_state* s, int mat) {
    /* density.metal:50: `materialId == 400 ? texForesite : texFe` */
    return mat == 400 ? s->eos_forsterite : s->eos_fe;
}

/* --------------------------------------------------------------- gravity */
/* poles.h:18-39 index map; third-order terms exist at P = 3 */
enum {
    M = 0, X = 1, Y = 2, Z = 3, XX = 4, XY = 5, XZ = 6, YY = 7, YZ = 8, ZZ = 9, NT = 10,
    XXX = 10, XXY = 11, XXZ = 12, XYY = 13, XYZ = 14, XZZ = 15, YYY = 16, YYZ = 17, YZZ = 18, ZZZ = 19, NT3 = 20
};
/* record strides of the oracle's buffers (the first 24 / 16 floats are the
 * sph_capi.h P = 2 layouts; P = 3 appends the third-order terms) */
constexpr int MP_REC = 40;
constexpr int LC_REC = 32;

struct Multipole { /* poles.h:46-55 */
    f3 pos, mn, mx;
    float size;
    float e[NT3];
    float power[4];
    float minGrav;
    bool active;
};
struct Local { /* poles.h:58-61 */
    f3 pos;
    float e[NT3];
    /* interaction account (orc_gravity_account only; not in the reference):
     * particles and mass of every source node M2L'd into this expansion along
     * the path from the root, and the part of that count that comes from
     * accepting a node which contains the expansion's own node */
    int64_t acct_n = 0;
    double acct_m = 0.0;
    int64_t acct_excess_n = 0;
    double acct_excess_m = 0.0;
    int self_node = -1;
};

Multipole load_mp(const float* rec, int order = 2) {
    Multipole m;
    m.pos = mk(rec[0], rec[1], rec[2]);
    m.size = rec[3];
    m.mn = mk(rec[4], rec[5], rec[6]);
    m.minGrav = rec[7];
    m.mx = mk(rec[8], rec[9], rec[10]);
    m.active = rec[11] != 0.f;
    for (int i = 0; i < NT; ++i) m.e[i] = rec[12 + i];
    for (int i = NT; i < NT3; ++i) m.e[i] = order > 2 ? rec[24 + i - NT] : 0.f;
    m.power[0] = rec[12];
    m.power[1] = rec[23];
    m.power[2] = rec[22];
    m.power[3] = order > 2 ? rec[34] : 0.f;
    return m;
}
void store_mp(float* rec, const Multipole& m, int order = 2) {
    rec[0] = m.pos.x; rec[1] = m.pos.y; rec[2] = m.pos.z; rec[3] = m.size;
    rec[4] = m.mn.x; rec[5] = m.mn.y; rec[6] = m.mn.z; rec[7] = m.minGrav;
    rec[8] = m.mx.x; rec[9] = m.mx.y; rec[10] = m.mx.z; rec[11] = m.active ? 1.f : 0.f;
    for (int i = 0; i < NT; ++i) rec[12 + i] = m.e[i];
    rec[22] = m.power[2];
    rec[23] = m.power[1];
    /* power[0] is stored as e[M] (addPowers sets power[0] = expansion[0]) */
    if (order > 2) {
        for (int i = NT; i < NT3; ++i) rec[24 + i - NT] = m.e[i];
        rec[34] = m.power[3];
    }
}
Local load_lc(const float* rec, int order = 2) {
    Local l;
    l.pos = mk(rec[0], rec[1], rec[2]);
    for (int i = 0; i < NT; ++i) l.e[i] = rec[4 + i];
    for (int i = NT; i < NT3; ++i) l.e[i] = order > 2 ? rec[16 + i - NT] : 0.f;
    return l;
}
void store_lc(float* rec, const Local& l, int order = 2) {
    rec[0] = l.pos.x; rec[1] = l.pos.y; rec[2] = l.pos.z; rec[3] = 0.f;
    for (int i = 0; i < NT; ++i) rec[4 + i] = l.e[i];
    rec[14] = 0.f; rec[15] = 0.f;
    if (order > 2)
        for (int i = NT; i < NT3; ++i) rec[16 + i - NT] = l.e[i];
}

/* helpers.metal:12-27 (rows used at P <= 3) */
inline int binomial(int n, int k) {
    static const int c[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
    return c[n][k];
}
/* helpers.metal:30-61 */
inline float integer_powf(float x, unsigned n) {
    switch (n) {
        case 0: return 1.f;
        case 1: return x;
        case 2: return x * x;
        default: return powf(x, (float)n);
    }
}
/* helpers.metal:63-89.  At P = 3, as written, the weights (1 / 3) and
 * (1 / 6) are integer divisions (0) and three lines assign instead of add,
 * so power[3] = |XXX| (helpers.metal:76-88); `fix_p3` sums the norm. */
inline void add_powers(Multipole& mp, int order = 2, bool fix_p3 = false) {
    mp.power[0] = mp.e[0];
    mp.power[1] = 0;
    float p2 = mp.e[ZZ] * mp.e[ZZ];
    p2 += 0.5f * mp.e[YZ] * mp.e[YZ];
    p2 += mp.e[YY] * mp.e[YY];
    p2 += 0.5f * mp.e[XZ] * mp.e[XZ];
    p2 += 0.5f * mp.e[XY] * mp.e[XY];
    p2 += 0.5f * mp.e[XX] * mp.e[XX];
    mp.power[2] = sqrtf(p2);
    mp.power[3] = 0.f;
    if (order < 3) return;
    const float third = fix_p3 ? 1.f / 3.f : 0.f, sixth = fix_p3 ? 1.f / 6.f : 0.f;
    float p3 = mp.e[ZZZ] * mp.e[ZZZ];
    p3 += third * mp.e[YZZ] * mp.e[YZZ];
    p3 += third * mp.e[YYZ] * mp.e[YYZ];
    p3 = fix_p3 ? p3 + mp.e[YYY] * mp.e[YYY] : mp.e[YYY] * mp.e[YYY];
    p3 += third * mp.e[XZZ] * mp.e[XZZ];
    p3 += sixth * mp.e[XYZ] * mp.e[XYZ];
    p3 += third * mp.e[XYY] * mp.e[XYY];
    p3 += third * mp.e[XXZ] * mp.e[XXZ];
    p3 += third * mp.e[XXY] * mp.e[XXY];
    p3 = fix_p3 ? p3 + mp.e[XXX] * mp.e[XXX] : mp.e[XXX] * mp.e[XXX];
    mp.power[3] = sqrtf(p3);
}

/* accept.metal:14-55.  r^p is integer_powf(r2, p / 2): at P = 3, as
 * written, r^2 instead of r^3 (`fix_p3`: r^3). */
inline bool m2l_accept(const Multipole& A, const Multipole& B, float eta, int p = 2, bool fix_p3 = false) {
    const float rho_A = A.size;
    const float rho_B = B.size;
    const float r2 = len2(A.pos - B.pos);
    const float rho_sum = rho_A + rho_B;
    if (rho_sum * rho_sum > r2) return false;
    const float rho_max = std::max(rho_A, rho_B);
    float E_BA_term = 0.f;
    for (int n = 0; n <= p; ++n) E_BA_term += binomial(p, n) * B.power[n] * integer_powf(rho_A, p - n);
    E_BA_term *= 8.f;
    if (rho_A + rho_B > 0.f) {
        E_BA_term *= rho_max;
        E_BA_term /= (rho_A + rho_B);
    }
    const float r_to_p = (fix_p3 && p == 3) ? r2 * sqrtf(r2) : integer_powf(r2, (p / 2));
    float f_MAC_inv = r2;
    const float min_a_grav = A.minGrav;
    return E_BA_term < eta * min_a_grav * r_to_p * f_MAC_inv;
}

/* derivatives.metal:12-57 */
inline float D_soft_1(float u) {
    float phi = -3.f * u + 15.f;
    phi = phi * u - 28.f; phi = phi * u + 21.f; phi = phi * u;
    phi = phi * u - 7.f; phi = phi * u; phi = phi * u + 3.f;
    return phi;
}
inline float D_soft_2(float u) {
    float phi = -21.f * u + 90.f;
    phi = phi * u - 140.f; phi = phi * u + 84.f; phi = phi * u;
    phi = phi * u - 14.f; phi = phi * u;
    return phi;
}
inline float D_soft_3(float u) {
    float phi = -105.f * u + 360.f;
    phi = phi * u - 420.f; phi = phi * u + 168.f; phi = phi * u; phi = phi * u;
    return phi;
}
inline float D_soft_4(float u) {
    float phi = -315.f * u + 720.f;
    phi = phi * u - 420.f; phi = phi * u; phi = phi * u;
    return phi;
}
/* derivatives.metal:59-156; `d` has NT3 entries (third order when order 3) */
inline void derivatives(f3 vec, float eps, float* d, int order = 2) {
    float Dt_1, Dt_2, Dt_3, Dt_4 = 0.f;
    float r2 = len2(vec);
    float r = sqrtf(r2);
    float r_inv = r > 1e-6f ? 1 / r : 0;
    float eps2 = eps * eps;
    if (r2 < eps2) {
        const float eps_inv = 1.f / eps;
        const float u = r * eps_inv;
        Dt_1 = eps_inv * D_soft_1(u);
        const float eps_inv2 = eps_inv * eps_inv;
        Dt_2 = eps_inv2 * D_soft_2(u);
        const float eps_inv3 = eps_inv2 * eps_inv;
        Dt_3 = eps_inv3 * D_soft_3(u);
        const float eps_inv4 = eps_inv3 * eps_inv;
        Dt_4 = eps_inv4 * D_soft_4(u);
    } else {
        Dt_1 = r_inv;
        Dt_2 = -1.f * Dt_1 * r_inv;
        Dt_3 = -3.f * Dt_2 * r_inv;
        Dt_4 = -5.f * Dt_3 * r_inv;
    }
    const float rx_r = vec.x * r_inv, ry_r = vec.y * r_inv, rz_r = vec.z * r_inv;
    const float rx_r2 = rx_r * rx_r, ry_r2 = ry_r * ry_r, rz_r2 = rz_r * rz_r;
    d[M] = Dt_1;
    d[X] = rx_r * Dt_2;
    d[Y] = ry_r * Dt_2;
    d[Z] = rz_r * Dt_2;
    Dt_2 *= r_inv;
    d[XX] = rx_r2 * Dt_3 + Dt_2;
    d[XY] = rx_r * ry_r * Dt_3;
    d[XZ] = rx_r * rz_r * Dt_3;
    d[YY] = ry_r2 * Dt_3 + Dt_2;
    d[YZ] = ry_r * rz_r * Dt_3;
    d[ZZ] = rz_r2 * Dt_3 + Dt_2;
    for (int i = NT; i < NT3; ++i) d[i] = 0.f;
    if (order < 3) return;
    Dt_3 *= r_inv;
    const float rx_r3 = rx_r2 * rx_r, ry_r3 = ry_r2 * ry_r, rz_r3 = rz_r2 * rz_r;
    d[XXX] = rx_r3 * Dt_4 + 3.f * rx_r * Dt_3;
    d[YYY] = ry_r3 * Dt_4 + 3.f * ry_r * Dt_3;
    d[ZZZ] = rz_r3 * Dt_4 + 3.f * rz_r * Dt_3;
    d[XXY] = rx_r2 * ry_r * Dt_4 + ry_r * Dt_3;
    d[XXZ] = rx_r2 * rz_r * Dt_4 + rz_r * Dt_3;
    d[XYY] = ry_r2 * rx_r * Dt_4 + rx_r * Dt_3;
    d[YYZ] = ry_r2 * rz_r * Dt_4 + rz_r * Dt_3;
    d[XZZ] = rz_r2 * rx_r * Dt_4 + rx_r * Dt_3;
    d[YZZ] = rz_r2 * ry_r * Dt_4 + ry_r * Dt_3;
    d[XYZ] = rx_r * ry_r * rz_r * Dt_4;
}

/* M2L.metal:12-108 */
inline Local M2L(f3 x, const Multipole& mp, float eps, int order = 2) {
    f3 r = x - mp.pos;
    float D[NT3];
    derivatives(r, eps, D, order);
    Local l;
    const float M_000 = mp.e[M];
    l.e[M] = M_000 * D[M];
    l.e[X] = M_000 * D[X];
    l.e[Y] = M_000 * D[Y];
    l.e[Z] = M_000 * D[Z];
    l.e[M] += mp.e[XX] * D[XX] + mp.e[YY] * D[YY] + mp.e[ZZ] * D[ZZ];
    l.e[M] += mp.e[XY] * D[XY] + mp.e[XZ] * D[XZ] + mp.e[YZ] * D[YZ];
    l.e[XX] = M_000 * D[XX];
    l.e[YY] = M_000 * D[YY];
    l.e[ZZ] = M_000 * D[ZZ];
    l.e[XY] = M_000 * D[XY];
    l.e[XZ] = M_000 * D[XZ];
    l.e[YZ] = M_000 * D[YZ];
    for (int i = NT; i < NT3; ++i) l.e[i] = 0.f;
    if (order < 3) return l;
    /* M2L.metal:58-107 */
    const float* q = mp.e;
    l.e[M] += q[XXX] * D[XXX] + q[YYY] * D[YYY] + q[ZZZ] * D[ZZZ];
    l.e[M] += q[XXY] * D[XXY] + q[XXZ] * D[XXZ] + q[XYY] * D[XYY];
    l.e[M] += q[YYZ] * D[YYZ] + q[XZZ] * D[XZZ] + q[YZZ] * D[YZZ];
    l.e[M] += q[XYZ] * D[XYZ];
    l.e[X] += q[XX] * D[XXX] + q[YY] * D[XYY] + q[ZZ] * D[XZZ];
    l.e[X] += q[XY] * D[XXY] + q[XZ] * D[XXZ] + q[YZ] * D[XYZ];
    l.e[Y] += q[XX] * D[XXY] + q[YY] * D[YYY] + q[ZZ] * D[YZZ];
    l.e[Y] += q[XY] * D[XYY] + q[XZ] * D[XYZ] + q[YZ] * D[YYZ];
    l.e[Z] += q[XX] * D[XXZ] + q[YY] * D[YYZ] + q[ZZ] * D[ZZZ];
    l.e[Z] += q[XY] * D[XYZ] + q[XZ] * D[XZZ] + q[YZ] * D[YZZ];
    for (int i = NT; i < NT3; ++i) l.e[i] = M_000 * D[i];
    return l;
}

/* L2L.metal:12-88 */
inline Local transform_local(const Local& L, f3 r, int order = 2) {
    Local n;
    n.pos = L.pos + r;
    n.e[M] = L.e[M];
    n.e[M] += r.x * L.e[X] + r.y * L.e[Y] + r.z * L.e[Z];
    n.e[X] = L.e[X];
    n.e[Y] = L.e[Y];
    n.e[Z] = L.e[Z];
    n.e[M] += 0.5f * (r.x * r.x * L.e[XX] + r.y * r.y * L.e[YY] + r.z * r.z * L.e[ZZ]);
    n.e[M] += r.x * r.y * L.e[XY] + r.x * r.z * L.e[XZ] + r.y * r.z * L.e[YZ];
    n.e[X] += r.x * L.e[XX] + r.y * L.e[XY] + r.z * L.e[XZ];
    n.e[Y] += r.x * L.e[XY] + r.y * L.e[YY] + r.z * L.e[YZ];
    n.e[Z] += r.x * L.e[XZ] + r.y * L.e[YZ] + r.z * L.e[ZZ];
    for (int i = XX; i <= ZZ; ++i) n.e[i] = L.e[i];
    for (int i = NT; i < NT3; ++i) n.e[i] = 0.f;
    if (order < 3) return n;
    /* L2L.metal:37-86 */
    const float* e = L.e;
    n.e[M] += 0.166666667f * r.x * r.x * r.x * e[XXX] + 0.166666667f * r.y * r.y * r.y * e[YYY] +
              0.166666667f * r.z * r.z * r.z * e[ZZZ];
    n.e[M] += 0.5f * r.x * r.x * r.y * e[XXY] + 0.5f * r.x * r.x * r.z * e[XXZ] + 0.5f * r.x * r.y * r.y * e[XYY];
    n.e[M] += 0.5f * r.y * r.y * r.z * e[YYZ] + 0.5f * r.x * r.z * r.z * e[XZZ] + 0.5f * r.y * r.z * r.z * e[YZZ];
    n.e[M] += r.x * r.y * r.z * e[XYZ];
    n.e[X] += 0.5f * r.x * r.x * e[XXX] + 0.5f * r.y * r.y * e[XYY] + 0.5f * r.z * r.z * e[XZZ];
    n.e[X] += r.x * r.y * e[XXY] + r.x * r.z * e[XXZ] + r.y * r.z * e[XYZ];
    n.e[Y] += 0.5f * r.x * r.x * e[XXY] + 0.5f * r.y * r.y * e[YYY] + 0.5f * r.z * r.z * e[YZZ];
    n.e[Y] += r.x * r.y * e[XYY] + r.x * r.z * e[XYZ] + r.y * r.z * e[YYZ];
    n.e[Z] += 0.5f * r.x * r.x * e[XXZ] + 0.5f * r.y * r.y * e[YYZ] + 0.5f * r.z * r.z * e[ZZZ];
    n.e[Z] += r.x * r.y * e[XYZ] + r.x * r.z * e[XZZ] + r.y * r.z * e[YZZ];
    n.e[XX] += r.x * e[XXX] + r.y * e[XXY] + r.z * e[XXZ];
    n.e[YY] += r.x * e[XYY] + r.y * e[YYY] + r.z * e[YYZ];
    n.e[ZZ] += r.x * e[XZZ] + r.y * e[YZZ] + r.z * e[ZZZ];
    n.e[XY] += r.x * e[XXY] + r.y * e[XYY] + r.z * e[XYZ];
    n.e[XZ] += r.x * e[XXZ] + r.y * e[XYZ] + r.z * e[XZZ];
    n.e[YZ] += r.x * e[XYZ] + r.y * e[YYZ] + r.z * e[YZZ];
    for (int i = NT; i < NT3; ++i) n.e[i] = e[i];
    return n;
}

/* L2P.metal:12-32 */
inline f3 L2P(const Local& L, f3 pos, int order = 2) {
    f3 acc = mk(0, 0, 0);
    f3 r = pos - L.pos;
    acc += mk(L.e[X], L.e[Y], L.e[Z]);
    acc.x += r.x * L.e[XX] + r.y * L.e[XY] + r.z * L.e[XZ];
    acc.y += r.x * L.e[XY] + r.y * L.e[YY] + r.z * L.e[YZ];
    acc.z += r.x * L.e[XZ] + r.y * L.e[YZ] + r.z * L.e[ZZ];
    if (order < 3) return acc;
    const float* e = L.e;
    acc.x += 0.5f * r.x * r.x * e[XXX] + 0.5f * r.y * r.y * e[XYY] + 0.5f * r.z * r.z * e[XZZ];
    acc.x += r.x * r.y * e[XXY] + r.x * r.z * e[XXZ] + r.y * r.z * e[XYZ];
    acc.y += 0.5f * r.x * r.x * e[XXY] + 0.5f * r.y * r.y * e[YYY] + 0.5f * r.z * r.z * e[YZZ];
    acc.y += r.x * r.y * e[XYY] + r.x * r.z * e[XYZ] + r.y * r.z * e[YYZ];
    acc.z += 0.5f * r.x * r.x * e[XXZ] + 0.5f * r.y * r.y * e[YYZ] + 0.5f * r.z * r.z * e[ZZZ];
    acc.z += r.x * r.y * e[XYZ] + r.x * r.z * e[XZZ] + r.y * r.z * e[YZZ];
    return acc;
}

/* M2M.metal:12-73.  As written the ZZ shift multiplies by the dipole
 * expansion[Z] instead of the mass (M2M.metal:29). */
inline Multipole transform_multipole(const Multipole& mp, f3 r, bool fix_zz, int order = 2) {
    Multipole n = mp;
    n.pos = mp.pos + r;
    n.e[M] = mp.e[M];
    n.e[X] = 0;
    n.e[Y] = 0;
    n.e[Z] = 0;
    n.e[XX] = mp.e[XX] + 0.5f * r.x * r.x * mp.e[M];
    n.e[XY] = mp.e[XY] + r.x * r.y * mp.e[M];
    n.e[XZ] = mp.e[XZ] + r.x * r.z * mp.e[M];
    n.e[YY] = mp.e[YY] + 0.5f * r.y * r.y * mp.e[M];
    n.e[YZ] = mp.e[YZ] + r.y * r.z * mp.e[M];
    n.e[ZZ] = mp.e[ZZ] + 0.5f * r.z * r.z * (fix_zz ? mp.e[M] : mp.e[Z]);
    for (int i = NT; i < NT3; ++i) n.e[i] = 0.f;
    if (order < 3) return n;
    /* M2M.metal:31-71 */
    const float* e = mp.e;
    const float m0 = e[M];
    n.e[ZZZ] = e[ZZZ] + r.z * e[ZZ] + 0.166666667f * r.z * r.z * r.z * m0;
    n.e[YZZ] = e[YZZ] + r.z * e[YZ] + r.y * e[ZZ] + 0.5f * r.y * r.z * r.z * m0;
    n.e[YYZ] = e[YYZ] + r.z * e[YY] + r.y * e[YZ] + 0.5f * r.y * r.y * r.z * m0;
    n.e[YYY] = e[YYY] + r.y * e[YY] + 0.166666667f * r.y * r.y * r.y * m0;
    n.e[XZZ] = e[XZZ] + r.z * e[XZ] + r.x * e[ZZ] + 0.5f * r.x * r.z * r.z * m0;
    n.e[XYZ] = e[XYZ] + r.z * e[XY] + r.y * e[XZ] + r.x * e[YZ] + r.x * r.y * r.z * m0;
    n.e[XYY] = e[XYY] + r.y * e[XY] + r.x * e[YY] + 0.5f * r.x * r.y * r.y * m0;
    n.e[XXZ] = e[XXZ] + r.z * e[XX] + r.x * e[XZ] + 0.5f * r.x * r.x * r.z * m0;
    n.e[XXY] = e[XXY] + r.y * e[XX] + r.x * e[XY] + 0.5f * r.x * r.x * r.y * m0;
    n.e[XXX] = e[XXX] + r.x * e[XX] + 0.166666667f * r.x * r.x * r.x * m0;
    return n;
}

struct GravCtx {
    const sph_params* prm;
    const orc_state* s;
    const orc_tree* t;
    orc_gravity_buffers* g;
    int order;      /* expansion order P (Parameters.h) */
    int mac_order;  /* order of the acceptance test's error estimate */
    bool fix_p3;    /* SPH_FIX_ORDER3 */
    GravCtx(const sph_params* p, const orc_state* st, const orc_tree* tr, orc_gravity_buffers* gb)
        : prm(p), s(st), t(tr), g(gb), order(p->expansion_order == 3 ? 3 : 2),
          mac_order(p->mac_order == 2 ? 2 : (p->expansion_order == 3 ? 3 : 2)),
          fix_p3((p->fixes & SPH_FIX_ORDER3) != 0) {}
    Multipole mp_at(int dataPointer) const { return load_mp(g->multipoles + MP_REC * (int64_t)dataPointer, order); }
    struct Account* acct = nullptr;   /* test-only interaction account, see orc_gravity_account */
};

/* Test-only interaction account of the down pass.  Per node (by data
 * pointer): the particles and mass of its subtree (FP64) and its parent; per
 * node the account carried into its local expansion; per particle the final
 * account.  `plant` puts a deliberate mistake into the traversal so that a
 * test can show the account detects it: 1 = skip the last entry of every
 * non-empty input list, 2 = M2L the first accepted node of every leaf twice. */
struct Account {
    std::vector<int64_t> node_n;
    std::vector<double> node_m;
    std::vector<int> parent;
    std::vector<int64_t> local_n, local_excess_n;
    std::vector<double> local_m, local_excess_m;
    int64_t* out_n;
    double* out_m;
    int64_t* out_excess_n;
    double* out_excess_m;
    std::atomic<int64_t> unexpected{0};
    int plant = 0;
};

/* P2M.metal:12-107 */
Multipole P2M(const GravCtx& c, int treePointer) {
    const int* tree = c.t->octree;
    const orc_state* s = c.s;
    Multipole mp;
    int nParticles = tree[treePointer];
    int start = treePointer + 2;
    int end = start + nParticles;
    mp.pos = mk(0, 0, 0);
    for (int i = 0; i < NT3; ++i) mp.e[i] = 0;
    mp.minGrav = kMaxFloat;
    mp.active = false;
    mp.mx = mk(-kMaxFloat, -kMaxFloat, -kMaxFloat);
    mp.mn = mk(kMaxFloat, kMaxFloat, kMaxFloat);
    const int gt = *s->global_time;
    for (int i = start; i < end; i++) {
        int p = tree[i];
        float p_mass = s->m[p];
        f3 p_position = ld3(s->pos, p);
        int isActive = c.prm->activity_rule == SPH_ACTIVE_README ? (s->next_active[p] <= gt) : (gt <= s->next_active[p]);
        if (!mp.active && isActive) mp.active = true;
        mp.mx = fmax3(mp.mx, p_position);
        mp.mn = fmin3(mp.mn, p_position);
        mp.e[M] += p_mass;
        mp.pos += p_position * p_mass;
        if (isActive) mp.minGrav = std::min(mp.minGrav, s->gravnorm[p]);
    }
    for (int i = start; i < end; i++) s->active[tree[i]] = mp.active;
    mp.pos = mp.pos / mp.e[M];
    for (int i = start; i < end; i++) {
        int p = tree[i];
        float p_mass = s->m[p];
        f3 r = ld3(s->pos, p) - mp.pos;
        mp.e[X] -= r.x * p_mass;
        mp.e[Y] -= r.y * p_mass;
        mp.e[Z] -= r.z * p_mass;
        mp.e[XX] += p_mass * r.x * r.x;
        mp.e[XY] += p_mass * r.x * r.y;
        mp.e[XZ] += p_mass * r.x * r.z;
        mp.e[YY] += p_mass * r.y * r.y;
        mp.e[YZ] += p_mass * r.y * r.z;
        mp.e[ZZ] += p_mass * r.z * r.z;
        if (c.order > 2) { /* P2M.metal:70-81 */
            mp.e[XXX] -= p_mass * r.x * r.x * r.x;
            mp.e[XXY] -= p_mass * r.x * r.x * r.y;
            mp.e[XXZ] -= p_mass * r.x * r.x * r.z;
            mp.e[XYY] -= p_mass * r.x * r.y * r.y;
            mp.e[XYZ] -= p_mass * r.x * r.y * r.z;
            mp.e[XZZ] -= p_mass * r.x * r.z * r.z;
            mp.e[YYY] -= p_mass * r.y * r.y * r.y;
            mp.e[YYZ] -= p_mass * r.y * r.y * r.z;
            mp.e[YZZ] -= p_mass * r.y * r.z * r.z;
            mp.e[ZZZ] -= p_mass * r.z * r.z * r.z;
        }
    }
    mp.e[XX] *= 0.5f;
    mp.e[YY] *= 0.5f;
    mp.e[ZZ] *= 0.5f;
    if (c.order > 2) { /* P2M.metal:89-99 */
        mp.e[XXX] *= 0.166666667f;
        mp.e[YYY] *= 0.166666667f;
        mp.e[ZZZ] *= 0.166666667f;
        mp.e[XXY] *= 0.5f;
        mp.e[XXZ] *= 0.5f;
        mp.e[XYY] *= 0.5f;
        mp.e[XZZ] *= 0.5f;
        mp.e[YYZ] *= 0.5f;
        mp.e[YZZ] *= 0.5f;
    }
    f3 dims = mp.mx - mp.mn;
    mp.size = std::max(std::max(dims.x, dims.y), dims.z);
    add_powers(mp, c.order, c.fix_p3);
    return mp;
}

/* M2M.metal:75-160 */
Multipole M2M(const GravCtx& c, uint32_t index, int treePointer) {
    const int* tree = c.t->octree;
    const bool fix_zz = c.prm->fixes & SPH_FIX_M2M_ZZ;
    const bool fix_range = c.prm->fixes & SPH_FIX_SYNC_RANGE;
    Multipole mp;
    mp.minGrav = kMaxFloat;
    mp.pos = mk(0, 0, 0);
    mp.active = false;
    for (int i = 0; i < NT3; ++i) mp.e[i] = 0;
    float mass = 0;
    mp.mx = mk(-kMaxFloat, -kMaxFloat, -kMaxFloat);
    mp.mn = mk(kMaxFloat, kMaxFloat, kMaxFloat);
    int start = treePointer + 2;
    int end = start + 8;
    bool allLeaves = true;
    for (int i = start; i < end; i++) {
        int childPointer = tree[i];
        if (childPointer == -1) continue;
        int childDataPointer = tree[childPointer + 1];
        c.g->parent_index[childDataPointer] = index;
        Multipole childMp = c.mp_at(childDataPointer);
        mp.minGrav = std::min(childMp.minGrav, mp.minGrav);
        mass += childMp.e[M];
        mp.pos += childMp.pos * childMp.e[M];
        if (!mp.active && childMp.active) mp.active = true;
        if (allLeaves && tree[childPointer] == 0) allLeaves = false;
        mp.mx = fmax3(mp.mx, childMp.mx);
        mp.mn = fmin3(mp.mn, childMp.mn);
    }
    f3 dims = mp.mx - mp.mn;
    mp.size = std::max(std::max(dims.x, dims.y), dims.z);
    mp.pos = mp.pos / mass;
    for (int i = start; i < end; i++) {
        int childPointer = tree[i];
        if (childPointer == -1) continue;
        int childDataPointer = tree[childPointer + 1];
        Multipole childMp = c.mp_at(childDataPointer);
        f3 r = mp.pos - childMp.pos;
        Multipole tr = transform_multipole(childMp, r, fix_zz, c.order);
        for (int j = 0; j < NT3; ++j) mp.e[j] += tr.e[j];
        if (allLeaves && mp.active && !childMp.active) {
            c.g->multipoles[MP_REC * (int64_t)childDataPointer + 11] = 1.f;
            int nParticles = tree[childPointer];
            int childStart = childPointer + 2;
            int childEnd = (fix_range ? childStart : start) + nParticles;
            for (int j = childStart; j < childEnd; j++) c.s->active[tree[j]] = true;
        }
    }
    add_powers(mp, c.order, c.fix_p3);
    return mp;
}

/* Books an accepted source node into the account of `local`.  A source that
 * contains the expansion's own node (the reference's acceptance test can
 * accept an ancestor, accept.metal:17-25) counts that node's particles a
 * second time: that part is kept apart as the excess.  Only the direct
 * parent can be such a source (a list holds its node itself and nodes
 * disjoint from it); any other containing source is counted as unexpected. */
inline void book_accept(const GravCtx& c, Local& local, int src) {
    Account& a = *c.acct;
    local.acct_n += a.node_n[src];
    local.acct_m += a.node_m[src];
    const int self = local.self_node;
    if (self < 0) return;
    for (int q = self; q >= 0; q = a.parent[q]) {
        if (q != src) continue;
        if (a.parent[self] == src) {
            local.acct_excess_n += a.node_n[self];
            local.acct_excess_m += a.node_m[self];
        } else {
            a.unexpected++;
        }
        break;
    }
}

inline bool check_and_add(const GravCtx& c, const Multipole& mp, Local& local, int toScan, bool forceAccept) {
    /* helpers.h:34-45 */
    int nodeDataPointer = c.t->octree[toScan + 1];
    Multipole nodeMp = c.mp_at(nodeDataPointer);
    if (forceAccept || m2l_accept(mp, nodeMp, c.prm->gravity_eta, c.mac_order, c.fix_p3)) {
        Local nl = M2L(local.pos, nodeMp, c.prm->gravity_softening, c.order);
        for (int j = 0; j < NT3; ++j) local.e[j] += nl.e[j];
        if (c.acct) book_accept(c, local, nodeDataPointer);
        return true;
    }
    return false;
}

/* helpers.h:83-123 */
void recursive_scan(const GravCtx& c, const Multipole& mp, Local& local, int toScan, int treePointer) {
    const int* tree = c.t->octree;
    const int R = c.prm->gravity_max_recursion;
    if (check_and_add(c, mp, local, toScan, tree[toScan] != 0)) return;
    int indexes[8], ends[8];
    unsigned currentDepth = 0;
    indexes[0] = toScan + 2;
    ends[0] = indexes[0] + 8;
    int count = 0;
    while (indexes[0] < ends[0] && count < 2000) {
        count++;
        int nodePointer = tree[indexes[currentDepth]];
        if (nodePointer == treePointer || nodePointer == -1) {
            indexes[currentDepth]++;
            while (indexes[currentDepth] == ends[currentDepth] && currentDepth != 0) {
                currentDepth--;
                indexes[currentDepth]++;
            }
            continue;
        }
        if (!check_and_add(c, mp, local, nodePointer, tree[nodePointer] != 0 || (int)currentDepth == R - 1)) {
            currentDepth++;
            indexes[currentDepth] = nodePointer + 2;
            ends[currentDepth] = indexes[currentDepth] + 8;
            continue;
        }
        indexes[currentDepth]++;
        while (indexes[currentDepth] == ends[currentDepth] && currentDepth != 0) {
            currentDepth--;
            indexes[currentDepth]++;
        }
    }
}

/* helpers.h:125-147.  With SPH_FIX_NEAR_FIELD a rejected leaf (an entry or a
 * child) is pushed like a branch instead of being force-accepted. */
void non_recursive_scan(const GravCtx& c, int32_t* out, const Multipole& mp, Local& local, uint32_t& outputPointer,
                        uint32_t maxOutputPointer, int toScan, int treePointer) {
    const int* tree = c.t->octree;
    const bool near = c.prm->fixes & SPH_FIX_NEAR_FIELD;
    if (check_and_add(c, mp, local, toScan, false)) return;
    if (near && tree[toScan] != 0) {
        if (outputPointer == maxOutputPointer) check_and_add(c, mp, local, toScan, true);
        else out[outputPointer++] = toScan;
        return;
    }
    int start = toScan + 2;
    int end = start + 8;
    for (int j = start; j < end; j++) {
        int childPointer = tree[j];
        if (childPointer == treePointer || childPointer == -1) continue;
        const bool force = (!near && tree[childPointer] != 0) || outputPointer == maxOutputPointer;
        if (!check_and_add(c, mp, local, childPointer, force)) {
            out[outputPointer] = childPointer;
            outputPointer++;
        }
    }
}

/* SPH_FIX_NEAR_FIELD for a leaf: accept by the MAC, else open branches down to
 * leaves, and collect the rejected source leaves for the P2P sum. */
void near_scan(const GravCtx& c, const Multipole& mp, Local& local, int toScan, int treePointer,
               std::vector<int>& p2p_leaves) {
    const int* tree = c.t->octree;
    std::vector<int> stack{toScan};
    while (!stack.empty()) {
        const int node = stack.back();
        stack.pop_back();
        if (node == treePointer) continue;
        if (check_and_add(c, mp, local, node, false)) continue;
        if (tree[node] != 0) {
            p2p_leaves.push_back(node);
            continue;
        }
        for (int j = node + 9; j >= node + 2; --j)   /* children in order 0..7 */
            if (tree[j] != -1) stack.push_back(tree[j]);
    }
}

/* -d(potential)/dr of a unit mass for the potential whose derivatives M2L
 * uses (derivatives.metal:12-57 with support eps): the near-field form sums
 * every pair with this one kernel so that a pair's force does not depend on
 * whether it is reached by P2P or by M2L. */
inline float dphi_dr_m2l_kernel(float r, float eps) {
    if (r >= eps) return 1.f / (r * r);
    const float eps_inv = 1.f / eps;
    const float u = r * eps_inv;
    return -(eps_inv * eps_inv) * D_soft_2(u);
}

/* helpers.h:47-81; `p2p_leaves` (near-field form only) adds the particles of
 * rejected source leaves to the intra-leaf sum. */
void local_to_leaf(const GravCtx& c, const Local& local, int treePointer, int nParticles,
                   const std::vector<int>& p2p_leaves) {
    const int* tree = c.t->octree;
    const orc_state* s = c.s;
    const bool fixed = c.prm->fixes & SPH_FIX_SOFTENING;
    const bool near = c.prm->fixes & SPH_FIX_NEAR_FIELD;
    int start = treePointer + 2;
    int end = start + nParticles;
    for (int i = start; i < end; i++) {
        int pp = tree[i];
        if (!s->active[pp]) continue;
        f3 x_i = ld3(s->pos, pp);
        f3 a = L2P(local, x_i, c.order);
        for (int j = start; j < end; j++) {
            if (i == j) continue;
            int p_j = tree[j];
            f3 x_ij = x_i - ld3(s->pos, p_j);
            const float r = len(x_ij);
            if (r == 0) continue;
            const f3 r_hat = x_ij / r;
            const float f = near ? dphi_dr_m2l_kernel(r, c.prm->gravity_softening)
                                 : dphi_dr(x_ij, c.prm->gravity_softening, fixed);
            a -= (s->m[p_j] * f) * r_hat;
        }
        for (int leaf : p2p_leaves) {
            for (int j = leaf + 2; j < leaf + 2 + tree[leaf]; j++) {
                int p_j = tree[j];
                f3 x_ij = x_i - ld3(s->pos, p_j);
                const float r = len(x_ij);
                if (r == 0) continue;
                const f3 r_hat = x_ij / r;
                a -= (s->m[p_j] * dphi_dr_m2l_kernel(r, c.prm->gravity_softening)) * r_hat;
            }
        }
        st3(s->acc, pp, c.prm->gravity_g * a);
        s->gravnorm[pp] = len(a);
        if (c.acct) {
            Account& ac = *c.acct;
            int64_t n_acc = local.acct_n + (nParticles - 1);
            double m_acc = local.acct_m;
            for (int j = start; j < end; j++)
                if (j != i) m_acc += (double)s->m[tree[j]];
            for (int leaf : p2p_leaves) {
                n_acc += tree[leaf];
                for (int j = leaf + 2; j < leaf + 2 + tree[leaf]; j++) m_acc += (double)s->m[tree[j]];
            }
            ac.out_n[pp] = n_acc;
            ac.out_m[pp] = m_acc;
            ac.out_excess_n[pp] = local.acct_excess_n;
            ac.out_excess_m[pp] = local.acct_excess_m;
        }
    }
}

/* L2L.metal:90-104.  A missing child (-1) makes the reference read
 * treeStructure[0] and overwrite the root's local (never read afterwards);
 * the restatement skips it. */
void L2L(const GravCtx& c, const Local& local, int treePointer) {
    const int* tree = c.t->octree;
    for (int j = treePointer + 2; j < treePointer + 10; j++) {
        int nodePointer = tree[j];
        if (nodePointer == -1) continue;
        int ndp = tree[nodePointer + 1];
        Multipole nodeMp = c.mp_at(ndp);
        if (!nodeMp.active) continue;
        f3 r = nodeMp.pos - local.pos;
        store_lc(c.g->locals + LC_REC * (int64_t)ndp, transform_local(local, r, c.order), c.order);
        if (c.acct) {
            Account& ac = *c.acct;
            ac.local_n[ndp] = local.acct_n;
            ac.local_m[ndp] = local.acct_m;
            ac.local_excess_n[ndp] = local.acct_excess_n;
            ac.local_excess_m[ndp] = local.acct_excess_m;
        }
    }
}

/* helpers.h:15-32 */
void zero_first_branches(const GravCtx& c, int32_t* out) {
    const int* tree = c.t->octree;
    for (int i = 2; i < 10; i++) {
        int nodePointer = tree[i];
        if (nodePointer == -1) continue;
        int ndp = tree[nodePointer + 1];
        Local l;
        l.pos = c.mp_at(ndp).pos;
        for (int k = 0; k < NT3; ++k) l.e[k] = 0;
        store_lc(c.g->locals + LC_REC * (int64_t)ndp, l, c.order);
    }
    out[0] = 0;
    out[1] = -1;
}

/* gravity.metal:64-143 */
void down_pass_node(const GravCtx& c, uint32_t index, int treePointer, const int32_t* in, int32_t* out) {
    const int* tree = c.t->octree;
    const uint32_t MAXU = (uint32_t)c.prm->max_unchecked_pointers;
    if (treePointer == 0) {
        if (tree[0] != 0) {
            /* a root that is itself a leaf (n <= MAX_CHILDREN_IN_LEAF):
             * zeroFirstBranches (helpers.h:15-32) would read its particle
             * indices as child pointers.  Restated as the GPU path handles it:
             * the intra-leaf sum with a zero local expansion at the root. */
            Local root;
            root.pos = c.mp_at(tree[1]).pos;
            root.self_node = tree[1];
            for (int k = 0; k < NT3; ++k) root.e[k] = 0;
            std::vector<int> none;
            local_to_leaf(c, root, 0, tree[0], none);
            out[0] = -1;
            return;
        }
        zero_first_branches(c, out);
        return;
    }
    int dataPointer = tree[treePointer + 1];
    Multipole mp = c.mp_at(dataPointer);
    if (!mp.active) return;
    uint32_t outputPointer = index * MAXU;
    uint32_t maxOutputPointer = (index + 1) * MAXU;
    int nParticles = tree[treePointer];
    if (nParticles == 0) {
        out[outputPointer] = treePointer;
        outputPointer++;
    }
    uint32_t parentIndex = c.g->parent_index[dataPointer];
    uint32_t startScan = parentIndex * MAXU;
    uint32_t endScan = (parentIndex + 1) * MAXU;
    Local local = load_lc(c.g->locals + LC_REC * (int64_t)dataPointer, c.order);
    local.self_node = dataPointer;
    uint32_t scanEnd = endScan;
    int repeatFirst = 0;
    if (c.acct) {
        Account& ac = *c.acct;
        local.acct_n = ac.local_n[dataPointer];
        local.acct_m = ac.local_m[dataPointer];
        local.acct_excess_n = ac.local_excess_n[dataPointer];
        local.acct_excess_m = ac.local_excess_m[dataPointer];
        if (ac.plant == 1) { /* planted mistake: drop the last list entry */
            uint32_t e = startScan;
            while (e < endScan && in[e] != -1) ++e;
            if (e > startScan) scanEnd = e - 1;
        }
        if (ac.plant == 2 && nParticles != 0) repeatFirst = 1; /* planted mistake: a leaf scans its first entry twice */
    }
    const bool near = c.prm->fixes & SPH_FIX_NEAR_FIELD;
    std::vector<int> p2p_leaves;
    for (uint32_t i = startScan; i < scanEnd; i++) {
        int toScan = in[i];
        if (toScan == -1) break;
        for (int rep = 0; rep <= (i == startScan ? repeatFirst : 0); ++rep) {
            if (nParticles != 0) {
                if (near) near_scan(c, mp, local, toScan, treePointer, p2p_leaves);
                else recursive_scan(c, mp, local, toScan, treePointer);
            } else {
                non_recursive_scan(c, out, mp, local, outputPointer, maxOutputPointer, toScan, treePointer);
            }
        }
    }
    if (outputPointer < maxOutputPointer && nParticles == 0) out[outputPointer] = -1;
    if (nParticles == 0) L2L(c, local, treePointer);
    else local_to_leaf(c, local, treePointer, nParticles, p2p_leaves);
}

/* octree.cpp:14-109 */
int build_recursive(const float* pos4, const std::vector<int>& subset, int level, int maxLeafSize, int maxDepth,
                    std::vector<int>& data, std::vector<std::vector<int>>& levels, int& nodeValues) {
    if ((int)subset.size() <= maxLeafSize || level == maxDepth) {
        int nodeStart = (int)data.size();
        if (level >= (int)levels.size()) levels.resize(level + 1);
        levels[level].push_back(nodeStart);
        data.push_back((int)subset.size());
        data.push_back(nodeValues);
        nodeValues += 1;
        for (int pid : subset) data.push_back(pid);
        return nodeStart;
    }
    int nodeStart = (int)data.size();
    for (int i = 0; i < 10; i++) data.push_back(-1);
    data[nodeStart + 0] = 0;
    data[nodeStart + 1] = nodeValues;
    nodeValues += 1;
    if (level >= (int)levels.size()) levels.resize(level + 1);
    levels[level].push_back(nodeStart);
    f3 center = mk(0, 0, 0);
    for (int pid : subset) center += ld3(pos4, pid);
    center = center / (float)subset.size();
    std::vector<std::vector<int>> child(8);
    for (int pid : subset) {
        f3 p = ld3(pos4, pid);
        int oct = 0;
        oct |= (p.x > center.x) ? 1 : 0;
        oct |= (p.y > center.y) ? 2 : 0;
        oct |= (p.z > center.z) ? 4 : 0;
        child[oct].push_back(pid);
    }
    for (int c = 0; c < 8; c++) {
        if (!child[c].empty()) {
            int idx = build_recursive(pos4, child[c], level + 1, maxLeafSize, maxDepth, data, levels, nodeValues);
            data[nodeStart + 2 + c] = idx;
        } else {
            data[nodeStart + 2 + c] = -1;
        }
    }
    return nodeStart;
}

} /* namespace */

/* ======================================================================= */
extern "C" {

uint32_t orc_morton3d(uint32_t x, uint32_t y, uint32_t z) { return morton(x, y, z); }

/* sort.metal:15-44 */
void orc_hash(const sph_params* prm, const orc_state* s) {
    const float cellSize = prm->cell_width;
    const int cellsPerDim = 1 << prm->cell_power;
#pragma omp parallel for schedule(static)
    for (int index = 0; index < s->n; ++index) {
        f3 normalised = ld3(s->pos, index) / cellSize;
        int px = (int)floorf(normalised.x) % cellsPerDim;
        int py = (int)floorf(normalised.y) % cellsPerDim;
        int pz = (int)floorf(normalised.z) % cellsPerDim;
        if (px < 0) px += cellsPerDim;
        if (py < 0) py += cellsPerDim;
        if (pz < 0) pz += cellsPerDim;
        s->cell_array[2 * index] = morton((uint32_t)px, (uint32_t)py, (uint32_t)pz);
        s->cell_array[2 * index + 1] = (uint32_t)index;
    }
}

/* sort.metal:46-129 with the ping-pong of Compute.cpp:276-312.  The four
 * 8-bit passes leave the result in cell_array (an even number of passes). */
void orc_radix_sort(const sph_params* prm, const orc_state* s) {
    (void)prm;
    const int n = s->n;
    const int B = 256;
    const int nBlocks = n / B + 1;
    std::vector<uint32_t> other(2 * (size_t)n), particleOffsets(n), bucketHist((size_t)nBlocks * 256), bucketOffset(256);
    uint32_t* bufs[2] = {s->cell_array, other.data()};
    for (int it = 0; it < 4; ++it) {
        const uint32_t* src = bufs[it % 2];
        uint32_t* dst = bufs[(it + 1) % 2];
        const uint32_t shift = it * 8;
        /* hist */
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nBlocks; ++b) {
            uint32_t localHist[256];
            for (int i = 0; i < 256; i++) localHist[i] = 0;
            for (int i = B * b; i < std::min(B * (b + 1), n); i++) {
                uint32_t d = (src[2 * i] >> shift) & 255u;
                particleOffsets[i] = localHist[d];
                localHist[d]++;
            }
            for (int i = 0; i < 256; i++) bucketHist[(size_t)b * 256 + i] = localHist[i];
        }
        /* scan */
#pragma omp parallel for schedule(static)
        for (int d = 0; d < 256; ++d) {
            uint32_t count = 0;
            for (int i = 0; i < nBlocks; i++) {
                count += bucketHist[(size_t)i * 256 + d];
                bucketHist[(size_t)i * 256 + d] = count;
            }
            if (d < 255) bucketOffset[d + 1] = count;
            else bucketOffset[0] = 0;
        }
        /* sum */
        uint32_t count = 0;
        for (int i = 0; i < 256; i++) {
            count += bucketOffset[i];
            bucketOffset[i] = count;
        }
        /* sort */
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            uint32_t d = (src[2 * i] >> shift) & 255u;
            int block = i / B;
            uint32_t bOff = bucketOffset[d];
            uint32_t blockOffset = block == 0 ? 0 : bucketHist[(size_t)(block - 1) * 256 + d];
            uint32_t dstIndex = bOff + particleOffsets[i] + blockOffset;
            dst[2 * dstIndex] = src[2 * i];
            dst[2 * dstIndex + 1] = src[2 * i + 1];
        }
    }
}

/* sort.metal:131-148: initialise + atomic min/max of the sorted slot. */
void orc_find_cell_positions(const sph_params* prm, const orc_state* s) {
    const int nc = n_cells(prm);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; ++c) {
        s->cell_start[c] = 0xFFFFFFFFu;
        s->cell_end[c] = 0;
    }
    for (int i = 0; i < s->n; ++i) {
        uint32_t cell = s->cell_array[2 * i];
        s->cell_start[cell] = std::min(s->cell_start[cell], (uint32_t)i);
        s->cell_end[cell] = std::max(s->cell_end[cell], (uint32_t)i);
    }
}

/* Compute::sort (Compute.cpp:265-326) */
void orc_sort(const sph_params* prm, const orc_state* s) {
    orc_hash(prm, s);
    orc_radix_sort(prm, s);
    orc_find_cell_positions(prm, s);
}

float orc_W(float r, float h1) { return W(r, h1); }
void orc_gradW(const float* x, float r, float r1, float h1, float* out) {
    f3 g = gradW(mk(x[0], x[1], x[2]), r, r1, h1);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
float orc_dW_dh(float r, float h1) { return dW_dh(r, h1); }
float orc_dphi_dr(const float* x, float h, int fixed) { return dphi_dr(mk(x[0], x[1], x[2]), h, fixed != 0); }
void orc_cell_range(const sph_params* prm, const float* p, float h, int32_t* out) {
    CellRange r = cell_range(mk(p[0], p[1], p[2]), prm->cell_width, h, 1 << prm->cell_power);
    for (int k = 0; k < 3; ++k) { out[k] = r.mn[k]; out[3 + k] = r.mx[k]; }
}
void orc_eos_sample(const sph_params* prm, const float* table, int32_t res, float u, float rho, float* out2) {
    eos_sample(prm, table, res, u, rho, &out2[0], &out2[1]);
}

/* density.metal:12-239, one thread per sorted slot. */
/* The stopping test of the iteration after a scan with relative change eps
 * (eps_prev: the previous scan's, 3e38 before the second).  Rule 0:
 * density.metal:73,131-136, successive iterates within eta.  Rule 1 (not in
 * the reference; sph_params.density_stop): the estimated distance from the
 * fixed point, eps / (1 - k) with k = eps / eps_prev (one ulp of slack on
 * each), within eta, or eps at the rounding floor; never stops on a first
 * scan or while k >= SPH_NR_MAX_CONTRACTION. */
static bool nr_goes_on(int rule, float eps, float eps_prev, float eta) {
    if (rule == 0) return eps > eta;
    if (eps <= SPH_NR_NOISE_FLOOR) return false;
    if (eps_prev > 1.0e38f) return true;
    const float e = eps + SPH_NR_EPS_ULP; /* one ulp of slack on each eps */
    const float k = e / (eps_prev - SPH_NR_EPS_ULP);
    if (k >= SPH_NR_MAX_CONTRACTION) return true;
    return e > eta * (1.f - k);
}

void orc_density(const sph_params* prm, const orc_state* s) {
    const int n = s->n;
    const float cellSize = prm->cell_width;
    const float eta = prm->density_eta;
    const int maxIt = prm->max_nr_iterations;
    const int NE = prm->n_neighbours_estim;
    const bool fix_cache = prm->fixes & SPH_FIX_DENSITY_CACHE;
    /* thread 0: *dt = DT0 * DT_MIN (density.metal:34-36) */
    *s->dt = (int)(prm->dt0 * (1.f / dt_min1(prm)));
#pragma omp parallel
    {
        std::vector<uint32_t> neighbourIndex(NE > 0 ? NE : 1);
        std::vector<float> rs(NE > 0 ? NE : 1);
#pragma omp for schedule(dynamic, 256)
        for (int index = 0; index < n; ++index) {
            int ind = (int)s->cell_array[2 * index + 1];
            const float* table = eos_table_for(s, s->mat[ind]);
            if (!s->active[ind]) {
                eos_sample(prm, table, s->eos_res, s->u[ind], s->rho[ind], &s->pressure[ind], &s->soundspeed[ind]);
                if (s->nr_iters) s->nr_iters[ind] = 0;
                continue;
            }
            f3 position = ld3(s->pos, ind);
            float h_i = s->h[ind];
            int max_h = 0;
            float h2x4 = 4 * h_i * h_i;
            float h1 = 1 / h_i;
            float mass_i = s->m[ind];
            float density = 0;
            float eps = 1;
            int count = 0;
            int nNeighbours = 0;
            float eps_prev = 3.0e38f;
            bool goOn = prm->density_stop ? true : eps > eta;
            while (goOn && count < maxIt) {
                density = 0;
                if (h_i < max_h && nNeighbours < NE) {
                    for (int i = 0; i < nNeighbours; i++) {
                        const uint32_t j = neighbourIndex[i];
                        float r = rs[i];
                        if (r > h_i + h_i) continue;
                        float W_ij = W(r, h1);
                        density += W_ij * s->m[j];
                    }
                } else {
                    CellRange range = cell_range(position, cellSize, h_i, 1 << prm->cell_power);
                    nNeighbours = 0;
                    for (int x = range.mn[0]; x <= range.mx[0]; x++)
                        for (int y = range.mn[1]; y <= range.mx[1]; y++)
                            for (int z = range.mn[2]; z <= range.mx[2]; z++) {
                                uint32_t ci = cell_index(prm, x, y, z);
                                uint32_t start = s->cell_start[ci];
                                if (start == 0xFFFFFFFFu) continue;
                                uint32_t end = s->cell_end[ci] + 1;
                                for (uint32_t k = start; k < end; k++) {
                                    const uint32_t j = s->cell_array[2 * k + 1];
                                    f3 x_ij = ld3(s->pos, j) - position;
                                    float r2 = len2(x_ij);
                                    if (r2 > h2x4) continue;
                                    float r = sqrtf(r2);
                                    if (nNeighbours < NE) {
                                        neighbourIndex[nNeighbours] = j;
                                        rs[nNeighbours] = r;
                                        nNeighbours++;
                                    }
                                    float W_ij = W(r, h1);
                                    density += W_ij * s->m[j];
                                }
                            }
                }
                float newH = std::min(prm->resolution_eta * powf(mass_i / density, 1.f / 3), prm->max_smoothing_length);
                newH = 0.5f * (h_i + newH);
                eps = fabsf(newH - h_i) / h_i;
                goOn = nr_goes_on(prm->density_stop, eps, eps_prev, eta);
                eps_prev = eps;
                if (goOn) {
                    max_h = (int)std::fmax((float)max_h, h_i);
                    h_i = newH;
                    h1 = 1 / h_i;
                    h2x4 = 4 * h_i * h_i;
                    count++;
                }
            }
            if (s->nr_iters) s->nr_iters[ind] = count;
            float density_h = 0;
            float velDiv = 0;
            f3 velCurl = mk(0, 0, 0);
            f3 v_i = ld3(s->vel, ind);
            if (nNeighbours < NE) {
                for (int i = 0; i < nNeighbours; i++) {
                    const uint32_t j = neighbourIndex[i];
                    float r = rs[i];
                    if (r > h_i + h_i) continue;
                    float mass = s->m[j];
                    density_h += mass * dW_dh(r, h1);
                    if (r > 1e-12f) {
                        float r1 = 1 / r;
                        f3 v_ij = ld3(s->vel, j) - v_i;
                        f3 x_ij = ld3(s->pos, j) - position;
                        f3 gW = gradW(x_ij, r, r1, h1);
                        velDiv += mass * dot(v_ij, gW);
                        velCurl += mass * cross(v_ij, gW);
                    }
                }
            } else {
                /* density.metal:173-199: `density` is not reset before this
                 * rescan, so it ends up counted twice (fixed form resets it). */
                if (fix_cache) density = 0;
                CellRange range = cell_range(position, cellSize, h_i, 1 << prm->cell_power);
                for (int x = range.mn[0]; x <= range.mx[0]; x++)
                    for (int y = range.mn[1]; y <= range.mx[1]; y++)
                        for (int z = range.mn[2]; z <= range.mx[2]; z++) {
                            uint32_t ci = cell_index(prm, x, y, z);
                            uint32_t start = s->cell_start[ci];
                            if (start == 0xFFFFFFFFu) continue;
                            uint32_t end = s->cell_end[ci] + 1;
                            for (uint32_t k = start; k < end; k++) {
                                const uint32_t j = s->cell_array[2 * k + 1];
                                f3 x_ij = ld3(s->pos, j) - position;
                                float r2 = len2(x_ij);
                                if (r2 > h2x4) continue;
                                float r = sqrtf(r2);
                                float W_ij = W(r, h1);
                                float mass = s->m[j];
                                density += W_ij * mass;
                                density_h += mass * dW_dh(r, h1);
                                if (r > 1e-12f) {
                                    float r1 = 1 / r;
                                    f3 v_ij = ld3(s->vel, j) - v_i;
                                    f3 gW = gradW(x_ij, r, r1, h1);
                                    velDiv += mass * dot(v_ij, gW);
                                    velCurl += mass * cross(v_ij, gW);
                                }
                            }
                        }
            }
            velDiv = fabsf(velDiv) / density;
            float absCurl = len(velCurl) / density;
            s->h[ind] = h_i;
            s->rho[ind] = density;
            s->gradterm[ind] = 1 / (1 + (h_i / (3 * density)) * density_h);
            float Pp, cc;
            eos_sample(prm, table, s->eos_res, s->u[ind], density, &Pp, &cc);
            s->pressure[ind] = Pp;
            s->soundspeed[ind] = cc;
            s->balsara[ind] = velDiv / (velDiv + absCurl + 1e-4f * (cc / h_i));
        }
    }
}

/* accelerations.metal:12-141 */
void orc_acceleration(const sph_params* prm, const orc_state* s) {
    const int n = s->n;
    const float cellSize = prm->cell_width;
    const float dtmin = 1.f / dt_min1(prm);
    const int gt = *s->global_time;
    int dtGlobal = *s->dt;
#pragma omp parallel for schedule(dynamic, 256) reduction(min : dtGlobal)
    for (int index = 0; index < n; ++index) {
        int ind = (int)s->cell_array[2 * index + 1];
        if (!s->active[ind]) continue;
        float f_i = s->gradterm[ind];
        float P_i = s->pressure[ind];
        float rho_i = s->rho[ind];
        float fact_i = (f_i * P_i / (rho_i * rho_i));
        f3 x_i = ld3(s->pos, ind);
        f3 v_i = ld3(s->vel, ind);
        float c_i = s->soundspeed[ind];
        float h_i = s->h[ind];
        float B_i = s->balsara[ind];
        float hi1 = 1 / h_i;
        float hi2x4 = h_i * h_i * 4;
        f3 dv_dt = mk(0, 0, 0);
        float du_dt = 0;
        float dhdt = 0;
        CellRange range = cell_range(x_i, cellSize, std::min(h_i, prm->max_smoothing_length), 1 << prm->cell_power);
        float v_sigi = 1e-12f;
        for (int x = range.mn[0]; x <= range.mx[0]; x++)
            for (int y = range.mn[1]; y <= range.mx[1]; y++)
                for (int z = range.mn[2]; z <= range.mx[2]; z++) {
                    uint32_t ci = cell_index(prm, x, y, z);
                    uint32_t start = s->cell_start[ci];
                    if (start == 0xFFFFFFFFu) continue;
                    uint32_t end = s->cell_end[ci] + 1;
                    for (uint32_t k = start; k < end; k++) {
                        const uint32_t j = s->cell_array[2 * k + 1];
                        f3 x_j = ld3(s->pos, j);
                        f3 x_ij = x_i - x_j;
                        float r2 = len2(x_ij);
                        float h_j = s->h[j];
                        float hj2x4 = h_j * h_j * 4;
                        if ((r2 > hi2x4 && r2 > hj2x4) || r2 < 1e-12f) continue;
                        float r = sqrtf(r2);
                        float r1 = 1 / r;
                        float m_j = s->m[j];
                        float f_j = s->gradterm[j];
                        float P_j = s->pressure[j];
                        float rho_j = s->rho[j];
                        float fact_j = (f_j * P_j / (rho_j * rho_j));
                        f3 v_j = ld3(s->vel, j);
                        float c_j = s->soundspeed[j];
                        float hj1 = 1 / h_j;
                        f3 v_ij = v_i - v_j;
                        float rho_ij = 0.5f * (rho_i + rho_j);
                        const float v_dot_r = dot(v_ij, x_ij);
                        float mu_ij = v_dot_r < 0 ? v_dot_r * r1 : 0;
                        float v_sigij = c_i + c_j - 3 * mu_ij;
                        v_sigi = std::max(v_sigij, v_sigi);
                        float B_j = s->balsara[j];
                        float B_ij = 0.5f * (B_i + B_j);
                        float nu_ij = -0.5f * prm->viscosity_alpha * B_ij * mu_ij * v_sigij / rho_ij;
                        f3 gW_i = gradW(x_ij, r, r1, hi1);
                        f3 gW_j = gradW(x_ij, r, r1, hj1);
                        f3 gW_ij = 0.5f * (gW_i + gW_j);
                        f3 t_i = fact_i * gW_i;
                        f3 t_j = fact_j * gW_j;
                        f3 t_visc = nu_ij * gW_ij;
                        dv_dt -= m_j * (t_i + t_j + t_visc);
                        du_dt += m_j * (fact_i * dot(v_ij, gW_i) + 0.5f * nu_ij * dot(v_ij, gW_ij));
                        dhdt += (m_j / rho_j) * dot(v_ij, gW_i);
                    }
                }
        f3 a = ld3(s->acc, ind) + dv_dt;
        st3(s->acc, ind, a);
        s->du[ind] = du_dt;
        s->dhdt[ind] = dhdt;
        float goaldt = 2.0f * prm->cfl * prm->gamma * h_i / v_sigi;
        /* the corrected form also bounds |dh/dt| dt / h by SPH_DLNH_LIMIT */
        if (prm->fixes & SPH_FIX_INACTIVE_DRIFT) goaldt = std::min(goaldt, SPH_DLNH_LIMIT * h_i / fabsf(dhdt));
        int r = 0;
        float dtCandidate = prm->dt0;
        while (dtCandidate > goaldt && r < prm->r_max) {
            dtCandidate *= 0.5f;
            r++;
        }
        int integerDt = (int)ceilf(dtmin * dtCandidate);
        s->next_active[ind] = gt + integerDt;
        dtGlobal = std::min(dtGlobal, integerDt);
    }
    *s->dt = dtGlobal;
}

/* step.metal:12-46.  With a rung cap (max_rungs > 0) each active particle's
 * tick step is clamped to dt_min << (max_rungs - 1) once the global minimum is
 * known (a new knob; 0 reproduces the reference). */
void orc_step(const sph_params* prm, const orc_state* s) {
    const int dtInt = *s->dt;
    const float dt = dt_min1(prm) * (float)dtInt;
    const int gt = *s->global_time;
    const int cap = prm->max_rungs > 0 ? (dtInt << (prm->max_rungs - 1)) : 0;
#pragma omp parallel for schedule(static)
    for (int index = 0; index < s->n; ++index) {
        f3 velocity = ld3(s->vel, index);
        const f3 acceleration = ld3(s->acc, index);
        float h_i = s->h[index];
        float h1 = 1 / h_i;
        float dhdt = s->dhdt[index];
        float rate = h1 * dhdt;
        if (prm->fixes & SPH_FIX_INACTIVE_DRIFT) {
            /* the corrected form: inactive particles keep their last log-rate */
            if (s->active[index]) s->dlnh[index] = rate;
            else rate = s->dlnh[index];
        }
        s->h[index] = std::min(h_i * expf(rate * dt), prm->max_smoothing_length);
        s->rho[index] *= expf(-3 * rate * dt);
        st3(s->vel, index, velocity + dt * acceleration);
        s->u[index] += dt * s->du[index];
        st3(s->pos, index, ld3(s->pos, index) + (dt * velocity + 0.5f * acceleration * dt * dt));
        if (cap > 0 && s->active[index] && s->next_active[index] - gt > cap) s->next_active[index] = gt + cap;
    }
    *s->global_time += dtInt;
}

int orc_build_octree(const float* pos4, int32_t n, int32_t max_leaf, int32_t max_depth, orc_tree* out) {
    std::vector<int> data;
    std::vector<std::vector<int>> levels;
    int nodeValues = 0;
    std::memset(out, 0, sizeof(*out));
    if (n <= 0) return 0;
    /* not in octree.cpp: a non-finite coordinate would make every mean split
     * NaN and put all particles into one depth-limited leaf */
    for (int64_t i = 0; i < 4 * (int64_t)n; ++i)
        if ((i & 3) != 3 && !std::isfinite(pos4[i])) return -1;
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    build_recursive(pos4, all, 0, max_leaf, max_depth, data, levels, nodeValues);
    out->n_ints = (int64_t)data.size();
    out->octree = (int32_t*)std::malloc(sizeof(int32_t) * data.size());
    std::memcpy(out->octree, data.data(), sizeof(int32_t) * data.size());
    out->n_levels = (int32_t)levels.size();
    out->level_counts = (int32_t*)std::malloc(sizeof(int32_t) * levels.size());
    size_t total = 0;
    for (auto& l : levels) total += l.size();
    out->level_nodes = (int32_t*)std::malloc(sizeof(int32_t) * (total ? total : 1));
    size_t off = 0;
    for (size_t l = 0; l < levels.size(); ++l) {
        out->level_counts[l] = (int32_t)levels[l].size();
        std::memcpy(out->level_nodes + off, levels[l].data(), sizeof(int32_t) * levels[l].size());
        off += levels[l].size();
    }
    out->node_count = nodeValues;
    return 0;
}

/* Octree over Morton keys (the build's GPU tree, not the reference's): the
 * positions' bounding cube is quantised to 2^10 cells per axis (FP64, no
 * fused operations), each particle gets the 30-bit key x | y << 1 | z << 2
 * interleaved, particles are sorted by (key, index), and a node at level L
 * holds the particles sharing the key's top 3 L bits.  A node is a leaf when
 * it holds at most max_leaf particles or L == max_depth; children are the
 * octants of level L + 1 in key order.  The flat array, the pre-order node
 * numbering and the level lists are laid out as octree.cpp lays out its own
 * tree. */
namespace {
uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

int morton_build(const std::vector<uint32_t>& key, const std::vector<int>& idx, int s, int e, int level,
                 int maxLeaf, int maxDepth, std::vector<int>& data, std::vector<std::vector<int>>& levels,
                 int& nodeValues) {
    const int nodeStart = (int)data.size();
    if (level >= (int)levels.size()) levels.resize(level + 1);
    levels[level].push_back(nodeStart);
    if (e - s <= maxLeaf || level == maxDepth) {
        data.push_back(e - s);
        data.push_back(nodeValues++);
        for (int j = s; j < e; ++j) data.push_back(idx[j]);
        return nodeStart;
    }
    for (int i = 0; i < 10; i++) data.push_back(-1);
    data[nodeStart] = 0;
    data[nodeStart + 1] = nodeValues++;
    const int shift = 30 - 3 * (level + 1);
    int j = s;
    for (int c = 0; c < 8; ++c) {
        const int b = j;
        while (j < e && (int)((key[j] >> shift) & 7u) == c) ++j;
        if (j > b) data[nodeStart + 2 + c] = morton_build(key, idx, b, j, level + 1, maxLeaf, maxDepth, data, levels,
                                                          nodeValues);
    }
    return nodeStart;
}
}  // namespace

int orc_build_octree_morton(const float* pos4, int32_t n, int32_t max_leaf, int32_t max_depth, orc_tree* out) {
    std::memset(out, 0, sizeof(*out));
    if (n <= 0) return 0;
    if (max_depth < 0 || max_depth > 10 || max_leaf < 1) return -2;
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = pos4[a];
    for (int64_t i = 0; i < (int64_t)n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float v = pos4[4 * i + a];
            if (!std::isfinite(v)) return -1;
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
        }
    float ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(ext > 0.f)) ext = 1.f;
    const double scale = 1024.0 / (double)ext;
    std::vector<uint32_t> key0(n);
    for (int i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            const double t = ((double)pos4[4 * i + a] - (double)lo[a]) * scale;
            const int v = (int)std::floor(t);
            q[a] = (uint32_t)std::min(std::max(v, 0), 1023);
        }
        key0[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key0[a] < key0[b]; });
    std::vector<uint32_t> key(n);
    for (int i = 0; i < n; ++i) key[i] = key0[idx[i]];
    std::vector<int> data;
    std::vector<std::vector<int>> levels;
    int nodeValues = 0;
    morton_build(key, idx, 0, n, 0, max_leaf, max_depth, data, levels, nodeValues);
    out->n_ints = (int64_t)data.size();
    out->octree = (int32_t*)std::malloc(sizeof(int32_t) * data.size());
    std::memcpy(out->octree, data.data(), sizeof(int32_t) * data.size());
    out->n_levels = (int32_t)levels.size();
    out->level_counts = (int32_t*)std::malloc(sizeof(int32_t) * levels.size());
    size_t total = 0;
    for (auto& l : levels) total += l.size();
    out->level_nodes = (int32_t*)std::malloc(sizeof(int32_t) * total);
    size_t off = 0;
    for (size_t l = 0; l < levels.size(); ++l) {
        out->level_counts[l] = (int32_t)levels[l].size();
        std::memcpy(out->level_nodes + off, levels[l].data(), sizeof(int32_t) * levels[l].size());
        off += levels[l].size();
    }
    out->node_count = nodeValues;
    return 0;
}

void orc_free_tree(orc_tree* t) {
    std::free(t->octree);
    std::free(t->level_nodes);
    std::free(t->level_counts);
    std::memset(t, 0, sizeof(*t));
}

/* gravity.metal:36-62 over levels deepest..0 (Compute.cpp:330-351) */
void orc_gravity_up(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g) {
    GravCtx c{prm, s, t, g};
    std::vector<int64_t> off(t->n_levels + 1, 0);
    for (int l = 0; l < t->n_levels; ++l) off[l + 1] = off[l] + t->level_counts[l];
    for (int l = t->n_levels - 1; l >= 0; --l) {
        const int32_t* pointers = t->level_nodes + off[l];
#pragma omp parallel for schedule(dynamic, 64)
        for (int index = 0; index < t->level_counts[l]; ++index) {
            int treePointer = pointers[index];
            int nParticles = t->octree[treePointer];
            int dataPointer = t->octree[treePointer + 1];
            Multipole mp = nParticles == 0 ? M2M(c, (uint32_t)index, treePointer) : P2M(c, treePointer);
            store_mp(g->multipoles + MP_REC * (int64_t)dataPointer, mp, c.order);
        }
    }
}

/* gravity.metal:64-143 over levels 0..deepest with the list ping-pong of
 * Compute.cpp:353-376 */
static void gravity_down(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g,
                         Account* acct) {
    GravCtx c{prm, s, t, g};
    c.acct = acct;
    std::vector<int64_t> off(t->n_levels + 1, 0);
    for (int l = 0; l < t->n_levels; ++l) off[l + 1] = off[l] + t->level_counts[l];
    for (int l = 0; l < t->n_levels; ++l) {
        const int32_t* pointers = t->level_nodes + off[l];
        const int32_t* in = (l % 2 == 0) ? g->list_a : g->list_b;
        int32_t* out = (l % 2 == 0) ? g->list_b : g->list_a;
#pragma omp parallel for schedule(dynamic, 16)
        for (int index = 0; index < t->level_counts[l]; ++index)
            down_pass_node(c, (uint32_t)index, pointers[index], in, out);
    }
}

void orc_gravity_down(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g) {
    gravity_down(prm, s, t, g, nullptr);
}

/* Subtree particle counts and FP64 masses and the parent of every node, by a
 * walk of the flat tree (octree.cpp:14-141 layout). */
static void account_nodes(const orc_state* s, const orc_tree* t, int treePointer, int parent, Account& a) {
    const int* tree = t->octree;
    const int dp = tree[treePointer + 1];
    a.parent[dp] = parent;
    const int np = tree[treePointer];
    if (np != 0) {
        a.node_n[dp] = np;
        double m = 0.0;
        for (int j = treePointer + 2; j < treePointer + 2 + np; ++j) m += (double)s->m[tree[j]];
        a.node_m[dp] = m;
        return;
    }
    int64_t n = 0;
    double m = 0.0;
    for (int j = treePointer + 2; j < treePointer + 10; ++j) {
        const int child = tree[j];
        if (child == -1) continue;
        account_nodes(s, t, child, dp, a);
        n += a.node_n[tree[child + 1]];
        m += a.node_m[tree[child + 1]];
    }
    a.node_n[dp] = n;
    a.node_m[dp] = m;
}

int orc_gravity_account(const sph_params* prm, const orc_state* s, const orc_tree* t, orc_gravity_buffers* g,
                        int plant, int64_t* count_out, double* mass_out, int64_t* excess_count_out,
                        double* excess_mass_out) {
    Account a;
    const size_t K = (size_t)t->node_count;
    a.node_n.assign(K, 0);
    a.node_m.assign(K, 0.0);
    a.parent.assign(K, -1);
    a.local_n.assign(K, 0);
    a.local_excess_n.assign(K, 0);
    a.local_m.assign(K, 0.0);
    a.local_excess_m.assign(K, 0.0);
    a.out_n = count_out;
    a.out_m = mass_out;
    a.out_excess_n = excess_count_out;
    a.out_excess_m = excess_mass_out;
    a.plant = plant;
    for (int32_t i = 0; i < s->n; ++i) {
        count_out[i] = -1;
        mass_out[i] = 0.0;
        excess_count_out[i] = 0;
        excess_mass_out[i] = 0.0;
    }
    account_nodes(s, t, 0, -1, a);
    gravity_down(prm, s, t, g, &a);
    return (int)a.unexpected.load();
}

void orc_pole_m2l(float* local16, const float* mp24, float eps) {
    Local l = load_lc(local16);
    Local add = M2L(l.pos, load_mp(mp24), eps);
    for (int j = 0; j < NT; ++j) l.e[j] += add.e[j];
    store_lc(local16, l);
}

void orc_pole_l2l(const float* local16, const float* r3, float* out16) {
    store_lc(out16, transform_local(load_lc(local16), mk(r3[0], r3[1], r3[2])));
}

void orc_pole_l2p(const float* local16, const float* x3, float* acc3) {
    f3 a = L2P(load_lc(local16), mk(x3[0], x3[1], x3[2]));
    acc3[0] = a.x; acc3[1] = a.y; acc3[2] = a.z;
}

int orc_pole_mac(const float* A24, const float* B24, float eta) {
    return m2l_accept(load_mp(A24), load_mp(B24), eta) ? 1 : 0;
}

void orc_pole_shift_multipole(const float* mp24, const float* r3, int fix_zz, float* out24) {
    Multipole m = transform_multipole(load_mp(mp24), mk(r3[0], r3[1], r3[2]), fix_zz != 0);
    add_powers(m);
    store_mp(out24, m);
}

void orc_pole_derivatives(const float* r3, float eps, float* d10) {
    float d[NT3];
    derivatives(mk(r3[0], r3[1], r3[2]), eps, d);
    for (int i = 0; i < NT; ++i) d10[i] = d[i];
}

/* P = 3 records (40-float multipoles, 32-float locals) */
void orc_pole3_m2l(float* local32, const float* mp40, float eps) {
    Local l = load_lc(local32, 3);
    Local add = M2L(l.pos, load_mp(mp40, 3), eps, 3);
    for (int j = 0; j < NT3; ++j) l.e[j] += add.e[j];
    store_lc(local32, l, 3);
}

void orc_pole3_l2l(const float* local32, const float* r3, float* out32) {
    store_lc(out32, transform_local(load_lc(local32, 3), mk(r3[0], r3[1], r3[2]), 3), 3);
}

void orc_pole3_l2p(const float* local32, const float* x3, float* acc3) {
    f3 a = L2P(load_lc(local32, 3), mk(x3[0], x3[1], x3[2]), 3);
    acc3[0] = a.x; acc3[1] = a.y; acc3[2] = a.z;
}

int orc_pole3_mac(const float* A40, const float* B40, float eta, int fix_p3) {
    return m2l_accept(load_mp(A40, 3), load_mp(B40, 3), eta, 3, fix_p3 != 0) ? 1 : 0;
}

void orc_pole3_shift_multipole(const float* mp40, const float* r3, int fix_zz, int fix_p3, float* out40) {
    Multipole m = transform_multipole(load_mp(mp40, 3), mk(r3[0], r3[1], r3[2]), fix_zz != 0, 3);
    add_powers(m, 3, fix_p3 != 0);
    store_mp(out40, m, 3);
}

void orc_pole3_derivatives(const float* r3, float eps, float* d20) {
    derivatives(mk(r3[0], r3[1], r3[2]), eps, d20, 3);
}

void orc_direct_sum_fp64(const float* pos4, const float* m, int32_t n, const int32_t* targets, int32_t n_targets,
                         double G, double softening, double* acc3_out, double* phi_out) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int t = 0; t < n_targets; ++t) {
        const int i = targets[t];
        const double xi = pos4[4 * i], yi = pos4[4 * i + 1], zi = pos4[4 * i + 2];
        double ax = 0, ay = 0, az = 0, phi = 0;
        const double hh = softening;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = (double)pos4[4 * j] - xi, dy = (double)pos4[4 * j + 1] - yi, dz = (double)pos4[4 * j + 2] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            const double r = std::sqrt(r2);
            double f, p;
            if (hh < 0 && r < -hh) {
                /* the M2L softening (derivatives.metal:12-57), support |hh| */
                const double e = -hh, u = r / e;
                const double d1 = ((((((-3 * u + 15) * u - 28) * u + 21) * u) * u - 7) * u) * u + 3;
                const double d2 = (((((-21 * u + 90) * u - 140) * u + 84) * u) * u - 14) * u;
                f = -d2 / (e * e);
                p = -d1 / e;
            } else if (hh > 0 && r < 2 * hh) {
                const double q = r / hh, h1 = 1 / hh;
                if (q <= 1) {
                    f = h1 * h1 * (4.0 / 3.0 * q - 6.0 / 5.0 * q * q * q + 0.5 * q * q * q * q);
                    p = h1 * (2.0 / 3.0 * q * q - 3.0 / 10.0 * q * q * q * q + 1.0 / 10.0 * q * q * q * q * q - 7.0 / 5.0);
                } else {
                    f = h1 * h1 * (8.0 / 3.0 * q - 3 * q * q + 6.0 / 5.0 * q * q * q - 1.0 / 6.0 * q * q * q * q - 1.0 / (15.0 * q * q));
                    p = h1 * (4.0 / 3.0 * q * q - q * q * q + 3.0 / 10.0 * q * q * q * q - 1.0 / 30.0 * q * q * q * q * q - 8.0 / 5.0 + 1.0 / (15.0 * q));
                }
            } else {
                f = 1.0 / r2;
                p = -1.0 / r;
            }
            const double w = (double)m[j] * f / r;
            ax += w * dx;
            ay += w * dy;
            az += w * dz;
            phi += (double)m[j] * p;
        }
        acc3_out[3 * t] = G * ax;
        acc3_out[3 * t + 1] = G * ay;
        acc3_out[3 * t + 2] = G * az;
        if (phi_out) phi_out[t] = G * phi;
    }
}

} /* extern "C" */
