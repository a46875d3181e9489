/*
 * Snapshot.cpp -- HDF5 initial conditions and snapshots in DataStruct layout
 * (hdfHandler.cpp:12-254), checkpoint / resume of a context, and their C
 * entry points (sph_host.h).
 */
#include <cstdlib>
#include <exception>
#include <cstring>
#include <string>
#include <vector>

#include "Compute.hpp"
#include "Hdf5.hpp"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg, int code = SPH_E_ARG) {
    g_err = msg;
    return code;
}

const char* kGroup = "PartType0";
const char* kCkpt = "Checkpoint";

bool read_vec3(h5::Reader& r, const std::string& name, std::vector<float>& out4, size_t& n, std::string& err) {
    std::vector<float> flat;
    std::vector<uint64_t> dims;
    if (!r.readFloat(std::string(kGroup) + "/" + name, flat, dims)) {
        err = r.error();
        return false;
    }
    /* hdfHandler.cpp:41-50 */
    if (dims.size() != 2 || dims[1] != 3) {
        err = std::string(kGroup) + "/" + name + ": a vector dataset needs shape (n, 3)";
        return false;
    }
    n = (size_t)dims[0];
    out4.assign(4 * n, 0.f);
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) out4[4 * i + a] = flat[3 * i + a];
    return true;
}

bool read_scalar(h5::Reader& r, const std::string& group, const std::string& name, std::vector<float>& out, size_t n,
                 std::string& err) {
    std::vector<uint64_t> dims;
    if (!r.readFloat(group + "/" + name, out, dims)) {
        err = r.error();
        return false;
    }
    if (out.size() != n) {
        err = group + "/" + name + " holds " + std::to_string(out.size()) + " values for " + std::to_string(n) +
              " particles";
        return false;
    }
    return true;
}

bool read_ints(h5::Reader& r, const std::string& group, const std::string& name, std::vector<int32_t>& out, size_t n,
               std::string& err) {
    std::vector<uint64_t> dims;
    if (!r.readInt32(group + "/" + name, out, dims)) {
        err = r.error();
        return false;
    }
    if (out.size() != n) {
        err = group + "/" + name + " holds " + std::to_string(out.size()) + " values for " + std::to_string(n) +
              " particles";
        return false;
    }
    return true;
}

void add_vec3(h5::Writer& w, const std::string& group, const std::string& name, const float* v4, size_t n) {
    std::vector<float> flat(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) flat[3 * i + a] = v4[4 * i + a];
    w.addFloat(group, name, {(uint64_t)n, 3}, flat.data());
}

/* writeHDFFile's datasets (hdfHandler.cpp:232-248) */
void add_part_type0(h5::Writer& w, const DataStruct& d) {
    const size_t n = d.masses.size();
    add_vec3(w, kGroup, "Coordinates", d.positions.data(), n);
    w.addFloat(kGroup, "Densities", {d.densities.size()}, d.densities.data());
    w.addFloat(kGroup, "Entropies", {d.entropies.size()}, d.entropies.data());
    w.addFloat(kGroup, "InternalEnergies", {d.internalEnergy.size()}, d.internalEnergy.data());
    w.addFloat(kGroup, "Masses", {d.masses.size()}, d.masses.data());
    w.addInt32(kGroup, "MaterialIDs", {d.materialIDs.size()}, d.materialIDs.data());
    w.addFloat(kGroup, "Potentials", {d.potentials.size()}, d.potentials.data());
    w.addFloat(kGroup, "Pressures", {d.pressures.size()}, d.pressures.data());
    w.addFloat(kGroup, "SmoothingLengths", {d.smoothingLengths.size()}, d.smoothingLengths.data());
    add_vec3(w, kGroup, "Velocities", d.velocities.data(), n);
    std::vector<int32_t> ids(n);
    for (size_t i = 0; i < n; ++i) ids[i] = (int32_t)i;
    w.addInt32(kGroup, "ParticleIDs", {(uint64_t)n}, ids.data());
}

/* the per-particle float fields a checkpoint carries beyond DataStruct */
struct CkptField {
    int field;
    const char* name;
};
const CkptField kFloatFields[] = {
    {SPH_F_GRADTERM, "GradTerms"}, {SPH_F_SOUNDSPEED, "SoundSpeeds"}, {SPH_F_DU, "InternalEnergyRates"},
    {SPH_F_BALSARA, "BalsaraSwitches"}, {SPH_F_DHDT, "SmoothingLengthRates"}, {SPH_F_GRAVNORM, "GravityNorms"},
    {SPH_F_DLNH, "SmoothingLengthLogRates"},
};

}  // namespace

void setHdfError(const std::string& msg) { g_err = msg; }

bool readHDFFile(const std::string& path, DataStruct& d, std::string& err) {
    h5::Reader r;
    if (!r.open(path)) {
        err = r.error();
        return false;
    }
    size_t n = 0, nv = 0;
    d = DataStruct();
    if (!read_vec3(r, "Coordinates", d.positions, n, err)) return false;
    if (!read_scalar(r, kGroup, "Densities", d.densities, n, err)) return false;
    if (!read_scalar(r, kGroup, "InternalEnergies", d.internalEnergy, n, err)) return false;
    if (!read_scalar(r, kGroup, "Masses", d.masses, n, err)) return false;
    if (!read_ints(r, kGroup, "MaterialIDs", d.materialIDs, n, err)) return false;
    if (!read_scalar(r, kGroup, "Pressures", d.pressures, n, err)) return false;
    if (!read_scalar(r, kGroup, "SmoothingLengths", d.smoothingLengths, n, err)) return false;
    if (!read_vec3(r, "Velocities", d.velocities, nv, err)) return false;
    if (nv != n) {
        err = "Velocities holds " + std::to_string(nv) + " rows for " + std::to_string(n) + " particles";
        return false;
    }
    if (n == 0) {
        err = path + " holds no particles";
        return false;
    }
    return true;
}

bool writeHDFFile(const std::string& path, const DataStruct& d, std::string& err) {
    const size_t n = d.masses.size();
    if (d.positions.size() != 4 * n || d.velocities.size() != 4 * n) {
        err = "positions / velocities do not hold 4 floats per particle";
        return false;
    }
    h5::Writer w;
    add_part_type0(w, d);
    return w.write(path, err);
}

int writeCheckpoint(sph_ctx* ctx, const std::string& path, std::string& err) {
    const int n = sph_num_particles(ctx);
    if (n <= 0) {
        err = "no context";
        return SPH_E_ARG;
    }
    DataStruct d;
    d.positions.resize(4 * (size_t)n);
    d.velocities.resize(4 * (size_t)n);
    d.densities.resize(n);
    d.internalEnergy.resize(n);
    d.masses.resize(n);
    d.materialIDs.resize(n);
    d.pressures.resize(n);
    d.smoothingLengths.resize(n);
    int rc = SPH_OK;
    auto dl = [&](int field, void* dst) {
        if (rc == SPH_OK) rc = sph_download(ctx, field, dst);
    };
    dl(SPH_F_POS, d.positions.data());
    dl(SPH_F_VEL, d.velocities.data());
    dl(SPH_F_RHO, d.densities.data());
    dl(SPH_F_U, d.internalEnergy.data());
    dl(SPH_F_MASS, d.masses.data());
    dl(SPH_F_MAT, d.materialIDs.data());
    dl(SPH_F_PRESSURE, d.pressures.data());
    dl(SPH_F_H, d.smoothingLengths.data());
    std::vector<float> acc(4 * (size_t)n);
    dl(SPH_F_ACC, acc.data());
    std::vector<std::vector<float>> extra;
    for (const auto& f : kFloatFields) {
        extra.emplace_back(n);
        dl(f.field, extra.back().data());
    }
    std::vector<uint8_t> active(n);
    std::vector<int32_t> next(n), gt(1), dt(1);
    dl(SPH_F_ACTIVE, active.data());
    dl(SPH_F_NEXT_ACTIVE, next.data());
    dl(SPH_F_GLOBAL_TIME, gt.data());
    dl(SPH_F_DT, dt.data());
    if (rc != SPH_OK) {
        err = std::string("download: ") + sph_last_error(ctx);
        return rc;
    }
    h5::Writer w;
    add_part_type0(w, d);
    add_vec3(w, kCkpt, "Accelerations", acc.data(), (size_t)n);
    for (size_t k = 0; k < extra.size(); ++k) w.addFloat(kCkpt, kFloatFields[k].name, {(uint64_t)n}, extra[k].data());
    std::vector<int32_t> act(active.begin(), active.end());
    w.addInt32(kCkpt, "Active", {(uint64_t)n}, act.data());
    w.addInt32(kCkpt, "NextActiveTimes", {(uint64_t)n}, next.data());
    w.addInt32(kCkpt, "GlobalTime", {1}, gt.data());
    w.addInt32(kCkpt, "GlobalStep", {1}, dt.data());
    return w.write(path, err) ? SPH_OK : SPH_E_ARG;
}

int readCheckpoint(sph_ctx* ctx, const std::string& path, std::string& err) {
    const int n = sph_num_particles(ctx);
    if (n <= 0) {
        err = "no context";
        return SPH_E_ARG;
    }
    DataStruct d;
    if (!readHDFFile(path, d, err)) return SPH_E_ARG;
    if (d.masses.size() != (size_t)n) {
        err = path + " holds " + std::to_string(d.masses.size()) + " particles, the context " + std::to_string(n);
        return SPH_E_ARG;
    }
    h5::Reader r;
    if (!r.open(path)) {
        err = r.error();
        return SPH_E_ARG;
    }
    std::vector<float> acc;
    size_t na = 0;
    std::vector<uint64_t> dims;
    {
        std::vector<float> flat;
        if (!r.readFloat(std::string(kCkpt) + "/Accelerations", flat, dims) || dims.size() != 2 || dims[0] != (uint64_t)n ||
            dims[1] != 3) {
            err = r.error().empty() ? "bad Checkpoint/Accelerations" : r.error();
            return SPH_E_ARG;
        }
        na = (size_t)n;
        acc.assign(4 * na, 0.f);
        for (size_t i = 0; i < na; ++i)
            for (int a = 0; a < 3; ++a) acc[4 * i + a] = flat[3 * i + a];
    }
    std::vector<std::vector<float>> extra;
    for (const auto& f : kFloatFields) {
        extra.emplace_back();
        if (!read_scalar(r, kCkpt, f.name, extra.back(), (size_t)n, err)) return SPH_E_ARG;
    }
    std::vector<int32_t> act, next, gt, dt;
    if (!read_ints(r, kCkpt, "Active", act, (size_t)n, err) || !read_ints(r, kCkpt, "NextActiveTimes", next, n, err) ||
        !read_ints(r, kCkpt, "GlobalTime", gt, 1, err) || !read_ints(r, kCkpt, "GlobalStep", dt, 1, err))
        return SPH_E_ARG;
    std::vector<uint8_t> active(act.begin(), act.end());
    int rc = SPH_OK;
    auto ul = [&](int field, const void* src) {
        if (rc == SPH_OK) rc = sph_upload(ctx, field, src);
    };
    ul(SPH_F_POS, d.positions.data());
    ul(SPH_F_VEL, d.velocities.data());
    ul(SPH_F_ACC, acc.data());
    ul(SPH_F_RHO, d.densities.data());
    ul(SPH_F_U, d.internalEnergy.data());
    ul(SPH_F_MASS, d.masses.data());
    ul(SPH_F_MAT, d.materialIDs.data());
    ul(SPH_F_PRESSURE, d.pressures.data());
    ul(SPH_F_H, d.smoothingLengths.data());
    for (size_t k = 0; k < extra.size(); ++k) ul(kFloatFields[k].field, extra[k].data());
    ul(SPH_F_ACTIVE, active.data());
    ul(SPH_F_NEXT_ACTIVE, next.data());
    ul(SPH_F_GLOBAL_TIME, gt.data());
    ul(SPH_F_DT, dt.data());
    if (rc != SPH_OK) err = std::string("upload: ") + sph_last_error(ctx);
    return rc;
}

extern "C" {

const char* sph_hdf_last_error(void) { return g_err.c_str(); }

static int read_hdf_impl(const char* path, sph_particles* out) {
    if (!path || !out) return fail("null argument");
    std::memset(out, 0, sizeof(*out));
    DataStruct d;
    std::string err;
    if (!readHDFFile(path, d, err)) return fail(err);
    const size_t n = d.masses.size();
    out->n = (int32_t)n;
    auto dup = [](const void* src, size_t bytes) {
        void* p = std::malloc(bytes ? bytes : 1);
        if (p && bytes) std::memcpy(p, src, bytes);
        return p;
    };
    out->pos4 = (float*)dup(d.positions.data(), 16 * n);
    out->vel4 = (float*)dup(d.velocities.data(), 16 * n);
    out->rho = (float*)dup(d.densities.data(), 4 * n);
    out->u = (float*)dup(d.internalEnergy.data(), 4 * n);
    out->m = (float*)dup(d.masses.data(), 4 * n);
    out->mat = (int32_t*)dup(d.materialIDs.data(), 4 * n);
    out->p = (float*)dup(d.pressures.data(), 4 * n);
    out->h = (float*)dup(d.smoothingLengths.data(), 4 * n);
    if (!out->pos4 || !out->vel4 || !out->rho || !out->u || !out->m || !out->mat || !out->p || !out->h) {
        sph_particles_free(out);
        return fail("out of host memory", SPH_E_CAPACITY);
    }
    return SPH_OK;
}

void sph_particles_free(sph_particles* p) {
    if (!p) return;
    void* arrays[] = {p->pos4, p->vel4, p->rho, p->u, p->m, p->mat, p->p, p->h};
    for (void* a : arrays) std::free(a);
    std::memset(p, 0, sizeof(*p));
}

static int write_hdf_impl(const char* path, const sph_ic_view* ic, const float* potentials) {
    if (!path || !ic || ic->n < 0 || !ic->pos4 || !ic->vel4 || !ic->rho || !ic->u || !ic->m || !ic->mat || !ic->p ||
        !ic->h)
        return fail("null argument");
    const size_t n = (size_t)ic->n;
    DataStruct d;
    d.positions.assign(ic->pos4, ic->pos4 + 4 * n);
    d.velocities.assign(ic->vel4, ic->vel4 + 4 * n);
    d.densities.assign(ic->rho, ic->rho + n);
    d.internalEnergy.assign(ic->u, ic->u + n);
    d.masses.assign(ic->m, ic->m + n);
    d.materialIDs.assign(ic->mat, ic->mat + n);
    d.pressures.assign(ic->p, ic->p + n);
    d.smoothingLengths.assign(ic->h, ic->h + n);
    if (potentials) d.potentials.assign(potentials, potentials + n);
    std::string err;
    if (!writeHDFFile(path, d, err)) return fail(err);
    return SPH_OK;
}

static int checkpoint_write_impl(sph_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail("null argument");
    std::string err;
    const int rc = writeCheckpoint(ctx, path, err);
    if (rc != SPH_OK) fail(err, rc);
    return rc;
}

static int checkpoint_read_impl(sph_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail("null argument");
    std::string err;
    const int rc = readCheckpoint(ctx, path, err);
    if (rc != SPH_OK) fail(err, rc);
    return rc;
}

/* no C++ exception crosses the C boundary */
int sph_read_hdf(const char* path, sph_particles* out) {
    try {
        return read_hdf_impl(path, out);
    } catch (const std::exception& e) {
        return fail(std::string("sph_read_hdf: ") + e.what(), SPH_E_CAPACITY);
    }
}

int sph_write_hdf(const char* path, const sph_ic_view* ic, const float* potentials) {
    try {
        return write_hdf_impl(path, ic, potentials);
    } catch (const std::exception& e) {
        return fail(std::string("sph_write_hdf: ") + e.what(), SPH_E_CAPACITY);
    }
}

int sph_checkpoint_write(sph_ctx* ctx, const char* path) {
    try {
        return checkpoint_write_impl(ctx, path);
    } catch (const std::exception& e) {
        return fail(std::string("sph_checkpoint_write: ") + e.what(), SPH_E_CAPACITY);
    }
}

int sph_checkpoint_read(sph_ctx* ctx, const char* path) {
    try {
        return checkpoint_read_impl(ctx, path);
    } catch (const std::exception& e) {
        return fail(std::string("sph_checkpoint_read: ") + e.what(), SPH_E_CAPACITY);
    }
}

}  // extern "C"
