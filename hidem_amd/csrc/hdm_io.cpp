// HDF5 checkpoint and particle-file IO (include/hidem/hdm_io.h), host only,
// built into libhidem_io.so against the image's HDF5 (1.10).
#include "../../include/hidem/hdm_io.h"

#include <hdf5.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

// Owns an HDF5 handle and closes it with the matching call.
struct H5Handle {
    hid_t id = H5I_INVALID_HID;
    herr_t (*closer)(hid_t) = nullptr;
    H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
    H5Handle(const H5Handle &) = delete;
    H5Handle &operator=(const H5Handle &) = delete;
    ~H5Handle() {
        if (id >= 0 && closer) closer(id);
    }
    bool ok() const { return id >= 0; }
};

struct Err : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(herr_t e, const std::string &what) {
    if (e < 0) throw Err(what);
}

template <class T> hid_t native();
template <> hid_t native<double>() { return H5T_NATIVE_DOUBLE; }
template <> hid_t native<float>() { return H5T_NATIVE_FLOAT; }
template <> hid_t native<int32_t>() { return H5T_NATIVE_INT32; }
template <> hid_t native<int64_t>() { return H5T_NATIVE_INT64; }
template <> hid_t native<uint8_t>() { return H5T_NATIVE_UINT8; }

// File types as the reference's FP64 build writes them.
template <class T> hid_t file_type();
template <> hid_t file_type<double>() { return H5T_IEEE_F64LE; }
template <> hid_t file_type<int32_t>() { return H5T_STD_I32LE; }
template <> hid_t file_type<int64_t>() { return H5T_STD_I64LE; }
template <> hid_t file_type<uint8_t>() { return H5T_STD_U8LE; }

template <class T>
void write_attr(hid_t target, const char *name, const T *v, hsize_t n, bool scalar = false) {
    H5Handle space(scalar ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!space.ok()) throw Err(std::string("dataspace for attribute ") + name);
    H5Handle a(H5Acreate2(target, name, file_type<T>(), space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) throw Err(std::string("attribute ") + name);
    check(H5Awrite(a.id, native<T>(), v), std::string("writing attribute ") + name);
}

// A 1-D dataset of n values, deflate level 4 in chunks of up to 1M values
// when n > 0 (the reference's [HDF5] compression setting).
template <class T>
void write_data(hid_t group, const std::string &name, const T *v, hsize_t n) {
    H5Handle space(H5Screate_simple(1, &n, nullptr), H5Sclose);
    H5Handle dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    if (n > 0) {
        const hsize_t chunk = std::min<hsize_t>(n, (hsize_t)1 << 20);
        check(H5Pset_chunk(dcpl.id, 1, &chunk), "chunk");
        check(H5Pset_deflate(dcpl.id, 4), "deflate");
    }
    H5Handle d(H5Dcreate2(group, name.c_str(), file_type<T>(), space.id, H5P_DEFAULT, dcpl.id, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) throw Err("dataset " + name);
    if (n > 0) check(H5Dwrite(d.id, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, v), "writing " + name);
}

hsize_t data_size(hid_t group, const std::string &name) {
    H5Handle d(H5Dopen2(group, name.c_str(), H5P_DEFAULT), H5Dclose);
    if (!d.ok()) throw Err("no dataset " + name);
    H5Handle s(H5Dget_space(d.id), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(s.id);
    if (n < 0) throw Err("extent of " + name);
    return (hsize_t)n;
}

// Reads dataset `name` (n values, converted to T by HDF5) into out.
template <class T>
void read_data(hid_t group, const std::string &name, T *out, hsize_t n) {
    H5Handle d(H5Dopen2(group, name.c_str(), H5P_DEFAULT), H5Dclose);
    if (!d.ok()) throw Err("no dataset " + name);
    H5Handle s(H5Dget_space(d.id), H5Sclose);
    if ((hsize_t)H5Sget_simple_extent_npoints(s.id) != n) throw Err("size of " + name);
    if (n > 0) check(H5Dread(d.id, native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, out), "reading " + name);
}

bool has_link(hid_t loc, const char *name) { return H5Lexists(loc, name, H5P_DEFAULT) > 0; }

// Sum of spann_sizes of a group (the ranks' counts; HDF5IO::readSpan).
int64_t span_total(hid_t group) {
    H5Handle a(H5Aopen(group, "spann_sizes", H5P_DEFAULT), H5Aclose);
    if (!a.ok()) throw Err("no spann_sizes");
    H5Handle s(H5Aget_space(a.id), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(s.id);
    std::vector<int64_t> v((size_t)std::max<hssize_t>(n, 0));
    if (n > 0) check(H5Aread(a.id, H5T_NATIVE_INT64, v.data()), "reading spann_sizes");
    int64_t t = 0;
    for (int64_t x : v) t += x;
    return t;
}

void write_span(hid_t group, int64_t count) {
    const int64_t sizes[1] = {count}, offsets[1] = {0};
    write_attr<int64_t>(group, "spann_sizes", sizes, 1);
    write_attr<int64_t>(group, "spann_offsets", offsets, 1);
}

// Silences HDF5's own error printing for the lifetime of the object.
struct QuietH5 {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    QuietH5() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

}  // namespace

extern "C" {

const char *hdm_io_last_error(void) { return g_error.c_str(); }

int hdm_checkpoint_write(const char *path, int32_t timestep, const hdm_particles_host *p, const hdm_links_host *l) {
    if (!path || !p || p->n < 0 || (l && l->n < 0)) return fail(HDM_E_INVALID, "hdm_checkpoint_write: bad arguments");
    const int64_t n = p->n, L = l ? l->n : 0;
    if (n > 0 && (!p->xCord || !p->yCord || !p->zCord || !p->xTheta || !p->yTheta || !p->zTheta || !p->xDiff ||
                  !p->yDiff || !p->zDiff || !p->uDiff || !p->vDiff || !p->wDiff || !p->radius))
        return fail(HDM_E_INVALID, "hdm_checkpoint_write: missing particle field");
    if (L > 0 && (!l->localPartIndA || !l->localPartIndB || !l->state || !l->initialLen))
        return fail(HDM_E_INVALID, "hdm_checkpoint_write: missing link field");
    for (int64_t k = 0; k < L; k++)
        if (l->localPartIndA[k] < 0 || l->localPartIndA[k] >= n || l->localPartIndB[k] < 0 || l->localPartIndB[k] >= n)
            return fail(HDM_E_INVALID, "hdm_checkpoint_write: link end out of range");
    QuietH5 quiet;
    try {
        H5Handle f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        if (!f.ok()) throw Err(std::string("cannot create ") + path);
        write_attr<int32_t>(f.id, "timeStep", &timestep, 1, true);
        {
            // one domain: the particles' bounding box (H5Checkpoint::writeDomains)
            double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            const double *c[3] = {p->xCord, p->yCord, p->zCord};
            for (int d = 0; d < 3 && n > 0; d++) {
                lo[d] = hi[d] = c[d][0];
                for (int64_t i = 1; i < n; i++) {
                    lo[d] = std::min(lo[d], c[d][i]);
                    hi[d] = std::max(hi[d], c[d][i]);
                }
            }
            H5Handle g(H5Gcreate2(f.id, "Domains", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            const char *names[12] = {"domainUpperX", "domainUpperY", "domainUpperZ", "domainLowerX", "domainLowerY",
                                     "domainLowerZ", "domainUpperHaloX", "domainUpperHaloY", "domainUpperHaloZ",
                                     "domainLowerHaloX", "domainLowerHaloY", "domainLowerHaloZ"};
            for (int k = 0; k < 12; k++) {
                const double v = (k % 6) < 3 ? hi[k % 3] : lo[k % 3];
                write_attr<double>(g.id, names[k], &v, 1);
            }
            const int32_t owner = 0;
            write_attr<int32_t>(g.id, "domainOwner", &owner, 1);
        }
        {
            H5Handle g(H5Gcreate2(f.id, "Particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            const double *d[13] = {p->xCord, p->yCord, p->zCord, p->xTheta, p->yTheta, p->zTheta, p->xDiff,
                                   p->yDiff, p->zDiff, p->uDiff, p->vDiff, p->wDiff, p->radius};
            for (int k = 0; k < 13; k++) write_data<double>(g.id, "field_" + std::to_string(k), d[k], (hsize_t)n);
            std::vector<int32_t> owners((size_t)n, 0);
            std::vector<uint8_t> type((size_t)n, 0);
            std::vector<int64_t> keys((size_t)n);
            for (int64_t i = 0; i < n; i++) {
                keys[(size_t)i] = i;
                if (p->type) type[(size_t)i] = p->type[i];
            }
            write_data<int32_t>(g.id, "field_13", owners.data(), (hsize_t)n);
            write_data<uint8_t>(g.id, "field_14", type.data(), (hsize_t)n);
            write_data<int64_t>(g.id, "field_15", keys.data(), (hsize_t)n);
            write_span(g.id, n);
        }
        {
            H5Handle g(H5Gcreate2(f.id, "Links", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            std::vector<int32_t> owners((size_t)L, 0), num((size_t)L);
            std::vector<int64_t> keys((size_t)L), ka((size_t)L), kb((size_t)L);
            std::unordered_map<int32_t, int32_t> perA;
            for (int64_t k = 0; k < L; k++) {
                const int32_t a = l->localPartIndA[k];
                const int32_t r = perA[a]++;
                num[(size_t)k] = r;
                keys[(size_t)k] = (int64_t)a * 256 + r;
                ka[(size_t)k] = a;
                kb[(size_t)k] = l->localPartIndB[k];
            }
            write_data<int32_t>(g.id, "field_0", owners.data(), (hsize_t)L);
            write_data<uint8_t>(g.id, "field_1", L ? l->state : nullptr, (hsize_t)L);
            write_data<int64_t>(g.id, "field_2", keys.data(), (hsize_t)L);
            write_data<int64_t>(g.id, "field_3", ka.data(), (hsize_t)L);
            write_data<int64_t>(g.id, "field_4", kb.data(), (hsize_t)L);
            write_data<double>(g.id, "field_5", L ? l->initialLen : nullptr, (hsize_t)L);
            write_data<int32_t>(g.id, "field_6", num.data(), (hsize_t)L);
            if (L > 0 && l->maxLoad) write_data<double>(g.id, "hdm_maxLoad", l->maxLoad, (hsize_t)L);
            write_span(g.id, L);
        }
        {
            H5Handle g(H5Gcreate2(f.id, "LinkSideData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            std::vector<double> zeros;
            for (int c = 0; c < 6; c++) {
                const double *v = L ? l->CT[c] : nullptr;
                if (L > 0 && !v) {
                    zeros.assign((size_t)(2 * L), 0.0);
                    v = zeros.data();
                }
                write_data<double>(g.id, "side_field_" + std::to_string(7 + c), v, (hsize_t)(2 * L));
            }
            write_span(g.id, 2 * L);
        }
        return HDM_OK;
    } catch (const std::exception &e) {
        return fail(HDM_E_INVALID, std::string("hdm_checkpoint_write: ") + e.what());
    }
}

int hdm_checkpoint_info(const char *path, int64_t *nParticles, int64_t *nLinks, int32_t *timestep) {
    if (!path) return fail(HDM_E_INVALID, "hdm_checkpoint_info: no path");
    QuietH5 quiet;
    try {
        H5Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!f.ok()) throw Err(std::string("cannot open ") + path);
        H5Handle gp(H5Gopen2(f.id, "Particles", H5P_DEFAULT), H5Gclose);
        if (!gp.ok()) throw Err("no Particles group");
        if (nParticles) *nParticles = span_total(gp.id);
        if (nLinks) {
            *nLinks = 0;
            if (has_link(f.id, "Links")) {
                H5Handle gl(H5Gopen2(f.id, "Links", H5P_DEFAULT), H5Gclose);
                *nLinks = span_total(gl.id);
            }
        }
        if (timestep) {
            H5Handle a(H5Aopen(f.id, "timeStep", H5P_DEFAULT), H5Aclose);
            if (!a.ok()) throw Err("no timeStep attribute");
            check(H5Aread(a.id, H5T_NATIVE_INT32, timestep), "reading timeStep");
        }
        return HDM_OK;
    } catch (const std::exception &e) {
        return fail(HDM_E_INVALID, std::string("hdm_checkpoint_info: ") + e.what());
    }
}

int hdm_checkpoint_read(const char *path, hdm_particles_host *p, hdm_links_host *l) {
    if (!path || !p) return fail(HDM_E_INVALID, "hdm_checkpoint_read: bad arguments");
    int64_t n = 0, L = 0;
    int rc = hdm_checkpoint_info(path, &n, &L, nullptr);
    if (rc != HDM_OK) return rc;
    if (p->n != n || (l && l->n != L)) return fail(HDM_E_INVALID, "hdm_checkpoint_read: sizes differ from the file");
    QuietH5 quiet;
    try {
        H5Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!f.ok()) throw Err(std::string("cannot open ") + path);
        std::vector<int64_t> pkeys((size_t)n);
        {
            H5Handle g(H5Gopen2(f.id, "Particles", H5P_DEFAULT), H5Gclose);
            double *d[13] = {p->xCord, p->yCord, p->zCord, p->xTheta, p->yTheta, p->zTheta, p->xDiff,
                             p->yDiff, p->zDiff, p->uDiff, p->vDiff, p->wDiff, p->radius};
            for (int k = 0; k < 13; k++) {
                if (!d[k] && n > 0) throw Err("missing particle array");
                read_data<double>(g.id, "field_" + std::to_string(k), d[k], (hsize_t)n);
            }
            std::vector<uint8_t> type((size_t)n);
            read_data<uint8_t>(g.id, "field_14", type.data(), (hsize_t)n);
            if (p->type) std::memcpy(p->type, type.data(), (size_t)n);
            read_data<int64_t>(g.id, "field_15", pkeys.data(), (hsize_t)n);
            double *fo[6] = {p->xForce, p->yForce, p->zForce, p->uForce, p->vForce, p->wForce};
            for (double *v : fo)
                if (v) std::fill(v, v + n, 0.0);
        }
        if (l && L > 0) {
            std::unordered_map<int64_t, int32_t> index;
            index.reserve((size_t)n * 2);
            for (int64_t i = 0; i < n; i++)
                if (!index.emplace(pkeys[(size_t)i], (int32_t)i).second) throw Err("duplicate particle global key");
            H5Handle g(H5Gopen2(f.id, "Links", H5P_DEFAULT), H5Gclose);
            std::vector<int64_t> ka((size_t)L), kb((size_t)L);
            read_data<uint8_t>(g.id, "field_1", l->state, (hsize_t)L);
            read_data<int64_t>(g.id, "field_3", ka.data(), (hsize_t)L);
            read_data<int64_t>(g.id, "field_4", kb.data(), (hsize_t)L);
            read_data<double>(g.id, "field_5", l->initialLen, (hsize_t)L);
            for (int64_t k = 0; k < L; k++) {
                auto a = index.find(ka[(size_t)k]), b = index.find(kb[(size_t)k]);
                if (a == index.end() || b == index.end()) throw Err("dangling link (end key not among the particles)");
                l->localPartIndA[k] = a->second;
                l->localPartIndB[k] = b->second;
            }
            if (l->maxLoad) {
                if (!has_link(g.id, "hdm_maxLoad")) throw Err("no per-link maxLoad in the file (pass maxLoad = NULL)");
                read_data<double>(g.id, "hdm_maxLoad", l->maxLoad, (hsize_t)L);
            }
            H5Handle gs(H5Gopen2(f.id, "LinkSideData", H5P_DEFAULT), H5Gclose);
            if (!gs.ok()) throw Err("no LinkSideData group");
            for (int c = 0; c < 6; c++) {
                if (!l->CT[c]) throw Err("missing CT array");
                read_data<double>(gs.id, "side_field_" + std::to_string(7 + c), l->CT[c], (hsize_t)(2 * L));
            }
        }
        return HDM_OK;
    } catch (const std::exception &e) {
        return fail(HDM_E_INVALID, std::string("hdm_checkpoint_read: ") + e.what());
    }
}

int hdm_particle_file_count(const char *path, const char *xName, int64_t *n) {
    if (!path || !n) return fail(HDM_E_INVALID, "hdm_particle_file_count: bad arguments");
    QuietH5 quiet;
    try {
        H5Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!f.ok()) throw Err(std::string("cannot open ") + path);
        *n = (int64_t)data_size(f.id, xName ? xName : "xCord");
        return HDM_OK;
    } catch (const std::exception &e) {
        return fail(HDM_E_INVALID, std::string("hdm_particle_file_count: ") + e.what());
    }
}

int hdm_particle_file_read(const char *path, const char *xName, const char *yName, const char *zName,
                           const char *rName, hdm_particles_host *p) {
    if (!path || !p || p->n < 0 || (p->n > 0 && (!p->xCord || !p->yCord || !p->zCord || !p->radius)))
        return fail(HDM_E_INVALID, "hdm_particle_file_read: bad arguments");
    QuietH5 quiet;
    try {
        H5Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!f.ok()) throw Err(std::string("cannot open ") + path);
        const char *names[4] = {xName ? xName : "xCord", yName ? yName : "yCord", zName ? zName : "zCord",
                                rName ? rName : "radius"};
        double *out[4] = {p->xCord, p->yCord, p->zCord, p->radius};
        std::vector<float> v((size_t)p->n);
        for (int k = 0; k < 4; k++) {
            // FileParticleGenerator reads every column as float
            read_data<float>(f.id, names[k], v.data(), (hsize_t)p->n);
            for (int64_t i = 0; i < p->n; i++) out[k][i] = (double)v[(size_t)i];
        }
        double *zero[15] = {p->xTheta, p->yTheta, p->zTheta, p->xDiff,  p->yDiff,  p->zDiff,  p->uDiff, p->vDiff,
                            p->wDiff,  p->xForce, p->yForce, p->zForce, p->uForce, p->vForce, p->wForce};
        for (double *z : zero)
            if (z) std::fill(z, z + p->n, 0.0);
        if (p->type) std::memset(p->type, 0, (size_t)p->n);
        return HDM_OK;
    } catch (const std::exception &e) {
        return fail(HDM_E_INVALID, std::string("hdm_particle_file_read: ") + e.what());
    }
}

}  // extern "C"
