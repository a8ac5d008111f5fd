// Snapshot files in the layout of the reference's t_Hdf5ForOhfield
// (src/hdf5/hdf5_for_ohfield.f90:27-76, src/hdf5/hdf5.f90:46-185): one HDF5
// file, one group, one 3-D FP64 dataset per write of the global shape given at
// open, each write filling the hyperslab of its local block.  The reference
// calls the Fortran HDF5 API with Fortran-order shapes (x first); the same
// file seen through the C API has the dimensions reversed (z first) with x
// fastest in memory, which is how the datasets are created here.
//
// HDF5 (1.10 or later) is loaded at run time with dlopen, so that libohpic.so
// itself has no dependency on it: OHPIC_HDF5_LIB names the library file, else
// libhdf5.so.103 / libhdf5.so on the loader path, then /opt/conda/lib.  Calls
// into HDF5 are serialised by one mutex (the library need not be built
// thread-safe).
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "ohpic.h"
#include "ohpic_internal.hpp"

namespace {

// ---- the part of the HDF5 1.10 C API used here (H5public.h, H5Ipublic.h,
// H5Fpublic.h, H5Ppublic.h, H5Spublic.h, H5Tpublic.h) ----------------------
using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kH5P_DEFAULT = 0;
constexpr hid_t kH5S_ALL = 0;
constexpr hid_t kH5E_DEFAULT = 0;
constexpr unsigned kH5F_ACC_RDONLY = 0x0000u;
constexpr unsigned kH5F_ACC_TRUNC = 0x0002u;
constexpr int kH5S_SELECT_SET = 0;

struct H5Api {
    void *handle = nullptr;
    std::string err;
    herr_t (*open)();
    herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
    herr_t (*eset_auto2)(hid_t, void *, void *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate2)(hid_t, const char *, hid_t, hid_t, hid_t);
    hid_t (*gopen2)(hid_t, const char *, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *,
                                const hsize_t *, const hsize_t *);
    int (*sget_simple_extent_ndims)(hid_t);
    int (*sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen2)(hid_t, const char *, hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dclose)(hid_t);
    hid_t *native_double;  // H5T_NATIVE_DOUBLE_g, valid after H5open
};

std::mutex g_h5_mu;  // serialises every HDF5 call and the loading
H5Api g_h5;
bool g_h5_tried = false;

template <class F>
bool sym(void *h, const char *name, F *out, std::string &err)
{
    void *p = dlsym(h, name);
    if (!p) {
        err = std::string("HDF5 library lacks ") + name;
        return false;
    }
    *out = reinterpret_cast<F>(p);
    return true;
}

// Loads HDF5 once per process; returns nullptr (with g_h5.err) if absent.
// Caller holds g_h5_mu.
H5Api *h5()
{
    if (g_h5_tried) return g_h5.handle ? &g_h5 : nullptr;
    g_h5_tried = true;
    std::vector<std::string> names;
    if (const char *e = std::getenv("OHPIC_HDF5_LIB")) names.push_back(e);
    for (const char *n : {"libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                          "/opt/conda/lib/libhdf5.so"})
        names.push_back(n);
    void *h = nullptr;
    for (const auto &n : names)
        if ((h = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) {
        g_h5.err = "HDF5 library not found (set OHPIC_HDF5_LIB to libhdf5.so)";
        return nullptr;
    }
    H5Api a;
    std::string &err = g_h5.err;
    bool ok = sym(h, "H5open", &a.open, err) &&
              sym(h, "H5get_libversion", &a.get_libversion, err) &&
              sym(h, "H5Eset_auto2", &a.eset_auto2, err) &&
              sym(h, "H5Fcreate", &a.fcreate, err) && sym(h, "H5Fopen", &a.fopen, err) &&
              sym(h, "H5Fclose", &a.fclose, err) && sym(h, "H5Gcreate2", &a.gcreate2, err) &&
              sym(h, "H5Gopen2", &a.gopen2, err) && sym(h, "H5Gclose", &a.gclose, err) &&
              sym(h, "H5Screate_simple", &a.screate_simple, err) &&
              sym(h, "H5Sselect_hyperslab", &a.sselect_hyperslab, err) &&
              sym(h, "H5Sget_simple_extent_ndims", &a.sget_simple_extent_ndims, err) &&
              sym(h, "H5Sget_simple_extent_dims", &a.sget_simple_extent_dims, err) &&
              sym(h, "H5Sclose", &a.sclose, err) && sym(h, "H5Dcreate2", &a.dcreate2, err) &&
              sym(h, "H5Dopen2", &a.dopen2, err) && sym(h, "H5Dget_space", &a.dget_space, err) &&
              sym(h, "H5Dwrite", &a.dwrite, err) && sym(h, "H5Dread", &a.dread, err) &&
              sym(h, "H5Dclose", &a.dclose, err) &&
              sym(h, "H5T_NATIVE_DOUBLE_g", &a.native_double, err);
    unsigned maj = 0, min = 0, rel = 0;
    if (ok && (a.open() < 0 || a.get_libversion(&maj, &min, &rel) < 0 || maj != 1 || min < 10)) {
        // hid_t is 64-bit from 1.10 on; older libraries have another ABI
        char buf[128];
        snprintf(buf, sizeof buf, "HDF5 %u.%u.%u: 1.10 or later is needed", maj, min, rel);
        err = buf;
        ok = false;
    }
    if (!ok) {
        dlclose(h);
        return nullptr;
    }
    a.eset_auto2(kH5E_DEFAULT, nullptr, nullptr);  // errors come back as status codes
    a.handle = h;
    a.err.clear();
    g_h5 = a;
    return &g_h5;
}

int fail_snap(ohpic_snapshot *s, int code, const std::string &msg);

}  // namespace

struct ohpic_snapshot {
    hid_t file = -1;
    hid_t group = -1;
    int64_t global_shape[3] = {0, 0, 0};   // Fortran order: x, y, z
    int64_t global_offset[3] = {0, 0, 0};
    std::string err;
};

namespace {

int fail_snap(ohpic_snapshot *s, int code, const std::string &msg)
{
    if (s) s->err = msg;
    ohpic::set_global_error(msg.c_str());
    return code;
}

// One dataset of the global shape with the local block written at its offset
// (hdf5.f90:140-185).  Shapes and offsets in Fortran order; caller holds g_h5_mu.
int write_block(ohpic_snapshot *s, H5Api *a, const char *name, const double *values,
                const int64_t ls[3], const int64_t lo[3])
{
    const hsize_t gdims[3] = {(hsize_t)s->global_shape[2], (hsize_t)s->global_shape[1],
                              (hsize_t)s->global_shape[0]};
    const hsize_t ldims[3] = {(hsize_t)ls[2], (hsize_t)ls[1], (hsize_t)ls[0]};
    const hsize_t start[3] = {(hsize_t)lo[2], (hsize_t)lo[1], (hsize_t)lo[0]};
    hid_t fspace = a->screate_simple(3, gdims, nullptr);
    if (fspace < 0) return fail_snap(s, OHPIC_EIO, "H5Screate_simple failed");
    const hid_t dtype = *a->native_double;
    hid_t dset = a->dcreate2(s->group, name, dtype, fspace, kH5P_DEFAULT, kH5P_DEFAULT,
                             kH5P_DEFAULT);
    a->sclose(fspace);
    if (dset < 0)
        return fail_snap(s, OHPIC_EIO, std::string("cannot create dataset '") + name +
                                           "' (does it exist already?)");
    hid_t mspace = a->screate_simple(3, ldims, nullptr);
    hid_t fsel = a->dget_space(dset);
    int rc = OHPIC_OK;
    if (mspace < 0 || fsel < 0 ||
        a->sselect_hyperslab(fsel, kH5S_SELECT_SET, start, nullptr, ldims, nullptr) < 0 ||
        a->dwrite(dset, dtype, mspace, fsel, kH5P_DEFAULT, values) < 0)
        rc = fail_snap(s, OHPIC_EIO, std::string("writing dataset '") + name + "' failed");
    if (fsel >= 0) a->sclose(fsel);
    if (mspace >= 0) a->sclose(mspace);
    if (a->dclose(dset) < 0 && rc == OHPIC_OK)
        rc = fail_snap(s, OHPIC_EIO, std::string("closing dataset '") + name + "' failed");
    return rc;
}

bool valid_name(const char *name) { return name && name[0] != '\0'; }

}  // namespace

extern "C" {

int ohpic_snapshot_open(const char *filename, const char *group, const int64_t global_shape[3],
                        const int64_t global_offset[3], ohpic_snapshot **out)
{
    if (!out) return fail_snap(nullptr, OHPIC_EINVAL, "NULL argument");
    *out = nullptr;
    if (!valid_name(filename) || !valid_name(group) || !global_shape)
        return fail_snap(nullptr, OHPIC_EINVAL, "snapshot: empty file or group name, or NULL shape");
    for (int k = 0; k < 3; ++k)
        if (global_shape[k] < 1 || (global_offset && global_offset[k] < 0))
            return fail_snap(nullptr, OHPIC_EINVAL, "snapshot: global shape must be positive");
    std::lock_guard<std::mutex> lk(g_h5_mu);
    H5Api *a = h5();
    if (!a) return fail_snap(nullptr, OHPIC_ENOLIB, g_h5.err);
    auto *s = new ohpic_snapshot;
    for (int k = 0; k < 3; ++k) {
        s->global_shape[k] = global_shape[k];
        s->global_offset[k] = global_offset ? global_offset[k] : 0;
    }
    // new_HDF5File(filename, 'w', comm): create, truncating (hdf5.f90:65-66);
    // create_group in 'w' mode (hdf5.f90:84-85)
    s->file = a->fcreate(filename, kH5F_ACC_TRUNC, kH5P_DEFAULT, kH5P_DEFAULT);
    if (s->file < 0) {
        delete s;
        return fail_snap(nullptr, OHPIC_EIO, std::string("cannot create '") + filename + "'");
    }
    s->group = a->gcreate2(s->file, group, kH5P_DEFAULT, kH5P_DEFAULT, kH5P_DEFAULT);
    if (s->group < 0) {
        a->fclose(s->file);
        delete s;
        return fail_snap(nullptr, OHPIC_EIO, std::string("cannot create group '") + group + "'");
    }
    *out = s;
    return OHPIC_OK;
}

int ohpic_snapshot_write_host(ohpic_snapshot *s, const char *dataset, const double *values,
                              const int64_t local_shape[3], const int64_t local_offset[3])
{
    if (!s || !values || !local_shape || !valid_name(dataset))
        return fail_snap(s, OHPIC_EINVAL, "snapshot write: NULL argument or empty dataset name");
    if (s->file < 0) return fail_snap(s, OHPIC_ESTATE, "snapshot is closed");
    for (int k = 0; k < 3; ++k) {
        const int64_t o = local_offset ? local_offset[k] : 0;
        if (local_shape[k] < 1 || o < 0 || o + local_shape[k] > s->global_shape[k])
            return fail_snap(s, OHPIC_EINVAL, "snapshot write: local block outside the global shape");
    }
    const int64_t zero[3] = {0, 0, 0};
    std::lock_guard<std::mutex> lk(g_h5_mu);
    return write_block(s, &g_h5, dataset, values, local_shape, local_offset ? local_offset : zero);
}

int ohpic_snapshot_write(ohpic_snapshot *s, const char *dataset, ohpic_ctx *ctx, int which)
{
    if (!s || !ctx || !valid_name(dataset))
        return fail_snap(s, OHPIC_EINVAL, "snapshot write: NULL argument or empty dataset name");
    if (which == OHPIC_FIELD_EB)
        return fail_snap(s, OHPIC_EINVAL, "snapshot write: one component per dataset (EX, EY or EZ)");
    ohpic_config cfg;
    int rc = ohpic_get_config(ctx, &cfg);
    if (rc) return rc;
    // the block of one subdomain covering the whole periodic box: nodes 0..n
    // per axis, node n carrying node 0 (hdf5_for_ohfield.f90:53-75)
    const int32_t lb[3] = {0, 0, 0};
    const int32_t ub[3] = {cfg.nx, cfg.ny, cfg.nz};
    const int64_t ls[3] = {(int64_t)cfg.nx + 1, (int64_t)cfg.ny + 1, (int64_t)cfg.nz + 1};
    std::vector<double> host((size_t)(ls[0] * ls[1] * ls[2]));
    if ((rc = ohpic_get_field(ctx, which, host.data(), 1, lb, ub, 0))) return rc;
    return ohpic_snapshot_write_host(s, dataset, host.data(), ls, nullptr);
}

int ohpic_snapshot_close(ohpic_snapshot *s)
{
    if (!s) return OHPIC_OK;
    int rc = OHPIC_OK;
    {
        std::lock_guard<std::mutex> lk(g_h5_mu);
        if (s->group >= 0 && g_h5.gclose(s->group) < 0)
            rc = fail_snap(nullptr, OHPIC_EIO, "closing the snapshot group failed");
        if (s->file >= 0 && g_h5.fclose(s->file) < 0)
            rc = fail_snap(nullptr, OHPIC_EIO, "closing the snapshot file failed");
    }
    delete s;
    return rc;
}

const char *ohpic_snapshot_last_error(const ohpic_snapshot *s)
{
    if (!s) return ohpic_last_error(nullptr);
    static thread_local std::string copy;
    copy = s->err;
    return copy.c_str();
}

int ohpic_snapshot_read(const char *filename, const char *group, const char *dataset,
                        double *values, int64_t capacity, int64_t shape[3])
{
    if (!valid_name(filename) || !valid_name(group) || !valid_name(dataset) || !shape)
        return fail_snap(nullptr, OHPIC_EINVAL, "snapshot read: NULL or empty argument");
    std::lock_guard<std::mutex> lk(g_h5_mu);
    H5Api *a = h5();
    if (!a) return fail_snap(nullptr, OHPIC_ENOLIB, g_h5.err);
    hid_t f = a->fopen(filename, kH5F_ACC_RDONLY, kH5P_DEFAULT);
    if (f < 0) return fail_snap(nullptr, OHPIC_EIO, std::string("cannot open '") + filename + "'");
    int rc = OHPIC_OK;
    hid_t g = a->gopen2(f, group, kH5P_DEFAULT);
    hid_t d = g >= 0 ? a->dopen2(g, dataset, kH5P_DEFAULT) : -1;
    hid_t sp = d >= 0 ? a->dget_space(d) : -1;
    hsize_t dims[3] = {0, 0, 0};
    if (sp < 0) {
        rc = fail_snap(nullptr, OHPIC_EIO, std::string("no dataset '") + group + "/" + dataset + "'");
    } else if (a->sget_simple_extent_ndims(sp) != 3 ||
               a->sget_simple_extent_dims(sp, dims, nullptr) != 3) {
        rc = fail_snap(nullptr, OHPIC_EINVAL, "snapshot read: dataset is not 3-D");
    } else {
        for (int k = 0; k < 3; ++k) shape[k] = (int64_t)dims[2 - k];
        const int64_t count = shape[0] * shape[1] * shape[2];
        if (values) {
            if (capacity < count)
                rc = fail_snap(nullptr, OHPIC_EINVAL, "snapshot read: buffer too small");
            else if (a->dread(d, *a->native_double, kH5S_ALL, kH5S_ALL, kH5P_DEFAULT, values) < 0)
                rc = fail_snap(nullptr, OHPIC_EIO, "snapshot read: H5Dread failed");
        }
    }
    if (sp >= 0) a->sclose(sp);
    if (d >= 0) a->dclose(d);
    if (g >= 0) a->gclose(g);
    a->fclose(f);
    return rc;
}

}  // extern "C"
