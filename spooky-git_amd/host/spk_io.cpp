// spk_io.cpp -- HDF5 snapshots in the reference layout (include/spk_io.h),
// through the HDF5 C API.
#include "spk_io.h"

#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <sys/stat.h>

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

// HDF5 prints its error stack by default; the status codes carry it here
struct Quiet {
    Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};
const Quiet g_quiet;

struct File {
    hid_t id = -1;
    ~File() {
        if (id >= 0) H5Fclose(id);
    }
};

int write_1d(hid_t file, const char *name, hid_t type, const void *data, hsize_t n) {
    hid_t sp = H5Screate_simple(1, &n, nullptr);
    if (sp < 0) return fail("H5Screate_simple(%s)", name);
    hid_t ds = H5Dcreate2(file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
        H5Sclose(sp);
        return fail("H5Dcreate2(%s)", name);
    }
    const herr_t e = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    H5Sclose(sp);
    return e < 0 ? fail("H5Dwrite(%s)", name) : 0;
}

int read_1d(hid_t file, const char *name, hid_t type, void *data, hsize_t n) {
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) return fail("dataset '%s' not found", name);
    hid_t sp = H5Dget_space(ds);
    hsize_t dims[8] = {0};
    const int rank = sp >= 0 ? H5Sget_simple_extent_ndims(sp) : -1;
    if (rank >= 1 && rank <= 8) H5Sget_simple_extent_dims(sp, dims, nullptr);
    hsize_t total = 1;
    for (int r = 0; r < rank; r++) total *= dims[r];
    if (sp >= 0) H5Sclose(sp);
    if (rank < 0 || total != n) {
        H5Dclose(ds);
        return fail("dataset '%s' has %llu values, %llu expected", name, (unsigned long long)total,
                    (unsigned long long)n);
    }
    const herr_t e = H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    return e < 0 ? fail("H5Dread(%s)", name) : 0;
}

}  // namespace

extern "C" const char *spkio_last_error(void) { return g_err.c_str(); }

extern "C" int spkio_write_snapshot(const char *path, int nx, int ny, int nz, int nfields, const char *const *names,
                                    const double *const *fields, double t_start, double t_save, double t_lastsnap,
                                    double t_lastvar, int step) {
    if (!path || nx < 1 || ny < 1 || nz < 1 || nfields < 0 || (nfields > 0 && (!names || !fields)))
        return fail("spkio_write_snapshot: bad arguments");
    File f;
    f.id = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f.id < 0) return fail("cannot create %s", path);
    if (write_1d(f.id, "t_start", H5T_NATIVE_DOUBLE, &t_start, 1) || write_1d(f.id, "t_save", H5T_NATIVE_DOUBLE, &t_save, 1) ||
        write_1d(f.id, "t_lastsnap", H5T_NATIVE_DOUBLE, &t_lastsnap, 1) ||
        write_1d(f.id, "t_lastvar", H5T_NATIVE_DOUBLE, &t_lastvar, 1) || write_1d(f.id, "step", H5T_NATIVE_INT32, &step, 1))
        return -1;
    const hsize_t n = (hsize_t)nx * ny * nz;
    for (int i = 0; i < nfields; i++) {
        if (!names[i] || !fields[i]) return fail("spkio_write_snapshot: field %d missing", i);
        if (write_1d(f.id, names[i], H5T_NATIVE_DOUBLE, fields[i], n)) return -1;
    }
    if (H5Fflush(f.id, H5F_SCOPE_LOCAL) < 0) return fail("H5Fflush(%s)", path);
    return 0;
}

extern "C" int spkio_read_header(const char *path, double *t_save, double *t_lastsnap, double *t_lastvar, int *step) {
    if (!path || !t_save || !t_lastsnap || !t_lastvar || !step) return fail("spkio_read_header: bad arguments");
    File f;
    f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f.id < 0) return fail("cannot open %s", path);
    if (read_1d(f.id, "t_save", H5T_NATIVE_DOUBLE, t_save, 1) || read_1d(f.id, "t_lastsnap", H5T_NATIVE_DOUBLE, t_lastsnap, 1) ||
        read_1d(f.id, "t_lastvar", H5T_NATIVE_DOUBLE, t_lastvar, 1) || read_1d(f.id, "step", H5T_NATIVE_INT32, step, 1))
        return -1;
    return 0;
}

extern "C" int spkio_read_field(const char *path, const char *name, double *out, size_t n) {
    if (!path || !name || !out) return fail("spkio_read_field: bad arguments");
    File f;
    f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f.id < 0) return fail("cannot open %s", path);
    return read_1d(f.id, name, H5T_NATIVE_DOUBLE, out, (hsize_t)n);
}

extern "C" int spkio_last_snapshot(const char *data_dir) {
    if (!data_dir) return -1;
    int last = -1;
    for (int i = 0; i < 5000000; i++) {
        char name[32];
        std::snprintf(name, sizeof name, "snap%04d.h5", i);
        std::string p = std::string(data_dir) + "/" + name;
        struct stat st;
        if (stat(p.c_str(), &st) != 0) break;
        last = i;
    }
    return last;
}
