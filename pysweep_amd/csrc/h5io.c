/*
 * h5io.c -- HDF5 output and input of pysweep_amd through the HDF5 C library
 * (built only where the library is installed; pysweep_amd.hdf5io falls back
 * to .npz files otherwise).
 *
 * Output layout of the reference (core/io.py:78-90): datasets blocksize (1,)
 * int64, share (1,) float64, exid (len) float64, globals (len) float64,
 * clocktime (1,) float64 and data (nt, nv, nx, ny) of the solver dtype.
 * Input: the initial condition is dataset "data" (core/solver.py:89-101),
 * converted to float64 on read.
 */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

static char g_err[512];

static int fail(const char *what, const char *name)
{
    snprintf(g_err, sizeof g_err, "%s%s%s", what, name ? ": " : "", name ? name : "");
    return 1;
}

const char *ps_h5_error(void) { return g_err; }

/* Types: 0 float64, 1 float32, 2 int64. */
static hid_t mem_type(int t)
{
    return t == 0 ? H5T_NATIVE_DOUBLE : t == 1 ? H5T_NATIVE_FLOAT : H5T_NATIVE_LLONG;
}

static hid_t file_type(int t)
{
    return t == 0 ? H5T_IEEE_F64LE : t == 1 ? H5T_IEEE_F32LE : H5T_STD_I64LE;
}

/* Write n datasets into a new file (truncating): names[i] of type types[i],
 * rank ranks[i], dimensions taken in order from dims, contents data[i]
 * (C order).  Returns 0 on success. */
int ps_h5_write(const char *path, int n, const char *const *names, const int *types, const int *ranks,
                const long long *dims, const void *const *data)
{
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("cannot create", path);
    int rc = 0, d = 0;
    for (int i = 0; i < n && !rc; ++i) {
        hsize_t h[8];
        if (ranks[i] < 0 || ranks[i] > 8) {
            rc = fail("rank out of range", names[i]);
            break;
        }
        for (int r = 0; r < ranks[i]; ++r) h[r] = (hsize_t)dims[d++];
        hid_t sp = H5Screate_simple(ranks[i], h, NULL);
        hid_t ds = H5Dcreate2(f, names[i], file_type(types[i]), sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (sp < 0 || ds < 0)
            rc = fail("cannot create dataset", names[i]);
        else if (H5Dwrite(ds, mem_type(types[i]), H5S_ALL, H5S_ALL, H5P_DEFAULT, data[i]) < 0)
            rc = fail("cannot write dataset", names[i]);
        if (ds >= 0) H5Dclose(ds);
        if (sp >= 0) H5Sclose(sp);
    }
    if (H5Fclose(f) < 0 && !rc) rc = fail("cannot close", path);
    return rc;
}

/* Rank and dimensions (up to 8) of a dataset. */
int ps_h5_dims(const char *path, const char *name, int *rank, long long *dims)
{
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open", path);
    int rc = 0;
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds < 0) {
        rc = fail("no dataset", name);
    } else {
        hid_t sp = H5Dget_space(ds);
        hsize_t h[8];
        int r = H5Sget_simple_extent_ndims(sp);
        if (r < 0 || r > 8) {
            rc = fail("unsupported rank", name);
        } else {
            H5Sget_simple_extent_dims(sp, h, NULL);
            *rank = r;
            for (int i = 0; i < r; ++i) dims[i] = (long long)h[i];
        }
        H5Sclose(sp);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

/* Read a dataset, converted to float64, into out (sized by ps_h5_dims). */
int ps_h5_read_f64(const char *path, const char *name, double *out)
{
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open", path);
    int rc = 0;
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds < 0)
        rc = fail("no dataset", name);
    else if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        rc = fail("cannot read dataset", name);
    if (ds >= 0) H5Dclose(ds);
    H5Fclose(f);
    return rc;
}
