/* Snapshot files in the layout of src/mhd_write_h5.f90 through the HDF5 C API. */
#include "../../include/chimera_h5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>

static __thread char h5_err[512] = "";

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(h5_err, sizeof(h5_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *chimera_h5_last_error(void) { return h5_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* one float64 dataset of rank 1 or 2 under loc */
static int write_dataset(hid_t loc, const char *name, int rank, const hsize_t *dims,
                         const double *data) {
    hid_t space = H5Screate_simple(rank, dims, NULL);
    if (space < 0) return fail(-2, "H5Screate_simple failed for %s", name);
    hid_t dset = H5Dcreate2(loc, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT,
                            H5P_DEFAULT);
    if (dset < 0) {
        H5Sclose(space);
        return fail(-2, "H5Dcreate2 failed for %s", name);
    }
    herr_t e = H5Dwrite(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(dset);
    H5Sclose(space);
    return e < 0 ? fail(-2, "H5Dwrite failed for %s", name) : 0;
}

int chimera_h5_write_snapshots(const char *path, int nsnap, int nx, int ny, const double *rho,
                               const double *P, const double *Bx, const double *By,
                               const double *Vx, const double *Vy, const double *times,
                               double gamma, double M_s, double beta) {
    if (!path || nsnap < 0 || nx < 1 || ny < 1 || (nsnap > 0 && (!rho || !P || !Bx || !By ||
                                                                  !Vx || !Vy || !times)))
        return fail(-1, "invalid argument");
    quiet();
    hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return fail(-2, "Error opening file %s", path);
    /* groups in the reference's creation order (mhd_write_h5.f90:44-51) */
    static const char *field_names[6] = {"rho", "P", "Bx", "By", "Vx", "Vy"};
    const double *fields[6] = {rho, P, Bx, By, Vx, Vy};
    hid_t groups[8];
    int rc = 0, made = 0;
    for (int g = 0; g < 6; ++g, ++made)
        if ((groups[g] = H5Gcreate2(file, field_names[g], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)) < 0)
            break;
    if (made == 6 && (groups[6] = H5Gcreate2(file, "time", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)) >= 0) ++made;
    if (made == 7 && (groups[7] = H5Gcreate2(file, "parameters", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)) >= 0) ++made;
    if (made != 8) rc = fail(-2, "Error creating groups");
    /* SNAPSHOT1..n per field (mhd_write_h5.f90:58-145) */
    const hsize_t dims2[2] = {(hsize_t)ny, (hsize_t)nx};
    const size_t plane = (size_t)nx * (size_t)ny;
    char name[32];
    for (int s = 0; rc == 0 && s < nsnap; ++s) {
        snprintf(name, sizeof(name), "SNAPSHOT%d", s + 1);
        for (int g = 0; rc == 0 && g < 6; ++g)
            rc = write_dataset(groups[g], name, 2, dims2, fields[g] + (size_t)s * plane);
    }
    /* time/sim_time and the scalar parameters (mhd_write_h5.f90:148-172) */
    const hsize_t tdim[1] = {(hsize_t)nsnap}, one[1] = {1};
    if (rc == 0) rc = write_dataset(groups[6], "sim_time", 1, tdim, times);
    if (rc == 0) rc = write_dataset(groups[7], "gamma", 1, one, &gamma);
    if (rc == 0) rc = write_dataset(groups[7], "M_s", 1, one, &M_s);
    if (rc == 0) rc = write_dataset(groups[7], "beta", 1, one, &beta);
    for (int g = 0; g < made; ++g) H5Gclose(groups[g]);
    if (H5Fclose(file) < 0 && rc == 0) rc = fail(-2, "H5Fclose failed");
    return rc;
}

int chimera_h5_group_count(const char *path, const char *group, int *count) {
    if (!path || !group || !count) return fail(-1, "invalid argument");
    quiet();
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return fail(-2, "cannot open %s", path);
    hid_t g = H5Gopen2(file, group, H5P_DEFAULT);
    int rc = 0;
    if (g < 0) {
        rc = fail(-3, "no group %s", group);
    } else {
        H5G_info_t info;
        if (H5Gget_info(g, &info) < 0) rc = fail(-2, "H5Gget_info failed for %s", group);
        else *count = (int)info.nlinks;
        H5Gclose(g);
    }
    H5Fclose(file);
    return rc;
}

int chimera_h5_dataset_dims(const char *path, const char *dataset, int *rank, long long dims[2]) {
    if (!path || !dataset || !rank || !dims) return fail(-1, "invalid argument");
    quiet();
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return fail(-2, "cannot open %s", path);
    hid_t d = H5Dopen2(file, dataset, H5P_DEFAULT);
    int rc = 0;
    if (d < 0) {
        rc = fail(-3, "no dataset %s", dataset);
    } else {
        hid_t sp = H5Dget_space(d);
        hsize_t dd[8] = {0};
        int r = H5Sget_simple_extent_dims(sp, dd, NULL);
        if (r < 0 || r > 2) {
            rc = fail(-2, "dataset %s has rank %d", dataset, r);
        } else {
            *rank = r;
            dims[0] = (long long)dd[0];
            dims[1] = r > 1 ? (long long)dd[1] : 0;
        }
        H5Sclose(sp);
        H5Dclose(d);
    }
    H5Fclose(file);
    return rc;
}

int chimera_h5_read_dataset(const char *path, const char *dataset, double *out, long long n) {
    if (!path || !dataset || !out) return fail(-1, "invalid argument");
    int rank;
    long long dims[2];
    int rc = chimera_h5_dataset_dims(path, dataset, &rank, dims);
    if (rc) return rc;
    long long total = rank == 1 ? dims[0] : dims[0] * dims[1];
    if (total != n) return fail(-1, "dataset %s holds %lld values, buffer %lld", dataset, total, n);
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return fail(-2, "cannot open %s", path);
    hid_t d = H5Dopen2(file, dataset, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(file);
        return fail(-3, "no dataset %s", dataset);
    }
    if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        rc = fail(-2, "H5Dread failed for %s", dataset);
    H5Dclose(d);
    H5Fclose(file);
    return rc;
}
