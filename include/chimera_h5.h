/* chimera_h5.h -- snapshot files in the layout of CHIMERA's HDF5 writer.
 *
 * libchimera_h5.so (plain C over the HDF5 1.10 C library) writes and reads the file
 * that src/mhd_write_h5.f90:12-177 (write_prims_to_hdf5) produces:
 *
 *   /rho /P /Bx /By /Vx /Vy   groups, each with datasets SNAPSHOT1 .. SNAPSHOTn
 *                             (float64, C dims {ny, nx}: a Fortran reader sees (nx, ny),
 *                             the reference's (i, j) cell order)
 *   /time/sim_time            float64 {n}
 *   /parameters/gamma, M_s, beta   float64 {1}
 *
 * P is the thermal pressure (src/main.f90:283).  Every function returns 0 on success
 * or a negative code; chimera_h5_last_error() returns the message.  No counterpart
 * in the reference's C interface: the reference writes through the gfortran HDF5
 * module, which cannot be linked here (its .mod is a gfortran-7 file).
 */
#ifndef CHIMERA_H5_H
#define CHIMERA_H5_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces write_prims_to_hdf5 (src/mhd_write_h5.f90:12-177).  The six field arrays
 * are [nsnap][ny][nx] contiguous (snapshot s of field f at f + s*nx*ny); the file is
 * created or truncated. */
int chimera_h5_write_snapshots(const char *path, int nsnap, int nx, int ny, const double *rho,
                               const double *P, const double *Bx, const double *By,
                               const double *Vx, const double *Vy, const double *times,
                               double gamma, double M_s, double beta);

/* Readers (for checks and post-processing): number of links in a group, rank and
 * dims of a dataset, and a whole float64 dataset into out[n]. */
int chimera_h5_group_count(const char *path, const char *group, int *count);
int chimera_h5_dataset_dims(const char *path, const char *dataset, int *rank, long long dims[2]);
int chimera_h5_read_dataset(const char *path, const char *dataset, double *out, long long n);

const char *chimera_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CHIMERA_H5_H */
