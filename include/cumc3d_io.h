/*
 * cumc3d_io.h - C-ABI of libcumc3d_io.so, the writer of the reference's output
 * files: the HDF5 profile of print_hydroprofile (solver/mhd/dataio.f90:7-249)
 * and the div B log line of FIND_DIVB (solver/mhd/mhd_module.f90:199).
 *
 * Host code only (no GPU), linked against the HDF5 C library, so the step
 * library libcumc3d_hip.so carries no HDF5 dependency.  The arrays come from
 * cumc3d_profile() (include/cumc3d_hip.h), which packs them on the device in
 * print_hydroprofile's order, or from the reference's own module arrays.
 * Every entry point returns 0 on success, non-zero on failure
 * (cumc3d_io_last_error() gives the message).
 */
#ifndef CUMC3D_IO_H
#define CUMC3D_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* The scalars and coordinates of one profile file. */
typedef struct cumc3d_profile_meta {
    int nx, ny, nz;                 /* parameter.h:70-72 */
    int n_dim;                      /* written as the dataset "dimension" */
    int coordinate_flag;            /* written as "coordinate" = coordinate_flag + 1 */
    double time;                    /* global_time, written as "time" */
    const double *xF, *yF, *zF;     /* face coordinates xF(-3:nx+3) etc. (n+7 values);
                                       xF(-1:nx+1) is written as "x-interface" */
} cumc3d_profile_meta;

/* print_hydroprofile (dataio.f90:36-246): create (truncate) the HDF5 file
 * `path` with the datasets, in this order,
 *   time (1), dimension (1), coordinate (1),
 *   x-interface (nx+3), y-interface (ny+3), z-interface (nz+3),
 *   primitive  = prim(1:5, 0:nx+1, 0:ny+1, 0:nz+1)
 *   epsilon    = epsilon(0:nx+1, 0:ny+1, 0:nz+1)
 *   bfield     = prim(ibx:ibz, -1:nx+1, -1:ny+1, -1:nz+1)
 * all native double.  The array arguments hold those sections in Fortran
 * column-major order; as the reference's Fortran HDF5 calls do, the file
 * stores the dimensions reversed (primitive has C shape (nz+2, ny+2, nx+2, 5)),
 * which readers undo with a transpose (analysis/plot/plotprofile-1d.py:41). */
int cumc3d_io_write_profile(const char *path, const cumc3d_profile_meta *meta, const double *primitive,
                            const double *epsilon, const double *bfield);

/* The file name print_hydroprofile forms (dataio.f90:31-34):
 * "<dir>/rkiter-<n_iter>-nm.hdf5" (the reference uses dir "./outfile"). */
int cumc3d_io_profile_name(const char *dir, int n_iter, char *buf, int len);

/* Read one dataset of an HDF5 file back (for checks and for analysis without
 * h5py): cumc3d_io_dataset_shape gives its rank and C-order dimensions (up to
 * 8); cumc3d_io_read_dataset copies its values as doubles into out. */
int cumc3d_io_dataset_shape(const char *path, const char *name, int *rank, long long *dims);
int cumc3d_io_read_dataset(const char *path, const char *name, double *out, long long capacity);
/* Names of the file's root-group objects in name order, separated by '\n'. */
int cumc3d_io_list(const char *path, char *buf, int len);

/* One REAL*8 value as flang's list-directed output writes it: shortest
 * round-trip digits; F editing without a leading zero when the value rounded
 * to one significant digit has decimal exponent (0.d x 10^e) 0..15, else
 * 1P E editing with an exponent of at least two digits ("4.5E-02", "1.E+15");
 * "NaN", "Inf", "-Inf". */
int cumc3d_io_format_real(double x, char *buf, int len);
/* FIND_DIVB's log record WRITE(999,*) global_time, maxDivB
 * (mhd_module.f90:199): " <time> <maxdivb>\n". */
int cumc3d_io_format_divb(double time, double maxdivb, char *buf, int len);
/* Append that record to path (truncate = 1 starts the file afresh, as
 * OPENFILE_MHD's OPEN does at the start of a run, mhd_module.f90:98-105). */
int cumc3d_io_append_divb(const char *path, double time, double maxdivb, int truncate);

/* Message of the last error on this thread (empty if none). */
const char *cumc3d_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
