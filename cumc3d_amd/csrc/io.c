/*
 * libcumc3d_io.so: the reference's output files (include/cumc3d_io.h).
 *
 * print_hydroprofile (solver/mhd/dataio.f90:7-249) writes nine datasets
 * through HDF5's Fortran API, which hands the C library every dimension list
 * reversed; this writer calls the C API with the reversed lists directly, so
 * the files are the same.  FIND_DIVB's log record (mhd_module.f90:199) is
 * flang list-directed output of two REAL*8 values, restated in
 * cumc3d_io_format_real.
 *
 * Host code only; compiled with gcc against the HDF5 C library.
 */
#include "../../include/cumc3d_io.h"

#include <hdf5.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char g_err[512];

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return 1;
}

const char *cumc3d_io_last_error(void) { return g_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* ------------------------------------------------------------------------ */
/* print_hydroprofile                                                        */
/* ------------------------------------------------------------------------ */

static int write_dataset(hid_t file, const char *name, int rank, const hsize_t *dims, const double *data)
{
    hid_t space = H5Screate_simple(rank, dims, NULL);
    if (space < 0) return fail("H5Screate_simple failed for %s", name);
    hid_t dset = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (dset < 0) {
        H5Sclose(space);
        return fail("H5Dcreate2 failed for %s", name);
    }
    herr_t w = H5Dwrite(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(dset);
    H5Sclose(space);
    return w < 0 ? fail("H5Dwrite failed for %s", name) : 0;
}

int cumc3d_io_write_profile(const char *path, const cumc3d_profile_meta *m, const double *primitive,
                            const double *epsilon, const double *bfield)
{
    g_err[0] = 0;
    if (!path || !m || !primitive || !epsilon || !bfield || !m->xF || !m->yF || !m->zF)
        return fail("null argument");
    if (m->nx < 1 || m->ny < 1 || m->nz < 1) return fail("grid sizes must be >= 1");
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("cannot create %s", path);
    const hsize_t nx = (hsize_t)m->nx, ny = (hsize_t)m->ny, nz = (hsize_t)m->nz;
    const double dim = (double)m->n_dim, coord = (double)(m->coordinate_flag + 1);
    const hsize_t one[1] = {1};
    const hsize_t xi[1] = {nx + 3}, yi[1] = {ny + 3}, zi[1] = {nz + 3};
    /* Fortran dims (5, nx+2, ny+2, nz+2) etc., reversed as the Fortran API passes them */
    const hsize_t pd[4] = {nz + 2, ny + 2, nx + 2, 5};
    const hsize_t ed[3] = {nz + 2, ny + 2, nx + 2};
    const hsize_t bd[4] = {nz + 3, ny + 3, nx + 3, 3};
    int rc = 0;
    /* dataio.f90:36-233, in the reference's order; xF(-1) is entry 2 of xF(-3:nx+3) */
    rc = rc ? rc : write_dataset(f, "time", 1, one, &m->time);
    rc = rc ? rc : write_dataset(f, "dimension", 1, one, &dim);
    rc = rc ? rc : write_dataset(f, "coordinate", 1, one, &coord);
    rc = rc ? rc : write_dataset(f, "x-interface", 1, xi, m->xF + 2);
    rc = rc ? rc : write_dataset(f, "y-interface", 1, yi, m->yF + 2);
    rc = rc ? rc : write_dataset(f, "z-interface", 1, zi, m->zF + 2);
    rc = rc ? rc : write_dataset(f, "primitive", 4, pd, primitive);
    rc = rc ? rc : write_dataset(f, "epsilon", 3, ed, epsilon);
    rc = rc ? rc : write_dataset(f, "bfield", 4, bd, bfield);
    if (H5Fclose(f) < 0 && !rc) rc = fail("H5Fclose failed for %s", path);
    return rc;
}

int cumc3d_io_profile_name(const char *dir, int n_iter, char *buf, int len)
{
    g_err[0] = 0;
    if (!dir || !buf || len < 1) return fail("null argument");
    int n = snprintf(buf, (size_t)len, "%s/rkiter-%d-nm.hdf5", dir, n_iter);
    return (n < 0 || n >= len) ? fail("buffer too small") : 0;
}

/* ------------------------------------------------------------------------ */
/* reading back                                                              */
/* ------------------------------------------------------------------------ */

int cumc3d_io_dataset_shape(const char *path, const char *name, int *rank, long long *dims)
{
    g_err[0] = 0;
    if (!path || !name || !rank || !dims) return fail("null argument");
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open %s", path);
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        return fail("no dataset %s in %s", name, path);
    }
    hid_t s = H5Dget_space(d);
    int r = H5Sget_simple_extent_ndims(s);
    int rc = 0;
    if (r < 0 || r > 8) rc = fail("unsupported rank %d", r);
    else {
        hsize_t hd[8];
        H5Sget_simple_extent_dims(s, hd, NULL);
        *rank = r;
        for (int i = 0; i < r; ++i) dims[i] = (long long)hd[i];
    }
    H5Sclose(s);
    H5Dclose(d);
    H5Fclose(f);
    return rc;
}

int cumc3d_io_read_dataset(const char *path, const char *name, double *out, long long capacity)
{
    int rank;
    long long dims[8];
    if (cumc3d_io_dataset_shape(path, name, &rank, dims)) return 1;
    long long n = 1;
    for (int i = 0; i < rank; ++i) n *= dims[i];
    if (!out || n > capacity) return fail("output buffer holds %lld values, dataset %s has %lld", capacity, name, n);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    herr_t r = H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Dclose(d);
    H5Fclose(f);
    return r < 0 ? fail("H5Dread failed for %s", name) : 0;
}

typedef struct { char *buf; int len, used, overflow; } list_acc;

static herr_t list_cb(hid_t g, const char *name, const H5L_info_t *info, void *data)
{
    (void)g;
    (void)info;
    list_acc *a = (list_acc *)data;
    int n = (int)strlen(name);
    if (a->used + n + 2 > a->len) {
        a->overflow = 1;
        return 1;
    }
    memcpy(a->buf + a->used, name, (size_t)n);
    a->used += n;
    a->buf[a->used++] = '\n';
    a->buf[a->used] = 0;
    return 0;
}

int cumc3d_io_list(const char *path, char *buf, int len)
{
    g_err[0] = 0;
    if (!path || !buf || len < 1) return fail("null argument");
    quiet();
    buf[0] = 0;
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open %s", path);
    list_acc a = {buf, len, 0, 0};
    hsize_t idx = 0;
    H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, &idx, list_cb, &a);
    H5Fclose(f);
    return a.overflow ? fail("buffer too small") : 0;
}

/* ------------------------------------------------------------------------ */
/* flang list-directed REAL*8                                                */
/* ------------------------------------------------------------------------ */

/* Exact decimal expansion of v > 0: digits (801 significant) and the
 * exponent e with v = d0.d1d2... x 10^e.  A long double holding a double's
 * rounding-interval endpoint is a dyadic rational with at most about 750
 * significant decimal digits, so glibc's conversion here is exact. */
static int expand(long double v, char *dig, int cap)
{
    char s[900];
    snprintf(s, sizeof s, "%.800Le", v);
    int n = 0;
    const char *q = s;
    for (; *q && *q != 'e'; ++q)
        if (*q >= '0' && *q <= '9' && n < cap) dig[n++] = *q;
    dig[n] = 0;
    return atoi(q + 1);
}

/* floor(v / 10^k) for v given by expand() (at most 18 digits), and whether
 * the division leaves a remainder */
static unsigned long long floor_scaled(const char *dig, int e, int k, int *rem)
{
    int nint = e - k + 1;   /* digits before the point of v / 10^k */
    unsigned long long m = 0;
    for (int i = 0; i < nint; ++i) m = m * 10 + (unsigned long long)(dig[i] - '0');
    *rem = 0;
    for (int i = nint > 0 ? nint : 0; dig[i]; ++i)
        if (dig[i] != '0') { *rem = 1; break; }
    return m;
}

/* The digits flang's runtime prints for x > 0 (finite), found as it finds
 * them: the fewest significant digits p for which some p-digit decimal lies
 * strictly inside x's rounding interval (midpoints to the neighbouring
 * doubles, endpoints excluded), and of the n such decimals the one at index
 * floor((n-1)/2) counting up -- not the nearest (0x3FB1111111111111 prints
 * 6.666666666666666E-02).  d[0..n), e: x ~ d0.d1d2... x 10^e. */
static void shortest(double x, char *d, int *n, int *e)
{
    const double below = nextafter(x, 0.0), above = nextafter(x, INFINITY);
    const long double X = x;
    const long double lo = (X + (long double)below) / 2.0L;
    const long double hi = isinf(above) ? X + (X - (long double)below) / 2.0L : (X + (long double)above) / 2.0L;
    static _Thread_local char dlo[820], dhi[820];
    const int elo = expand(lo, dlo, 810), ehi = expand(hi, dhi, 810);
    for (int p = 1; p <= 17; ++p) {
        for (int ee = ehi; ee >= ehi - 1; --ee) {
            const int k = ee - p + 1;
            if (elo - k + 1 > 18 || ehi - k + 1 > 18) continue;
            int rlo, rhi;
            unsigned long long mmin = floor_scaled(dlo, elo, k, &rlo) + 1;
            unsigned long long mhi = floor_scaled(dhi, ehi, k, &rhi);
            unsigned long long mmax = rhi ? mhi : mhi - 1;
            unsigned long long p10 = 1;
            for (int i = 1; i < p; ++i) p10 *= 10;
            if (mmin < p10) mmin = p10;
            if (mmax > p10 * 10 - 1) mmax = p10 * 10 - 1;
            if (mhi == 0 && !rhi) continue;
            if (mmax < mmin) continue;
            unsigned long long m = mmin + (mmax - mmin) / 2;
            char t[24];
            int nt = snprintf(t, sizeof t, "%llu", m);
            while (nt > 1 && t[nt - 1] == '0') --nt;
            memcpy(d, t, (size_t)nt);
            *n = nt;
            *e = ee;
            return;
        }
    }
    /* not reached: 17 digits always fit inside the interval */
    snprintf(dlo, sizeof dlo, "%.16e", x);
    *n = 0;
    for (const char *q = dlo; *q && *q != 'e'; ++q)
        if (*q >= '0' && *q <= '9') d[(*n)++] = *q;
    *e = atoi(strchr(dlo, 'e') + 1);
}

int cumc3d_io_format_real(double x, char *buf, int len)
{
    g_err[0] = 0;
    if (!buf || len < 32) return fail("buffer must hold at least 32 characters");
    char out[64];
    int k = 0;
    if (isnan(x)) {
        strcpy(out, "NaN");
    } else if (isinf(x)) {
        strcpy(out, x < 0 ? "-Inf" : "Inf");
    } else {
        if (signbit(x)) out[k++] = '-';
        double a = fabs(x);
        if (a == 0.0) {
            out[k++] = '0';
            out[k++] = '.';
            out[k] = 0;
        } else {
            /* the F/E choice: decimal exponent of the value rounded to one
             * significant digit, in the 0.d x 10^e convention */
            char s1[16];
            snprintf(s1, sizeof s1, "%.0e", a);
            int expo = atoi(strchr(s1, 'e') + 1) + 1;
            char d[24];
            int n, e;
            shortest(a, d, &n, &e);
            if (expo < 0 || expo > 15) {
                out[k++] = d[0];
                out[k++] = '.';
                for (int i = 1; i < n; ++i) out[k++] = d[i];
                k += snprintf(out + k, sizeof out - (size_t)k, "E%c%02d", e < 0 ? '-' : '+', e < 0 ? -e : e);
            } else {
                int point = e + 1;   /* digits before the decimal point */
                if (point <= 0) {
                    out[k++] = '.';
                    for (int i = 0; i < -point; ++i) out[k++] = '0';
                    for (int i = 0; i < n; ++i) out[k++] = d[i];
                } else if (point >= n) {
                    for (int i = 0; i < n; ++i) out[k++] = d[i];
                    for (int i = n; i < point; ++i) out[k++] = '0';
                    out[k++] = '.';
                } else {
                    for (int i = 0; i < point; ++i) out[k++] = d[i];
                    out[k++] = '.';
                    for (int i = point; i < n; ++i) out[k++] = d[i];
                }
                out[k] = 0;
            }
        }
    }
    if ((int)strlen(out) >= len) return fail("buffer too small");
    strcpy(buf, out);
    return 0;
}

int cumc3d_io_format_divb(double time, double maxdivb, char *buf, int len)
{
    char a[64], b[64];
    if (cumc3d_io_format_real(time, a, sizeof a) || cumc3d_io_format_real(maxdivb, b, sizeof b)) return 1;
    if (!buf) return fail("null argument");
    int n = snprintf(buf, (size_t)len, " %s %s\n", a, b);
    return (n < 0 || n >= len) ? fail("buffer too small") : 0;
}

int cumc3d_io_append_divb(const char *path, double time, double maxdivb, int truncate)
{
    char line[160];
    if (!path) return fail("null argument");
    if (cumc3d_io_format_divb(time, maxdivb, line, sizeof line)) return 1;
    FILE *fp = fopen(path, truncate ? "w" : "a");
    if (!fp) return fail("cannot open %s", path);
    int ok = fputs(line, fp) >= 0;
    ok = (fclose(fp) == 0) && ok;
    return ok ? 0 : fail("write to %s failed", path);
}
