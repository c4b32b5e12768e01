/*
 * dc_yell.c -- DISCUS "Yell 1.0" HDF5 output (lib_f90/hdf_write.f90:12-326,
 * gen_hdf5_write), written with the HDF5 C library.
 *
 * Dataset names, types, shapes and the order of writing follow
 * gen_hdf5_write; lib_f90/hdf5_read.f90:59-688 is the reader that consumes
 * them (format must read back as "Yell 1.0", hdf5_read.f90:227-255).
 */
#include <hdf5.h>
#include <stdlib.h>
#include <string.h>

#include "diffuse_yell.h"

/* The values cube in HDF5 C order (inc1, inc2, inc3) with the top index
 * fastest: the transpose of hdf_write.f90:117-133 (values(k,j,i) =
 * qvalues(i,j,k) on Fortran dims (inc3, inc2, inc1)).  valmax > 0 rescales
 * by valmax * v / max(v) when max(v) >= 0 (hdf_write.f90:135-140); an
 * all-non-positive cube with maximum 0 gives NaN/-Inf there, as in DISCUS. */
static int write_cube(hid_t file, const char *name, const dc_yell_desc *d, const double *v) {
    const size_t n1 = (size_t)d->inc[0], n2 = (size_t)d->inc[1], n3 = (size_t)d->inc[2];
    const size_t n = n1 * n2 * n3;
    double *buf = (double *)malloc(n * sizeof(double));
    if (!buf) return DC_YELL_ERR_IO;
    for (size_t i = 0; i < n1; ++i)
        for (size_t j = 0; j < n2; ++j)
            for (size_t k = 0; k < n3; ++k) buf[(i * n2 + j) * n3 + k] = v[i + n1 * (j + n2 * k)];
    if (d->valmax > 0.0) {
        double mx = buf[0];
        for (size_t x = 1; x < n; ++x)
            if (buf[x] > mx) mx = buf[x];
        if (mx >= 0.0)
            for (size_t x = 0; x < n; ++x) buf[x] = d->valmax * buf[x] / mx;
    }
    const hsize_t dims[3] = {(hsize_t)n1, (hsize_t)n2, (hsize_t)n3};
    int st = DC_YELL_ERR_IO;
    const hid_t sp = H5Screate_simple(3, dims, NULL);
    if (sp >= 0) {
        const hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds >= 0) {
            if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0) st = DC_YELL_OK;
            if (H5Dclose(ds) < 0) st = DC_YELL_ERR_IO;
        }
        H5Sclose(sp);
    }
    free(buf);
    return st;
}

/* n doubles stored big-endian (H5T_IEEE_F64BE), written from native memory */
static int write_f64be(hid_t file, const char *name, const double *v, hsize_t n) {
    int st = DC_YELL_ERR_IO;
    const hid_t sp = H5Screate_simple(1, &n, NULL);
    if (sp < 0) return st;
    const hid_t ds = H5Dcreate2(file, name, H5T_IEEE_F64BE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds >= 0) {
        if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0) st = DC_YELL_OK;
        if (H5Dclose(ds) < 0) st = DC_YELL_ERR_IO;
    }
    H5Sclose(sp);
    return st;
}

/* format: scalar, variable-length, NUL-terminated (hdf_write.f90:155-163) */
static int write_format(hid_t file) {
    int st = DC_YELL_ERR_IO;
    const hid_t t = H5Tcopy(H5T_C_S1);
    if (t < 0) return st;
    const hid_t sp = H5Screate(H5S_SCALAR);
    if (sp >= 0 && H5Tset_size(t, H5T_VARIABLE) >= 0 && H5Tset_strpad(t, H5T_STR_NULLTERM) >= 0) {
        const hid_t ds = H5Dcreate2(file, "format", t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds >= 0) {
            const char *w = "Yell 1.0";
            if (H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, &w) >= 0) st = DC_YELL_OK;
            if (H5Dclose(ds) < 0) st = DC_YELL_ERR_IO;
        }
    }
    if (sp >= 0) H5Sclose(sp);
    H5Tclose(t);
    return st;
}

/* is_direct: scalar 8-bit big-endian integer, 0 or 1 (hdf_write.f90:168-179) */
static int write_is_direct(hid_t file, int value) {
    int st = DC_YELL_ERR_IO;
    const signed char v = (value == 15 || value == 14) ? 1 : 0;   /* val_3DPDF, val_pdf (qval.f90:16-17) */
    const hid_t sp = H5Screate(H5S_SCALAR);
    if (sp < 0) return st;
    const hid_t ds = H5Dcreate2(file, "is_direct", H5T_STD_I8BE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds >= 0) {
        if (H5Dwrite(ds, H5T_STD_I8BE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) >= 0) st = DC_YELL_OK;
        if (H5Dclose(ds) < 0) st = DC_YELL_ERR_IO;
    }
    H5Sclose(sp);
    return st;
}

/* PROGRAM: one 9-byte C string "DISCUS60" + NUL (hdf_write.f90:268-283) */
static int write_program(hid_t file) {
    int st = DC_YELL_ERR_IO;
    const hid_t t = H5Tcopy(H5T_C_S1);
    if (t < 0) return st;
    const hsize_t one = 1;
    const hid_t sp = H5Screate_simple(1, &one, NULL);
    if (sp >= 0 && H5Tset_size(t, 9) >= 0) {
        const hid_t ds = H5Dcreate2(file, "PROGRAM", t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds >= 0) {
            const char w[9] = {'D', 'I', 'S', 'C', 'U', 'S', '6', '0', '\0'};
            if (H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, w) >= 0) st = DC_YELL_OK;
            if (H5Dclose(ds) < 0) st = DC_YELL_ERR_IO;
        }
    }
    if (sp >= 0) H5Sclose(sp);
    H5Tclose(t);
    return st;
}

int dc_write_yell(const char *path, const dc_yell_desc *d, const double *qvalues, const double *sigma) {
    if (!path || !d || !qvalues) return DC_YELL_ERR_ARG;
    for (int a = 0; a < 3; ++a) {
        if (d->inc[a] < 1) return DC_YELL_ERR_ARG;
        if (d->extr[a] < 1 || d->extr[a] > 3) return DC_YELL_ERR_ARG;
    }
    if ((double)d->inc[0] * (double)d->inc[1] * (double)d->inc[2] * 8.0 > 1e15) return DC_YELL_ERR_ARG;
    if (H5open() < 0) return DC_YELL_ERR_IO;
    /* the library's error stack stays quiet; failures come back as status codes */
    H5E_auto2_t efunc = NULL;
    void *edata = NULL;
    H5Eget_auto2(H5E_DEFAULT, &efunc, &edata);
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);

    int st = DC_YELL_ERR_IO;
    const hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file >= 0) {
        /* column j of vi(3,3) is vi + 3 j; eck(:,1) is eck[0..2] */
        const double steps[3] = {d->vi[d->extr[0] - 1], d->vi[3 + d->extr[1] - 1], d->vi[6 + d->extr[2] - 1]};
        const double cell[6] = {d->a0[0], d->a0[1], d->a0[2], d->win[0], d->win[1], d->win[2]};
        st = write_cube(file, "data", d, qvalues);
        if (!st) st = write_format(file);
        if (!st) st = write_is_direct(file, d->value);
        if (!st) st = write_f64be(file, "lower_limits", d->eck, 3);
        if (!st) st = write_f64be(file, "step_sizes", steps, 3);
        if (!st) st = write_f64be(file, "step_sizes_abs", d->vi, 3);
        if (!st) st = write_f64be(file, "step_sizes_ord", d->vi + 3, 3);
        if (!st) st = write_f64be(file, "step_sizes_top", d->vi + 6, 3);
        if (!st) st = write_f64be(file, "unit_cell", cell, 6);
        if (!st) st = write_program(file);
        if (!st && sigma) st = write_cube(file, "sigma", d, sigma);
        if (H5Fclose(file) < 0) st = DC_YELL_ERR_IO;
    }
    H5Eset_auto2(H5E_DEFAULT, efunc, edata);
    return st;
}
