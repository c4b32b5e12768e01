/*
 * diffuse_yell.h -- the DISCUS "Yell 1.0" HDF5 output file
 * (lib_f90/hdf_write.f90:12-326, gen_hdf5_write; called by DISCUS's output
 * menu for `outfile type hdf5`, discus/prog/graphic.f90:606-636).
 *
 * A separate library (libdiffuse_yell.so) linked against HDF5, as DISCUS
 * builds hdf_write.f90 or its stub hdf_write_no.f90 depending on whether HDF5
 * is present.  The values are the per-pixel qval numbers of the run
 * (dc_qval / dc_qval_dev of diffuse_turbo.h), already multiplied by the
 * output scale as graphic.f90:607-617 does.
 *
 * File layout (what hdf_write.f90 writes and lib_f90/hdf5_read.f90 reads):
 *   data            double, C-order dims (inc1, inc2, inc3): data[i][j][k] = qvalues(i,j,k)
 *   format          scalar variable-length NUL-terminated string "Yell 1.0"
 *   is_direct       scalar big-endian 8-bit integer, 1 for value 14 (PDF) or 15 (3DPDF)
 *   lower_limits    3 x big-endian double: eck(:,1)
 *   step_sizes      3 x big-endian double: vi(extr_abs,1), vi(extr_ord,2), vi(extr_top,3)
 *   step_sizes_abs  3 x big-endian double: vi(:,1)   (and _ord vi(:,2), _top vi(:,3))
 *   unit_cell       6 x big-endian double: a, b, c, alpha, beta, gamma
 *   PROGRAM         1 x 9-byte NUL-terminated string "DISCUS60"
 *   sigma           like data, when sigma values are given
 * With valmax > 0, data (and sigma) are scaled by valmax / max(values) when
 * max(values) >= 0, as hdf_write.f90:135-140 does.
 */
#ifndef DIFFUSE_YELL_H
#define DIFFUSE_YELL_H

#ifdef __cplusplus
extern "C" {
#endif

#define DC_YELL_OK 0
#define DC_YELL_ERR_IO (-2)      /* HDF5 could not open or write the file (ier_num -2, ER_IO) */
#define DC_YELL_ERR_ARG (-1000)  /* null pointer, size or index out of range */

typedef struct dc_yell_desc {
    int inc[3];        /* out_inc: points along abscissa, ordinate, top (>= 1) */
    double eck[12];    /* out_eck(3,4), column-major; the corner is eck(:,1) */
    double vi[9];      /* out_vi(3,3), column-major; vi(:,1) is the abscissa step */
    int extr[3];       /* out_extr_abs, _ord, _top: which of h, k, l (1..3) */
    double a0[3];      /* cr_a0: cell lengths */
    double win[3];     /* cr_win: cell angles (degrees) */
    int value;         /* the qval code of the values (14 PDF and 15 3DPDF mark is_direct) */
    double valmax;     /* > 0: rescale to this maximum; <= 0: write as given */
} dc_yell_desc;

/* gen_hdf5_write (hdf_write.f90:12-326).  qvalues and sigma (NULL: no sigma
 * dataset) are inc(1) x inc(2) x inc(3), Fortran order (abscissa fastest).
 * An existing file is replaced. */
int dc_write_yell(const char *path, const dc_yell_desc *desc, const double *qvalues, const double *sigma);

#ifdef __cplusplus
}
#endif

#endif
