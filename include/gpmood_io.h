/*
 * gpmood_io.h -- the reference's HDF5 file formats (host code, libgpmood_io.so).
 *
 * Snapshots, restarts, per-step diagnostics and NN training datasets in the
 * layouts the reference writes and reads (gp-mood-code/output.f90,
 * reader.f90), so that its plotting scripts (plot/plotter_hdf5.py,
 * plot/plotter_diag.py), its dataset loader (neural_network/utils.py) and its
 * restart path (main.f90:133-138) read this library's files, and this
 * library reads the reference's.  Separate from libgpmood.so so that the step
 * library has no HDF5 dependency; needs the HDF5 C library (1.10).
 *
 * Layout notes.  The reference writes from Fortran, whose dimensions HDF5
 * stores reversed: a field dims = [lf, nf] is a C dataset of shape (nf, lf),
 * row n-1 holding cells (1..lf, n).  Strings are 1-D datasets of one-byte
 * H5T_C_S1 elements, one per character.  Scalars are 1-element 1-D datasets.
 *
 * Every function returns 0 on success and a negative GPM_IO_E* code on
 * failure.
 */
#ifndef GPMOOD_IO_H
#define GPMOOD_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPM_IO_OK 0
#define GPM_IO_EINVAL -1  /* invalid argument, or a problem/method without a name (output.f90:13-69) */
#define GPM_IO_EFILE -2   /* HDF5 could not create, open, read or write the file */

/* The metadata compute_metadata (output.f90:13-72) derives from the
 * `parameters` module. */
typedef struct gpm_io_meta {
    int32_t lf, nf;
    int32_t problem;          /* the reference's problem id (constants.f90:27-34) */
    int32_t method;           /* FOG 1, GP_MOOD 2, POL_MOOD 200, NN_GP_MOOD 203, NN_GP_MOOD_CC 204 */
    int32_t dmp;              /* DMP switch: GP_MOOD without DMP is named GP_MOOD_noDMP */
    double cfl;
    const char *nn_filename;  /* NN_filename (may be empty) */
} gpm_io_meta;

/* The stem `file` of compute_metadata:
 *   output_<problem>_<method>_CFL_<cfl>_<lf>_<nf>
 * and the snapshot name write_output uses, <stem>_<niter + 100000>.h5.
 * Pass niter < 0 for the stem alone.  buf receives a NUL-terminated name. */
int gpm_io_file_name(const gpm_io_meta *meta, int32_t niter, char *buf, int32_t bufsize);

/* write_output (output.f90:74-143): primitive rho, ux, uy, p of the interior
 * of U (double U[nf+8][lf+8][4], the reference's ghosted layout; primitives by
 * conservative_to_primitive, physics.f90:162-175), the order masks ordr and
 * ordr_priori (int32 [nf][lf]; with ordr_priori NULL every cell holds
 * mord_if_null, the CellGPO_priori of a method without an a priori predictor,
 * mod_FE.f90:49), the strings CFL, problem, method, NN_filename, and lf, nf,
 * niter, time. */
int gpm_io_write_output(const char *path, const gpm_io_meta *meta, const double *U, const int32_t *ordr,
                        const int32_t *ordr_priori, int32_t mord_if_null, int32_t niter, double t);

/* read (reader.f90:12-120): the interior of U from a snapshot's rho, ux, uy,
 * p by primitive_to_conservative (physics.f90:146-160) into double
 * U[nf+8][lf+8][4] (ghost cells untouched), and t and niter.  The file's
 * fields must be nf x lf. */
int gpm_io_read_restart(const char *path, int32_t lf, int32_t nf, double *U, double *t, int32_t *niter);

/* write_diagnostic (output.f90:145-214): n steps of the float32 series time,
 * pct_detected_cell, pct_detected_cell_a_priori, pct_detected_cell_a_posteriori
 * and the int32 series steps_NN_produced_NAN, steps_NN_sample (each may be
 * NULL for zeros), the strings, and lf, nf, count_steps_NN_produced_NAN,
 * iter_0, niter. */
int gpm_io_write_diagnostic(const char *path, const gpm_io_meta *meta, int32_t n, const float *time,
                            const float *pct, const float *pct_priori, const float *pct_posteriori,
                            const int32_t *steps_nn_nan, const int32_t *steps_nn_sample,
                            int32_t count_steps_nn_nan, int32_t iter_0, int32_t niter);

/* write_NN_dataset_ (output.f90:216-300): `size` samples, inputs float
 * [size][L] and labels float [size][2] (one row per sample), the strings,
 * and lf, nf, n_overwrite, NR0, NR1. */
int gpm_io_write_nn_dataset(const char *path, const gpm_io_meta *meta, int32_t size, int32_t L, const float *inputs,
                            const float *labels, int32_t n_overwrite, int32_t nr0, int32_t nr1);

#ifdef __cplusplus
}
#endif

#endif /* GPMOOD_IO_H */
