/* Snapshot and restart files in CRUNA's on-disk format (HDF5), for states
 * that the device loop produces and resumes from.  Host-only library
 * (libcruna_io.so, HDF5 C API); it replaces src/io/hdf5_single.f90's
 * store/load for the file types named below and helper/io_helper.f90's
 * get_fname for their names.
 *
 * Return values: 0 on success, -1 on failure; cruna_io_last_error() gives
 * the message of the last failure of this thread. */
#ifndef CRUNA_IO_H
#define CRUNA_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference layout constants (src/parameter/parameter.f90:9-12): the params
 * dataset holds CRUNA_IO_NPARAMS fixed-length rows of CRUNA_IO_PARAM_LEN
 * blank-padded characters ("name=v1;v2;v3;v4", create_parameter_list,
 * parameter.f90:790-813). */
#define CRUNA_IO_NPARAMS 512
#define CRUNA_IO_PARAM_LEN (384 * 5)

/* get_fname (helper/io_helper.f90:20-436) for type "data_direct_snapshot",
 * "data_adjoint_snapshot" (..._<block:02>_<image:05>_<ns:02>_<nt:07>),
 * "data_direct", "data_adjoint" (..._<block>_<image>_<ns>) and
 * "restartfile"; workdir as io.workdirectory (null: "."); ".h5" appended as
 * store/load do.  Fails on other types or when cap is too small. */
int cruna_io_fname(char *out, size_t cap, const char *type, const char *workdir, int world_block,
                   int block_image, int ns, int nt);

/* store_4d_rk (src/io/hdf5_single.f90:659-774), default path: a new file
 * (truncating) with dataset "data" = data(d1,d2,d3,d4) as native double in
 * Fortran order and dataset "params" = the nparams rows given (at most
 * CRUNA_IO_NPARAMS, each shorter than CRUNA_IO_PARAM_LEN), the rest blank. */
int cruna_io_store_4d(const char *fname, const double *data, const int dims[4],
                      const char *const *params, int nparams);

/* load_4d_rk (hdf5_single.f90:982-1053): dataset "data" into data(dims);
 * the file's dataset must have exactly these dimensions. */
int cruna_io_load_4d(const char *fname, double *data, const int dims[4]);

/* store_1d_i (hdf5_single.f90:175-299) as cruna_hpc.f90:259 uses it for the
 * restart file: dataset "data" = (ns, nt) as native int, plus "params". */
int cruna_io_store_restart(const char *fname, int ns, int nt, const char *const *params,
                           int nparams);

/* load_1d_i (hdf5_single.f90:1449-1518) of the restart file (cruna_hpc.f90:122): ns, nt. */
int cruna_io_load_restart(const char *fname, int *ns, int *nt);

/* Number of non-blank rows of "params" and row i (trailing blanks removed). */
int cruna_io_param_count(const char *fname, int *count);
int cruna_io_param_row(const char *fname, int i, char *out, size_t cap);

const char *cruna_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
