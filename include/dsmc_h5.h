/* dsmc_h5.h -- HDF5 output of the DSMC state in the reference's file layout.
 *
 * Replaces the output surface of CUDADSMC/src/main.cu:202-383 (initial frame)
 * and main.cu:526-576 (one frame per print) with its helpers
 * hdf5Helpers.cu:13-163: a file holding group "/atomData" whose datasets are
 * rank 3, (rows, cols, t), chunked (rows, cols, 1), with t unlimited and
 * extended by one frame per write.  Host code only (links libhdf5); the
 * library is libdsmc_h5.so beside libdsmc.so.
 */
#ifndef DSMC_H5_H
#define DSMC_H5_H

#include <stddef.h>

#include "dsmc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dsmc_h5 dsmc_h5;

/* Element types of a dataset (H5T_NATIVE_DOUBLE, H5T_NATIVE_INT,
 * H5T_NATIVE_HBOOL). */
enum { DSMC_H5_DOUBLE = 0, DSMC_H5_INT = 1, DSMC_H5_BOOL = 2 };

/* createHDF5File (hdf5Helpers.cu:46-63): a new file (truncating an existing
 * one) with the group `group` (e.g. "/atomData"). */
dsmc_status dsmc_h5_create(const char *filename, const char *group, dsmc_h5 **out);

/* createHDF5Handle + intialiseHDF5File (hdf5Helpers.cu:13-44, 64-112): dataset
 * `name` (full path, e.g. "/atomData/positions") of shape (rows, cols, 0). */
dsmc_status dsmc_h5_dataset(dsmc_h5 *h, const char *name, long long rows, long long cols,
                            int type);

/* writeHDF5File (hdf5Helpers.cu:115-163): appends one frame, rows x cols
 * elements in C order (row-major), and flushes the file. */
dsmc_status dsmc_h5_append(dsmc_h5 *h, const char *name, const void *data);

/* One frame of the thirteen datasets of main.cu:526-576 from a context:
 * positions, velocities, evapPos, evapVel (n x 3 double), psiUp, psiDn (n x 2
 * double), collisionCount, atomCount (cells + 1 int), simuatedTime (1 double),
 * atomNumber (1 int), cellID, atomID (n int), atomIsSpinUp (n bool).  n is the
 * context's capacity (the reference's numberOfAtoms at creation); rows past
 * the alive atoms are zero.  atomIsSpinUp is DSMC_FIELD_SPIN_UP.  The datasets are created by the first call.
 * Fields that the context's numerical form does not hold are written as
 * zeros: the spinor of an FP32 run is the Bloch vector's spinor (phase of
 * psiUp zero), per-cell counts exist in the FP64 form only. */
dsmc_status dsmc_h5_write_frame(dsmc_h5 *h, dsmc_ctx *ctx, double time);

/* Flushes and closes the file. */
void dsmc_h5_close(dsmc_h5 *h);

/* Reading back (for the checkers and tests): the dims (rows, cols, t) of a
 * dataset and its whole contents in C order, converted to `type`. */
dsmc_status dsmc_h5_dims(const char *filename, const char *name, long long dims[3]);
dsmc_status dsmc_h5_read(const char *filename, const char *name, int type, void *out,
                         size_t bytes);

/* Message of the last failing call of this library (thread-local). */
const char *dsmc_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DSMC_H5_H */
