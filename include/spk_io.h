/*
 * spk_io.h -- C ABI of the snapshot files of a SPOOKY run (libspkio.so).
 *
 * A snapshot is an HDF5 file data/snapNNNN.h5 with the reference's layout
 * (src/io/hdf5_io.cu:19-106 writes it, :110-252 reads it):
 *   t_start, t_save, t_lastsnap, t_lastvar : float64, shape (1,)
 *   step                                   : int32,   shape (1,)
 *   vx vy vz bx by bz th (those present)   : float64, shape (NX*NY*NZ,),
 *       real-space values in C order [NX][NY][NZ] (no padding).
 * Used by the driver (spooky) and, through ctypes, by the initial-condition
 * writer that replaces the h5py script create_initial_conditions.py.
 * Every call returns 0 or a negative status; spkio_last_error() has the text.
 */
#ifndef SPK_IO_H
#define SPK_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *spkio_last_error(void);

/* Writes (truncates) a snapshot: nfields real fields of n = nx*ny*nz values,
 * fields[f] names names[f]. */
int spkio_write_snapshot(const char *path, int nx, int ny, int nz, int nfields, const char *const *names,
                         const double *const *fields, double t_start, double t_save, double t_lastsnap,
                         double t_lastvar, int step);

/* Scalars of a snapshot (InputOutput::ReadDataFile, hdf5_io.cu:176-190). */
int spkio_read_header(const char *path, double *t_save, double *t_lastsnap, double *t_lastvar, int *step);

/* One field: out[n], n must equal the dataset length. */
int spkio_read_field(const char *path, const char *name, double *out, size_t n);

/* Index of the last of the consecutive files snap0000.h5, snap0001.h5, ...
 * in data_dir (fileCounter, hdf5_io.cu:255-276), -1 when there is none. */
int spkio_last_snapshot(const char *data_dir);

#ifdef __cplusplus
}
#endif
#endif /* SPK_IO_H */
