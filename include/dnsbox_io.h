/*
 * dnsbox_io.h — state files of the reference (host code, no GPU), in libdnsbox_h5.so.
 *
 * stateNNNN.h5 holds three datasets u, v, w: the physical velocity in FP64 with Fortran shape
 * (Nx, Ny, Nz), i.e. the C-order dataspace {Nz, Ny, Nx}, x fastest.  The arrays passed here use
 * that byte order, the layout of dnsbox_get_state_physical / dnsbox_set_state_physical
 * (include/dnsbox.h).  Built against the HDF5 C library (1.10); where that library is absent the
 * library is not built and the Python front-end raises when a state file is asked for.
 *
 * Every function returns 0 or a negative DNSBOX_IO_E_* code; dnsbox_h5_last_error() gives the
 * message (thread-local).
 */
#ifndef DNSBOX_IO_H
#define DNSBOX_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define DNSBOX_IO_E_ARG (-1)  /* invalid argument, or a dataset of another shape */
#define DNSBOX_IO_E_IO  (-5)  /* the file cannot be created, opened, read or written */

const char* dnsbox_h5_last_error(void);
/* io_save_state (io.f90:100-278): create (truncate) path with datasets u, v, w */
int dnsbox_h5_write_state(const char* path, int nx, int ny, int nz, const double* u, const double* v,
                          const double* w);
/* the grid of a state file: Fortran shape (nx, ny, nz) of its dataset u */
int dnsbox_h5_state_dims(const char* path, int* nx, int* ny, int* nz);
/* io_load_state (io.f90:455-626): read datasets u, v, w, which must have Fortran shape (nx, ny, nz) */
int dnsbox_h5_read_state(const char* path, int nx, int ny, int nz, double* u, double* v, double* w);

#ifdef __cplusplus
}
#endif
#endif /* DNSBOX_IO_H */
