/* coffee_h5.h — C ABI of the snapshot / field-output files (libcoffee_h5.so).
 *
 * HDF5 files laid out as the reference's data exporter writes them on one
 * rank (src/utils/data_exporter.cpp:134-259, src/utils/hdf_wrapper_impl.hpp):
 * one dataset per array, named by the caller ("Ex", "B0x", "P", ...), whose
 * file dimensions are the array's extent in reverse order (a Fortran-order
 * array of extent {x, y, z} is the C-order dataset [z][y][x]), and scalar
 * datasets ("step" as uint32, "time" in the Scalar type).  Host buffers
 * only: the caller copies device arrays to host first (sim_data's
 * sync_to_host), as the reference does.
 *
 * Every entry point returns 0 on success and a negative code on failure
 * (never throws, never exits); coffee_h5_last_error() gives the message of
 * the calling thread's last failure.  Replaces Coffee::H5File and
 * hdf_create (src/utils/hdf_wrapper.h, hdf_wrapper.cpp:10-89).
 */
#ifndef COFFEE_H5_H
#define COFFEE_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define COFFEE_H5_OK 0
#define COFFEE_H5_ERR_INVALID (-1) /* bad argument */
#define COFFEE_H5_ERR_IO (-2)      /* the HDF5 library reported an error */
#define COFFEE_H5_ERR_SHAPE (-3)   /* dataset rank, extent or type differs from the request */

/* Element types (h5datatype<T>() of hdf_wrapper_impl.hpp). */
#define COFFEE_H5_F32 1
#define COFFEE_H5_F64 2
#define COFFEE_H5_U32 3
#define COFFEE_H5_I32 4

typedef struct coffee_h5_file coffee_h5_file;

/* Create (truncate) a file: hdf_create(filename, H5CreateMode::trunc). */
int coffee_h5_create(const char *path, coffee_h5_file **out);

/* Open an existing file, read-only (writable = 0) or read-write:
 * H5File(filename, H5OpenMode::read / rw). */
int coffee_h5_open(const char *path, int writable, coffee_h5_file **out);

/* Close and free the handle (NULL is a no-op): H5File::close. */
int coffee_h5_close(coffee_h5_file *f);

/* Write a whole array of `ndim` (1..3) dimensions, extent in Fortran order
 * (ext[0] varies fastest): H5File::write(const multi_array<T>&, name)
 * (hdf_wrapper_impl.hpp:29-49); on one rank write_parallel writes the same. */
int coffee_h5_write_array(coffee_h5_file *f, const char *name, const void *host, int type,
                          int ndim, const unsigned long long ext[3]);

/* Read a whole dataset into `host`; its rank, extent (Fortran order) and
 * element class and size must equal the request (COFFEE_H5_ERR_SHAPE
 * otherwise): H5File::read_subset over the full extent on one rank
 * (hdf_wrapper_impl.hpp:135-170). */
int coffee_h5_read_array(coffee_h5_file *f, const char *name, void *host, int type, int ndim,
                         const unsigned long long ext[3]);

/* Scalar datasets: H5File::write(T, name) and read_scalar<T>(name)
 * (hdf_wrapper_impl.hpp:8-27).  Reading converts to the requested type as
 * HDF5 does. */
int coffee_h5_write_scalar(coffee_h5_file *f, const char *name, int type, const void *value);
int coffee_h5_read_scalar(coffee_h5_file *f, const char *name, int type, void *value);

/* 1 if the file holds a dataset of that name, 0 if not, < 0 on error. */
int coffee_h5_has(coffee_h5_file *f, const char *name);

/* Message of the calling thread's last failure. */
const char *coffee_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* COFFEE_H5_H */
