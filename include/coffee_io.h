/* coffee_io.h — device side of the field output (part of libcoffee_hip.so).
 *
 * Same conventions as coffee_hip.h: device pointers, Fortran order with
 * guards, asynchronous on `stream`, 0 on success, negative codes and
 * coffee_last_error(NULL) on failure.
 */
#ifndef COFFEE_IO_H
#define COFFEE_IO_H

#include "coffee_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sample every d-th interior cell of one array of `grid` (from the first
 * interior cell) and interpolate it from `stagger` (bit a set = cell-centred
 * along axis a) to the cell centre, in the array's Scalar type, into the
 * float array `dst` of extent ext[a] = N[a] / d where N[a] > d, else N[a]
 * (N = dims - 2 guard), Fortran order.  3D grids: multi_array::downsample
 * with interpolate (src/data/multi_array.cpp:56-73,
 * src/cuda/data/multi_array.cu:52-71, src/algorithms/interpolation.h:9-34);
 * 2D grids (dims[2] == 1): downsample2d with interpolate2d
 * (multi_array.cu:73-92, interpolation.h:36-50).  An axis with
 * 1 < N[a] <= d keeps its N[a] cells at stride 1 (the reference steps by d
 * there, beyond the array).  1D grids are rejected, as the reference leaves
 * them unimplemented. */
int coffee_downsample(const coffee_grid *grid, const void *src, int scalar_bytes, int stagger,
                      int d, float *dst, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* COFFEE_IO_H */
