#if !defined(ARRAY_H)
#define ARRAY_H

#include <stddef.h>
#include "sbfs_config.h"
#include "domain.h"

/*
 * A two-dimensional field with ghost / halo cells.  Same layout as the
 * reference's array_t (include/array.h:10-23).
 *
 *   size      bytes per element (8 for double, 16 for a 2-vector, 24 for
 *             a normal + intercept triple)
 *   nadds     extra cells on each side: nadds[0] = {left, right} in x,
 *             nadds[1] = {bottom, top} in y
 *   datasize  total bytes of `data`
 *   data      x-fastest storage of (Nx+nadds[0][0]+nadds[0][1]) x
 *             (Ny+nadds[1][0]+nadds[1][1]) elements
 *
 * In this library `data` is a DEVICE pointer (HBM of the current GPU).  Use
 * sbfs_array_download / sbfs_array_upload (sbfs.h) to move it.  Like the
 * reference, the solver swaps `data` pointers between arrays (the scalar
 * potentials and the Runge-Kutta sources; on the fused path also the vof and
 * the pressure with scratch buffers), so callers must not cache `data`
 * across calls.
 */
typedef struct {
  size_t size;
  int (* nadds)[2];
  size_t datasize;
  void * data;
} array_t;

#endif /* ARRAY_H */
