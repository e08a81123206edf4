/* fs3d.h -- C ABI of the 3D MAC grid on gfx950 (libfs2d.so).
 *
 * The reference's 3D classes (include/MACGrid3D.h, include/Array3D.h) are
 * grid storage and sampling, not yet used by its simulator (README.md:10, 25).
 * This ABI keeps their data layout and arithmetic and evaluates them on the
 * GPU over whole grids and point sets:
 *   - a grid holds u ((nx+1) x ny x nz), v (nx x (ny+1) x nz) and
 *     w (nx x ny x (nz+1)) in FP32; element (i, j, k) of an array with
 *     dimensions NX x NY x NZ is at [k*NY*NX + j*NX + i] (Array3D.h:32-37);
 *   - positions are in grid units, as MACGrid3D::velInterp takes them;
 *   - every call is ordered on the grid's HIP stream; calls that return host
 *     data wait for it.
 * Status codes and element types are those of include/fs2d.h.
 */
#ifndef FS3D_H
#define FS3D_H

#include <stddef.h>

#include "fs2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs3d_grid fs3d_grid;

typedef enum { FS3D_U = 0, FS3D_V = 1, FS3D_W = 2 } fs3d_field;

/* How MACGrid3D::velInterp samples (Array3D::triCubic, Array3D.h:75-126):
 *   FS3D_SAMPLE_REFERENCE  as written: taps read as (*this)(zp, yp, xp), i.e.
 *       element xp*NY*NX + yp*NX + zp with xp clamped to [0, NX-1] and zp to
 *       [0, NZ-1] (:106-121), and the sum clamped to >= 0 (:123).  For some
 *       array shapes a tap's element lies past the end of the array; such a
 *       sample is NaN and counted (fs3d_range_errors), where the reference
 *       reads out of bounds;
 *   FS3D_SAMPLE_VELOCITY   taps read as (xp, yp, zp), signed sum: the
 *       Catmull-Rom interpolation of a velocity component. */
typedef enum { FS3D_SAMPLE_REFERENCE = 0, FS3D_SAMPLE_VELOCITY = 1 } fs3d_sample;

/* MACGrid3D::create (MACGrid3D.h:16-23); the arrays start at zero (the
 * reference leaves them uninitialised).  nx, ny, nz >= 1. */
fs2d_status fs3d_create(int nx, int ny, int nz, fs3d_grid** out);
/* MACGrid3D::free (MACGrid3D.h:25-29) */
void fs3d_destroy(fs3d_grid* g);
/* Error message of the last failing call on g (of the last failed create when g is NULL). */
const char* fs3d_last_error(const fs3d_grid* g);
/* The grid's HIP stream (hipStream_t). */
void* fs3d_stream(fs3d_grid* g);

/* Element count of u, v or w. */
size_t fs3d_field_count(const fs3d_grid* g, fs3d_field f);
/* Whole-array copies of u, v, w (the public Array3D members, MACGrid3D.h:13),
 * FS2D_F32 or FS2D_F64 host data of exactly fs3d_field_count elements. */
fs2d_status fs3d_set_field(fs3d_grid* g, fs3d_field f, fs2d_dtype t, const void* host, size_t count);
fs2d_status fs3d_get_field(fs3d_grid* g, fs3d_field f, fs2d_dtype t, void* host, size_t count);
/* Device pointer of u, v or w (FP32), for zero-copy use on the grid's stream. */
float* fs3d_device_field(fs3d_grid* g, fs3d_field f);

/* MACGrid3D::velInterp (MACGrid3D.h:143-163) at n points: pos and out hold
 * x, y, z per point (3 n doubles, host memory); FP64 arithmetic on the FP32
 * grid.  Returns FS2D_ERR_ARG with the count in the message if any sample of
 * FS3D_SAMPLE_REFERENCE left the array (those samples are NaN). */
fs2d_status fs3d_vel_interp(fs3d_grid* g, const double* pos, size_t n, double* out, fs3d_sample mode);
/* The same on device memory (3 n floats each, FP32 arithmetic), queued on the
 * grid's stream without waiting; samples that left the array are NaN and are
 * added to the grid's range-error counter. */
fs2d_status fs3d_vel_interp_device(fs3d_grid* g, const float* pos, size_t n, float* out, fs3d_sample mode);
/* Samples counted as out of range since the last call (waits for the stream; resets the counter). */
long fs3d_range_errors(fs3d_grid* g);

/* MACGrid3D::velDiv (MACGrid3D.h:165-167) at every cell: nx*ny*nz values
 * in cell order, host memory, FS2D_F32 or FS2D_F64. */
fs2d_status fs3d_vel_div(fs3d_grid* g, fs2d_dtype t, void* host, size_t count);
/* MACGrid3D::vel (MACGrid3D.h:73-78) at every cell: 3 nx*ny*nz values, x, y, z per cell. */
fs2d_status fs3d_vel_cell(fs3d_grid* g, fs2d_dtype t, void* host, size_t count);
/* MACGrid3D::velU / velV / velW (MACGrid3D.h:80-141) at every face of
 * component `which`, in that array's order: 3 fs3d_field_count(which) values. */
fs2d_status fs3d_vel_face(fs3d_grid* g, fs3d_field which, fs2d_dtype t, void* host, size_t count);
/* Device-memory forms of the three above (FP32 out), queued on the grid's stream. */
fs2d_status fs3d_vel_div_device(fs3d_grid* g, float* out);
fs2d_status fs3d_vel_cell_device(fs3d_grid* g, float* out);
fs2d_status fs3d_vel_face_device(fs3d_grid* g, fs3d_field which, float* out);

/* Array3D arithmetic (Array3D.h:39-72) on n-element FP32 device arrays, on
 * the grid's stream.  The reference's loops run j to NY inclusive (Array3D.h:41,
 * 50, 58, 67), past each slab and past the array; these cover the n elements.
 *   innerProduct (:55-62), FP64 accumulation, fixed order (deterministic);
 *   infiniteNorm (:64-72);
 *   setMultiplyAdd (:48-53): dst = a + b c (dst may alias a or c);
 *   operator+= (:39-46) is setMultiplyAdd(dst, dst, 1, rhs). */
fs2d_status fs3d_array_inner_product(fs3d_grid* g, const float* a, const float* b, size_t n, double* out);
fs2d_status fs3d_array_infinite_norm(fs3d_grid* g, const float* a, size_t n, double* out);
fs2d_status fs3d_array_multiply_add(fs3d_grid* g, float* dst, const float* a, double b, const float* c, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* FS3D_H */
