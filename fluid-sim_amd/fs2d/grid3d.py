"""ctypes binding of include/fs3d.h: the reference's 3D MAC grid
(include/MACGrid3D.h, include/Array3D.h) on the GPU.

Arrays are numpy arrays of shape (NZ, NY, NX) in C order, i.e. element
(i, j, k) at [k, j, i], the reference's [k*NY*NX + j*NX + i].  Positions are
(n, 3) arrays of x, y, z in grid units, as MACGrid3D::velInterp takes them.
"""
import ctypes

import numpy as np

from . import F32, F64, OK, FS2DError, load

U, V, W = 0, 1, 2
SAMPLE_REFERENCE, SAMPLE_VELOCITY = 0, 1


class MACGrid3D:
    """MACGrid3D::create (MACGrid3D.h:16-23) / free (:25-29) on the GPU."""

    def __init__(self, nx, ny, nz):
        self.L = load()
        h = ctypes.c_void_p()
        st = self.L.fs3d_create(nx, ny, nz, ctypes.byref(h))
        if st != OK:
            raise FS2DError(st, self.L.fs3d_last_error(None).decode())
        self.h = h
        self.nx, self.ny, self.nz = nx, ny, nz

    def _check(self, st):
        if st != OK:
            raise FS2DError(st, self.L.fs3d_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.fs3d_destroy(self.h)
            self.h = None

    free = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, f):
        """(NZ, NY, NX) of u, v or w."""
        return (self.nz + (f == W), self.ny + (f == V), self.nx + (f == U))

    def set(self, f, a):
        a = np.ascontiguousarray(a)
        assert a.shape == self.shape(f), (a.shape, self.shape(f))
        t = F64 if a.dtype == np.float64 else F32
        a = a if t == F64 else a.astype(np.float32)
        self._check(self.L.fs3d_set_field(self.h, f, t, a.ctypes.data, a.size))

    def get(self, f, dtype=np.float32):
        out = np.empty(self.shape(f), dtype=dtype)
        self._check(self.L.fs3d_get_field(self.h, f, F64 if dtype == np.float64 else F32, out.ctypes.data, out.size))
        return out

    def device_field(self, f):
        return self.L.fs3d_device_field(self.h, f)

    def vel_interp(self, pos, mode=SAMPLE_REFERENCE):
        """MACGrid3D::velInterp (MACGrid3D.h:143-163) at every row of pos; FP64."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(pos)
        self._check(self.L.fs3d_vel_interp(self.h, pos.ctypes.data, len(pos), out.ctypes.data, mode))
        return out

    def vel_div(self, dtype=np.float32):
        """MACGrid3D::velDiv (MACGrid3D.h:165-167) at every cell, shape (nz, ny, nx)."""
        out = np.empty((self.nz, self.ny, self.nx), dtype=dtype)
        self._check(self.L.fs3d_vel_div(self.h, F64 if dtype == np.float64 else F32, out.ctypes.data, out.size))
        return out

    def vel_cell(self, dtype=np.float32):
        """MACGrid3D::vel (MACGrid3D.h:73-78) at every cell, shape (nz, ny, nx, 3)."""
        out = np.empty((self.nz, self.ny, self.nx, 3), dtype=dtype)
        self._check(self.L.fs3d_vel_cell(self.h, F64 if dtype == np.float64 else F32, out.ctypes.data, out.size))
        return out

    def vel_face(self, which, dtype=np.float32):
        """MACGrid3D::velU / velV / velW (MACGrid3D.h:80-141) at every face of `which`."""
        out = np.empty(self.shape(which) + (3,), dtype=dtype)
        self._check(self.L.fs3d_vel_face(self.h, which, F64 if dtype == np.float64 else F32, out.ctypes.data,
                                         out.size))
        return out

    # Array3D arithmetic (Array3D.h:39-72) on device arrays of n FP32 elements
    def inner_product(self, a_dev, b_dev, n):
        r = ctypes.c_double()
        self._check(self.L.fs3d_array_inner_product(self.h, a_dev, b_dev, n, ctypes.byref(r)))
        return r.value

    def infinite_norm(self, a_dev, n):
        r = ctypes.c_double()
        self._check(self.L.fs3d_array_infinite_norm(self.h, a_dev, n, ctypes.byref(r)))
        return r.value

    def multiply_add(self, dst_dev, a_dev, b, c_dev, n):
        self._check(self.L.fs3d_array_multiply_add(self.h, dst_dev, a_dev, b, c_dev, n))
