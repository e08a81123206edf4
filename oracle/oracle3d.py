"""ctypes wrapper over the 3D part of oracle/liboracle.so (fs3d_oracle.cpp):
the CPU restatement of include/Array3D.h and include/MACGrid3D.h in FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's checks, always
as the checker.  Nothing in the product package imports it.

Arrays are numpy (NZ, NY, NX) in C order: element (i, j, k) at [k, j, i].
"""
import ctypes

import numpy as np

from . import oracle as _o

REFERENCE, VELOCITY = 0, 1

_sig_done = False


def lib():
    global _sig_done
    L = _o.lib()
    if not _sig_done:
        vp, i, l, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.o3_vel_interp.restype = l
        L.o3_vel_interp.argtypes = [vp, vp, vp, i, i, i, vp, l, vp, i]
        L.o3_tricubic.restype = l
        L.o3_tricubic.argtypes = [vp, i, i, i, vp, l, vp, i]
        L.o3_vel_div.argtypes = [vp, vp, vp, i, i, i, vp]
        L.o3_vel_cell.argtypes = [vp, vp, vp, i, i, i, vp]
        L.o3_vel_face.argtypes = [vp, vp, vp, i, i, i, i, vp]
        L.o3_inner_product.restype = d
        L.o3_inner_product.argtypes = [vp, vp, l]
        L.o3_infinite_norm.restype = d
        L.o3_infinite_norm.argtypes = [vp, l]
        L.o3_multiply_add.argtypes = [vp, vp, d, vp, l]
        _sig_done = True
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _mac(u, v, w):
    u, v, w = _f64(u), _f64(v), _f64(w)
    nz, ny, nx1 = u.shape
    nx = nx1 - 1
    assert v.shape == (nz, ny + 1, nx) and w.shape == (nz + 1, ny, nx), (u.shape, v.shape, w.shape)
    return u, v, w, nx, ny, nz


def vel_interp(u, v, w, pos, mode=REFERENCE):
    """MACGrid3D::velInterp at each row of pos; returns (values (n, 3), out-of-range sample count)."""
    u, v, w, nx, ny, nz = _mac(u, v, w)
    pos = _f64(pos).reshape(-1, 3)
    out = np.empty_like(pos)
    bad = lib().o3_vel_interp(u.ctypes.data, v.ctypes.data, w.ctypes.data, nx, ny, nz, pos.ctypes.data, len(pos),
                              out.ctypes.data, mode)
    return out, bad


def tricubic(a, pos, mode=REFERENCE):
    """Array3D::triCubic of one (NZ, NY, NX) array at each row of pos."""
    a = _f64(a)
    NZ, NY, NX = a.shape
    pos = _f64(pos).reshape(-1, 3)
    out = np.empty(len(pos))
    bad = lib().o3_tricubic(a.ctypes.data, NX, NY, NZ, pos.ctypes.data, len(pos), out.ctypes.data, mode)
    return out, bad


def vel_div(u, v, w):
    u, v, w, nx, ny, nz = _mac(u, v, w)
    out = np.empty((nz, ny, nx))
    lib().o3_vel_div(u.ctypes.data, v.ctypes.data, w.ctypes.data, nx, ny, nz, out.ctypes.data)
    return out


def vel_cell(u, v, w):
    u, v, w, nx, ny, nz = _mac(u, v, w)
    out = np.empty((nz, ny, nx, 3))
    lib().o3_vel_cell(u.ctypes.data, v.ctypes.data, w.ctypes.data, nx, ny, nz, out.ctypes.data)
    return out


def vel_face(u, v, w, which):
    u, v, w, nx, ny, nz = _mac(u, v, w)
    shape = (nz + (which == 2), ny + (which == 1), nx + (which == 0))
    out = np.empty(shape + (3,))
    lib().o3_vel_face(u.ctypes.data, v.ctypes.data, w.ctypes.data, nx, ny, nz, which, out.ctypes.data)
    return out


def inner_product(a, b):
    a, b = _f64(a), _f64(b)
    return lib().o3_inner_product(a.ctypes.data, b.ctypes.data, a.size)


def infinite_norm(a):
    a = _f64(a)
    return lib().o3_infinite_norm(a.ctypes.data, a.size)


def multiply_add(a, b, c):
    a, c = _f64(a), _f64(c)
    out = np.empty_like(a)
    lib().o3_multiply_add(out.ctypes.data, a.ctypes.data, float(b), c.ctypes.data, a.size)
    return out
