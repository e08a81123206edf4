"""The 3D MAC grid on the GPU (include/fs3d.h) against the oracle
(oracle/fs3d_oracle.cpp), through the C ABI.

Tolerances: the FP64 sampling API evaluates in FP64 on the FP32 grid, so it
agrees with the oracle (fed the same FP32 values) to 1e-12; the FP32 device
API and the FP32 stencils are held to 1e-5 and 1e-6 relative L2.  Samples the
reference would read out of bounds are NaN and counted identically."""
import ctypes

import numpy as np
import pytest

import fs2d
from fs2d import grid3d as G
from oracle import oracle3d as O3

pytestmark = pytest.mark.gpu

HIP = None


def hip():
    global HIP
    if HIP is None:
        HIP = ctypes.CDLL("libamdhip64.so.7")  # the runtime libfs2d.so is linked to
        HIP.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        HIP.hipFree.argtypes = [ctypes.c_void_p]
        HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        HIP.hipDeviceSynchronize.argtypes = []
    return HIP


class Dev:
    """A device buffer holding a copy of a numpy array."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(self.a.nbytes, 4)) == 0
        assert hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        hip().hipFree(self.p)


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def random_grid(nx, ny, nz, seed, lo=-1.0):
    r = np.random.default_rng(seed)
    g = G.MACGrid3D(nx, ny, nz)
    f = [r.uniform(lo, 1.0, g.shape(c)).astype(np.float32) for c in (G.U, G.V, G.W)]
    for c in (G.U, G.V, G.W):
        g.set(c, f[c])
    return g, f


@pytest.mark.parametrize("dims", [(19, 13, 11), (16, 16, 16), (9, 21, 30)])
def test_vel_interp_fp64_matches_oracle(dims):
    nx, ny, nz = dims
    g, (u, v, w) = random_grid(nx, ny, nz, 1)
    r = np.random.default_rng(2)
    p = r.uniform(-1.5, np.array(dims) + 1.5, (4000, 3))
    for mode in (G.SAMPLE_VELOCITY, G.SAMPLE_REFERENCE):
        want, bad = O3.vel_interp(u, v, w, p, mode)
        if bad:
            with pytest.raises(fs2d.FS2DError) as e:
                g.vel_interp(p, mode)
            assert e.value.status == fs2d.ERR_ARG and ("%d of 4000" % bad) in str(e.value)
            # the same call on the valid rows only
            ok = ~np.isnan(want).any(axis=1)
            p2, want = p[ok], want[ok]
        else:
            p2 = p
        got = g.vel_interp(p2, mode)
        err = np.abs(got - want).max()
        print(dims, "mode", mode, "out-of-range samples", bad, "max |gpu - oracle|", err)
        assert err < 1e-12
    g.close()


def test_vel_interp_device_fp32_and_range_count():
    nx, ny, nz = 24, 20, 24
    g, (u, v, w) = random_grid(nx, ny, nz, 3)
    r = np.random.default_rng(4)
    p = r.uniform(-1.0, np.array([nx, ny, nz]) + 1.0, (20000, 3)).astype(np.float32)
    dp, dout = Dev(p), Dev(np.zeros_like(p))
    for mode in (G.SAMPLE_VELOCITY, G.SAMPLE_REFERENCE):
        assert g.L.fs3d_vel_interp_device(g.h, dp.p, len(p), dout.p, mode) == 0
        got = dout.get()
        want, bad = O3.vel_interp(u, v, w, p.astype(np.float64), mode)
        assert g.L.fs3d_range_errors(g.h) == bad
        nan_g, nan_o = np.isnan(got), np.isnan(want)
        np.testing.assert_array_equal(nan_g, nan_o)
        e = rel_l2(got[~nan_o], want[~nan_o])
        print("mode", mode, "FP32 device rel-L2", e, "out-of-range", bad)
        assert e < 1e-5
    assert g.L.fs3d_range_errors(g.h) == 0  # reset by the read
    dp.free(); dout.free(); g.close()


@pytest.mark.parametrize("dims", [(19, 13, 11), (8, 8, 8)])
def test_stencils_match_oracle(dims):
    g, (u, v, w) = random_grid(*dims, seed=5)
    got, want = g.vel_div(np.float64), O3.vel_div(u, v, w)
    scale = np.linalg.norm(np.abs(np.diff(u, axis=2)) + np.abs(np.diff(v, axis=1)) + np.abs(np.diff(w, axis=0)))
    assert np.linalg.norm(got - want) / scale < 1e-6
    assert rel_l2(g.vel_cell(), O3.vel_cell(u, v, w)) < 1e-6
    for which in (G.U, G.V, G.W):
        assert rel_l2(g.vel_face(which), O3.vel_face(u, v, w, which)) < 1e-6
    g.close()


def test_array_arithmetic_matches_oracle():
    g, (u, v, w) = random_grid(30, 20, 10, 6)
    g2, (u2, _, _) = random_grid(30, 20, 10, 7)
    du, du2 = g.device_field(G.U), g2.device_field(G.U)
    n = u.size
    ip = g.inner_product(du, du2, n)
    assert abs(ip - O3.inner_product(u, u2)) < 1e-12 * n
    assert g.inner_product(du, du2, n) == ip  # deterministic
    assert g.infinite_norm(g.device_field(G.W), w.size) == O3.infinite_norm(w)
    g.multiply_add(du, du, -0.5, du2, n)  # u = u - 0.5 u2, in place
    want = O3.multiply_add(u, -0.5, u2).astype(np.float32)
    np.testing.assert_array_equal(g.get(G.U), want)
    g.close(); g2.close()


def test_divergence_free_field_at_256_cubed():
    # u depends on (y, z) only, v on (x, z), w on (x, y): every face difference is exactly zero
    n = 256
    g = G.MACGrid3D(n, n, n)
    r = np.random.default_rng(8)
    fu = r.normal(size=(n, n, 1)).astype(np.float32)
    fv = r.normal(size=(n, 1, n)).astype(np.float32)
    fw = r.normal(size=(1, n, n)).astype(np.float32)
    g.set(G.U, np.broadcast_to(fu, g.shape(G.U)))
    g.set(G.V, np.broadcast_to(fv, g.shape(G.V)))
    g.set(G.W, np.broadcast_to(fw, g.shape(G.W)))
    d = g.vel_div()
    assert np.count_nonzero(d) == 0
    # cell averages of a field constant along its own axis are the face values themselves
    c = g.vel_cell()
    np.testing.assert_array_equal(c[..., 0], np.broadcast_to(fu, (n, n, n)))
    g.close()


def test_errors():
    L = fs2d.load()
    h = ctypes.c_void_p()
    assert L.fs3d_create(0, 4, 4, ctypes.byref(h)) == fs2d.ERR_CONFIG
    assert b"each side" in L.fs3d_last_error(None)
    g = G.MACGrid3D(4, 5, 6)
    bad = np.zeros(7, np.float32)
    assert L.fs3d_set_field(g.h, G.U, fs2d.F32, bad.ctypes.data, bad.size) == fs2d.ERR_ARG
    assert L.fs3d_vel_face(g.h, 3, fs2d.F32, bad.ctypes.data, bad.size) == fs2d.ERR_ARG
    assert L.fs3d_vel_interp(g.h, bad.ctypes.data, 1, bad.ctypes.data, 7) == fs2d.ERR_ARG
    assert L.fs3d_field_count(g.h, G.W) == 4 * 5 * 7
    g.close()
