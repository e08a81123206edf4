"""fs2d_project_noalloc (the projection that allocates nothing) against
fs2d_run_stage(FS2D_STAGE_APPLY_PROJECTION) on a handle with the same state.

Both run the same kernels with the same launch shapes, so the pressures and
iteration counts must be identical bit for bit.  Two calls in a row, the
second on a smaller grid, check that a call starts clean from the static
scratch the previous call left behind.
"""
import ctypes

import numpy as np
import pytest

import fs2d
from oracle import oracle as O

pytestmark = pytest.mark.gpu


class ProjectionArgs(ctypes.Structure):
    _fields_ = [
        ("size_x", ctypes.c_int), ("size_y", ctypes.c_int),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("rho", ctypes.c_double),
        ("tau", ctypes.c_double), ("sigma", ctypes.c_double), ("pcg_tol", ctypes.c_double),
        ("pcg_max_iters", ctypes.c_int), ("flags", ctypes.c_int),
        ("cell", ctypes.c_void_p), ("phi", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
        ("p", ctypes.c_void_p), ("stream", ctypes.c_void_p),
    ]


def lib():
    L = fs2d.load()
    L.fs2d_project_noalloc.restype = ctypes.c_int
    L.fs2d_project_noalloc.argtypes = [ctypes.POINTER(ProjectionArgs), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_double)]
    return L


def hip():
    H = ctypes.CDLL("libamdhip64.so.7")  # the runtime libfs2d.so is linked to
    vp = ctypes.c_void_p
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [vp]
    return H


def device_field(g, f):
    p = ctypes.c_void_p()
    assert g.L.fs2d_device_field(g.h, f, ctypes.byref(p)) == fs2d.OK
    return p.value


SOLVER = dict(tau=0.97, sigma=0.25, tol=1e-6, max_iters=500)


def scene(n, mode, steps):
    """A dam break stepped a few times, then advanced to just before its
    projection (level set, transfer or advection, gravity)."""
    g = fs2d.Sim(O.dam_break(n, n), k=2, dt=0.005, dx=1.0 / n, mode=mode, alpha=0.05, **SOLVER)
    for _ in range(steps):
        g.step()
    g.run_stage(fs2d.STAGE_LEVEL_SET)
    g.run_stage(fs2d.STAGE_P2G if mode == fs2d.PICFLIP else fs2d.STAGE_SL)
    g.run_stage(fs2d.STAGE_GRAVITY)
    return g


def args_for(g, n, out_ptr):
    # the state's device arrays as they are now (stages swap u, v buffers)
    return ProjectionArgs(
        size_x=n, size_y=n, dt=0.005, dx=1.0 / n, rho=997.0, tau=SOLVER["tau"], sigma=SOLVER["sigma"],
        pcg_tol=SOLVER["tol"], pcg_max_iters=SOLVER["max_iters"], flags=0,
        cell=device_field(g, fs2d.F_CELL), phi=device_field(g, fs2d.F_PHI), u=device_field(g, fs2d.F_U),
        v=device_field(g, fs2d.F_V), p=out_ptr, stream=g.L.fs2d_stream(g.h))


def test_noalloc_projection_matches_run_stage():
    L, H = lib(), hip()
    # the entry point runs the PCG as four launches per iteration; the handle
    # runs the first case as two (fused_pcg), the second as four
    for n, mode, steps, fused in ((160, fs2d.PICFLIP, 4, 1), (72, fs2d.SEMILAGRANGIAN, 3, 0)):
        print("case n=%d mode=%d fused_pcg=%d" % (n, mode, fused), flush=True)
        g = scene(n, mode, steps)
        g.set_tuning(pcg_grid=0, fused_pcg=fused)
        out = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(out), n * n * 4) == 0
        try:
            a = args_for(g, n, out.value)
            iters, res = ctypes.c_int(-1), ctypes.c_double(-1.0)
            st = L.fs2d_project_noalloc(ctypes.byref(a), ctypes.byref(iters), ctypes.byref(res))
            assert st == fs2d.OK, L.fs2d_last_error(None).decode()
            p_na = np.empty((n, n), dtype=np.float32)
            assert H.hipMemcpy(p_na.ctypes.data, out.value, n * n * 4, 2) == 0  # device to host
        finally:
            H.hipFree(out)
        g.run_stage(fs2d.STAGE_PROJECT)
        s = g.stats()
        print("n=%d: noalloc %d iterations (res %.2e), handle %d (res %.2e)"
              % (n, iters.value, res.value, s["pcg_iters"], s["pcg_residual"]))
        assert iters.value == s["pcg_iters"] > 2
        assert res.value == s["pcg_residual"] <= 1e-6
        np.testing.assert_array_equal(p_na, g.get(fs2d.F_P))
        g.close()


def test_noalloc_rejects_large_grid():
    L, H = lib(), hip()
    n = 1024
    g = scene(64, fs2d.SEMILAGRANGIAN, 1)
    out = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(out), 64 * 64 * 4) == 0
    try:
        a = args_for(g, 64, out.value)
        a.size_x = a.size_y = n  # rejected on the host before any device work
        st = L.fs2d_project_noalloc(ctypes.byref(a), None, None)
        assert st == fs2d.ERR_ARG
        assert b"static scratch" in L.fs2d_last_error(None)
    finally:
        H.hipFree(out)
    g.close()
