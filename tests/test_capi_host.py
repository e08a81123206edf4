"""C ABI checks that need no GPU: the library loads, exports every function
include/fs2d.h declares, fills the reference's defaults, and rejects invalid
configurations before touching the device."""
import ctypes
import subprocess

import numpy as np
import pytest

import fs2d
from oracle import oracle as O


def test_library_exports_every_header_function():
    L = fs2d.load()
    names = fs2d.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", fs2d.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported
    # nothing else leaks out of the C ABI
    assert {n for n in exported if n.startswith(("fs2d_", "fs3d_"))} == set(names)
    assert any(n.startswith("fs3d_") for n in names)


def test_config_defaults_are_the_reference_constants():
    c = fs2d.default_config()
    assert (c.tau, c.sigma, c.pcg_tol, c.pcg_max_iters) == (0.999, 0.25, 1e-12, 200)  # FluidSim2D.cpp:369-370,429,453
    assert (c.rho, c.gravity_x, c.gravity_y) == (997.0, 0.0, -9.81)                   # FluidSim2D.h:80-82
    assert (c.dt, c.dx, c.seed) == (0.0001, 0.001, 1)


def _create(cells, **kw):
    c = fs2d.default_config()
    cells = np.ascontiguousarray(cells, dtype=np.uint8)
    c.size_y, c.size_x = cells.shape
    c.dt, c.dx = 0.01, 0.1
    c.initial_values = cells.ctypes.data
    for k, v in kw.items():
        setattr(c, k, v)
    h = ctypes.c_void_p()
    st = fs2d.load().fs2d_create(ctypes.byref(c), ctypes.byref(h))
    return st, h


@pytest.mark.parametrize("bad", [
    "open_border", "bad_flag", "tiny", "neg_dx", "bad_mode", "neg_iters", "null_cells"])
def test_invalid_configs_rejected_without_device(bad):
    cells = O.dam_break(16, 16)
    kw = {}
    if bad == "open_border":
        cells[0, 5] = O.EMPTY   # the reference would read out of bounds here
    elif bad == "bad_flag":
        cells[5, 5] = 7
    elif bad == "tiny":
        cells = cells[:2, :2]
    elif bad == "neg_dx":
        kw["dx"] = -1.0
    elif bad == "bad_mode":
        kw["mode"] = 5
    elif bad == "neg_iters":
        kw["pcg_max_iters"] = -1
    st, h = _create(cells, **kw) if bad != "null_cells" else (None, None)
    if bad == "null_cells":
        c = fs2d.default_config()
        c.size_x = c.size_y = 16
        h = ctypes.c_void_p()
        st = fs2d.load().fs2d_create(ctypes.byref(c), ctypes.byref(h))
    assert st == fs2d.ERR_CONFIG
    assert not h.value
    assert fs2d.load().fs2d_last_error(None)


def test_status_strings_and_kernel_names():
    assert fs2d.status_string(fs2d.ERR_NAN) == "FS2D_ERR_NAN"
    L = fs2d.load()
    names = [L.fs2d_kernel_name(k).decode() for k in range(L.fs2d_kernel_count())]
    assert "pcg_forward_solve" in names and "sl_advect" in names
    assert L.fs2d_kernel_name(-1) is None


def test_variant_flag_combinations_rejected_without_device():
    cells = O.dam_break(16, 16)
    # the particle-cell variant without the no-ghost variant would divide phi = 0 by phi = 0
    st, _ = _create(cells, flags=fs2d.FLAG_PARTICLE_CELLS)
    assert st == fs2d.ERR_CONFIG
    st, _ = _create(cells, flags=fs2d.FLAG_PARTICLE_CELLS | fs2d.FLAG_NO_GHOST_PRESSURE | fs2d.FLAG_STATIC_CELLS)
    assert st == fs2d.ERR_CONFIG
    st, _ = _create(cells, flags=64)
    assert st == fs2d.ERR_CONFIG
