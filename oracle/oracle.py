"""ctypes wrapper over oracle/liboracle.so, the CPU restatement of fluid-sim's
2D step (src/FluidSim2D.cpp, include/{FluidSim2D,MACGrid2D,Array2D}.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, always as the checker.  Nothing in the product
package imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EMPTY, FLUID, SOLID = 0, 1, 2
MODE_SL, MODE_PIC = 0, 1
(ST_INIT, ST_LEVELSET, ST_P2G, ST_SL, ST_GRAVITY, ST_SOLID, ST_PROJECT,
 ST_UPDATEVEL, ST_G2P, ST_ADVECT) = range(10)
F_U, F_V, F_NEWU, F_NEWV, F_P, F_PHI, F_DENS, F_CELL, F_POS, F_VEL = range(10)
F_ADIAG, F_AX, F_AY, F_RHS, F_PRECON = range(10, 15)


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("water_volume", ctypes.c_double), ("total_energy", ctypes.c_double),
        ("particle_total_energy", ctypes.c_double), ("current_time", ctypes.c_double),
        ("last_residual", ctypes.c_double), ("cfl_max", ctypes.c_double),
        ("orig_water_volume", ctypes.c_double),
        ("last_iters", ctypes.c_int), ("hit_max", ctypes.c_int),
        ("nan_flag", ctypes.c_int), ("n_particles", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, d, d, d, d, d, i, d, vp, ctypes.c_uint, i, i]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_options.argtypes = [vp, d, d, d, i, i, i]
        L.orc_set_params.argtypes = [vp, d, d, d, d]
        L.orc_step.argtypes = [vp]
        L.orc_run_stage.argtypes = [vp, i]
        L.orc_field_size.restype = ctypes.c_long
        L.orc_field_size.argtypes = [vp, i]
        L.orc_get_field.argtypes = [vp, i, vp, ctypes.c_long]
        L.orc_set_field.argtypes = [vp, i, vp, ctypes.c_long]
        L.orc_get_stats.argtypes = [vp, ctypes.POINTER(OracleStats)]
        L.orc_apply_precon.argtypes = [vp, vp, vp]
        L.orc_assemble_factor.argtypes = [vp]
        L.orc_level_set_parts.argtypes = [vp, i]
        L.orc_tricubic.restype = d
        L.orc_tricubic.argtypes = [vp, i, i, d, d]
        L.orc_extrapolate.argtypes = [vp, vp, i, i]
        L.orc_rand_stream.argtypes = [ctypes.c_uint, vp, i]
        L.orc_set_variants.argtypes = [vp, i, i, i]
        L.orc_adjoint_step.restype = i
        L.orc_adjoint_step.argtypes = [vp] * 7
        L.orc_set_threads.argtypes = [i]
        L.orc_max_threads.restype = i
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """FluidSim2D restated on the CPU in FP64 (one instance = one simulation)."""

    def __init__(self, cells, k=2, dt=0.005, dx=0.01, rho=997.0, gravity=(0.0, -9.81),
                 mode=MODE_PIC, alpha=1.0, seed=1, level_set=True, tau=0.999,
                 sigma=0.25, tol=1e-12, max_iters=200, sl_inplace=True, density=False,
                 uniform_seeding=False, rk2=False):
        cells = np.ascontiguousarray(cells, dtype=np.uint8)
        self.ny, self.nx = cells.shape
        self.L = lib()
        self.h = self.L.orc_create(self.nx, self.ny, k, dt, dx, rho, gravity[0], gravity[1],
                                   mode, alpha, _ptr(cells), seed, int(level_set), int(uniform_seeding))
        self._variants = dict(particle_cells=False, ghost=True, rk2=rk2)
        if rk2:
            self.set_variants()
        self.set_options(tau=tau, sigma=sigma, tol=tol, max_iters=max_iters,
                         sl_inplace=sl_inplace, density=density)

    def set_options(self, tau=0.999, sigma=0.25, tol=1e-12, max_iters=200, sl_inplace=True,
                    density=False):
        self.L.orc_set_options(self.h, tau, sigma, tol, max_iters, int(sl_inplace), int(density))

    def set_variants(self, particle_cells=None, ghost=None, rk2=None):
        """The reference's compile-time variants: cells from particles instead
        of the level set (USE_LEVEL_SET off), no ghost-fluid pressure
        (USE_GHOST_PRESSURE off); and RK2 particle advection (no reference
        counterpart).  Arguments left None keep their current setting."""
        for k, v in (("particle_cells", particle_cells), ("ghost", ghost), ("rk2", rk2)):
            if v is not None:
                self._variants[k] = v
        vv = self._variants
        self.L.orc_set_variants(self.h, int(vv["particle_cells"]), int(vv["ghost"]), int(vv["rk2"]))

    def set_params(self, dt, gravity, alpha):
        self.L.orc_set_params(self.h, dt, gravity[0], gravity[1], alpha)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self):
        self.L.orc_step(self.h)

    def run_stage(self, st):
        self.L.orc_run_stage(self.h, st)

    def _shape(self, f):
        n = self.L.orc_field_size(self.h, f)
        if f in (F_U, F_NEWU):
            return (self.ny, self.nx + 1)
        if f in (F_V, F_NEWV):
            return (self.ny + 1, self.nx)
        if f in (F_POS, F_VEL):
            return (n // 2, 2)
        return (self.ny, self.nx)

    def get(self, f):
        n = self.L.orc_field_size(self.h, f)
        dt = np.uint8 if f == F_CELL else np.float64
        out = np.empty(n, dtype=dt)
        if n and self.L.orc_get_field(self.h, f, _ptr(out), n) != 0:
            raise RuntimeError("orc_get_field failed")
        return out.reshape(self._shape(f))

    def set(self, f, arr):
        dt = np.uint8 if f == F_CELL else np.float64
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        if self.L.orc_set_field(self.h, f, _ptr(a), a.size) != 0:
            raise RuntimeError("orc_set_field: wrong size for field %d" % f)

    def stats(self):
        st = OracleStats()
        self.L.orc_get_stats(self.h, ctypes.byref(st))
        return {n: getattr(st, n) for n, _ in OracleStats._fields_}

    def assemble_factor(self):
        self.L.orc_assemble_factor(self.h)

    def level_set_parts(self, which):
        self.L.orc_level_set_parts(self.h, which)

    def adjoint_step(self, ub, vb, db):
        """Adjoint of one step at the current state: the adjoints of u, v and
        density after the step in, those before it out (FP64)."""
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        vb = np.ascontiguousarray(vb, dtype=np.float64)
        db = np.ascontiguousarray(db, dtype=np.float64)
        outs = [np.zeros_like(ub), np.zeros_like(vb), np.zeros_like(db)]
        if self.L.orc_adjoint_step(self.h, _ptr(ub), _ptr(vb), _ptr(db), *[_ptr(o) for o in outs]) != 0:
            raise ValueError("adjoint needs SL mode, static flags without EMPTY cells, no particles, "
                             "out-of-place advection")
        return outs

    def apply_precon(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        self.L.orc_apply_precon(self.h, _ptr(r), _ptr(z))
        return z


def tricubic(data, x, y):
    data = np.ascontiguousarray(data, dtype=np.float64)
    ny, nx = data.shape
    return lib().orc_tricubic(_ptr(data), nx, ny, x, y)


def extrapolate(data, mask):
    data = np.array(data, dtype=np.float64, order="C")
    mask = np.array(mask, dtype=np.uint32, order="C")
    ny, nx = data.shape
    lib().orc_extrapolate(_ptr(data), _ptr(mask), nx, ny)
    return data, mask


def rand_stream(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_rand_stream(seed, _ptr(out), n)
    return out


def set_threads(n):
    lib().orc_set_threads(n)


# ---------------------------------------------------------------------------
# Scene generators (row-major [j, i] flags).
# ---------------------------------------------------------------------------

def walls(nx, ny):
    c = np.full((ny, nx), EMPTY, dtype=np.uint8)
    c[0, :] = c[-1, :] = SOLID
    c[:, 0] = c[:, -1] = SOLID
    return c


def demo_triangle(n):
    """demo/App.cpp:145-158: SOLID border, FLUID where i + j < 3n/4."""
    c = walls(n, n)
    j, i = np.mgrid[0:n, 0:n]
    c[(c != SOLID) & (i + j < n * 3 // 4)] = FLUID
    return c


def dam_break(nx, ny, wfrac=0.4, hfrac=0.6):
    """Fluid block over the left wfrac width and bottom hfrac height, inside walls."""
    c = walls(nx, ny)
    c[1:int(hfrac * ny), 1:int(wfrac * nx)] = FLUID
    return c


def all_fluid(nx, ny):
    c = walls(nx, ny)
    c[1:-1, 1:-1] = FLUID
    return c
