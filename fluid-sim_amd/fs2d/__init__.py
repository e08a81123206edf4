"""ctypes binding of the C ABI in include/fs2d.h (lib/libfs2d.so).

The Python side is plumbing for tests and benchmarks; the product is the C ABI
and the C++ mirror in fluid-sim_amd/cpp/FluidSim2D.hpp.  Importing this module
never falls back to a CPU path: if the HIP library is missing, load() raises.
"""
import ctypes
import os
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # fluid-sim_amd/
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libfs2d.so")
HEADER = os.path.join(ROOT, "include", "fs2d.h")

OK, WARN_MAXITER, WARN_CFL = 0, 1, 2
ERR_NAN, ERR_OOM, ERR_CONFIG, ERR_ARG, ERR_DEVICE, ERR_HANG = -1, -2, -3, -4, -5, -6
SEMILAGRANGIAN, PICFLIP = 0, 1
EMPTY, FLUID, SOLID = 0, 1, 2
(STAGE_INIT, STAGE_LEVEL_SET, STAGE_P2G, STAGE_SL, STAGE_GRAVITY, STAGE_SOLID,
 STAGE_PROJECT, STAGE_UPDATE_VELOCITY, STAGE_G2P, STAGE_ADVECT) = range(10)
(F_U, F_V, F_NEWU, F_NEWV, F_P, F_PHI, F_DENSITY, F_CELL,
 F_PARTICLE_POS, F_PARTICLE_VEL, F_PARTICLE_CELL, F_PARTICLE_ORDER) = range(12)
F32, F64, U8, I32 = 0, 1, 2, 3
FLAG_STATIC_CELLS, FLAG_DENSITY, FLAG_PARTICLE_CELLS, FLAG_NO_GHOST_PRESSURE = 1, 2, 4, 8
FLAG_UNIFORM_SEEDING, FLAG_RK2 = 16, 32


class Config(ctypes.Structure):
    _fields_ = [
        ("size_x", ctypes.c_int), ("size_y", ctypes.c_int),
        ("particles_per_cell_sqrt", ctypes.c_int),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("rho", ctypes.c_double),
        ("gravity_x", ctypes.c_double), ("gravity_y", ctypes.c_double),
        ("mode", ctypes.c_int), ("pic_flip_alpha", ctypes.c_double),
        ("initial_values", ctypes.c_void_p),
        ("tau", ctypes.c_double), ("sigma", ctypes.c_double), ("pcg_tol", ctypes.c_double),
        ("pcg_max_iters", ctypes.c_int), ("seed", ctypes.c_uint), ("flags", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("water_volume", ctypes.c_double), ("total_energy", ctypes.c_double),
        ("particle_total_energy", ctypes.c_double), ("orig_water_volume", ctypes.c_double),
        ("current_time", ctypes.c_double), ("pcg_residual", ctypes.c_double),
        ("cfl_max", ctypes.c_double), ("pcg_iters", ctypes.c_int), ("pcg_hit_max", ctypes.c_int),
        ("n_particles", ctypes.c_longlong), ("steps", ctypes.c_longlong),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("pcg_grid", ctypes.c_int), ("sweep_prefetch", ctypes.c_int), ("fused_pcg", ctypes.c_int)]


RENDER_BUFFERS = ("water_cells", "solid_cells", "cell_vel_lines", "pressure_cells", "pressure_values",
                  "particle_lines", "phi_values")
RENDER_COUNTS = ("water_cells", "solid_cells", "pressure_cells", "particles", "cell_vel_lines", "phi_values")


class RenderTargets(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in RENDER_BUFFERS]


class RenderCounts(ctypes.Structure):
    _fields_ = [(k, ctypes.c_long) for k in RENDER_COUNTS]


class FS2DError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (status_string(status), msg))
        self.status = status


def build():
    """Compile lib/libfs2d.so for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-B", "-C", PKG], check=True)


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError("%s is not built: run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    sigs = {
        "fs2d_config_init": (i, [ctypes.POINTER(Config)]),
        "fs2d_create": (i, [ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "fs2d_destroy": (None, [vp]),
        "fs2d_step": (i, [vp]),
        "fs2d_run_stage": (i, [vp, i]),
        "fs2d_synchronize": (i, [vp]),
        "fs2d_field_count": (sz, [vp, i]),
        "fs2d_get_field": (i, [vp, i, i, vp, sz]),
        "fs2d_set_field": (i, [vp, i, i, vp, sz]),
        "fs2d_set_particles": (i, [vp, i, vp, sz]),
        "fs2d_device_field": (i, [vp, i, ctypes.POINTER(vp)]),
        "fs2d_stream": (vp, [vp]),
        "fs2d_set_gravity": (i, [vp, d, d]),
        "fs2d_set_pic_flip_alpha": (i, [vp, d]),
        "fs2d_set_dt": (i, [vp, d]),
        "fs2d_set_solver": (i, [vp, d, d, d, i]),
        "fs2d_get_stats": (i, [vp, ctypes.POINTER(Stats)]),
        "fs2d_avg_pressure": (i, [vp, i, ctypes.POINTER(d)]),
        "fs2d_max_velocity": (i, [vp, ctypes.POINTER(d)]),
        "fs2d_set_graph": (i, [vp, i]),
        "fs2d_get_tuning": (i, [vp, ctypes.POINTER(Tuning)]),
        "fs2d_set_tuning": (i, [vp, ctypes.POINTER(Tuning)]),
        "fs2d_autotune": (i, [vp, i, ctypes.c_char_p, ctypes.POINTER(Tuning), ctypes.POINTER(i)]),
        "fs2d_batch_get_tuning": (i, [vp, ctypes.POINTER(Tuning)]),
        "fs2d_batch_set_tuning": (i, [vp, ctypes.POINTER(Tuning)]),
        "fs2d_batch_autotune": (i, [vp, i, ctypes.c_char_p, ctypes.POINTER(Tuning), ctypes.POINTER(i)]),
        "fs2d_project_noalloc": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(d)]),
        "fs2d_render_buffers": (i, [vp, ctypes.POINTER(RenderTargets), ctypes.POINTER(RenderCounts)]),
        "fs2d_set_timing": (i, [vp, i]),
        "fs2d_set_timing_classes": (i, [vp, ctypes.c_ulonglong]),
        "fs2d_kernel_count": (i, []),
        "fs2d_kernel_name": (ctypes.c_char_p, [i]),
        "fs2d_get_kernel_time": (i, [vp, i, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_longlong)]),
        "fs2d_reset_kernel_times": (i, [vp]),
        "fs2d_last_error": (ctypes.c_char_p, [vp]),
        "fs2d_batch_create": (i, [vp, i, ctypes.POINTER(vp)]),
        "fs2d_batch_destroy": (None, [vp]),
        "fs2d_batch_size": (i, [vp]),
        "fs2d_batch_sim": (vp, [vp, i]),
        "fs2d_batch_step": (i, [vp]),
        "fs2d_batch_run_stage": (i, [vp, i]),
        "fs2d_batch_set_graph": (i, [vp, i]),
        "fs2d_batch_set_timing": (i, [vp, i]),
        "fs2d_batch_get_kernel_time": (i, [vp, i, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_longlong)]),
        "fs2d_batch_reset_kernel_times": (i, [vp]),
        "fs2d_batch_last_error": (ctypes.c_char_p, [vp]),
        "fs2d_adjoint_create": (i, [vp, i, i, ctypes.POINTER(vp)]),
        "fs2d_adjoint_destroy": (None, [vp]),
        "fs2d_adjoint_forward": (i, [vp, i]),
        "fs2d_adjoint_backward": (i, [vp]),
        "fs2d_adjoint_step": (i, [vp]),
        "fs2d_adjoint_get_field": (i, [vp, i, i, vp, sz]),
        "fs2d_adjoint_set_field": (i, [vp, i, i, vp, sz]),
        "fs2d_adjoint_device_field": (i, [vp, i, ctypes.POINTER(vp)]),
        "fs2d_status_string": (ctypes.c_char_p, [i]),
        # include/fs3d.h
        "fs3d_create": (i, [i, i, i, ctypes.POINTER(vp)]),
        "fs3d_destroy": (None, [vp]),
        "fs3d_last_error": (ctypes.c_char_p, [vp]),
        "fs3d_stream": (vp, [vp]),
        "fs3d_field_count": (sz, [vp, i]),
        "fs3d_set_field": (i, [vp, i, i, vp, sz]),
        "fs3d_get_field": (i, [vp, i, i, vp, sz]),
        "fs3d_device_field": (vp, [vp, i]),
        "fs3d_vel_interp": (i, [vp, vp, sz, vp, i]),
        "fs3d_vel_interp_device": (i, [vp, vp, sz, vp, i]),
        "fs3d_range_errors": (ctypes.c_long, [vp]),
        "fs3d_vel_div": (i, [vp, i, vp, sz]),
        "fs3d_vel_cell": (i, [vp, i, vp, sz]),
        "fs3d_vel_face": (i, [vp, i, i, vp, sz]),
        "fs3d_vel_div_device": (i, [vp, vp]),
        "fs3d_vel_cell_device": (i, [vp, vp]),
        "fs3d_vel_face_device": (i, [vp, i, vp]),
        "fs3d_array_inner_product": (i, [vp, vp, vp, sz, ctypes.POINTER(d)]),
        "fs3d_array_infinite_norm": (i, [vp, vp, sz, ctypes.POINTER(d)]),
        "fs3d_array_multiply_add": (i, [vp, vp, vp, d, vp, sz]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def status_string(st):
    return load().fs2d_status_string(st).decode()


def default_config():
    c = Config()
    load().fs2d_config_init(ctypes.byref(c))
    return c


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_SHAPES = {
    F_U: lambda s: (s.ny, s.nx + 1), F_NEWU: lambda s: (s.ny, s.nx + 1),
    F_V: lambda s: (s.ny + 1, s.nx), F_NEWV: lambda s: (s.ny + 1, s.nx),
}


def make_config(cells, k=2, dt=0.005, dx=0.01, rho=997.0, gravity=(0.0, -9.81), mode=PICFLIP,
                alpha=1.0, tau=0.999, sigma=0.25, tol=1e-12, max_iters=200, seed=1,
                static_cells=False, density=False, particle_cells=False, ghost_pressure=True,
                uniform_seeding=False, rk2=False):
    """A Config for `cells` (uint8, row-major [j, i]); the caller keeps `cells`
    alive until the handle is created."""
    c = default_config()
    c.size_y, c.size_x = cells.shape
    c.particles_per_cell_sqrt = k
    c.dt, c.dx, c.rho = dt, dx, rho
    c.gravity_x, c.gravity_y = gravity
    c.mode, c.pic_flip_alpha = mode, alpha
    c.initial_values = cells.ctypes.data
    c.tau, c.sigma, c.pcg_tol, c.pcg_max_iters = tau, sigma, tol, max_iters
    c.seed = seed
    c.flags = ((FLAG_STATIC_CELLS if static_cells else 0) | (FLAG_DENSITY if density else 0) |
               (FLAG_PARTICLE_CELLS if particle_cells else 0) | (0 if ghost_pressure else FLAG_NO_GHOST_PRESSURE) |
               (FLAG_UNIFORM_SEEDING if uniform_seeding else 0) | (FLAG_RK2 if rk2 else 0))
    return c


class Sim:
    """One FluidSim2D on the GPU (a handle of the C ABI)."""

    def __init__(self, cells, **kw):
        L = load()
        cells = np.ascontiguousarray(cells, dtype=np.uint8)
        self.ny, self.nx = cells.shape
        c = make_config(cells, **kw)
        h = ctypes.c_void_p()
        st = L.fs2d_create(ctypes.byref(c), ctypes.byref(h))
        if st != OK:
            raise FS2DError(st, L.fs2d_last_error(None).decode())
        self.h = h
        self.L = L
        self._owned = True

    @classmethod
    def _member(cls, L, h, nx, ny):
        """A non-owning wrapper of a batch member's handle."""
        self = cls.__new__(cls)
        self.L, self.h, self.nx, self.ny, self._owned = L, ctypes.c_void_p(h), nx, ny, False
        return self

    def _check(self, st, allow=(OK,)):
        if st not in allow:
            raise FS2DError(st, self.L.fs2d_last_error(self.h).decode())
        return st

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.L.fs2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, allow=(OK, WARN_MAXITER, WARN_CFL)):
        return self._check(self.L.fs2d_step(self.h), allow)

    def run_stage(self, stage, allow=(OK, WARN_MAXITER, WARN_CFL)):
        return self._check(self.L.fs2d_run_stage(self.h, stage), allow)

    def synchronize(self):
        self._check(self.L.fs2d_synchronize(self.h))

    def shape(self, f):
        if f in _SHAPES:
            return _SHAPES[f](self)
        n = self.L.fs2d_field_count(self.h, f)
        if f in (F_PARTICLE_POS, F_PARTICLE_VEL):
            return (n // 2, 2)
        if f in (F_PARTICLE_CELL, F_PARTICLE_ORDER):
            return (n,)
        return (self.ny, self.nx)

    def get(self, f, dtype=np.float32):
        n = self.L.fs2d_field_count(self.h, f)
        if f == F_CELL:
            dtype, code = np.uint8, U8
        elif f in (F_PARTICLE_CELL, F_PARTICLE_ORDER):
            dtype, code = np.int32, I32
        else:
            code = F64 if np.dtype(dtype) == np.float64 else F32
        out = np.empty(n, dtype=dtype)
        self._check(self.L.fs2d_get_field(self.h, f, code, _ptr(out), n))
        return out.reshape(self.shape(f))

    def set(self, f, arr):
        if f == F_CELL:
            a, code = np.ascontiguousarray(arr, dtype=np.uint8).ravel(), U8
        else:
            a = np.ascontiguousarray(arr).ravel()
            if a.dtype == np.float64:
                code = F64
            else:
                a, code = a.astype(np.float32), F32
        self._check(self.L.fs2d_set_field(self.h, f, code, _ptr(a), a.size))

    def set_particles(self, pos):
        a = np.ascontiguousarray(pos, dtype=np.float64).ravel()
        self._check(self.L.fs2d_set_particles(self.h, F64, _ptr(a), a.size // 2))

    def set_gravity(self, gx, gy):
        self._check(self.L.fs2d_set_gravity(self.h, gx, gy))

    def set_alpha(self, a):
        self._check(self.L.fs2d_set_pic_flip_alpha(self.h, a))

    def set_dt(self, dt):
        self._check(self.L.fs2d_set_dt(self.h, dt))

    def set_solver(self, tau=0.999, sigma=0.25, tol=1e-12, max_iters=200):
        self._check(self.L.fs2d_set_solver(self.h, tau, sigma, tol, max_iters))

    def stats(self):
        s = Stats()
        self._check(self.L.fs2d_get_stats(self.h, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in Stats._fields_}

    def avg_pressure(self, fluid_only=False):
        d = ctypes.c_double()
        self._check(self.L.fs2d_avg_pressure(self.h, int(fluid_only), ctypes.byref(d)))
        return d.value

    def max_velocity(self):
        d = ctypes.c_double()
        self._check(self.L.fs2d_max_velocity(self.h, ctypes.byref(d)))
        return d.value

    def set_graph(self, on):
        """Step by replaying captured HIP graphs (fs2d_set_graph)."""
        self._check(self.L.fs2d_set_graph(self.h, int(on)))

    def render_buffers(self, **targets):
        """The demo renderer's per-frame buffers written to device addresses
        (fs2d_render_buffers); keyword arguments of RENDER_BUFFERS name the
        buffers to fill (device pointers as ints).  Returns the counts."""
        t = RenderTargets(*[targets.get(k) for k in RENDER_BUFFERS])
        c = RenderCounts()
        self._check(self.L.fs2d_render_buffers(self.h, ctypes.byref(t), ctypes.byref(c)))
        return {k: getattr(c, k) for k in RENDER_COUNTS}

    def tuning(self):
        t = Tuning()
        self._check(self.L.fs2d_get_tuning(self.h, ctypes.byref(t)))
        return {"pcg_grid": t.pcg_grid, "sweep_prefetch": t.sweep_prefetch, "fused_pcg": t.fused_pcg}

    def set_tuning(self, pcg_grid=0, sweep_prefetch=0, fused_pcg=None):
        """fused_pcg None keeps the handle's current choice."""
        if fused_pcg is None:
            fused_pcg = self.tuning()["fused_pcg"]
        t = Tuning(pcg_grid, sweep_prefetch, fused_pcg)
        self._check(self.L.fs2d_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, iters=24, cache=None):
        """Pick the fastest launch shapes (fs2d_autotune); returns the choice
        and whether it came from the cache file."""
        t, fc = Tuning(), ctypes.c_int()
        self._check(self.L.fs2d_autotune(self.h, iters, cache.encode() if cache else None, ctypes.byref(t),
                                         ctypes.byref(fc)))
        return {"pcg_grid": t.pcg_grid, "sweep_prefetch": t.sweep_prefetch, "fused_pcg": t.fused_pcg,
                "from_cache": bool(fc.value)}

    def set_timing(self, on, only=None):
        """Per-kernel HIP-event timing; `only` names the kernel classes to time
        (each timed launch adds two events to the stream)."""
        if only is None or not on:
            self._check(self.L.fs2d_set_timing(self.h, int(on)))
            return
        names = [self.L.fs2d_kernel_name(k).decode() for k in range(self.L.fs2d_kernel_count())]
        mask = 0
        for n in only:
            mask |= 1 << names.index(n)
        self._check(self.L.fs2d_set_timing_classes(self.h, mask))

    def kernel_times(self):
        out = {}
        for k in range(self.L.fs2d_kernel_count()):
            t, n = ctypes.c_double(), ctypes.c_longlong()
            self._check(self.L.fs2d_get_kernel_time(self.h, k, ctypes.byref(t), ctypes.byref(n)))
            out[self.L.fs2d_kernel_name(k).decode()] = (t.value, n.value)
        return out

    def reset_kernel_times(self):
        self._check(self.L.fs2d_reset_kernel_times(self.h))


class Batch:
    """An ensemble stepped together (fs2d_batch_* of the C ABI).  `members` is
    a list of (cells, kwargs) with the keyword arguments of Sim."""

    def __init__(self, members):
        L = load()
        self.L = L
        keep = []
        cfgs = (Config * len(members))()
        for k, (cells, kw) in enumerate(members):
            cells = np.ascontiguousarray(cells, dtype=np.uint8)
            keep.append(cells)
            cfgs[k] = make_config(cells, **kw)
        h = ctypes.c_void_p()
        st = L.fs2d_batch_create(cfgs, len(members), ctypes.byref(h))
        if st != OK:
            raise FS2DError(st, L.fs2d_batch_last_error(None).decode())
        self.h = h
        self.sims = [Sim._member(L, L.fs2d_batch_sim(h, k), c.shape[1], c.shape[0]) for k, c in enumerate(keep)]

    def _check(self, st, allow=(OK,)):
        if st not in allow:
            raise FS2DError(st, self.L.fs2d_batch_last_error(self.h).decode())
        return st

    def __len__(self):
        return len(self.sims)

    def close(self):
        if getattr(self, "h", None):
            for s in self.sims:
                s.h = None
            self.L.fs2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, allow=(OK, WARN_MAXITER, WARN_CFL)):
        return self._check(self.L.fs2d_batch_step(self.h), allow)

    def run_stage(self, stage, allow=(OK, WARN_MAXITER, WARN_CFL)):
        return self._check(self.L.fs2d_batch_run_stage(self.h, stage), allow)

    def set_graph(self, on):
        """Step by replaying captured HIP graphs (fs2d_batch_set_graph)."""
        self._check(self.L.fs2d_batch_set_graph(self.h, int(on)))

    def tuning(self):
        t = Tuning()
        self._check(self.L.fs2d_batch_get_tuning(self.h, ctypes.byref(t)))
        return {"pcg_grid": t.pcg_grid, "sweep_prefetch": t.sweep_prefetch, "fused_pcg": t.fused_pcg}

    def set_tuning(self, pcg_grid, sweep_prefetch, fused_pcg=None):
        """fused_pcg None keeps the batch's current choice."""
        if fused_pcg is None:
            fused_pcg = self.tuning()["fused_pcg"]
        t = Tuning(pcg_grid, sweep_prefetch, fused_pcg)
        self._check(self.L.fs2d_batch_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, iters=24, cache=None):
        t, fc = Tuning(), ctypes.c_int()
        self._check(self.L.fs2d_batch_autotune(self.h, iters, cache.encode() if cache else None, ctypes.byref(t),
                                               ctypes.byref(fc)))
        return {"pcg_grid": t.pcg_grid, "sweep_prefetch": t.sweep_prefetch, "fused_pcg": t.fused_pcg,
                "from_cache": bool(fc.value)}

    def set_timing(self, on):
        self._check(self.L.fs2d_batch_set_timing(self.h, int(on)))

    def kernel_times(self):
        out = {}
        for k in range(self.L.fs2d_kernel_count()):
            t, n = ctypes.c_double(), ctypes.c_longlong()
            self._check(self.L.fs2d_batch_get_kernel_time(self.h, k, ctypes.byref(t), ctypes.byref(n)))
            out[self.L.fs2d_kernel_name(k).decode()] = (t.value, n.value)
        return out

    def reset_kernel_times(self):
        self._check(self.L.fs2d_batch_reset_kernel_times(self.h))


class Adjoint:
    """Checkpointed reverse pass over rollouts of a Sim (fs2d_adjoint_* of the
    C ABI).  The Sim must be semi-Lagrangian with static cells, no EMPTY cell
    and no particles."""

    def __init__(self, sim, checkpoint_every=8, max_steps=64):
        self.sim = sim
        self.L = sim.L
        h = ctypes.c_void_p()
        sim._check(self.L.fs2d_adjoint_create(sim.h, checkpoint_every, max_steps, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.fs2d_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, nsteps, allow=(OK, WARN_MAXITER)):
        return self.sim._check(self.L.fs2d_adjoint_forward(self.h, nsteps), allow)

    def backward(self, allow=(OK, WARN_MAXITER)):
        return self.sim._check(self.L.fs2d_adjoint_backward(self.h), allow)

    def step(self, allow=(OK, WARN_MAXITER)):
        return self.sim._check(self.L.fs2d_adjoint_step(self.h), allow)

    def get(self, f, dtype=np.float32):
        n = self.L.fs2d_field_count(self.sim.h, f)
        code = F64 if np.dtype(dtype) == np.float64 else F32
        out = np.empty(n, dtype=dtype)
        self.sim._check(self.L.fs2d_adjoint_get_field(self.h, f, code, _ptr(out), n))
        return out.reshape(self.sim.shape(f))

    def set(self, f, arr):
        a = np.ascontiguousarray(arr).ravel()
        if a.dtype == np.float64:
            code = F64
        else:
            a, code = a.astype(np.float32), F32
        self.sim._check(self.L.fs2d_adjoint_set_field(self.h, f, code, _ptr(a), a.size))


def header_functions():
    """Names of the functions include/fs2d.h and include/fs3d.h declare."""
    import re
    names = set()
    for h in (HEADER, os.path.join(ROOT, "include", "fs3d.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b(fs[23]d_[a-z0-9_]+)\s*\(", text))
    return sorted(names)
