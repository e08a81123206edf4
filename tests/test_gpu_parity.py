"""GPU parity: every stage of the HIP path against the CPU oracle from an
identical state, plus teacher-forced whole steps.

Bars: cell flags, particle cell indices and the level set's
discrete decisions bit-exact; velocities, pressures and particle positions
within 1e-5 relative L2 of the FP64 oracle.  The GPU stores FP32; the oracle
is given the GPU's FP32 values (widened to FP64) so that both start from the
same numbers.  The semi-Lagrangian comparison uses the oracle's out-of-place
mode: the reference's in-place update (FluidSim2D.cpp:209-233) has no parallel
equivalent; its gap is measured in test_sl_inplace_gap.
"""
import numpy as np
import pytest

import fs2d
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def pair(cells, **kw):
    """A GPU sim and an oracle with the same configuration."""
    okw = dict(kw)
    static = okw.pop("static_cells", False)
    dens = okw.pop("density", False)
    g = fs2d.Sim(cells, static_cells=static, density=dens, **okw)
    o = O.Oracle(cells, level_set=not static, density=dens,
                 sl_inplace=False, **{k: v for k, v in okw.items() if k != "seed"})
    return g, o


def push(g, o, fields):
    """Copy GPU state into the oracle (FP32 values widened to FP64)."""
    for f in fields:
        if f == fs2d.F_CELL:
            o.set(O.F_CELL, g.get(fs2d.F_CELL))
        elif f == fs2d.F_PHI:
            o.set(O.F_PHI, g.get(fs2d.F_PHI, np.float64))
        elif f == fs2d.F_PARTICLE_POS:
            o.set(O.F_POS, g.get(f).astype(np.float64))
        elif f == fs2d.F_PARTICLE_VEL:
            o.set(O.F_VEL, g.get(f).astype(np.float64))
        else:
            o.set(f, g.get(f).astype(np.float64))


def check_particle_keys(g, dx):
    """Bit-exact particle cell keys and sort order: the GPU's keys equal the
    reference's rule x_cell = (int)(x / dx) in FP64 (FluidSim2D.cpp:766) on
    the GPU's own positions, and its gather order is the stable sort of those
    keys (ties by particle index)."""
    pos = g.get(fs2d.F_PARTICLE_POS).astype(np.float64)
    want = (pos[:, 1] / dx).astype(np.int64) * g.nx + (pos[:, 0] / dx).astype(np.int64)
    keys = g.get(fs2d.F_PARTICLE_CELL)
    np.testing.assert_array_equal(keys, want)
    np.testing.assert_array_equal(g.get(fs2d.F_PARTICLE_ORDER), np.argsort(want, kind="stable"))


def smooth_field(shape, rng, scale=1.0):
    from scipy.ndimage import gaussian_filter
    return (scale * gaussian_filter(rng.standard_normal(shape), 2.0, mode="nearest") * 4).astype(np.float32)


def settle(g, o, steps=3, mode_stages=None):
    """Run a few GPU steps so that velocities, levels and particles are generic."""
    for _ in range(steps):
        g.step()


SCENE64 = dict(k=2, dt=0.005, dx=1.0 / 64, mode=fs2d.SEMILAGRANGIAN)


# ---------------------------------------------------------------------------
def test_create_seeding_matches_oracle():
    cells = O.dam_break(64, 64)
    g, o = pair(cells, **SCENE64)
    pg = g.get(fs2d.F_PARTICLE_POS)
    po = o.get(O.F_POS)
    assert pg.shape == po.shape == (4 * 888, 2)
    np.testing.assert_array_equal(pg, po.astype(np.float32))


@pytest.mark.parametrize("n,mode", [(64, fs2d.SEMILAGRANGIAN), (96, fs2d.PICFLIP)])
def test_level_set_bit_exact(n, mode):
    cells = O.dam_break(n, n)
    g, o = pair(cells, k=2, dt=0.005, dx=1.0 / n, mode=mode)
    for it in range(3):
        push(g, o, [fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL, fs2d.F_U, fs2d.F_V, fs2d.F_CELL])
        g.run_stage(fs2d.STAGE_LEVEL_SET)
        o.run_stage(O.ST_LEVELSET)
        np.testing.assert_array_equal(g.get(fs2d.F_CELL), o.get(O.F_CELL))
        np.testing.assert_array_equal(g.get(fs2d.F_PHI, np.float64), o.get(O.F_PHI))
        sg, so = g.stats(), o.stats()
        assert sg["water_volume"] == pytest.approx(so["water_volume"], rel=1e-12)
        assert sg["particle_total_energy"] == pytest.approx(so["particle_total_energy"], rel=1e-9)
        assert sg["total_energy"] == pytest.approx(so["total_energy"], rel=1e-5)
        g.step()  # move on to a generic state


# Particle sets whose closest-particle passes first change nothing at pass 4
# (direction 0), pass 2 (direction 2) and, for the seeded dam break, pass 3
# (direction 3): the passes after it return at once and the adopted particles
# are read back from that pass's output, in its direction's order.
LS_EARLY_EXIT = {
    "five particles": [[0.9742, 0.6448], [0.2478, 0.4382], [0.9505, 0.8778], [0.827, 0.3978], [0.4934, 0.6679]],
    "one particle": [[0.8536, 0.0426]],
    "dam break": None,
}


def test_level_set_early_exit_bit_exact():
    """The closest-particle passes stop after the first pass that changes no
    cell (a fixed point of every later pass): flags and phi stay bit-exact
    against the oracle, which runs all 16 passes, whichever direction that
    pass had; and a batch whose members stop at different passes matches."""
    n = 48
    kw = dict(k=2, dt=0.005, dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05)
    members = []
    for name, pts in LS_EARLY_EXIT.items():
        g, o = pair(O.dam_break(n, n), **kw)
        if pts is not None:
            g.set_particles(np.array(pts, np.float64))
        a = g.get(fs2d.F_PARTICLE_POS).astype(np.float64)  # the GPU's FP32 positions
        o.set(O.F_POS, a)
        o.set(O.F_VEL, np.zeros_like(a))
        g.run_stage(fs2d.STAGE_LEVEL_SET)
        o.run_stage(O.ST_LEVELSET)
        np.testing.assert_array_equal(g.get(fs2d.F_CELL), o.get(O.F_CELL), err_msg=name)
        np.testing.assert_array_equal(g.get(fs2d.F_PHI, np.float64), o.get(O.F_PHI), err_msg=name)
        members.append((name, pts, g.get(fs2d.F_PHI, np.float64), g.get(fs2d.F_CELL)))
        g.close()
    batch = fs2d.Batch([(O.dam_break(n, n), kw) for _ in members])
    for m, (name, pts, _, _) in zip(batch.sims, members):
        if pts is not None:
            m.set_particles(np.array(pts, np.float64))
    batch.run_stage(fs2d.STAGE_LEVEL_SET)
    for m, (name, _, phi, cells) in zip(batch.sims, members):
        np.testing.assert_array_equal(m.get(fs2d.F_PHI, np.float64), phi, err_msg="batch " + name)
        np.testing.assert_array_equal(m.get(fs2d.F_CELL), cells, err_msg="batch " + name)
    batch.close()


def test_sl_advection_stage():
    n = 64
    cells = O.dam_break(n, n)
    g, o = pair(cells, **SCENE64)
    rng = np.random.default_rng(0)
    g.set(fs2d.F_U, smooth_field((n, n + 1), rng))
    g.set(fs2d.F_V, smooth_field((n + 1, n), rng))
    push(g, o, [fs2d.F_U, fs2d.F_V])
    g.run_stage(fs2d.STAGE_SL)
    o.run_stage(O.ST_SL)
    eu, ev = rel_l2(g.get(fs2d.F_U), o.get(O.F_U)), rel_l2(g.get(fs2d.F_V), o.get(O.F_V))
    print("SL rel-L2 u %.3e v %.3e" % (eu, ev))
    assert eu < REL and ev < REL


@pytest.mark.parametrize("nx,ny,cfl", [(100, 70, 0.9), (130, 45, 3.5), (64, 200, 6.0)])
def test_sl_advection_tiles_and_fast_flow(nx, ny, cfl):
    """The tiled stage on grids that are not multiples of its 64 x 16 tiles,
    at Courant numbers whose departure points stay in a tile's LDS window
    (0.9) and leave it (3.5, 6: those samples read the arrays), with density
    and dx not a power of two."""
    cells = O.all_fluid(nx, ny)
    dx = 0.013
    rng = np.random.default_rng(nx + ny)
    u0 = smooth_field((ny, nx + 1), rng).astype(np.float64)
    v0 = smooth_field((ny + 1, nx), rng).astype(np.float64)
    vmax = max(np.abs(u0).max(), np.abs(v0).max())
    dt = cfl * dx / vmax
    g, o = pair(cells, k=0, dt=dt, dx=dx, mode=fs2d.SEMILAGRANGIAN, static_cells=True, density=True,
                gravity=(0.0, 0.0))
    g.set(fs2d.F_U, u0.astype(np.float32))
    g.set(fs2d.F_V, v0.astype(np.float32))
    g.set(fs2d.F_DENSITY, np.abs(smooth_field((ny, nx), rng)))
    push(g, o, [fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY])
    g.run_stage(fs2d.STAGE_SL)
    o.run_stage(O.ST_SL)
    e = [rel_l2(g.get(f), o.get(of)) for f, of in ((fs2d.F_U, O.F_U), (fs2d.F_V, O.F_V),
                                                    (fs2d.F_DENSITY, O.F_DENS))]
    print("%dx%d CFL %.1f: SL rel-L2 u %.2e v %.2e density %.2e" % (nx, ny, cfl, *e))
    assert max(e) < REL


def test_density_advection_stage():
    n = 48
    cells = O.all_fluid(n, n)
    g, o = pair(cells, k=0, dt=0.01, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, static_cells=True, density=True)
    rng = np.random.default_rng(1)
    g.set(fs2d.F_U, smooth_field((n, n + 1), rng))
    g.set(fs2d.F_V, smooth_field((n + 1, n), rng))
    g.set(fs2d.F_DENSITY, np.abs(smooth_field((n, n), rng)))
    push(g, o, [fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY])
    g.run_stage(fs2d.STAGE_SL)
    o.run_stage(O.ST_SL)
    e = rel_l2(g.get(fs2d.F_DENSITY), o.get(O.F_DENS))
    print("density rel-L2 %.3e" % e)
    assert e < REL


def test_gravity_stage():
    n = 32
    g, o = pair(O.dam_break(n, n), k=2, dt=0.01, dx=1.0 / n)
    g.set_gravity(0.3, -9.81)
    o.set_params(0.01, (0.3, -9.81), 1.0)
    push(g, o, [fs2d.F_U, fs2d.F_V])
    g.run_stage(fs2d.STAGE_GRAVITY)
    o.run_stage(O.ST_GRAVITY)
    assert rel_l2(g.get(fs2d.F_U), o.get(O.F_U)) < 1e-7
    assert rel_l2(g.get(fs2d.F_V), o.get(O.F_V)) < 1e-7


def projection_case(n, mode, steps):
    cells = O.dam_break(n, n)
    g, o = pair(cells, k=2, dt=0.005, dx=1.0 / n, mode=mode, max_iters=1000)
    for _ in range(steps):
        g.step()
    g.run_stage(fs2d.STAGE_LEVEL_SET)
    if mode == fs2d.PICFLIP:
        g.run_stage(fs2d.STAGE_P2G)
    else:
        g.run_stage(fs2d.STAGE_SL)
    g.run_stage(fs2d.STAGE_GRAVITY)
    push(g, o, [fs2d.F_CELL, fs2d.F_PHI, fs2d.F_U, fs2d.F_V])
    return g, o


@pytest.mark.parametrize("mode", [fs2d.SEMILAGRANGIAN, fs2d.PICFLIP])
def test_projection_matched_iterations(mode):
    """Same algorithm, same iteration count: FP32 GPU vs FP64 oracle."""
    g, o = projection_case(64, mode, 4)
    for iters in (1, 5, 20):
        g.set_solver(tau=0.999, tol=0.0, max_iters=iters)
        o.set_options(tau=0.999, tol=0.0, max_iters=iters, sl_inplace=False)
        g.run_stage(fs2d.STAGE_PROJECT, allow=(fs2d.OK, fs2d.WARN_MAXITER))
        o.run_stage(O.ST_PROJECT)
        assert g.stats()["pcg_iters"] == o.stats()["last_iters"] == iters
        e = rel_l2(g.get(fs2d.F_P), o.get(O.F_P))
        print("iters %d p rel-L2 %.3e" % (iters, e))
        assert e < REL


@pytest.mark.parametrize("mode", [fs2d.SEMILAGRANGIAN, fs2d.PICFLIP])
def test_projection_converged_and_velocity_update(mode):
    """GPU to tol 1e-6 vs the oracle to 1e-12; then updateVelocity."""
    g, o = projection_case(64, mode, 4)
    g.set_solver(tau=0.97, tol=1e-6, max_iters=500)
    o.set_options(tau=0.97, tol=1e-12, max_iters=2000, sl_inplace=False)
    g.run_stage(fs2d.STAGE_PROJECT)
    o.run_stage(O.ST_PROJECT)
    sg = g.stats()
    ep = rel_l2(g.get(fs2d.F_P), o.get(O.F_P))
    print("converged: gpu iters %d (res %.2e), oracle iters %d, p rel-L2 %.3e"
          % (sg["pcg_iters"], sg["pcg_residual"], o.stats()["last_iters"], ep))
    assert sg["pcg_residual"] <= 1e-6
    assert ep < REL
    # velocity update from the same pressure
    push(g, o, [fs2d.F_P])
    g.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    o.run_stage(O.ST_UPDATEVEL)
    eu = rel_l2(g.get(fs2d.F_NEWU), o.get(O.F_NEWU))
    ev = rel_l2(g.get(fs2d.F_NEWV), o.get(O.F_NEWV))
    print("update velocity rel-L2 u %.3e v %.3e" % (eu, ev))
    assert eu < REL and ev < REL


def test_mic0_preconditioner_matches():
    """One preconditioner application (0 CG updates) equals the oracle's."""
    g, o = projection_case(48, fs2d.SEMILAGRANGIAN, 3)
    # max_iters=1, tol=0: p = alpha * M^-1 rhs
    g.set_solver(tau=0.999, tol=0.0, max_iters=1)
    o.set_options(tau=0.999, tol=0.0, max_iters=1, sl_inplace=False)
    g.run_stage(fs2d.STAGE_PROJECT, allow=(fs2d.OK, fs2d.WARN_MAXITER))
    o.run_stage(O.ST_PROJECT)
    e = rel_l2(g.get(fs2d.F_P), o.get(O.F_P))
    print("1-iteration p rel-L2 %.3e" % e)
    assert e < REL


def test_p2g_stage():
    n = 64
    cells = O.dam_break(n, n)
    g, o = pair(cells, k=2, dt=0.005, dx=1.0 / n, mode=fs2d.PICFLIP)
    for _ in range(3):
        g.step()
    rng = np.random.default_rng(3)
    m = g.stats()["n_particles"]
    g.set(fs2d.F_PARTICLE_VEL, rng.standard_normal((m, 2)).astype(np.float32))
    push(g, o, [fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL])
    g.run_stage(fs2d.STAGE_P2G)
    o.run_stage(O.ST_P2G)
    eu, ev = rel_l2(g.get(fs2d.F_U), o.get(O.F_U)), rel_l2(g.get(fs2d.F_V), o.get(O.F_V))
    print("P2G rel-L2 u %.3e v %.3e" % (eu, ev))
    assert eu < REL and ev < REL


def test_g2p_and_particle_advection():
    n = 64
    cells = O.dam_break(n, n)
    g, o = pair(cells, k=2, dt=0.005, dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05)
    for _ in range(3):
        g.step()
    g.run_stage(fs2d.STAGE_LEVEL_SET)
    g.run_stage(fs2d.STAGE_P2G)
    g.run_stage(fs2d.STAGE_GRAVITY)
    g.run_stage(fs2d.STAGE_PROJECT)
    g.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    push(g, o, [fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL, fs2d.F_U, fs2d.F_V, fs2d.F_NEWU, fs2d.F_NEWV])
    g.run_stage(fs2d.STAGE_G2P)
    o.run_stage(O.ST_G2P)
    e = rel_l2(g.get(fs2d.F_PARTICLE_VEL), o.get(O.F_VEL))
    print("G2P rel-L2 %.3e" % e)
    assert e < REL
    push(g, o, [fs2d.F_PARTICLE_VEL, fs2d.F_U, fs2d.F_V])
    g.run_stage(fs2d.STAGE_ADVECT)
    o.run_stage(O.ST_ADVECT)
    e = rel_l2(g.get(fs2d.F_PARTICLE_POS), o.get(O.F_POS))
    print("advect rel-L2 %.3e" % e)
    assert e < REL
    check_particle_keys(g, 1.0 / n)
    # against the oracle's FP64 positions the cells may differ only where the
    # two positions straddle a cell edge within FP32 accuracy
    pg = g.get(fs2d.F_PARTICLE_POS).astype(np.float64)
    po = o.get(O.F_POS)
    kg, ko = np.floor(pg * n).astype(int), np.floor(po * n).astype(int)
    diff = np.any(kg != ko, axis=1)
    near = np.any(np.abs(pg * n - np.round(pg * n)) < 1e-6 * n, axis=1)
    print("cell indices differing from the oracle's: %d of %d (%d particles within 1e-6 of an edge)"
          % (diff.sum(), len(pg), near.sum()))
    assert not (diff & ~near).any()


@pytest.mark.parametrize("mode", [fs2d.SEMILAGRANGIAN, fs2d.PICFLIP])
def test_teacher_forced_steps(mode):
    """Whole steps: the oracle restarts from the GPU's state every step."""
    n = 64
    cells = O.dam_break(n, n)
    g, o = pair(cells, k=2, dt=0.005, dx=1.0 / n, mode=mode, alpha=0.05, tau=0.97,
                tol=1e-6, max_iters=500)
    o.set_options(tau=0.97, tol=1e-12, max_iters=2000, sl_inplace=False)
    for step in range(6):
        push(g, o, [fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL, fs2d.F_U, fs2d.F_V, fs2d.F_CELL])
        g.step()
        o.step()
        np.testing.assert_array_equal(g.get(fs2d.F_CELL), o.get(O.F_CELL))
        eu, ev = rel_l2(g.get(fs2d.F_U), o.get(O.F_U)), rel_l2(g.get(fs2d.F_V), o.get(O.F_V))
        ex = rel_l2(g.get(fs2d.F_PARTICLE_POS), o.get(O.F_POS))
        ep = rel_l2(g.get(fs2d.F_P), o.get(O.F_P))
        print("step %d: u %.2e v %.2e pos %.2e p %.2e" % (step, eu, ev, ex, ep))
        assert ex < REL and eu < REL and ev < REL and ep < REL
        check_particle_keys(g, 1.0 / n)


def test_golden_semilag_rows_on_gpu():
    """The GPU step reproduces the reference's conservation rows 1-2
    (paper/data/conservation-semilag-state1.csv) within FP32 accuracy."""
    import json, os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "paper_conservation.json")))
    ref = gold["files"]["conservation-semilag-state1.csv"]
    g = fs2d.Sim(O.demo_triangle(128), k=2, dt=0.033, dx=0.02, mode=fs2d.SEMILAGRANGIAN)
    got = []
    for r in range(2):
        g.run_stage(fs2d.STAGE_LEVEL_SET)
        s = g.stats()
        got.append([s["water_volume"], s["total_energy"], s["particle_total_energy"]])
        if r == 0:
            for st in (fs2d.STAGE_SL, fs2d.STAGE_GRAVITY, fs2d.STAGE_PROJECT,
                       fs2d.STAGE_UPDATE_VELOCITY, fs2d.STAGE_ADVECT):
                g.run_stage(st, allow=(fs2d.OK, fs2d.WARN_MAXITER))
    print("gpu rows", got, "ref", ref[:2])
    np.testing.assert_allclose(np.array(got)[:, 0], np.array(ref[:2])[:, 0], rtol=1e-9)
    np.testing.assert_allclose(np.array(got)[:, 1:], np.array(ref[:2])[:, 1:], rtol=1e-6)


@pytest.mark.parametrize("nx,ny", [(5, 130), (40, 200), (1000, 65), (77, 64), (130, 5), (1024, 600)])
def test_projection_odd_shapes(nx, ny):
    """Matched-iteration PCG on shapes whose strips and columns do not fill the
    solver's tiles: a partial last strip (ny mod 64 != 0), a single strip,
    widths far from a multiple of 32 and a 5-wide grid (the hand-off rows then
    carry mostly padding columns).  The shapes also cover the SpMV / axpy
    workgroup counts: one entry per thread (77 x 64), the floor of 256
    (1000 x 65) and about 8 entries per thread (1024 x 600: 336 workgroups,
    so 336 w.s and |r| partials).  The narrow grid's fluid is kept short (16
    rows of 3 cells): a long thin column is a near-1D Laplacian whose condition
    number grows as its length squared (85 rows of one cell put the FP32 error
    at ~2e-4), and on which MIC(0) is nearly exact, so after a few iterations
    the FP32 residual is exhausted; 3 iterations run every sweep."""
    rng = np.random.default_rng(nx * 1000 + ny)
    cells = O.walls(nx, ny)
    top = 17 if nx < 8 else max(2, (2 * ny) // 3)
    cells[1:top, 1:nx - 1] = O.FLUID  # fluid below, air above (ghost terms)
    g, o = pair(cells, k=0, dt=0.005, dx=1.0 / max(nx, ny), mode=fs2d.SEMILAGRANGIAN, static_cells=True,
                gravity=(0.0, 0.0))
    g.set(fs2d.F_U, rng.standard_normal((ny, nx + 1)).astype(np.float32))
    g.set(fs2d.F_V, rng.standard_normal((ny + 1, nx)).astype(np.float32))
    push(g, o, [fs2d.F_CELL, fs2d.F_PHI, fs2d.F_U, fs2d.F_V])
    for iters in ((1, 3) if nx < 8 else (1, 7)):
        g.set_solver(tau=0.97, tol=0.0, max_iters=iters)
        o.set_options(tau=0.97, tol=0.0, max_iters=iters, sl_inplace=False)
        g.run_stage(fs2d.STAGE_PROJECT, allow=(fs2d.OK, fs2d.WARN_MAXITER))
        o.run_stage(O.ST_PROJECT)
        assert g.stats()["pcg_iters"] == o.stats()["last_iters"] == iters
        e = rel_l2(g.get(fs2d.F_P), o.get(O.F_P))
        print("%dx%d iters %d p rel-L2 %.3e" % (nx, ny, iters, e))
        assert np.isfinite(g.get(fs2d.F_P)).all()
        assert e < REL


@pytest.mark.parametrize("nx,ny,maxh", [(3000, 24, 5), (4200, 24, 5), (40, 8300, 5), (48, 48, 0)])
def test_extrapolation_wide_and_deep_layers(nx, ny, maxh):
    """updateVelocity + extrapolation (Array2D.h:552-591) where the layer-order
    pass leaves its fast path: layers wider than 2048 faces (3000 x 24), wider
    than the 4096 kept in LDS (4200 x 24), more than 8190 layers (40 x 8300),
    and no known face at all (48 x 48 without fluid).  Fluid fills the bottom
    rows to a random height per column, so the first layers are ragged."""
    rng = np.random.default_rng(nx * 7 + ny)
    cells = O.walls(nx, ny)
    if maxh:
        h = rng.integers(1, maxh + 1, size=nx)
        for i in range(1, nx - 1):
            cells[1:1 + h[i], i] = O.FLUID
    g, o = pair(cells, k=0, dt=0.005, dx=1.0 / max(nx, ny), mode=fs2d.SEMILAGRANGIAN, static_cells=True,
                gravity=(0.0, 0.0))
    g.set(fs2d.F_U, rng.standard_normal((ny, nx + 1)).astype(np.float32))
    g.set(fs2d.F_V, rng.standard_normal((ny + 1, nx)).astype(np.float32))
    g.set(fs2d.F_P, rng.standard_normal((ny, nx)).astype(np.float32))
    push(g, o, [fs2d.F_CELL, fs2d.F_PHI, fs2d.F_U, fs2d.F_V, fs2d.F_P])
    g.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    o.run_stage(O.ST_UPDATEVEL)
    for fg, fo in ((fs2d.F_NEWU, O.F_NEWU), (fs2d.F_NEWV, O.F_NEWV)):
        a, b = g.get(fg), o.get(fo)
        e = rel_l2(a, b)
        print("%dx%d field %d: rel-L2 %.3e, max abs %.3e" % (nx, ny, fg, e, np.abs(a - b).max()))
        assert np.isfinite(a).all()
        assert e < REL


def test_all_fluid_box_divergence_free_2048():
    """Size-independent property at the benchmark size: after the projection
    the discrete divergence of every FLUID cell is at the solver tolerance."""
    n = 2048
    cells = O.all_fluid(n, n)
    g = fs2d.Sim(cells, k=0, dt=1e-3, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, static_cells=True,
                 gravity=(0.0, 0.0), tau=0.97, tol=1e-6, max_iters=2000)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((n, n + 1)).astype(np.float32)
    v = rng.standard_normal((n + 1, n)).astype(np.float32)
    g.set(fs2d.F_U, u)
    g.set(fs2d.F_V, v)
    g.run_stage(fs2d.STAGE_PROJECT)
    s = g.stats()
    g.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    u2, v2 = g.get(fs2d.F_U).astype(np.float64), g.get(fs2d.F_V).astype(np.float64)
    div = (u2[:, 1:] - u2[:, :-1]) + (v2[1:, :] - v2[:-1, :])
    div0 = (u[:, 1:].astype(np.float64) - u[:, :-1]) + (v[1:, :].astype(np.float64) - v[:-1, :])
    fl = cells == O.FLUID
    print("2048^2: iters %d residual %.2e  max|div| %.3e (before %.3e)"
          % (s["pcg_iters"], s["pcg_residual"], np.abs(div[fl]).max(), np.abs(div0[fl]).max()))
    assert s["pcg_residual"] <= 1e-6
    assert np.abs(div[fl]).max() < 1e-4 * np.abs(div0[fl]).max()


def test_timing_classes():
    """fs2d_set_timing_classes times exactly the classes in its mask; timing
    leaves the results unchanged (bit for bit)."""
    n = 64
    cells = O.dam_break(n, n)
    runs = {}
    for mode in ("off", "fwd", "all"):
        g = fs2d.Sim(cells, k=2, dt=0.005, dx=1.0 / n, mode=fs2d.PICFLIP)
        g.step()
        g.reset_kernel_times()
        if mode == "fwd":
            g.set_timing(True, only=["pcg_forward_solve", "p2g"])
        elif mode == "all":
            g.set_timing(True)
        for _ in range(2):
            g.step()
        kt = g.kernel_times()
        timed = {k for k, (ms, launches) in kt.items() if launches > 0}
        if mode == "off":
            assert timed == set()
        elif mode == "fwd":
            assert timed == {"pcg_forward_solve", "p2g"}
            assert kt["pcg_forward_solve"][0] > 0
        else:
            assert {"pcg_forward_solve", "pcg_backward_solve", "p2g", "levelset_sweep"} <= timed
            # the p/r update and the SpMV run inside the solves' launches with fused_pcg
            separate = {"pcg_spmv_dot", "pcg_axpy_norm"}
            if g.tuning()["fused_pcg"]:
                assert not separate & timed
            else:
                assert separate <= timed
        runs[mode] = (g.get(fs2d.F_U), g.get(fs2d.F_P), g.get(fs2d.F_PARTICLE_POS))
        g.close()
    for mode in ("fwd", "all"):
        for a, b in zip(runs["off"], runs[mode]):
            np.testing.assert_array_equal(a, b)


def test_error_paths():
    n = 32
    g = fs2d.Sim(O.dam_break(n, n), k=2, dt=0.005, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN)
    # PCG cap: a warning status, not an abort
    g.set_solver(tol=0.0, max_iters=1)
    g.run_stage(fs2d.STAGE_LEVEL_SET)
    g.run_stage(fs2d.STAGE_GRAVITY)
    assert g.run_stage(fs2d.STAGE_PROJECT, allow=(fs2d.WARN_MAXITER,)) == fs2d.WARN_MAXITER
    # particles outside the interior box are rejected
    with pytest.raises(fs2d.FS2DError) as e:
        g.set_particles(np.array([[0.5 / n, 0.5]]))
    assert e.value.status == fs2d.ERR_ARG
    with pytest.raises(fs2d.FS2DError):
        g.set(fs2d.F_U, np.zeros(5, np.float32))
    # empty particle set: the level set marks nothing FLUID and rhs = 0 (p = 0)
    g.set_particles(np.zeros((0, 2)))
    g.set_solver()
    g.step()
    cell = g.get(fs2d.F_CELL)
    assert not (cell == fs2d.FLUID).any()
    assert not g.get(fs2d.F_P).any()
