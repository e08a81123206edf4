"""Captured steps (fs2d_set_graph, fs2d_batch_set_graph) against uncaptured
ones, bit for bit, and launches with more wavefront strips than the GPU holds
at once.

A captured step replays the stages up to a fixed PCG iteration budget (the
kernels of iterations after convergence return at once on the device), checks
convergence once on the host, finishes a solve that needed more iterations
with queued chunks, and replays the rest of the step.  The budget follows the
previous solve's count, so the first steps exercise the continuation; the
semi-Lagrangian stage swaps buffers, so successive steps alternate between two
captures; parameter changes force new captures.
"""
import numpy as np
import pytest

import fs2d
from fs2d import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = (fs2d.F_U, fs2d.F_V, fs2d.F_NEWU, fs2d.F_NEWV, fs2d.F_P, fs2d.F_CELL, fs2d.F_PARTICLE_POS,
          fs2d.F_PARTICLE_VEL, fs2d.F_DENSITY)
STATS = ("pcg_iters", "pcg_residual", "water_volume", "total_energy", "particle_total_energy", "cfl_max",
         "current_time", "steps")


def assert_same(a, b, what):
    for f in FIELDS:
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg="%s field %d" % (what, f))
    np.testing.assert_array_equal(a.get(fs2d.F_PHI, np.float64), b.get(fs2d.F_PHI, np.float64), err_msg=what)
    sa, sb = a.stats(), b.stats()
    for k in STATS:
        assert sa[k] == sb[k], (what, k, sa[k], sb[k])


def smoke_pair(n, seed):
    kw = dict(k=0, dt=0.01, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0), tau=0.97, tol=1e-6,
              max_iters=1000, static_cells=True, density=True)
    sims = [fs2d.Sim(scenes.all_fluid(n), **kw) for _ in range(2)]
    u, v, d = scenes.smoke_state(n, seed, sigma=2.0)
    for s in sims:
        s.set(fs2d.F_U, u)
        s.set(fs2d.F_V, v)
        s.set(fs2d.F_DENSITY, d)
    return sims


def pic_pair(n):
    kw = dict(k=2, dt=0.003, dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97, tol=1e-6, max_iters=500)
    return [fs2d.Sim(scenes.dam_break(n, 0.45, 0.6), **kw) for _ in range(2)]


@pytest.mark.parametrize("case", ["sl", "pic", "sl_dam_break"])
def test_graph_steps_bit_identical(case):
    if case == "sl":
        eager, graph = smoke_pair(96, 4)
    elif case == "pic":
        eager, graph = pic_pair(80)
    else:
        kw = dict(k=2, dt=0.005, dx=1.0 / 64, mode=fs2d.SEMILAGRANGIAN, tau=0.97, tol=1e-6, max_iters=500)
        eager, graph = [fs2d.Sim(O.dam_break(64, 64), **kw) for _ in range(2)]
    graph.set_graph(True)
    for step in range(7):
        if step == 4:  # a new dt: new captures
            for s in (eager, graph):
                s.set_dt(0.004)
        eager.step()
        graph.step()
        print("%s step %d: %d iterations" % (case, step, graph.stats()["pcg_iters"]))
        assert_same(eager, graph, "%s step %d" % (case, step))
    eager.close()
    graph.close()


def test_graph_budget_exceeded_continues_on_host():
    """A tighter tolerance after a loose one: the captured budget falls short
    and the solve continues with queued chunks."""
    eager, graph = smoke_pair(80, 5)
    graph.set_graph(True)
    for step, tol in enumerate((1e-2, 1e-2, 1e-7, 1e-7)):
        for s in (eager, graph):
            s.set_solver(tau=0.97, tol=tol, max_iters=1000)
        eager.step()
        graph.step()
        print("tol %g: %d iterations" % (tol, graph.stats()["pcg_iters"]))
        assert_same(eager, graph, "step %d" % step)


def test_graph_max_iters_reached():
    """A cap below the iterations needed: both report FS2D_WARN_MAXITER after
    the same work."""
    eager, graph = smoke_pair(64, 6)
    graph.set_graph(True)
    for s in (eager, graph):
        s.set_solver(tau=0.97, tol=1e-9, max_iters=12)
    for step in range(3):
        assert eager.step() == fs2d.WARN_MAXITER
        assert graph.step() == fs2d.WARN_MAXITER
        assert_same(eager, graph, "step %d" % step)


def pic_members(sides, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sides:
        w, h = rng.uniform(0.2, 0.7, size=2)
        kw = dict(k=2, dt=float(rng.choice([0.002, 0.003])), dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05,
                  tau=0.97, tol=1e-6, max_iters=500)
        out.append((scenes.dam_break(n, w, h), kw))
    return out


def test_graph_batch_bit_identical():
    members = pic_members([48, 32, 96, 64, 80], seed=2)
    eager = fs2d.Batch(members)
    graph = fs2d.Batch(members)
    graph.set_graph(True)
    singles = [fs2d.Sim(c, **kw) for c, kw in members]
    for step in range(5):
        eager.step()
        graph.step()
        for s in singles:
            s.step()
        for k in range(len(members)):
            assert_same(eager.sims[k], graph.sims[k], "step %d member %d (graph vs eager)" % (step, k))
            assert_same(singles[k], graph.sims[k], "step %d member %d (graph vs single)" % (step, k))
    # a member's particle count changes: new launch maps, a new capture
    rng = np.random.default_rng(9)
    n = 48
    pos = np.stack([rng.uniform(2.0, n - 3.0, 500), rng.uniform(2.0, n / 2, 500)], axis=1) / n
    for s in (eager.sims[0], graph.sims[0], singles[0]):
        s.set_particles(pos)
    for step in range(2):
        eager.step()
        graph.step()
        for s in singles:
            s.step()
        for k in range(len(members)):
            assert_same(eager.sims[k], graph.sims[k], "after edit step %d member %d" % (step, k))
            assert_same(singles[k], graph.sims[k], "after edit step %d member %d (single)" % (step, k))
    eager.close()
    graph.close()


def test_batch_more_strips_than_resident():
    """48 members of 48 x 704 cells: 528 strips in every batched wavefront
    launch, more than the GPU holds at once (about one sweep workgroup per CU
    on 256 CUs); strips wait only on strips taken earlier, so the launch
    drains.  Members must still step bit-identically to single steps."""
    n_x, n_y = 48, 704
    members = []
    for k in range(48):
        kw = dict(k=0, dt=0.01, dx=1.0 / n_y, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0), tau=0.97, tol=1e-6,
                  max_iters=400, static_cells=True, density=True)
        members.append((O.all_fluid(n_x, n_y), kw))
    batch = fs2d.Batch(members)
    batch.set_graph(True)
    rng = np.random.default_rng(12)
    states = []
    for k in range(len(members)):
        u = rng.standard_normal((n_y, n_x + 1)).astype(np.float32)
        v = rng.standard_normal((n_y + 1, n_x)).astype(np.float32)
        d = rng.random((n_y, n_x)).astype(np.float32)
        states.append((u, v, d))
        s = batch.sims[k]
        s.set(fs2d.F_U, u); s.set(fs2d.F_V, v); s.set(fs2d.F_DENSITY, d)
    batch.step()
    for k in (0, 17, 47):
        single = fs2d.Sim(members[k][0], **members[k][1])
        u, v, d = states[k]
        single.set(fs2d.F_U, u); single.set(fs2d.F_V, v); single.set(fs2d.F_DENSITY, d)
        single.step()
        for f in (fs2d.F_U, fs2d.F_V, fs2d.F_P, fs2d.F_DENSITY):
            np.testing.assert_array_equal(batch.sims[k].get(f), single.get(f), err_msg="member %d field %d" % (k, f))
        assert batch.sims[k].stats()["pcg_iters"] == single.stats()["pcg_iters"]
        single.close()
    batch.close()


def test_single_projection_full_chip_of_strips():
    """A 64 x 16384 grid, the tallest a handle takes: 256 strips of 64 rows in
    each solve launch, each workgroup holding 98 KB of LDS (one per CU), so the
    launch fills every CU of the chip.  Matched iterations against the
    oracle."""
    nx, ny = 64, 16384
    cells = O.walls(nx, ny)
    cells[1:ny - 1, 1:nx - 1] = O.FLUID
    kw = dict(k=0, dt=0.005, dx=1.0 / ny, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0))
    g = fs2d.Sim(cells, static_cells=True, **kw)
    o = O.Oracle(cells, level_set=False, sl_inplace=False, **kw)
    rng = np.random.default_rng(20000)
    u = rng.standard_normal((ny, nx + 1)).astype(np.float32)
    v = rng.standard_normal((ny + 1, nx)).astype(np.float32)
    g.set(fs2d.F_U, u)
    g.set(fs2d.F_V, v)
    o.set(O.F_U, u.astype(np.float64))
    o.set(O.F_V, v.astype(np.float64))
    o.set(O.F_PHI, g.get(fs2d.F_PHI, np.float64))
    g.set_solver(tau=0.97, tol=0.0, max_iters=3)
    o.set_options(tau=0.97, tol=0.0, max_iters=3, sl_inplace=False)
    g.run_stage(fs2d.STAGE_PROJECT, allow=(fs2d.OK, fs2d.WARN_MAXITER))
    o.run_stage(O.ST_PROJECT)
    assert g.stats()["pcg_iters"] == o.stats()["last_iters"] == 3
    a, b = g.get(fs2d.F_P).astype(np.float64), o.get(O.F_P)
    e = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("64x16384, 3 iterations: p rel-L2 %.3e" % e)
    assert e < 1e-5
