"""The reference's compile-time variants on the GPU (src/FluidSim2D.cpp:15-16):
USE_GHOST_PRESSURE undefined (FS2D_FLAG_NO_GHOST_PRESSURE) and USE_LEVEL_SET
undefined (FS2D_FLAG_PARTICLE_CELLS, which needs the former).  Teacher-forced
whole PIC steps against the oracle (flags bit-exact; u, v, p and particle
positions within 1e-5 relative L2 of the FP64 oracle per step), and batch
members against single steps, bit for bit.

The particle-cell variant marks cells from particle positions in metres, as
the reference does; dx = 1 makes metres cells, so the marking is exercised."""
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


def push(g, o):
    o.set(O.F_POS, g.get(fs2d.F_PARTICLE_POS).astype(np.float64))
    o.set(O.F_VEL, g.get(fs2d.F_PARTICLE_VEL).astype(np.float64))
    o.set(O.F_U, g.get(fs2d.F_U).astype(np.float64))
    o.set(O.F_V, g.get(fs2d.F_V).astype(np.float64))
    o.set(O.F_CELL, g.get(fs2d.F_CELL))


@pytest.mark.parametrize("particle_cells", [False, True])
def test_variant_teacher_forced_steps(particle_cells):
    n = 48
    dx = 1.0 if particle_cells else 1.0 / n
    dt = 0.05 if particle_cells else 0.004
    cells = O.dam_break(n, n)
    kw = dict(k=2, dt=dt, dx=dx, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97)
    g = fs2d.Sim(cells, tol=1e-6, max_iters=500, particle_cells=particle_cells, ghost_pressure=False, **kw)
    o = O.Oracle(cells, tol=1e-12, max_iters=2000, sl_inplace=False, **kw)
    o.set_variants(particle_cells=particle_cells, ghost=False)
    for step in range(5):
        push(g, o)
        g.step()
        o.step()
        np.testing.assert_array_equal(g.get(fs2d.F_CELL), o.get(O.F_CELL))
        eu, ev = rel_l2(g.get(fs2d.F_U), o.get(O.F_U)), rel_l2(g.get(fs2d.F_V), o.get(O.F_V))
        ex = rel_l2(g.get(fs2d.F_PARTICLE_POS), o.get(O.F_POS))
        ep = rel_l2(g.get(fs2d.F_P), o.get(O.F_P))
        print("step %d: u %.2e v %.2e pos %.2e p %.2e fluid %d" % (step, eu, ev, ex, ep,
                                                                    int((g.get(fs2d.F_CELL) == 1).sum())))
        assert ex < REL and eu < REL and ev < REL and ep < REL
    assert (g.get(fs2d.F_CELL) == fs2d.FLUID).sum() > 100


def test_variant_batch_members_match_single_steps():
    members = []
    for n, pc in ((40, True), (56, True)):
        kw = dict(k=2, dt=0.02, dx=1.0, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97, tol=1e-6, max_iters=500,
                  particle_cells=pc, ghost_pressure=False)
        members.append((O.dam_break(n, n), kw))
    batch = fs2d.Batch(members)
    singles = [fs2d.Sim(c, **kw) for c, kw in members]
    for step in range(3):
        batch.step()
        for s in singles:
            s.step()
        for k, (b, s) in enumerate(zip(batch.sims, singles)):
            for f in (fs2d.F_U, fs2d.F_V, fs2d.F_P, fs2d.F_CELL, fs2d.F_PARTICLE_POS):
                np.testing.assert_array_equal(b.get(f), s.get(f), err_msg="step %d member %d field %d" % (step, k, f))
    batch.close()
