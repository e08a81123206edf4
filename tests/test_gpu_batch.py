"""Ensembles (fs2d_batch_* of the C ABI): every member of a batch steps
bit-identically to the same simulation stepped alone with fs2d_step, whose
parity with the oracle test_gpu_parity.py establishes.

Members differ in size, fluid block, dt and PIC/FLIP blend, as in the ensemble
workload (PIC/FLIP dam breaks, fluid block width and height uniform in
[0.2, 0.7] of the side, 4 particles per cell, FLIP blend 0.95), so that they
converge after different numbers of PCG iterations and the masking of
converged members is exercised.
"""
import numpy as np
import pytest

import fs2d
from fs2d import scenes

pytestmark = pytest.mark.gpu

FIELDS = (fs2d.F_U, fs2d.F_V, fs2d.F_NEWU, fs2d.F_NEWV, fs2d.F_P, fs2d.F_PHI, fs2d.F_CELL,
          fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL)


def pic_members(sides, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sides:
        w, h = rng.uniform(0.2, 0.7, size=2)
        kw = dict(k=2, dt=float(rng.choice([0.002, 0.003])), dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05,
                  tau=0.97, tol=1e-6, max_iters=500)
        out.append((scenes.dam_break(n, w, h), kw))
    return out


def assert_same(a, b, what):
    for f in FIELDS:
        if f == fs2d.F_PHI:
            x, y = a.get(f, np.float64), b.get(f, np.float64)
        else:
            x, y = a.get(f), b.get(f)
        assert x.shape == y.shape, (what, f)
        np.testing.assert_array_equal(x, y, err_msg="%s field %d" % (what, f))
    sa, sb = a.stats(), b.stats()
    for k in ("pcg_iters", "water_volume", "total_energy", "particle_total_energy", "cfl_max", "current_time",
              "steps", "pcg_residual"):
        assert sa[k] == sb[k], (what, k, sa[k], sb[k])


def test_pic_batch_members_match_single_steps():
    members = pic_members([48, 32, 96, 64, 80], seed=2)
    batch = fs2d.Batch(members)
    singles = [fs2d.Sim(c, **kw) for c, kw in members]
    iters = set()
    for step in range(4):
        batch.step()
        for s in singles:
            s.step()
        for k, (b, s) in enumerate(zip(batch.sims, singles)):
            assert_same(b, s, "step %d member %d" % (step, k))
            iters.add(s.stats()["pcg_iters"])
    assert len(iters) > 1  # members converged after different iteration counts
    batch.close()


def test_sl_batch_members_match_single_steps():
    members = []
    for n, seed in ((64, 1), (40, 2), (96, 3)):
        kw = dict(k=0, dt=0.01, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0), tau=0.97, tol=1e-6,
                  max_iters=1000, static_cells=True, density=True)
        members.append((scenes.all_fluid(n), kw))
    batch = fs2d.Batch(members)
    singles = [fs2d.Sim(c, **kw) for c, kw in members]
    for k, (n, seed) in enumerate(((64, 1), (40, 2), (96, 3))):
        u, v, d = scenes.smoke_state(n, seed, sigma=2.0)
        for s in (batch.sims[k], singles[k]):
            s.set(fs2d.F_U, u); s.set(fs2d.F_V, v); s.set(fs2d.F_DENSITY, d)
    for step in range(3):
        batch.step()
        for s in singles:
            s.step()
        for k, (b, s) in enumerate(zip(batch.sims, singles)):
            for f in (fs2d.F_U, fs2d.F_V, fs2d.F_P, fs2d.F_DENSITY):
                np.testing.assert_array_equal(b.get(f), s.get(f), err_msg="step %d member %d field %d" % (step, k, f))
    batch.close()


def test_batch_stage_calls_and_member_edits():
    members = pic_members([40, 56], seed=5)
    batch = fs2d.Batch(members)
    singles = [fs2d.Sim(c, **kw) for c, kw in members]
    # a member's particles replaced through its own handle
    rng = np.random.default_rng(3)
    n = 40
    pos = np.stack([rng.uniform(2.0, n - 3.0, 300), rng.uniform(2.0, n / 2, 300)], axis=1) / n
    for s in (batch.sims[0], singles[0]):
        s.set_particles(pos)
    for st in (fs2d.STAGE_LEVEL_SET, fs2d.STAGE_P2G, fs2d.STAGE_GRAVITY, fs2d.STAGE_PROJECT,
               fs2d.STAGE_UPDATE_VELOCITY, fs2d.STAGE_G2P, fs2d.STAGE_ADVECT):
        batch.run_stage(st)
        for s in singles:
            s.run_stage(st)
    for k, (b, s) in enumerate(zip(batch.sims, singles)):
        assert_same(b, s, "stages member %d" % k)
    batch.step()
    for s in singles:
        s.step()
    for k, (b, s) in enumerate(zip(batch.sims, singles)):
        assert_same(b, s, "step member %d" % k)
    batch.close()


def test_batch_rejects_mixed_configs():
    a = (scenes.dam_break(32), dict(k=2, dt=0.002, dx=1 / 32, mode=fs2d.PICFLIP))
    b = (scenes.dam_break(32), dict(k=2, dt=0.002, dx=1 / 32, mode=fs2d.SEMILAGRANGIAN))
    with pytest.raises(fs2d.FS2DError) as e:
        fs2d.Batch([a, b])
    assert e.value.status == fs2d.ERR_CONFIG
    bad = scenes.dam_break(32)
    bad[0, 3] = fs2d.EMPTY  # open border
    with pytest.raises(fs2d.FS2DError) as e:
        fs2d.Batch([a, (bad, a[1])])
    assert e.value.status == fs2d.ERR_CONFIG and "member 1" in str(e.value)
