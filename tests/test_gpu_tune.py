"""Launch shapes that change no result (fs2d_set_tuning,
fs2d_batch_set_tuning) and the autotuner that picks them.

The PCG SpMV and axpy loop over a fixed list of virtual workgroups, each
writing its own partial sums, so the number of physical workgroups cannot
change a bit of the result; nor can whether the batched solves' loaders
prefetch.  Steps under every tuning must match bit for bit, and the autotuner
must record its choice and read it back.
"""
import numpy as np
import pytest

import fs2d
from fs2d import scenes

pytestmark = pytest.mark.gpu

FIELDS = (fs2d.F_U, fs2d.F_V, fs2d.F_P, fs2d.F_DENSITY, fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL, fs2d.F_CELL)


def smoke(n, seed):
    kw = dict(k=0, dt=0.01, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0), tau=0.97, tol=1e-6,
              max_iters=1000, static_cells=True, density=True)
    s = fs2d.Sim(scenes.all_fluid(n), **kw)
    u, v, d = scenes.smoke_state(n, seed, sigma=2.0)
    s.set(fs2d.F_U, u)
    s.set(fs2d.F_V, v)
    s.set(fs2d.F_DENSITY, d)
    return s


def test_single_tunings_bit_identical():
    n = 300  # pcg_blocks 256: tunings below it loop several virtual workgroups per workgroup
    ref = smoke(n, 7)
    for _ in range(3):
        ref.step()
    want = [ref.get(f) for f in FIELDS]
    it = ref.stats()["pcg_iters"]
    for g in (8, 64, 248):
        s = smoke(n, 7)
        s.set_tuning(pcg_grid=g)
        assert s.tuning()["pcg_grid"] == g
        for _ in range(3):
            s.step()
        for f, w in zip(FIELDS, want):
            np.testing.assert_array_equal(s.get(f), w, err_msg="pcg_grid %d field %d" % (g, f))
        assert s.stats()["pcg_iters"] == it
        s.close()


def pic_members(sides, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sides:
        w, h = rng.uniform(0.2, 0.7, size=2)
        kw = dict(k=2, dt=0.002, dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97, tol=1e-6, max_iters=500)
        out.append((scenes.dam_break(n, w, h), kw))
    return out


def test_batch_tunings_bit_identical():
    members = pic_members([64, 40, 96, 72], seed=3)
    ref = fs2d.Batch(members)
    for _ in range(3):
        ref.step()
    want = [[s.get(f) for f in FIELDS] for s in ref.sims]
    for g, pf in ((8, 0), (64, 1), (512, 0), (2048, 1)):
        b = fs2d.Batch(members)
        b.set_tuning(g, pf)
        assert b.tuning() == {"pcg_grid": g, "sweep_prefetch": pf}
        for _ in range(3):
            b.step()
        for k, s in enumerate(b.sims):
            for f, w in zip(FIELDS, want[k]):
                np.testing.assert_array_equal(s.get(f), w, err_msg="grid %d prefetch %d member %d field %d"
                                              % (g, pf, k, f))
        b.close()


def test_autotune_records_and_reuses(tmp_path):
    cache = str(tmp_path / "tune.txt")
    s = smoke(256, 8)
    s.step()
    first = s.autotune(iters=8, cache=cache)
    assert not first["from_cache"]
    assert first["pcg_grid"] % 8 == 0
    assert s.tuning()["pcg_grid"] == first["pcg_grid"]
    t = smoke(256, 9)
    t.step()
    again = t.autotune(iters=8, cache=cache)
    assert again["from_cache"] and again["pcg_grid"] == first["pcg_grid"]
    # the tuned handle keeps stepping like an untuned one
    u = smoke(256, 8)
    u.step()
    s.step()
    u.step()
    for f in FIELDS:
        np.testing.assert_array_equal(s.get(f), u.get(f))

    b = fs2d.Batch(pic_members([64, 48], seed=4))
    b.step()
    bt = b.autotune(iters=8, cache=cache)
    assert not bt["from_cache"] and bt["pcg_grid"] % 8 == 0
    assert b.tuning() == {"pcg_grid": bt["pcg_grid"], "sweep_prefetch": bt["sweep_prefetch"]}
    b2 = fs2d.Batch(pic_members([64, 48], seed=4))
    assert b2.autotune(iters=8, cache=cache)["from_cache"]
    print(open(cache).read())
