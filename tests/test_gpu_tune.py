"""Launch shapes that change no result (fs2d_set_tuning,
fs2d_batch_set_tuning) and the autotuner that picks them.

The PCG SpMV and axpy loop over a fixed list of virtual workgroups, each
writing its own partial sums, so the number of physical workgroups cannot
change a bit of the result; nor can whether the batched solves' loaders
prefetch, nor whether a single handle runs the SpMV inside the backward
solve's launch (fused_pcg: the same virtual workgroups, started as the
solve's strips finish, with sigma summed in the solve's loaders' order).
Steps under every tuning must match bit for bit, and the autotuner must
record its choice and read it back.
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
    assert ref.tuning()["fused_pcg"] == 0  # the default below 2^18 solver entries
    for _ in range(3):
        ref.step()
    want = [ref.get(f) for f in FIELDS]
    it = ref.stats()["pcg_iters"]
    for g, fu in ((8, 1), (64, 0), (248, 1), (0, 1)):
        s = smoke(n, 7)
        s.set_tuning(pcg_grid=g, fused_pcg=fu)
        assert s.tuning()["pcg_grid"] == g and s.tuning()["fused_pcg"] == fu
        for _ in range(3):
            s.step()
        for f, w in zip(FIELDS, want):
            np.testing.assert_array_equal(s.get(f), w, err_msg="pcg_grid %d fused %d field %d" % (g, fu, f))
        assert s.stats()["pcg_iters"] == it
        s.close()


def box(nx, ny, seed):
    """All-fluid nx x ny box with a smoothed random velocity (static cells)."""
    from scipy.ndimage import gaussian_filter
    c = scenes.walls(nx, ny)
    c[1:-1, 1:-1] = scenes.FLUID
    kw = dict(k=0, dt=0.01, dx=1.0 / max(nx, ny), mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0), tau=0.97,
              tol=1e-6, max_iters=150, static_cells=True, density=True)
    s = fs2d.Sim(c, **kw)
    rng = np.random.default_rng(seed)
    s.set(fs2d.F_U, gaussian_filter(rng.standard_normal((ny, nx + 1)), 4.0, mode="wrap").astype(np.float32))
    s.set(fs2d.F_V, gaussian_filter(rng.standard_normal((ny + 1, nx)), 4.0, mode="wrap").astype(np.float32))
    s.set(fs2d.F_DENSITY, np.ones((ny, nx), np.float32))
    return s


@pytest.mark.parametrize("nx,ny", [(64, 64), (1000, 130), (300, 300), (200, 700), (64, 16000), (16384, 64),
                                   (16000, 70), (5, 5), (129, 65)])
def test_fused_pcg_bit_identical_across_strip_counts(nx, ny):
    """1, 3, 5, 11 and 250 strips; the 250-strip grid launches 125 + 250 + 8
    workgroups of one CU each, more than the device holds at once, so
    workgroups wait on strips that started before them across rounds of
    residency.  A 16384-wide single strip puts every SpMV run and every p/r
    update run on one strip; 5 x 5 is the smallest launch (8 runs)."""
    a, b = box(nx, ny, 5), box(nx, ny, 5)
    a.set_tuning(pcg_grid=0, fused_pcg=1)
    b.set_tuning(pcg_grid=0, fused_pcg=0)
    for step in range(2):
        a.step()
        b.step()
        for f in FIELDS[:4]:
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg="%dx%d step %d field %d" % (nx, ny, step, f))
        assert a.stats()["pcg_iters"] == b.stats()["pcg_iters"]
    print("%dx%d: %d strips, iterations %d" % (nx, ny, (ny + 63) // 64, a.stats()["pcg_iters"]))
    a.close()
    b.close()


def test_fused_pcg_bit_identical_2048():
    """The benchmark's grid size, 8 steps (about 4000 PCG iterations, each a
    fused forward and a fused backward launch whose SpMV workgroups read z
    written through by strips on other XCDs within the launch): every field
    equals the separate launches' after every step."""
    n = 2048
    a, b = smoke(n, 11), smoke(n, 11)
    assert a.tuning()["fused_pcg"] == 1  # the default at this size
    b.set_tuning(pcg_grid=0, fused_pcg=0)
    total = 0
    for step in range(8):
        a.step()
        b.step()
        for f in FIELDS[:4]:
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg="step %d field %d" % (step, f))
        assert a.stats()["pcg_iters"] == b.stats()["pcg_iters"]
        total += a.stats()["pcg_iters"]
    print("2048^2: 8 steps, %d PCG iterations, bit-identical" % total)
    a.close()
    b.close()


def test_fused_pcg_pic_and_captured_steps():
    """A PIC/FLIP dam break (EMPTY cells, a changing matrix) stepped with the
    fused launch, uncaptured and from captured graphs, against separate
    launches."""
    n = 160
    kw = dict(k=2, dt=0.002, dx=1.0 / n, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97, tol=1e-6, max_iters=500)
    c = scenes.dam_break(n)
    sims = [fs2d.Sim(c, **kw) for _ in range(3)]
    for s in sims[:2]:
        s.set_tuning(pcg_grid=0, fused_pcg=1)
    sims[1].set_graph(True)
    sims[2].set_tuning(pcg_grid=0, fused_pcg=0)
    for step in range(4):
        for s in sims:
            s.step()
        for f in FIELDS[:3] + (fs2d.F_PARTICLE_POS, fs2d.F_PARTICLE_VEL, fs2d.F_CELL):
            want = sims[2].get(f)
            for s in sims[:2]:
                np.testing.assert_array_equal(s.get(f), want, err_msg="step %d field %d" % (step, f))
    for s in sims:
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
    assert ref.tuning()["fused_pcg"] == 0  # the batch's default
    for g, pf, fu in ((8, 0, 1), (64, 1, 0), (512, 0, 0), (2048, 1, 1)):
        b = fs2d.Batch(members)
        b.set_tuning(g, pf, fu)
        assert b.tuning() == {"pcg_grid": g, "sweep_prefetch": pf, "fused_pcg": fu}
        for _ in range(3):
            b.step()
        for k, s in enumerate(b.sims):
            for f, w in zip(FIELDS, want[k]):
                np.testing.assert_array_equal(s.get(f), w, err_msg="grid %d prefetch %d fused %d member %d field %d"
                                              % (g, pf, fu, k, f))
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
    assert b.tuning() == {"pcg_grid": bt["pcg_grid"], "sweep_prefetch": bt["sweep_prefetch"],
                          "fused_pcg": bt["fused_pcg"]}
    b2 = fs2d.Batch(pic_members([64, 48], seed=4))
    assert b2.autotune(iters=8, cache=cache)["from_cache"]
    print(open(cache).read())
