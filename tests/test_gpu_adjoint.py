"""GPU adjoint (fs2d_adjoint_* of the C ABI) against the FP64 oracle adjoint
and FP64 central finite differences.

Workload: the differentiable rollout, all-fluid semi-Lagrangian smoke with a
passive density, no gravity (in a closed all-fluid box gravity is a pure
gradient that the projection removes; with it the FP32 solve would have to
resolve a hydrostatic pressure of order rho g H, beyond FP32's attainable
1e-6 residual), dt = 0.01, loss = sum((d_T - target)^2) with the target
rendered from 8 blobs (seed 4), checkpoint every 8 steps.  Bars: one adjoint
step within 1e-4 relative L2 of the oracle's; the 64-step gradient within
1e-3 of FP64 finite differences along 16 random directions and of the
oracle's reverse pass; gradients bit-identical across checkpoint intervals
and repeated runs (the fixed-point scatter is order-independent), and at
512^2 bit-identical between the two-launch and four-launch PCG iterations.
"""
import numpy as np
import pytest

import fs2d
from fs2d import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def gpu_sim(n, dt=0.01, tol=1e-6):
    g = fs2d.Sim(scenes.all_fluid(n), k=0, dt=dt, dx=1.0 / n, mode=fs2d.SEMILAGRANGIAN, gravity=(0.0, 0.0),
                 tau=0.97, tol=tol, max_iters=4000, static_cells=True, density=True)
    u, v, d = scenes.smoke_state(n, 3, sigma=8.0 * n / 256)
    g.set(fs2d.F_U, u); g.set(fs2d.F_V, v); g.set(fs2d.F_DENSITY, d)
    g.run_stage(fs2d.STAGE_PROJECT)
    g.run_stage(fs2d.STAGE_UPDATE_VELOCITY)
    return g


def oracle_for(n, dt=0.01, tol=1e-14):
    return O.Oracle(scenes.all_fluid(n), k=0, dt=dt, dx=1.0 / n, mode=O.MODE_SL, gravity=(0.0, 0.0),
                    level_set=False, tau=0.97, tol=tol, max_iters=4000, sl_inplace=False, density=True)


def gstate(g):
    return [g.get(f) for f in (fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY)]


def oput(o, x):
    o.set(O.F_U, np.asarray(x[0], np.float64)); o.set(O.F_V, np.asarray(x[1], np.float64))
    o.set(O.F_DENS, np.asarray(x[2], np.float64))


def test_adjoint_step_matches_oracle():
    n = 64
    g = gpu_sim(n)
    for _ in range(3):
        g.step()
    x = gstate(g)
    o = oracle_for(n)
    oput(o, x)
    rng = np.random.default_rng(5)
    seeds = [rng.standard_normal(a.shape).astype(np.float32) for a in x]
    adj = fs2d.Adjoint(g, checkpoint_every=1, max_steps=1)
    for f, s in zip((fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY), seeds):
        adj.set(f, s)
    adj.step()
    got = [adj.get(f) for f in (fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY)]
    want = o.adjoint_step(*[s.astype(np.float64) for s in seeds])
    errs = [rel_l2(a, b) for a, b in zip(got, want)]
    assert max(errs) < 1e-4, errs
    # the state is unchanged by the adjoint step
    for a, b in zip(gstate(g), x):
        np.testing.assert_array_equal(a, b)
    adj.close()


def _gpu_gradient(n, steps, every, target, fused_pcg=None):
    g = gpu_sim(n)
    if fused_pcg is not None:
        g.set_tuning(fused_pcg=fused_pcg)
    x0 = gstate(g)
    adj = fs2d.Adjoint(g, checkpoint_every=every, max_steps=steps)
    adj.forward(steps)
    dT = g.get(fs2d.F_DENSITY)
    adj.set(fs2d.F_U, np.zeros_like(x0[0]))
    adj.set(fs2d.F_V, np.zeros_like(x0[1]))
    adj.set(fs2d.F_DENSITY, 2.0 * (dT - target))
    adj.backward()
    grad = [adj.get(f) for f in (fs2d.F_U, fs2d.F_V, fs2d.F_DENSITY)]
    final = gstate(g)
    adj.close()
    g.close()
    return x0, dT, grad, final


def test_adjoint_rollout_matches_fp64_finite_differences():
    n, steps = 32, 64
    target = scenes.target_density(n)
    x0, dT, grad, _ = _gpu_gradient(n, steps, 8, target)
    o = oracle_for(n)
    t64 = target.astype(np.float64)

    def loss(x):
        oput(o, x)
        for _ in range(steps):
            o.step()
        return float(((o.get(O.F_DENS) - t64) ** 2).sum())

    x64 = [a.astype(np.float64) for a in x0]
    rng = np.random.default_rng(7)
    eps = 1e-4
    errs = []
    for _ in range(16):
        dirs = [rng.standard_normal(a.shape) for a in x64[:2]]
        fd = (loss([x64[0] + eps * dirs[0], x64[1] + eps * dirs[1], x64[2]]) -
              loss([x64[0] - eps * dirs[0], x64[1] - eps * dirs[1], x64[2]])) / (2 * eps)
        ad = float((grad[0] * dirs[0]).sum() + (grad[1] * dirs[1]).sum())
        errs.append(abs(fd - ad) / abs(fd))
    assert max(errs) < 1e-3, errs

    # the FP64 oracle's own reverse pass
    states = []
    oput(o, x64)
    for _ in range(steps):
        states.append([o.get(O.F_U), o.get(O.F_V), o.get(O.F_DENS)])
        o.step()
    gb = [np.zeros_like(x64[0]), np.zeros_like(x64[1]), 2.0 * (o.get(O.F_DENS) - t64)]
    for t in reversed(range(steps)):
        oput(o, states[t])
        gb = o.adjoint_step(*gb)
    errs = [rel_l2(a, b) for a, b in zip(grad, gb)]
    assert max(errs) < 1e-3, errs


def test_adjoint_checkpointing_and_repeats_are_bit_identical():
    n, steps = 48, 24
    target = scenes.target_density(n)
    ref = _gpu_gradient(n, steps, 1, target)
    for every in (5, 8, 24):
        got = _gpu_gradient(n, steps, every, target)
        for a, b in zip(got[2], ref[2]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(got[3], ref[3]):  # the handle ends at the final state
            np.testing.assert_array_equal(a, b)
    again = _gpu_gradient(n, steps, 8, target)
    for a, b in zip(again[2], ref[2]):
        np.testing.assert_array_equal(a, b)


def test_adjoint_two_launch_pcg_is_bit_identical():
    """At 512^2 the forward and the adjoint Poisson solves take the two-launch
    PCG iteration by default; the rollout, its final state and the gradient
    are bit-identical to the four-launch iteration's."""
    n, steps = 512, 4
    target = scenes.target_density(n)
    g = gpu_sim(n)
    assert g.tuning()["fused_pcg"] == 1
    g.close()
    two = _gpu_gradient(n, steps, 2, target, fused_pcg=1)
    four = _gpu_gradient(n, steps, 2, target, fused_pcg=0)
    for a, b in zip([two[1]] + list(two[2]) + list(two[3]), [four[1]] + list(four[2]) + list(four[3])):
        np.testing.assert_array_equal(a, b)
    assert np.abs(two[2][2]).max() > 0


def test_autograd_rollout_matches_the_c_abi():
    # The rollout runs on the GPU through the C ABI; the harness's tensors
    # live in host memory.  (torch ships its own HIP runtime under the same
    # soname as /opt/rocm's, so torch GPU tensors need torch loaded first;
    # this process has loaded the library already.)
    import torch
    from fs2d.autograd import rollout
    n, steps = 32, 16
    target = scenes.target_density(n)
    x0, dT, grad, _ = _gpu_gradient(n, steps, 4, target)
    g = gpu_sim(n)
    adj = fs2d.Adjoint(g, checkpoint_every=4, max_steps=steps)
    dev = torch.device("cpu")
    u0, v0, d0 = (torch.tensor(a, device=dev, requires_grad=True) for a in x0)
    uT, vT, dT_t = rollout(adj, u0, v0, d0, steps)
    np.testing.assert_array_equal(dT_t.detach().cpu().numpy(), dT)
    loss = ((dT_t - torch.tensor(target, device=dev)) ** 2).sum()
    loss.backward()
    for t, want in zip((u0, v0, d0), grad):
        got = t.grad.cpu().numpy()
        assert rel_l2(got, want) < 1e-6  # the seed 2(d_T - target) is formed in torch here
    adj.close()


def test_adjoint_rejects_out_of_scope_handles():
    g = fs2d.Sim(O.dam_break(32, 32), k=2, dt=0.005, dx=1.0 / 32, mode=fs2d.PICFLIP)
    with pytest.raises(fs2d.FS2DError) as e:
        fs2d.Adjoint(g)
    assert e.value.status == fs2d.ERR_CONFIG
    g2 = gpu_sim(16)
    with pytest.raises(fs2d.FS2DError):
        fs2d.Adjoint(g2, checkpoint_every=0)
    adj = fs2d.Adjoint(g2, checkpoint_every=2, max_steps=4)
    with pytest.raises(fs2d.FS2DError):
        adj.backward()          # no forward yet
    with pytest.raises(fs2d.FS2DError):
        adj.forward(5)          # beyond max_steps
    adj.close()
