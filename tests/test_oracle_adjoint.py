"""The oracle's FP64 adjoint of the all-fluid semi-Lagrangian step, pinned
against central finite differences of the oracle's own forward step (the
check that fixes it independently of its formulas), on the 32 x 32 copy of
the differentiable-rollout workload: smoothed random velocity (seed 3, sigma
scaled to 1 cell), 16 density blobs, dt = 0.01, 16 random directions,
eps = 1e-4.  The forward solve runs to 1e-14 so that solver noise stays far
below the finite-difference step.
"""
import numpy as np
import pytest

from oracle import oracle as O
from fs2d import scenes

N = 32
EPS = 1e-4


def make(n=N, dt=0.01, gravity=(0.0, -9.81), tol=1e-14):
    o = O.Oracle(scenes.all_fluid(n), k=0, dt=dt, dx=1.0 / n, mode=O.MODE_SL, gravity=gravity,
                 level_set=False, tau=0.97, tol=tol, max_iters=4000, sl_inplace=False, density=True)
    u, v, d = scenes.smoke_state(n, 3, sigma=8.0 * n / 256)
    o.set(O.F_U, u); o.set(O.F_V, v); o.set(O.F_DENS, d)
    o.run_stage(O.ST_PROJECT)
    o.run_stage(O.ST_UPDATEVEL)
    return o


def state(o):
    return [o.get(O.F_U), o.get(O.F_V), o.get(O.F_DENS)]


def put(o, x):
    o.set(O.F_U, x[0]); o.set(O.F_V, x[1]); o.set(O.F_DENS, x[2])


def rollout(o, x, steps):
    put(o, x)
    for _ in range(steps):
        o.step()
    return state(o)


def reverse(o, x0, steps, seeds):
    states = []
    put(o, x0)
    for _ in range(steps):
        states.append(state(o))
        o.step()
    g = seeds
    for t in reversed(range(steps)):
        put(o, states[t])
        g = o.adjoint_step(*g)
    return g


@pytest.mark.parametrize("steps", [1, 64])
def test_oracle_adjoint_matches_finite_differences(steps):
    o = make()
    x0 = state(o)
    target = scenes.target_density(N).astype(np.float64)

    def loss(x):
        d = rollout(o, x, steps)[2]
        return float(((d - target) ** 2).sum())

    dT = rollout(o, x0, steps)[2]
    seeds = [np.zeros_like(x0[0]), np.zeros_like(x0[1]), 2.0 * (dT - target)]
    grad = reverse(o, x0, steps, seeds)
    rng = np.random.default_rng(7)
    errs = []
    for k in range(16):
        dirs = [rng.standard_normal(a.shape) for a in x0]
        dirs[2] *= 0.0 if k < 12 else 1.0   # velocity is the optimisation variable; 4 directions add density
        fd = (loss([a + EPS * b for a, b in zip(x0, dirs)]) - loss([a - EPS * b for a, b in zip(x0, dirs)])) / (2 * EPS)
        ad = sum(float((g * b).sum()) for g, b in zip(grad, dirs))
        errs.append(abs(fd - ad) / max(abs(fd), 1e-12))
    assert max(errs) < 1e-4, errs


def test_oracle_adjoint_is_the_transpose_of_the_linear_stages():
    # At zero velocity advection is the identity away from the clamps and the
    # step is linear in a perturbation of the velocity through the projection:
    # <J x, y> = <x, J^T y> for the step's Jacobian J (checked by directional
    # derivatives, one random pair).
    o = make(gravity=(0.0, 0.0))
    x0 = [np.zeros_like(a) for a in state(o)[:2]] + [state(o)[2]]
    rng = np.random.default_rng(11)
    y = [rng.standard_normal(a.shape) for a in x0]
    x = [rng.standard_normal(a.shape) for a in x0]
    x[2] *= 0
    jt_y = reverse(o, x0, 1, y)
    plus = rollout(o, [a + EPS * b for a, b in zip(x0, x)], 1)
    minus = rollout(o, [a - EPS * b for a, b in zip(x0, x)], 1)
    jx = [(p - m) / (2 * EPS) for p, m in zip(plus, minus)]
    lhs = sum(float((a * b).sum()) for a, b in zip(jx, y))
    rhs = sum(float((a * b).sum()) for a, b in zip(x, jt_y))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)


def test_oracle_adjoint_rejects_out_of_scope_states():
    o = O.Oracle(O.dam_break(16, 16), k=2, dt=0.01, dx=1.0 / 16, mode=O.MODE_SL)
    z = [np.zeros((16, 17)), np.zeros((17, 16)), np.zeros((16, 16))]
    with pytest.raises(ValueError):
        o.adjoint_step(*z)
