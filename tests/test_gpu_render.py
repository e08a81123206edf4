"""fs2d_render_buffers: the demo renderer's per-frame buffers
(FluidRenderer2D::updateBuffers, demo/FluidRenderer2D.cpp:435-486) written
into device buffers, checked against the same lists formed on the host from
the handle's fields:
  - FLUID / SOLID / non-zero-pressure cell positions: bit-exact, in the
    reference's row-major push order (FluidSim2D::iterate);
  - sigmoid(0.01 p) and sigmoid(100 phi): within 4 ulp of the FP32 logistic;
  - cell-centre and particle velocity lines: start points bit-exact; in a
    uniform velocity field (Catmull-Rom weights sum to 1) every end point is
    the start plus dt times that velocity, within FP32 rounding.
Device memory for the buffers comes from torch (plumbing only).
"""
import numpy as np
import pytest
import torch

import fs2d
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def logistic32(x):
    x = np.asarray(x, np.float32)
    return (np.float32(1.0) / (np.float32(1.0) + np.exp(-x))).astype(np.float32)


def render(sim, n_cells, n_lines, n_particles):
    dev = "cuda:0"
    bufs = {
        "water_cells": torch.full((n_cells, 2), -1.0, device=dev),
        "solid_cells": torch.full((n_cells, 2), -1.0, device=dev),
        "pressure_cells": torch.full((n_cells, 2), -1.0, device=dev),
        "pressure_values": torch.full((n_cells,), -1.0, device=dev),
        "cell_vel_lines": torch.full((n_lines, 4), -1.0, device=dev),
        "particle_lines": torch.full((max(n_particles, 1), 4), -1.0, device=dev),
        "phi_values": torch.full((n_cells,), -1.0, device=dev),
    }
    torch.cuda.synchronize()
    counts = sim.render_buffers(**{k: v.data_ptr() for k, v in bufs.items()})
    return counts, {k: v.cpu().numpy() for k, v in bufs.items()}


def host_cells(cells, dx, want):
    j, i = np.nonzero(want)
    return np.stack([(i * dx).astype(np.float32), (j * dx).astype(np.float32)], axis=1)


def test_render_buffers_pic_dam_break():
    n = 64
    dx = 1.0 / n
    sim = fs2d.Sim(O.dam_break(n, n), k=2, dt=0.005, dx=dx, mode=fs2d.PICFLIP, alpha=0.05, tau=0.97,
                   tol=1e-6, max_iters=500)
    for _ in range(4):
        sim.step()
    cells, p, phi = sim.get(fs2d.F_CELL), sim.get(fs2d.F_P), sim.get(fs2d.F_PHI, np.float64)
    pos = sim.get(fs2d.F_PARTICLE_POS)
    counts, b = render(sim, n * n, (n - 1) * (n - 1), len(pos))
    print("counts", counts)
    for name, want in (("water_cells", cells == O.FLUID), ("solid_cells", cells == O.SOLID),
                       ("pressure_cells", p != 0)):
        ref = host_cells(cells, dx, want)
        assert counts[name] == len(ref), name
        np.testing.assert_array_equal(b[name][:len(ref)], ref, err_msg=name)
        assert np.all(b[name][len(ref):] == -1.0), name  # nothing past the count
    assert counts["water_cells"] > 0 and counts["pressure_cells"] > 0
    pv = logistic32(np.float32(0.01) * p[p != 0])
    np.testing.assert_array_max_ulp(b["pressure_values"][:len(pv)], pv, maxulp=4)
    phv = logistic32((100.0 * phi).astype(np.float32)).ravel()
    np.testing.assert_array_max_ulp(b["phi_values"], phv, maxulp=4)
    assert counts["particles"] == len(pos)
    np.testing.assert_array_equal(b["particle_lines"][:, :2], pos)
    jj, ii = np.mgrid[0:n - 1, 0:n - 1]
    centres = np.stack([((ii + 0.5) * dx).astype(np.float32).ravel(), ((jj + 0.5) * dx).astype(np.float32).ravel()],
                       axis=1)
    assert counts["cell_vel_lines"] == (n - 1) * (n - 1)
    np.testing.assert_array_equal(b["cell_vel_lines"][:, :2], centres)
    sim.close()


def test_render_buffers_uniform_velocity():
    n = 48
    dx, dt = 1.0 / n, 0.01
    U, V = 0.3, -0.2
    sim = fs2d.Sim(O.dam_break(n, n), k=2, dt=dt, dx=dx, mode=fs2d.PICFLIP, alpha=0.05)
    sim.set(fs2d.F_U, np.full((n, n + 1), U, np.float32))
    sim.set(fs2d.F_V, np.full((n + 1, n), V, np.float32))
    pos = sim.get(fs2d.F_PARTICLE_POS)
    counts, b = render(sim, n * n, (n - 1) * (n - 1), len(pos))
    lines = b["cell_vel_lines"]
    np.testing.assert_allclose(lines[:, 2] - lines[:, 0], dt * U, rtol=0, atol=2e-7)
    np.testing.assert_allclose(lines[:, 3] - lines[:, 1], dt * V, rtol=0, atol=2e-7)
    pl = b["particle_lines"]
    np.testing.assert_allclose(pl[:, 2] - pl[:, 0], dt * U, rtol=0, atol=2e-7)
    np.testing.assert_allclose(pl[:, 3] - pl[:, 1], dt * V, rtol=0, atol=2e-7)
    # only the phi buffer: the cell lists are skipped and report no counts
    t = torch.full((n * n,), -1.0, device="cuda:0")
    torch.cuda.synchronize()
    c2 = sim.render_buffers(phi_values=t.data_ptr())
    assert c2["water_cells"] == 0 and c2["phi_values"] == n * n
    assert torch.all(t >= 0).item()
    sim.close()
