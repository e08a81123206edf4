"""Checks that pin the oracle's functions by what mathematics fixes, for the parts
that the reference-held golden rows leave open (PIC transfers, extrapolation
order, the MIC(0) factor and triangular solves, the Catmull-Rom weights).

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand) fails it.
"""
import ctypes
import ctypes.util

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------------------
# glibc rand(): the oracle's private generator equals the C library's stream.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 7, 12345])
def test_rand_stream_is_glibc(seed):
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(ctypes.c_uint(seed))
    ref = np.array([libc.rand() for _ in range(1000)], dtype=np.uint32)
    np.testing.assert_array_equal(O.rand_stream(seed, 1000), ref)


# ---------------------------------------------------------------------------
# Catmull-Rom (Array2D::triCubic): interpolatory at nodes and exact on
# quadratics away from clamped taps (Keys' a = -1/2 kernel reproduces degree 2).
# ---------------------------------------------------------------------------
def test_tricubic_nodes_and_quadratics():
    rng = np.random.default_rng(0)
    ny, nx = 9, 11
    data = rng.standard_normal((ny, nx))
    for j in range(1, ny - 2):
        for i in range(1, nx - 2):
            assert O.tricubic(data, i, j) == pytest.approx(data[j, i], abs=1e-14)
    c = rng.standard_normal(6)
    J, I = np.mgrid[0:ny, 0:nx].astype(float)
    q = c[0] + c[1] * I + c[2] * J + c[3] * I * I + c[4] * I * J + c[5] * J * J
    for _ in range(200):
        x = rng.uniform(1.0, nx - 2.0 - 1e-9)
        y = rng.uniform(1.0, ny - 2.0 - 1e-9)
        want = c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y
        assert O.tricubic(q, x, y) == pytest.approx(want, rel=1e-12, abs=1e-12)
    # outside the array the sampler returns 0 (Array2D.h:245-248); the index is
    # truncated, not floored, so -0.5 still lands in column 0
    assert O.tricubic(data, -1.5, 2.0) == 0.0
    assert O.tricubic(data, -0.5, 2.0) != 0.0
    assert O.tricubic(data, 3.0, ny + 0.5) == 0.0


# ---------------------------------------------------------------------------
# Extrapolation (Array2D::extrapolate): the FIFO BFS equals a layer-synchronous
# restatement (every entry of layer L is the mean of its neighbours of layer < L).
# ---------------------------------------------------------------------------
def layered_extrapolate(data, mask):
    data = data.copy()
    ny, nx = data.shape
    layer = np.where(mask == 0, 0, -1)
    nb = [(-1, 0), (1, 0), (0, -1), (0, 1)]  # same neighbour order
    L = 0
    while True:
        front = []
        for j in range(ny):
            for i in range(nx):
                if layer[j, i] == -1 and any(
                        0 <= i + di < nx and 0 <= j + dj < ny and layer[j + dj, i + di] == L
                        for di, dj in nb):
                    front.append((i, j))
        if not front:
            break
        L += 1
        for i, j in front:
            layer[j, i] = L
        for i, j in front:
            s, c = 0.0, 0
            for di, dj in nb:
                a, b = i + di, j + dj
                if 0 <= a < nx and 0 <= b < ny and 0 <= layer[b, a] < L:
                    s += data[b, a]
                    c += 1
            data[j, i] = 0 if c == 0 else s / c
    return data, layer


@pytest.mark.parametrize("seed", range(6))
def test_extrapolate_equals_layered_bfs(seed):
    rng = np.random.default_rng(seed)
    ny, nx = rng.integers(3, 14, size=2)
    data = rng.standard_normal((ny, nx))
    known = rng.random((ny, nx)) < rng.uniform(0.02, 0.5)
    mask = np.where(known, 0, np.uint32(0xFFFFFFFF)).astype(np.uint32)
    got, gmask = O.extrapolate(data, mask)
    want, layer = layered_extrapolate(data, mask)
    np.testing.assert_array_equal(got, want)
    reach = layer >= 0
    np.testing.assert_array_equal(gmask[reach], layer[reach].astype(np.uint32))


# ---------------------------------------------------------------------------
# Seeding: k^2 particles per FLUID cell, each in its own sub-cell.
# ---------------------------------------------------------------------------
def test_seeding_stratified():
    cells = O.dam_break(16, 16)
    o = O.Oracle(cells, k=3, dx=0.1)
    pos = o.get(O.F_POS)
    fl = np.argwhere(cells == O.FLUID)  # (j, i), row-major order like create()
    assert pos.shape[0] == 9 * len(fl)
    g = pos / 0.1
    for n, (j, i) in enumerate(fl):
        for e in range(9):
            x, y = g[9 * n + e]
            sx, sy = e % 3, e // 3
            assert i + sx / 3 - 1e-6 <= x < i + (sx + 1) / 3 + 1e-6
            assert j + sy / 3 - 1e-6 <= y < j + (sy + 1) / 3 + 1e-6


# ---------------------------------------------------------------------------
# Level set (LevelSet::constructFromParticles): one particle => every corner
# adopts it, phi = |corner - particle| - dr (closed form).  Two particles =>
# each corner adopts the nearer one.
# ---------------------------------------------------------------------------
def test_levelset_single_and_two_particles():
    n, dx = 12, 0.1
    o = O.Oracle(O.walls(n, n), k=2, dx=dx)
    for pts in ([(0.43, 0.61)], [(0.33, 0.31), (0.87, 0.92)]):
        P = np.array(pts)
        o.set(O.F_POS, P)
        o.set(O.F_VEL, np.zeros_like(P))
        o.level_set_parts(1)
        phi = o.get(O.F_PHI)
        J, I = np.mgrid[0:n, 0:n] * dx
        d = np.stack([np.sqrt((px - I) ** 2 + (py - J) ** 2) for px, py in P]).min(0) - 0.9 * dx
        np.testing.assert_allclose(phi, d, rtol=0, atol=1e-14)


def test_redistance_signs_and_eikonal_bound():
    """Redistancing keeps the sign pattern of the particle level set and the
    interior values satisfy |phi(c) - phi(n)| <= dx between 4-neighbours that
    are both inside (a discrete 1-Lipschitz bound of the eikonal update)."""
    cells = O.dam_break(32, 32)
    o = O.Oracle(cells, k=2, dx=1.0 / 32)
    o.level_set_parts(1)
    phi1 = o.get(O.F_PHI)
    o.level_set_parts(2)
    phi2 = o.get(O.F_PHI)
    inside1 = phi1 < 0
    # smoothing only lowers values, sweeps only touch negative cells
    assert np.all(phi2[~inside1] <= phi1[~inside1] + 1e-15)
    assert np.all(phi2[inside1] < 0)
    dx = 1.0 / 32
    core = np.zeros_like(inside1)
    core[3:-3, 3:-3] = True
    both = (phi2[:, 1:] < -2 * dx) & (phi2[:, :-1] < -2 * dx) & core[:, 1:] & core[:, :-1]
    assert np.all(np.abs(np.diff(phi2, axis=1))[both] <= dx * (1 + 1e-12))


# ---------------------------------------------------------------------------
# Projection: MIC(0) factor and triangular solves against dense algebra.
# M = Lt Lt^T with Lt = diag(1/P) + A_lower diag(P)  (FluidSim2D.cpp:364-421).
# ---------------------------------------------------------------------------
def dense_system(o):
    Ad, Ax, Ay, P = (o.get(f) for f in (O.F_ADIAG, O.F_AX, O.F_AY, O.F_PRECON))
    cell = o.get(O.F_CELL)
    ny, nx = Ad.shape
    idx = -np.ones((ny, nx), dtype=int)
    fl = np.argwhere(cell == O.FLUID)
    for n, (j, i) in enumerate(fl):
        idx[j, i] = n
    m = len(fl)
    A = np.zeros((m, m))
    Lt = np.zeros((m, m))
    for n, (j, i) in enumerate(fl):
        A[n, n] = Ad[j, i]
        Lt[n, n] = 1.0 / P[j, i]
        for (a, b, coef) in ((i + 1, j, Ax[j, i]), (i, j + 1, Ay[j, i])):
            k = idx[b, a]
            if k >= 0:
                A[n, k] = A[k, n] = coef
                Lt[k, n] = coef * P[j, i]   # lower-triangular entry of row k
    return A, Lt, fl


def pool_scene(n=10, seed=3):
    """Dam break with some EMPTY cells so that ghost-fluid terms appear."""
    cells = O.dam_break(n, n, 0.7, 0.6)
    o = O.Oracle(cells, k=2, dx=1.0 / n, dt=0.01, mode=O.MODE_SL)
    o.run_stage(O.ST_LEVELSET)
    rng = np.random.default_rng(seed)
    o.set(O.F_U, rng.standard_normal((n, n + 1)))
    o.set(O.F_V, rng.standard_normal((n + 1, n)))
    return o


@pytest.mark.parametrize("tau", [0.0, 0.999, 1.0])
def test_mic0_factor_properties(tau):
    o = pool_scene()
    o.set_options(tau=tau, sigma=0.0)   # sigma=0: safeguard off for the algebra check
    o.assemble_factor()
    A, Lt, fl = dense_system(o)
    M = Lt @ Lt.T
    assert np.allclose(A, A.T)
    nz = A != 0
    if tau == 0.0:      # IC(0): M equals A on A's sparsity pattern
        np.testing.assert_allclose(M[nz], A[nz], rtol=1e-12, atol=1e-9)
    if tau == 1.0:      # MIC(0): row sums of M equal row sums of A
        np.testing.assert_allclose(M.sum(1), A.sum(1), rtol=1e-10, atol=1e-9 * np.abs(A).max())
    # the preconditioner application solves M z = r
    rng = np.random.default_rng(1)
    r2 = np.zeros((o.ny, o.nx))
    rv = rng.standard_normal(len(fl))
    r2[fl[:, 0], fl[:, 1]] = rv
    z = o.apply_precon(r2)
    np.testing.assert_allclose(z[fl[:, 0], fl[:, 1]], np.linalg.solve(M, rv), rtol=1e-9, atol=1e-12)


def test_ghost_fluid_diagonal_closed_form():
    """A FLUID cell whose right neighbour is EMPTY gets scaleA*(1 - max(phi_r/phi_c, -1e3))
    on the diagonal; the left-neighbour EMPTY term is -scaleA*max(phi_l/phi_c, -1e3)
    (FluidSim2D.cpp:270-283)."""
    n = 8
    cells = O.walls(n, n)
    cells[1:7, 1:4] = O.FLUID
    o = O.Oracle(cells, k=1, dx=0.1, dt=0.02, rho=1000.0, level_set=False)
    phi = np.where(cells == O.FLUID, -0.25, 0.75)
    o.set(O.F_PHI, phi)
    o.assemble_factor()
    Ad = o.get(O.F_ADIAG)
    sA = 0.02 / (1000.0 * 0.01)
    # cell (3, 3): left FLUID, right EMPTY, below FLUID, above FLUID
    want = sA * 3 + sA * (1 - max(0.75 / -0.25, -1e3))
    assert Ad[3, 3] == pytest.approx(want, rel=1e-14)
    # a FLUID cell next to a SOLID wall: the wall contributes nothing
    assert Ad[3, 1] == pytest.approx(sA * 2 + sA * (1 if cells[3, 2] == O.FLUID else 0), rel=1e-14)


def test_projection_all_fluid_box_divergence_free():
    """All-fluid box inside walls: Neumann Poisson, A*1 = 0; after projection the
    discrete divergence of every FLUID cell vanishes to the solver tolerance."""
    n = 24
    cells = O.all_fluid(n, n)
    o = O.Oracle(cells, k=0, dx=1.0 / n, dt=0.01, mode=O.MODE_SL, level_set=False,
                 gravity=(0.0, 0.0), tol=1e-12, max_iters=1000)
    rng = np.random.default_rng(5)
    o.set(O.F_U, rng.standard_normal((n, n + 1)))
    o.set(O.F_V, rng.standard_normal((n + 1, n)))
    o.run_stage(O.ST_PROJECT)
    A, _, fl = dense_system(o)
    np.testing.assert_allclose(A.sum(1), 0.0, atol=1e-12 * np.abs(A).max())
    o.run_stage(O.ST_UPDATEVEL)
    u, v = o.get(O.F_U), o.get(O.F_V)
    div = (u[:, 1:] - u[:, :-1]) + (v[1:, :] - v[:-1, :])
    assert np.abs(div[cells == O.FLUID]).max() < 1e-8
    # walls are impermeable
    assert np.all(u[1:-1, 1] == 0) and np.all(v[1, 1:-1] == 0)


def test_projection_hydrostatic_pool():
    """Still water under gravity: after one projection the velocity in the
    pool's interior is zero (the pressure gradient balances dt*g)."""
    n = 24
    cells = O.walls(n, n)
    cells[1:12, 1:-1] = O.FLUID
    # static level set (flat surface half-way between cell rows 11 and 12)
    o = O.Oracle(cells, k=2, dx=1.0 / n, dt=0.01, mode=O.MODE_SL, max_iters=1000, level_set=False)
    o.run_stage(O.ST_GRAVITY)
    o.run_stage(O.ST_PROJECT)
    o.run_stage(O.ST_UPDATEVEL)
    v = o.get(O.F_V)
    cell = o.get(O.F_CELL)
    inner = np.zeros_like(cell, dtype=bool)
    inner[2:8, 2:-2] = True
    assert np.abs(v[:-1][inner & (cell == O.FLUID)]).max() < 1e-9


# ---------------------------------------------------------------------------
# Transfers: P2G partition of unity, G2P exactness on bilinear fields, PIC/FLIP
# limits (FluidSim2D.cpp:144-204, 552-568).
# ---------------------------------------------------------------------------
def test_p2g_uniform_velocity_gives_uniform_field():
    n = 16
    cells = O.dam_break(n, n)
    o = O.Oracle(cells, k=2, dx=1.0 / n, mode=O.MODE_PIC)
    m = o.stats()["n_particles"]
    o.set(O.F_VEL, np.tile([0.7, -0.3], (m, 1)))
    o.run_stage(O.ST_P2G)
    np.testing.assert_allclose(o.get(O.F_U), 0.7, rtol=1e-14)   # extrapolated everywhere
    np.testing.assert_allclose(o.get(O.F_V), -0.3, rtol=1e-14)


def tent_p2g(pts, vals, shape, offset):
    """P2G by its definition: face (i, j) sits at grid position (i, j) + offset
    (u faces at (i, j + 1/2), v faces at (i + 1/2, j)); a particle at grid
    position (X, Y) gives it the weight max(0, 1 - |X - i - ox|) *
    max(0, 1 - |Y - j - oy|).  Returns the weight sums and the normalised
    values (0 where no weight arrives)."""
    ny, nx = shape
    J, I = np.mgrid[0:ny, 0:nx].astype(np.float64)
    W = np.zeros(shape)
    S = np.zeros(shape)
    for (X, Y), val in zip(pts, vals):
        w = (np.maximum(0.0, 1.0 - np.abs(X - I - offset[0])) *
             np.maximum(0.0, 1.0 - np.abs(Y - J - offset[1])))
        W += w
        S += w * val
    out = np.where(W > 0, S / np.where(W > 0, W, 1.0), 0.0)
    return W, out


def test_p2g_tent_weights_closed_form():
    """transferVelocityToGrid (FluidSim2D.cpp:144-204, Array2D.h:361-376) on a
    few particles, against the tent-kernel definition and the layered BFS
    (pinned above) for the faces left unknown.  The particles are placed so
    that a half-cell shift of either face offset, or swapped tent weights
    (disp <-> 1 - disp), changes the known faces or their values:
      a, b : overlapping stencils with different velocities (normalised mix);
      c    : vx = 0 exactly, so its u faces hold 0.0 and are marked unknown
             (the reference's == 0.0 rule, :192-201), while its v faces are known;
      d    : on a u node (X = 6, Y - 1/2 = 7): one u face gets weight 1, its
             neighbours weight 0 and stay undivided and unknown.
    Valid particles lie in [dx, (N-1)dx) (seeding and clampPos), where the
    reference's truncating (int) equals floor and its index clamps never
    engage, so the tent kernel is the reference's rule there."""
    n, dx = 10, 0.125  # dx a power of two: x / dx is exact
    cells = O.dam_break(n, n)
    o = O.Oracle(cells, k=0, dx=dx, mode=O.MODE_PIC)
    grid_pos = np.array([[3.3, 4.8], [3.9, 5.15], [7.25, 2.6], [6.0, 7.5]])
    vel = np.array([[1.5, -0.7], [-0.4, 2.2], [0.0, 0.9], [0.6, -1.1]])
    o.set(O.F_POS, grid_pos * dx)
    o.set(O.F_VEL, vel)
    o.run_stage(O.ST_P2G)
    for f, comp, shape, offset in ((O.F_U, 0, (n, n + 1), (0.0, 0.5)),
                                   (O.F_V, 1, (n + 1, n), (0.5, 0.0))):
        W, val = tent_p2g(grid_pos, vel[:, comp], shape, offset)
        known = (W > 0) & (val != 0.0)
        mask = np.where(known, 0, np.uint32(0xFFFFFFFF)).astype(np.uint32)
        want, _ = layered_extrapolate(val, mask)
        got = o.get(f)
        np.testing.assert_allclose(got[known], val[known], rtol=1e-14, atol=0)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    # the cases the comment names are present
    Wu, vu = tent_p2g(grid_pos, vel[:, 0], (n, n + 1), (0.0, 0.5))
    assert Wu[7, 6] == 1.0 and Wu[7, 7] == 0.0 and Wu[8, 6] == 0.0
    assert ((Wu > 0) & (vu == 0.0)).sum() == 4      # particle c's u faces
    Wv, vv = tent_p2g(grid_pos, vel[:, 1], (n + 1, n), (0.5, 0.0))
    assert ((Wv > 0) & (vv != 0.0)).sum() > 4 * 3


def test_p2g_tent_weights_reject_shifted_offsets():
    """The closed form above separates the right face offsets from the
    plausible slips: every half-cell shift and the swapped weights give a
    different set of known faces or values for the same particles."""
    grid_pos = np.array([[3.3, 4.8], [3.9, 5.15]])
    vel = np.array([1.5, -0.4])
    shape = (10, 11)
    W0, v0 = tent_p2g(grid_pos, vel, shape, (0.0, 0.5))
    for off in ((0.5, 0.5), (0.0, 0.0), (0.5, 0.0), (-0.5, 0.5), (0.0, 1.0)):
        W1, v1 = tent_p2g(grid_pos, vel, shape, off)
        assert not (np.array_equal(W0 > 0, W1 > 0) and np.allclose(v0, v1))
    # swapped tent weights: mirror the fractional parts within each cell
    mirrored = np.floor(grid_pos) + 1.0 - (grid_pos - np.floor(grid_pos))
    W2, v2 = tent_p2g(mirrored, vel, shape, (0.0, 0.5))
    assert not np.allclose(v0, v2)


def test_g2p_bilinear_field_exact_and_flip_limit():
    n, dx = 16, 1.0 / 16
    cells = O.dam_break(n, n)
    for alpha in (1.0, 0.0, 0.3):
        o = O.Oracle(cells, k=2, dx=dx, mode=O.MODE_PIC, alpha=alpha)
        pos = o.get(O.F_POS)
        m = len(pos)
        v0 = np.random.default_rng(2).standard_normal((m, 2))
        o.set(O.F_VEL, v0)
        Ju, Iu = np.mgrid[0:n, 0:n + 1].astype(float)
        Jv, Iv = np.mgrid[0:n + 1, 0:n].astype(float)
        # new field: bilinear in grid coordinates; old field: a different bilinear
        fu = lambda I, J: 0.2 + 0.03 * I - 0.05 * J + 0.002 * I * J
        fv = lambda I, J: -0.1 + 0.01 * I + 0.04 * J - 0.003 * I * J
        o.set(O.F_NEWU, fu(Iu, Ju))
        o.set(O.F_NEWV, fv(Iv, Jv))
        o.set(O.F_U, 0.5 * fu(Iu, Ju))
        o.set(O.F_V, 0.5 * fv(Iv, Jv))
        o.run_stage(O.ST_G2P)
        x, y = pos[:, 0] / dx, pos[:, 1] / dx
        pu, pv = fu(x, y - 0.5), fv(x - 0.5, y)
        want = np.stack([alpha * pu + (1 - alpha) * (v0[:, 0] + 0.5 * pu),
                         alpha * pv + (1 - alpha) * (v0[:, 1] + 0.5 * pv)], 1)
        np.testing.assert_allclose(o.get(O.F_VEL), want, rtol=1e-12, atol=1e-13)
        # mac <- newMac afterwards
        np.testing.assert_array_equal(o.get(O.F_U), fu(Iu, Ju))


# ---------------------------------------------------------------------------
# Advection: RK3 weights sum to 1 and Catmull-Rom preserves constants, so a
# uniform velocity moves particles by c*dt and leaves a uniform grid unchanged.
# ---------------------------------------------------------------------------
def test_uniform_velocity_advection():
    n, dx, dt = 32, 1.0 / 32, 0.01
    cells = O.dam_break(n, n)
    o = O.Oracle(cells, k=2, dx=dx, dt=dt, mode=O.MODE_SL)
    o.set(O.F_U, np.full((n, n + 1), 0.4))
    o.set(O.F_V, np.full((n + 1, n), -0.2))
    p0 = o.get(O.F_POS)
    o.run_stage(O.ST_ADVECT)
    p1 = o.get(O.F_POS)
    lo, hi = (1 + 1e-3) * dx, (n - 1 - 1e-3) * dx
    want = np.clip(p0 + dt * np.array([0.4, -0.2]), lo, hi)
    np.testing.assert_allclose(p1, want, rtol=0, atol=1e-14)
    o.run_stage(O.ST_SL)
    np.testing.assert_allclose(o.get(O.F_U), 0.4, rtol=1e-14)
    np.testing.assert_allclose(o.get(O.F_V), -0.2, rtol=1e-14)


def test_sl_zero_velocity_is_identity_outofplace():
    n = 16
    o = O.Oracle(O.dam_break(n, n), k=2, dx=1.0 / n, mode=O.MODE_SL, sl_inplace=False)
    u = np.zeros((n, n + 1))
    v = np.zeros((n + 1, n))
    o.set(O.F_U, u)
    o.set(O.F_V, v)
    o.run_stage(O.ST_SL)
    assert not o.get(O.F_U).any() and not o.get(O.F_V).any()


# ---------------------------------------------------------------------------
# Options of the 512^2 PIC/FLIP dam-break workload: uniform seeding and RK2
# particle advection.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 5])
def test_uniform_seeding_is_libc_rand_over_whole_cells(seed):
    """Each of the k^2 particles of a FLUID cell (row-major cell order) at
    ((i + U) dx, (j + U) dx), U = rand()/RAND_MAX from the C library's own
    generator, x before y."""
    n, dx, k = 12, 0.05, 2
    cells = O.dam_break(n, n)
    o = O.Oracle(cells, k=k, dx=dx, seed=seed, uniform_seeding=True)
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(ctypes.c_uint(seed))
    want = []
    for j in range(n):
        for i in range(n):
            if cells[j, i] == O.FLUID:
                for _ in range(k * k):
                    ux = libc.rand() / 2147483647.0
                    uy = libc.rand() / 2147483647.0
                    want.append(((i + ux) * dx, (j + uy) * dx))
    got = o.get(O.F_POS)
    np.testing.assert_array_equal(got, np.array(want))
    # every particle inside its own cell, not confined to a sub-cell
    cell_ij = np.floor(got / dx).astype(int)
    frac = got / dx - cell_ij
    assert (frac.min(axis=0) < 0.25).all() and (frac.max(axis=0) > 0.75).all()


@pytest.mark.parametrize("rk2", [False, True])
def test_particle_advection_order_on_linear_field(rk2):
    """u = c x, v = 0 (Catmull-Rom reproduces a linear field exactly away
    from clamped taps): the midpoint rule moves x to x (1 + h + h^2/2),
    Ralston's RK3 to x (1 + h + h^2/2 + h^3/6), h = c dt."""
    n, dx, dt, c = 32, 1.0 / 32, 0.05, 2.0
    cells = O.all_fluid(n, n)
    o = O.Oracle(cells, k=0, dx=dx, dt=dt, rk2=rk2)
    i = np.arange(n + 1, dtype=np.float64)
    o.set(O.F_U, np.tile(c * i * dx, (n, 1)))
    o.set(O.F_V, np.zeros((n + 1, n)))
    x0 = np.array([[10.3 * dx, 12.7 * dx], [15.6 * dx, 9.2 * dx], [8.05 * dx, 20.5 * dx]])
    o.set(O.F_POS, x0)
    o.run_stage(O.ST_ADVECT)
    h = c * dt
    factor = 1 + h + h * h / 2 + (0.0 if rk2 else h ** 3 / 6)
    got = o.get(O.F_POS)
    np.testing.assert_allclose(got[:, 0], x0[:, 0] * factor, rtol=1e-13)
    np.testing.assert_allclose(got[:, 1], x0[:, 1], rtol=1e-13)
