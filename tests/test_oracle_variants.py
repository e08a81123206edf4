"""The reference's compile-time variants in the oracle, pinned by hand-derived
cases (src/FluidSim2D.cpp:15-16):

* USE_GHOST_PRESSURE undefined (:304-331, :486-530): a FLUID cell's diagonal
  counts its FLUID neighbours plus an EMPTY neighbour at +x and at +y only
  (the -x / -y EMPTY neighbours add nothing); a face between FLUID and EMPTY
  takes the plain pressure difference with p = 0 in the EMPTY cell.
* USE_LEVEL_SET undefined (:665-706): cells from (int) of particle positions
  *in metres*, then the serial bubble fill (i outer, j inner, in place).
"""
import numpy as np

from oracle import oracle as O


def pool(n=8):
    c = O.walls(n, n)
    return c


def test_no_ghost_diagonal_counts_plus_x_plus_y_empty_only():
    c = pool()
    c[3, 3] = O.FLUID          # alone: EMPTY on all four sides
    c[5, 2] = c[5, 3] = O.FLUID  # a horizontal pair
    dt, dx, rho = 0.01, 0.1, 997.0
    o = O.Oracle(c, k=0, dt=dt, dx=dx, rho=rho, mode=O.MODE_SL, level_set=False)
    o.set_variants(particle_cells=False, ghost=False)
    o.assemble_factor()
    A = o.get(O.F_ADIAG)
    sA = dt / (rho * dx * dx)
    assert np.isclose(A[3, 3], 2 * sA, rtol=0, atol=1e-15)  # +x and +y EMPTY
    # left of the pair: FLUID at +x, EMPTY at -x (nothing), +y EMPTY (1), -y EMPTY (nothing)
    assert np.isclose(A[5, 2], 2 * sA, rtol=0, atol=1e-15)
    # right of the pair: FLUID at -x (1), EMPTY at +x (1), +y EMPTY (1)
    assert np.isclose(A[5, 3], 3 * sA, rtol=0, atol=1e-15)
    assert np.isclose(o.get(O.F_AX)[5, 2], -sA, rtol=0, atol=1e-15)
    # with ghost pressure the same cells take the phi-ratio terms instead
    og = O.Oracle(c, k=0, dt=dt, dx=dx, rho=rho, mode=O.MODE_SL, level_set=False)
    og.assemble_factor()
    assert not np.isclose(og.get(O.F_ADIAG)[3, 3], 2 * sA)


def test_no_ghost_velocity_update_uses_plain_pressure_difference():
    c = pool()
    c[3, 3] = O.FLUID
    dt, dx, rho = 0.01, 0.1, 997.0
    o = O.Oracle(c, k=0, dt=dt, dx=dx, rho=rho, mode=O.MODE_SL, level_set=False)
    o.set_variants(particle_cells=False, ghost=False)
    u = np.full((8, 9), 0.25)
    v = np.full((9, 8), -0.5)
    p = np.zeros((8, 8))
    p[3, 3] = 7.0
    o.set(O.F_U, u); o.set(O.F_V, v); o.set(O.F_P, p)
    o.run_stage(O.ST_UPDATEVEL)
    nu, nv = o.get(O.F_NEWU), o.get(O.F_NEWV)
    scale = dt / (rho * dx)
    assert nu[3, 3] == 0.25 - scale * (7.0 - 0.0)   # left face: EMPTY | FLUID
    assert nu[3, 4] == 0.25 - scale * (0.0 - 7.0)   # right face: FLUID | EMPTY
    assert nv[3, 3] == -0.5 - scale * (7.0 - 0.0)
    assert nv[4, 3] == -0.5 - scale * (0.0 - 7.0)


def marked(particles, n=8, fluid_start=()):
    c = pool(n)
    for (i, j) in fluid_start:
        c[j, i] = O.FLUID
    o = O.Oracle(c, k=0, dt=0.01, dx=1.0, mode=O.MODE_PIC)
    o.set_variants(particle_cells=True, ghost=False)
    if particles:
        o.set(O.F_POS, np.array(particles, dtype=np.float64))
    o.run_stage(O.ST_LEVELSET)
    return o.get(O.F_CELL)


def test_particle_cells_use_positions_in_metres_and_clear_old_fluid():
    # (int)pos in metres: with dx = 1 that is the cell index; old FLUID cells vanish
    c = marked([(3.7, 2.2), (5.01, 5.99), (5.5, 5.5)], fluid_start=[(1, 1), (6, 1)])
    fluid = sorted(zip(*np.nonzero(c == O.FLUID)))
    assert fluid == [(2, 3), (5, 5)]  # (j, i)


def test_particle_cells_truncate_metres_not_cells():
    # the same particles with dx = 0.1: (int)0.37 = 0 -> cell (0, 0) is SOLID, nothing marked
    c = pool()
    o = O.Oracle(c, k=0, dt=0.01, dx=0.1, mode=O.MODE_PIC)
    o.set_variants(particle_cells=True, ghost=False)
    o.set(O.F_POS, np.array([(0.37, 0.22), (0.51, 0.59)]))
    o.run_stage(O.ST_LEVELSET)
    assert not (o.get(O.F_CELL) == O.FLUID).any()


def test_bubble_fill_four_fluid_and_three_fluid_one_solid():
    ring = [(3 + di, 3 + dj) for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1))]
    c = marked([(i + 0.5, j + 0.5) for i, j in ring])
    assert c[3, 3] == O.FLUID
    # against the wall: (1, 3) has SOLID at -x and FLUID at +x, (1, 2), (1, 4)
    c = marked([(i + 0.5, j + 0.5) for i, j in ((2, 3), (1, 2), (1, 4))])
    assert c[3, 1] == O.FLUID
    # two FLUID neighbours only: stays EMPTY
    c = marked([(i + 0.5, j + 0.5) for i, j in ((2, 3), (4, 3))])
    assert c[3, 3] == O.EMPTY


def test_bubble_fill_never_cascades():
    # Holes at (i, j) = (3, 3) and (3, 4), each with its other three neighbours
    # FLUID: serial order visits (3, 3) first, which sees the EMPTY (3, 4) and
    # stays; (3, 4) then sees the EMPTY (3, 3) and stays too.  A fill needs all
    # four neighbours FLUID/SOLID at its visit, so two adjacent holes never fill
    # and the in-place serial pass equals a pass over the flags before it.
    pts = [(2, 3), (4, 3), (3, 2), (2, 4), (4, 4), (3, 5)]
    c = marked([(i + 0.5, j + 0.5) for i, j in pts])
    assert c[3, 3] == O.EMPTY and c[4, 3] == O.EMPTY
    # the same holes along the wall: (1, 3) has SOLID at -x, (2, 3) is its hole neighbour
    pts = [(1, 2), (1, 4), (3, 3), (2, 2), (2, 4)]
    c = marked([(i + 0.5, j + 0.5) for i, j in pts])
    assert c[3, 1] == O.EMPTY and c[3, 2] == O.EMPTY
    # fill one hole with a particle: the other now sees 4 FLUID and fills
    c = marked([(i + 0.5, j + 0.5) for i, j in pts + [(1, 3)]])
    assert c[3, 2] == O.FLUID


def test_bubble_fill_on_random_flags_equals_a_parallel_pass():
    rng = np.random.default_rng(0)
    n = 24
    pts = []
    for j in range(1, n - 1):
        for i in range(1, n - 1):
            if rng.random() < 0.8:
                pts.append((i + 0.5, j + 0.5))
    c = marked(pts, n=n)
    # the marked flags before the fill, then one pass on them
    m = O.walls(n, n)
    for x, y in pts:
        m[int(y), int(x)] = O.FLUID
    want = m.copy()
    for j in range(1, n - 1):
        for i in range(1, n - 1):
            if m[j, i] != O.EMPTY:
                continue
            nb = [m[j, i - 1], m[j, i + 1], m[j - 1, i], m[j + 1, i]]
            f, sd = sum(x == O.FLUID for x in nb), sum(x == O.SOLID for x in nb)
            if f == 4 or (f == 3 and sd == 1):
                want[j, i] = O.FLUID
    np.testing.assert_array_equal(c, want)
