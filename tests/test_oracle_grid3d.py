"""The oracle's 3D grid (oracle/fs3d_oracle.cpp, restating include/Array3D.h and
include/MACGrid3D.h) pinned by what the mathematics and the data layout fix:
node values, the Catmull-Rom kernel, quadratic reproduction, the relation
between the reference's transposed tap order and the plain one, the staggered
offsets of velInterp, exact divergence and averages of linear fields, the
boundary quarter-sums, brute-force out-of-range taps, and numpy for the
Array3D arithmetic."""
import numpy as np
import pytest

from oracle import oracle3d as O3


def rng(seed=0):
    return np.random.default_rng(seed)


def keys_kernel(t):
    """Catmull-Rom (Keys, a = -1/2) in kernel form, independent of the
    reference's polynomial-in-d form."""
    t = abs(t)
    if t < 1:
        return 1.5 * t ** 3 - 2.5 * t ** 2 + 1
    if t < 2:
        return -0.5 * t ** 3 + 2.5 * t ** 2 - 4 * t + 2
    return 0.0


def test_nodes_velocity_mode_return_the_node():
    a = rng().uniform(-1, 1, (5, 6, 7))
    pts = np.array([[i, j, k] for k in range(5) for j in range(6) for i in range(7)], dtype=float)
    got, bad = O3.tricubic(a, pts, O3.VELOCITY)
    assert bad == 0
    np.testing.assert_array_equal(got, a[pts[:, 2].astype(int), pts[:, 1].astype(int), pts[:, 0].astype(int)])


def test_nodes_reference_mode_read_the_transposed_element_clamped_at_zero():
    # (*this)(zp, yp, xp) is element xp*NY*NX + yp*NX + zp of the flat array
    a = rng(1).uniform(-1, 1, (6, 4, 6))
    flat = a.ravel()
    NZ, NY, NX = a.shape
    pts = np.array([[x, y, z] for z in range(NZ) for y in range(NY) for x in range(NX)], dtype=float)
    got, bad = O3.tricubic(a, pts, O3.REFERENCE)
    assert bad == 0
    e = pts[:, 0].astype(int) * NY * NX + pts[:, 1].astype(int) * NX + pts[:, 2].astype(int)
    np.testing.assert_array_equal(got, np.maximum(flat[e], 0.0))


def test_weights_match_the_keys_kernel_in_each_axis():
    a = np.zeros((6, 6, 6))
    for d in (0.0, 0.1, 0.37, 0.5, 0.81, 0.999):
        for tap in range(4):
            a[:] = 0
            a[2, 2, 1 + tap] = 1.0  # x taps at 1..4 for x0 = 2
            got, _ = O3.tricubic(a, [[2 + d, 2.0, 2.0]], O3.VELOCITY)
            assert abs(got[0] - keys_kernel(2 + d - (1 + tap))) < 1e-13, (d, tap)


def test_quadratics_reproduced_in_the_interior():
    q = lambda s, c: c[0] + c[1] * s + c[2] * s * s
    cx, cy, cz = (1.0, 0.3, -0.02), (2.0, -0.1, 0.03), (0.5, 0.2, 0.01)
    NZ, NY, NX = 9, 8, 10
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    a = q(i, cx) * q(j, cy) * q(k, cz)
    p = rng(2).uniform([1, 1, 1], [NX - 3, NY - 3, NZ - 3], (200, 3))
    got, _ = O3.tricubic(a, p, O3.VELOCITY)
    want = q(p[:, 0], cx) * q(p[:, 1], cy) * q(p[:, 2], cz)
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_reference_mode_is_the_axis_swapped_velocity_mode_clamped():
    # NX == NZ: reading (zp, yp, xp) of A is reading (xp, yp, zp) of A with x and z swapped
    a = rng(3).uniform(-1, 1, (7, 5, 7))
    b = np.ascontiguousarray(a.transpose(2, 1, 0))
    p = rng(4).uniform(-0.9, 7.5, (500, 3))
    ref, bad = O3.tricubic(a, p, O3.REFERENCE)
    vel, _ = O3.tricubic(b, p, O3.VELOCITY)
    assert bad == 0
    np.testing.assert_allclose(ref, np.maximum(vel, 0.0), rtol=0, atol=1e-14)
    assert (ref == 0).sum() > 50 and (ref > 0).sum() > 50


def test_outside_the_array_is_zero_and_constants_are_kept_anywhere_inside():
    a = np.full((4, 5, 6), 2.5)
    out, _ = O3.tricubic(a, [[-1.0, 1, 1], [6.0, 1, 1], [1, 5.0, 1], [1, 1, 4.2], [1, -1.5, 1]], O3.VELOCITY)
    np.testing.assert_array_equal(out, 0.0)
    # (int) truncates: -0.5 -> 0 passes the range check; the cubic weights still sum to 1
    p = np.array([[-0.5, 1.2, 2.7], [5.9, 4.99, 3.999], [0.0, 0.0, 0.0], [-0.99, -0.99, -0.99]])
    for mode in (O3.VELOCITY, O3.REFERENCE):
        out, bad = O3.tricubic(np.full((6, 5, 6), 2.5), p, mode)
        assert bad == 0
        np.testing.assert_allclose(out, 2.5, rtol=1e-14)


def test_out_of_range_taps_match_brute_force():
    n = 6
    u = rng(5).uniform(0.1, 1, (n, n, n + 1))  # MACGrid3D u of an n^3 grid: NX = n+1, NZ = n
    NZ, NY, NX = u.shape
    p = rng(6).uniform(-0.5, n + 1, (400, 3))
    got, bad = O3.tricubic(u, p, O3.REFERENCE)
    want_bad = np.zeros(len(p), bool)
    for r, (x, y, z) in enumerate(p):
        xi, yi, zi = int(x), int(y), int(z)
        if not (0 <= xi < NX and 0 <= yi < NY and 0 <= zi < NZ):
            continue
        for kk in range(4):
            for jj in range(4):
                for ii in range(4):
                    xp = min(max(xi + ii - 1, 0), NX - 1)
                    yp = min(max(yi + jj - 1, 0), NY - 1)
                    zp = min(max(zi + kk - 1, 0), NZ - 1)
                    if xp * NY * NX + yp * NX + zp >= u.size:
                        want_bad[r] = True
    assert bad == want_bad.sum() > 0
    assert np.isnan(got[want_bad]).all() and not np.isnan(got[~want_bad]).any()


def staggered(nx, ny, nz, fu, fv, fw):
    """u, v, w sampled from functions of the face positions (grid units)."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx + 1), indexing="ij")
    u = fu(i, j + 0.5, k + 0.5)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny + 1), np.arange(nx), indexing="ij")
    v = fv(i + 0.5, j, k + 0.5)
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx), indexing="ij")
    w = fw(i + 0.5, j + 0.5, k)
    return u.astype(float), v.astype(float), w.astype(float)


def test_vel_interp_samples_each_component_at_its_face_offsets():
    q = lambda x, y, z: 1.0 + 0.2 * x - 0.1 * y + 0.05 * z + 0.01 * x * y - 0.02 * z * z
    nx, ny, nz = 12, 10, 9
    u, v, w = staggered(nx, ny, nz, q, lambda x, y, z: 2 * q(x, y, z), lambda x, y, z: -q(x, y, z))
    p = rng(7).uniform([2.5, 2.5, 2.5], [nx - 3, ny - 3, nz - 3], (300, 3))
    got, bad = O3.vel_interp(u, v, w, p, O3.VELOCITY)
    assert bad == 0
    want = q(p[:, 0], p[:, 1], p[:, 2])
    np.testing.assert_allclose(got, np.stack([want, 2 * want, -want], 1), rtol=1e-12, atol=1e-12)


def test_vel_div_and_cell_average_of_linear_fields():
    a, b, c = 3.0, -2.0, 0.5
    nx, ny, nz = 7, 6, 5
    u, v, w = staggered(nx, ny, nz, lambda x, y, z: a * x + 0 * y, lambda x, y, z: b * y + 0 * x,
                        lambda x, y, z: c * z + 0 * x)
    np.testing.assert_array_equal(O3.vel_div(u, v, w), a + b + c)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    want = np.stack([a * (i + 0.5), b * (j + 0.5), c * (k + 0.5)], -1)
    np.testing.assert_allclose(O3.vel_cell(u, v, w), want, rtol=1e-15)
    # random fields: the divergence is the sum of the three face differences
    r = rng(8)
    u, v, w = r.normal(size=u.shape), r.normal(size=v.shape), r.normal(size=w.shape)
    want = np.diff(u, axis=2) + np.diff(v, axis=1) + np.diff(w, axis=0)
    np.testing.assert_allclose(O3.vel_div(u, v, w), want, rtol=1e-13, atol=1e-13)


def test_face_velocities_interior_average_and_boundary_quarter_sums():
    nx, ny, nz = 5, 4, 6
    U0, V0, W0 = 1.5, -2.0, 0.25
    u, v, w = staggered(nx, ny, nz, lambda x, y, z: U0 + 0 * x, lambda x, y, z: V0 + 0 * x,
                        lambda x, y, z: W0 + 0 * x)
    fu = O3.vel_face(u, v, w, 0)
    np.testing.assert_allclose(fu[:, :, 1:nx], np.broadcast_to([U0, V0, W0], fu[:, :, 1:nx].shape))
    # first and last u faces: two of the four terms, still divided by 4
    for i in (0, nx):
        np.testing.assert_allclose(fu[:, :, i], np.broadcast_to([U0, V0 / 2, W0 / 2], fu[:, :, i].shape))
    fv = O3.vel_face(u, v, w, 1)
    np.testing.assert_allclose(fv[:, 1:ny], np.broadcast_to([U0, V0, W0], fv[:, 1:ny].shape))
    np.testing.assert_allclose(fv[:, 0], np.broadcast_to([U0 / 2, V0, W0 / 2], fv[:, 0].shape))
    fw = O3.vel_face(u, v, w, 2)
    np.testing.assert_allclose(fw[1:nz], np.broadcast_to([U0, V0, W0], fw[1:nz].shape))
    np.testing.assert_allclose(fw[nz], np.broadcast_to([U0 / 2, V0 / 2, W0], fw[nz].shape))
    # linear fields: interior transverse averages are the values at the face position
    u, v, w = staggered(nx, ny, nz, lambda x, y, z: x + 2 * y - z, lambda x, y, z: 3 * x - y + 2 * z,
                        lambda x, y, z: -x + y + 4 * z)
    fu = O3.vel_face(u, v, w, 0)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx + 1), indexing="ij")
    x, y, z = i, j + 0.5, k + 0.5
    want = np.stack([x + 2 * y - z, 3 * x - y + 2 * z, -x + y + 4 * z], -1)
    np.testing.assert_allclose(fu[:, :, 1:nx], want[:, :, 1:nx], rtol=1e-14, atol=1e-13)


def test_array_arithmetic_against_numpy():
    r = rng(9)
    a, b = r.normal(size=(4, 5, 6)), r.normal(size=(4, 5, 6))
    assert abs(O3.inner_product(a, b) - float(np.dot(a.ravel(), b.ravel()))) < 1e-12
    assert O3.infinite_norm(a) == np.abs(a).max()
    np.testing.assert_array_equal(O3.multiply_add(a, 0.75, b), a + 0.75 * b)
