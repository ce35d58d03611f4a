"""Pinning of the oracle's boundary-element functions (Lindholm.cpp, BEMprocessing.cpp).

  * solid angle against the independent Van Oosterom-Strackee atan2 formula,
    and the octant (pi/2) in closed form;
  * Lindholm weights against brute-force Gauss quadrature of the double-layer
    kernel -(1/4pi) N_i (y-x).n / |y-x|^3 over the triangle;
  * the assembled matrix: B 1 = -1 row by row (closed-surface double layer),
    strengthening the reference's aggregate test (BEMprocessing.cpp:440-445).
"""
import numpy as np
import pytest


def vos(x, a, b, c):
    r1, r2, r3 = a - x, b - x, c - x
    l1, l2, l3 = map(np.linalg.norm, (r1, r2, r3))
    num = abs(r1 @ np.cross(r2, r3))
    den = l1 * l2 * l3 + (r1 @ r2) * l3 + (r2 @ r3) * l1 + (r3 @ r1) * l2
    h = np.arctan2(num, den)
    return 2 * (h if h >= 0 else h + np.pi)


def test_solid_angle_octant(oracle_mod):
    o = oracle_mod
    om = o.solid_angle(np.zeros(3), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]))
    assert abs(om - np.pi / 2) < 1e-14


def test_solid_angle_vs_vos(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b, c = rng.normal(size=(3, 3))
        x = rng.normal(size=3) * 0.5
        ref = vos(x, a, b, c)
        got = oracle_mod.solid_angle(x, a, b, c)
        # acos form: relative precision ~1e-16 / Omega^2
        assert abs(got - ref) <= 1e-15 / max(ref, 1e-3) ** 2 + 1e-13


def _gauss_weights(x0, p1, p2, p3, nv, order=60):
    """-(1/4pi) int_T N_i(y) (y - x0).nv / |y - x0|^3 dS by a Duffy-mapped
    tensor Gauss-Legendre rule: y = p1 + u (p2 - p1) + u v (p3 - p2)."""
    g, w = np.polynomial.legendre.leggauss(order)
    u = 0.5 * (g + 1)
    wu = 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * U
    jac = np.linalg.norm(np.cross(p2 - p1, p3 - p2))
    Y = p1 + U[..., None] * (p2 - p1) + (U * V)[..., None] * (p3 - p2)
    d = Y - x0
    k = -(d @ nv) / (4 * np.pi * np.linalg.norm(d, axis=-1) ** 3)
    N = [1 - U, U * (1 - V), U * V]
    return np.array([np.sum(W * jac * Ni * k) for Ni in N])


@pytest.mark.parametrize("seed", range(6))
def test_lindholm_vs_quadrature(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    p1, p2, p3 = rng.uniform(-1, 1, size=(3, 3))
    n = np.cross(p2 - p1, p3 - p1)
    A = 0.5 * np.linalg.norm(n)
    nv = n / np.linalg.norm(n) * (1 if seed % 2 else -1)
    centroid = (p1 + p2 + p3) / 3
    x0 = centroid + rng.normal(size=3) * 1.5 + 2.0 * nv * (1 if seed % 3 else -1)
    L = oracle_mod.lindholm(x0, p1, p2, p3, A, nv)
    Q = _gauss_weights(x0, p1, p2, p3, nv)
    # the reference's acos solid angle carries a relative error ~1e-16/Omega^2
    # that the weights' cancellation amplifies; a wrong index or sign is O(1)
    om = vos(x0, p1, p2, p3)
    assert np.max(np.abs(L - Q)) <= (1e-12 + 1e-14 / om ** 2) * np.max(np.abs(Q))
    # weights of the three vertices sum to the signed solid angle / 4 pi
    assert abs(L.sum() + np.sign((centroid - x0) @ nv) * vos(x0, p1, p2, p3) / (4 * np.pi)) < 1e-12


def test_lindholm_zero_on_vertex(oracle_mod):
    p1, p2, p3 = np.eye(3)
    L = oracle_mod.lindholm(p2, p1, p2, p3, 0.866, np.ones(3) / np.sqrt(3))
    assert np.all(L == 0.0)


def test_bem_row_sums(oracle_mod):
    from tetmag_amd import mesh as M
    m = M.kuhn_cuboid(6, 5, 3, 2.0, 0.1, 4)
    o = oracle_mod.Oracle(m.xyz, m.fel)
    B = o.B
    # B_jj = Omega_j / 4pi - 1 and B 1 = -1 row-wise
    assert np.allclose(np.diag(B), o.sub / (4 * np.pi) - 1, atol=1e-15)
    assert np.max(np.abs(B.sum(axis=1) + 1.0)) < 1e-10
    # the reference's aggregate check: sum(B (-1)) = bnx within 5e-3 relative
    assert abs((B @ -np.ones(o.bnx)).sum() - o.bnx) / o.bnx < 5e-3
