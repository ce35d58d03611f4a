"""Pinning of the oracle's FEM setup (FEMprocessing.cpp, Boundary.cpp).

Each check is fixed by mathematics or by the reference's own self-test, not by
retyping the oracle's formulas:
  * the reference's gradient known-answer test (FEMprocessing.cpp:288-326),
  * the single-tetrahedron stiffness matrix in closed form,
  * exactness of P1 on linear fields (x^T K x = |a|^2 Vol),
  * operator identities grad_d = diag(1/V) tGrad_d^T, column sums of tGrad = 0,
  * boundary counts / areas / normals of a cuboid in closed form,
  * interior solid angles of corners (pi/2), edges (pi) and faces (2 pi).
"""
import numpy as np
import pytest

from conftest import rel_l2


def _mesh(nx=5, ny=4, nz=3, h=2.0, jitter=0.15, seed=3):
    from tetmag_amd import mesh as M
    return M.kuhn_cuboid(nx, ny, nz, h, jitter, seed)


@pytest.fixture(scope="module")
def orc(oracle_mod):
    m = _mesh()
    return oracle_mod.Oracle(m.xyz, m.fel), m


def _csr_dense(rp, col, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


def test_gradient_known_answer(orc):
    """FEMprocessing.cpp:288-326: grad of 0.5x - 17y + 13z + 42 is (0.5, -17, 13), tol 1e-10."""
    o, m = orc
    f = 0.5 * m.xyz[:, 0] - 17.0 * m.xyz[:, 1] + 13.0 * m.xyz[:, 2] + 42.0
    import pyoracle as O
    for val, c in ((o.gX, 0.5), (o.gY, -17.0), (o.gZ, 13.0)):
        g = O.csr_matvec(o.rowptr, o.col, val, f)
        assert np.max(np.abs(g - c)) <= 1e-10 * abs(c)


def test_gradient_known_answer_fails_for_wrong_sign(orc):
    """The KAT distinguishes grad from tGrad (a transposed operand fails it)."""
    o, m = orc
    import pyoracle as O
    f = 0.5 * m.xyz[:, 0] - 17.0 * m.xyz[:, 1] + 13.0 * m.xyz[:, 2] + 42.0
    g = O.csr_matvec(o.rowptr, o.col, o.tX / o.nodevol[np.repeat(np.arange(o.n), np.diff(o.rowptr))], f)
    assert np.max(np.abs(g - 0.5)) > 1e-3


def test_single_tet_stiffness_closed_form(oracle_mod):
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    expect = np.array([[3, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6.0
    for fel in (np.array([[0, 1, 2, 3]]), np.array([[1, 0, 2, 3]])):  # both orientations
        o = oracle_mod.Oracle(xyz, fel)
        K = _csr_dense(o.rowptr, o.col, o.K, 4)
        assert np.allclose(K, expect, atol=1e-15)
        assert np.allclose(o.nodevol, 1.0 / 24.0)
        # tGrad_x(i,j) = dN_i/dx * V/4 for every j
        TX = _csr_dense(o.rowptr, o.col, o.tX, 4)
        dNx = np.array([-1.0, 1.0, 0.0, 0.0])
        assert np.allclose(TX, np.outer(dNx, np.ones(4)) / 24.0, atol=1e-15)


def test_stiffness_properties(orc):
    o, m = orc
    n = o.n
    K = _csr_dense(o.rowptr, o.col, o.K, n)
    assert np.max(np.abs(K - K.T)) <= 1e-14 * np.max(np.abs(K))
    assert np.max(np.abs(K @ np.ones(n))) <= 1e-13 * np.max(np.abs(K))
    vol = o.elvol.sum()
    a = np.array([0.3, -1.2, 2.0])
    x = m.xyz @ a + 5.0
    assert abs(x @ K @ x - (a @ a) * vol) <= 1e-11 * (a @ a) * vol
    # positive semi-definite with a one-dimensional null space (connected mesh)
    w = np.linalg.eigvalsh(K)
    assert abs(w[0]) < 1e-12 * w[-1] and w[1] > 1e-8 * w[-1]


def test_operator_identities(orc):
    o, m = orc
    n = o.n
    for g, t in ((o.gX, o.tX), (o.gY, o.tY), (o.gZ, o.tZ)):
        G = _csr_dense(o.rowptr, o.col, g, n)
        T = _csr_dense(o.rowptr, o.col, t, n)
        assert np.max(np.abs(G - (T.T / o.nodevol[:, None]))) <= 1e-13 * np.max(np.abs(G))
        assert np.max(np.abs(T.sum(axis=0))) <= 1e-13 * np.max(np.abs(T))  # sum_i tGrad(i,j) = 0
    assert abs(o.nodevol.sum() - o.elvol.sum()) <= 1e-12 * o.elvol.sum()


def test_cuboid_volume_and_boundary_closed_form(oracle_mod):
    nx, ny, nz, h = 5, 4, 3, 2.0
    m = _mesh(nx, ny, nz, h)
    o = oracle_mod.Oracle(m.xyz, m.fel)
    a, b, c = nx * h, ny * h, nz * h
    assert abs(o.elvol.sum() - a * b * c) <= 1e-12 * a * b * c  # jitter moves interior nodes only
    assert o.nbel == 4 * (nx * ny + ny * nz + nx * nz)
    assert o.bnx == (nx + 1) * (ny + 1) * (nz + 1) - (nx - 1) * (ny - 1) * (nz - 1)
    S = 2 * (a * b + b * c + c * a)
    assert abs(o.face_area.sum() - S) <= 1e-12 * S
    assert abs(o.node_area.sum() - S) <= 1e-12 * S
    # closed surface: sum A_f n_f = 0; outward: n_f . (centroid_f - box centre) > 0
    assert np.max(np.abs((o.face_area[:, None] * o.face_normal).sum(axis=0))) <= 1e-12 * S
    cen = m.xyz[o.faces].mean(axis=1) - np.array([a, b, c]) / 2
    assert np.all(np.einsum("ij,ij->i", o.face_normal, cen) > 0)
    # faces are axis-aligned: normals are +-unit axes
    assert np.allclose(np.sort(np.abs(o.face_normal), axis=1), [0, 0, 1], atol=1e-14)


def test_interior_solid_angles(oracle_mod):
    nx, ny, nz, h = 4, 3, 2, 1.5
    m = _mesh(nx, ny, nz, h, jitter=0.1)
    o = oracle_mod.Oracle(m.xyz, m.fel)
    x = m.xyz[o.bnodes]
    on = [(np.isclose(x[:, d], 0) | np.isclose(x[:, d], L)).astype(int)
          for d, L in enumerate((nx * h, ny * h, nz * h))]
    k = on[0] + on[1] + on[2]  # number of bounding planes through the node
    expect = np.select([k == 3, k == 2, k == 1], [np.pi / 2, np.pi, 2 * np.pi])
    assert np.max(np.abs(o.sub - expect)) < 1e-12


def test_product_setup_matches_oracle(oracle_mod):
    """The product's host setup (inverse-Jacobian gradients, hashed faces,
    atan2 solid angles) reproduces the oracle's operators on a jittered mesh."""
    from tetmag_amd import mesh as M
    m = _mesh(7, 5, 4, 3.0, 0.12, 9)
    o = oracle_mod.Oracle(m.xyz, m.fel)
    ops = M.fem_operators(m)
    bnd = M.boundary(m)
    assert np.array_equal(ops.rowptr, o.rowptr) and np.array_equal(ops.col, o.col)
    for a, b in ((ops.K, o.K), (ops.gx, o.gX), (ops.gy, o.gY), (ops.gz, o.gZ), (ops.tgx, o.tX), (ops.tgy, o.tY),
                 (ops.tgz, o.tZ), (ops.nodevol, o.nodevol)):
        assert rel_l2(a, b) < 1e-14
    assert np.array_equal(bnd.bnodes, o.bnodes)
    assert rel_l2(bnd.node_area, o.node_area) < 1e-14
    assert rel_l2(bnd.node_normal, o.node_normal) < 1e-14
    assert rel_l2(bnd.solid_angle, o.sub) < 1e-13
    # same face set (product: ascending, oracle: descending lexicographic)
    assert np.array_equal(bnd.faces[::-1], o.faces)


def test_config_sizes():
    """Generated sizes of the two headline meshes."""
    from tetmag_amd import mesh as M
    import ctypes as C
    from tetmag_amd import _native as N
    n, t = C.c_int(), C.c_int()
    N.lib().tm_mesh_kuhn_cuboid_size(20, 20, 4, C.byref(n), C.byref(t))
    assert (n.value, t.value) == (2205, 9600)
    N.lib().tm_mesh_kuhn_cuboid_size(128, 128, 16, C.byref(n), C.byref(t))
    assert (n.value, t.value) == (282897, 1572864)
    assert 282897 - 127 * 127 * 15 == 40962
