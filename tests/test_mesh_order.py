"""Node renumbering for gather locality (tm_mesh_locality_order): a
permutation, operators equal up to the permutation, and shorter index
distances along the mesh edges than row-by-row numbering."""
import numpy as np
import scipy.sparse as sp

from tetmag_amd import mesh as M


def _csr(o, n):
    return sp.csr_matrix((o.K, o.col, o.rowptr), shape=(n, n))


def test_locality_order_is_a_permutation_and_keeps_the_operators():
    base = M.kuhn_cuboid(12, 10, 4, 3.0, 0.1, 5)
    perm = M.locality_order(base)
    n = base.n
    assert sorted(perm.tolist()) == list(range(n))
    mesh = M.renumber(base, perm)
    assert np.array_equal(mesh.xyz, base.xyz[perm])
    K0 = _csr(M.fem_operators(base), n)
    K1 = _csr(M.fem_operators(mesh), n)
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))  # (P x)[k] = x[perm[k]]
    assert abs(P @ K0 @ P.T - K1).max() < 1e-12 * abs(K0).max()
    b0, b1 = M.boundary(base), M.boundary(mesh)
    assert np.array_equal(np.sort(perm[b1.bnodes]), b0.bnodes)


def test_locality_order_shortens_edges():
    base = M.kuhn_cuboid(40, 40, 8, 3.0, 0.1, 5)
    mesh = M.renumber(base, M.locality_order(base))

    def spread(m):  # median over rows of the largest |i - j| in the row
        o = M.fem_operators(m)
        rows = np.repeat(np.arange(m.n), np.diff(o.rowptr))
        d = np.abs(o.col - rows)
        return np.median(np.maximum.reduceat(d, o.rowptr[:-1]))

    assert spread(mesh) < 0.5 * spread(base)


def test_locality_order_degenerate_inputs():
    one = M.Mesh(np.zeros((1, 3)), np.zeros((0, 4), np.int32))
    assert M.locality_order(one).tolist() == [0]
    flat = M.Mesh(np.stack([np.arange(5.0), np.zeros(5), np.zeros(5)], 1), np.zeros((0, 4), np.int32))
    assert M.locality_order(flat).tolist() == [0, 1, 2, 3, 4]
