"""Closed-form pins of the oracle functions that the physics tests leave open.

  * bulk DMI surface term: for a uniform m every interior row of tGrad sums
    to zero and a flat-face row sums to -S n (divergence theorem), so the
    field of EffFieldCalc.cpp:159-170 is H = -(D S / V)(m x n) on flat-face
    nodes and 0 inside; a flipped surface sign gives 3x that value, a dropped
    surface term 2x;
  * total field (TheLLG.cpp:238-254): uniform m, demag off, per-node Js
    (one node with Js = 0, where invJs = 0, TheLLG.cpp:110): exchange vanishes
    exactly, Heff = (2 Ku / Js)(k.m) k + Hext; a spin wave gives
    -(2A/Js) k^2 m at interior nodes;
  * Zhang-Li STT (TheLLG.cpp:201-220, 402-416): for a linear m (any mesh)
    the grad operators are exact, so U_c = sum_d eta_d dm_c/dx_d exactly at
    every node; the torque of m = x, U = y is (0, 1 + alpha beta,
    -(beta - alpha)), it vanishes for U || m, is normal to m and has norm
    sqrt((beta - alpha)^2 + (1 + alpha beta)^2) |U_perp|;
  * sampled rows of B equal the rows of the assembled matrix and sum to -1;
  * SA-AMG-preconditioned GMRES: against a dense direct solve, on both
    row-replaced matrices;
  * per-node applied fields enter Heff additively.
"""
import numpy as np
import pytest

from conftest import rel_l2

MU0 = 1.256637061435917295376e-6


def _oracle(oracle_mod, nx, ny, nz, h, jitter=0.0, seed=1):
    from tetmag_amd import mesh as M
    m = M.kuhn_cuboid(nx, ny, nz, h, jitter, seed)
    return oracle_mod.Oracle(m.xyz, m.fel), m


def _uniform(n, v):
    v = np.asarray(v, float)
    return np.concatenate([np.full(n, v[0]), np.full(n, v[1]), np.full(n, v[2])])


@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_bulk_dmi_uniform_m_surface_term(oracle_mod, jitter):
    nx, ny, nz, h = 6, 5, 4, 2.0
    o, mesh = _oracle(oracle_mod, nx, ny, nz, h, jitter, 3)
    n = o.n
    rng = np.random.default_rng(4)
    D = 3e6 * (1 + 0.5 * rng.random(n))  # per node, so D_i is pinned too
    mv = np.array([0.3, -0.5, 0.81])
    mv /= np.linalg.norm(mv)
    H = o.dmi(_uniform(n, mv), D).reshape(3, n).T
    x = mesh.xyz
    L = np.array([nx * h, ny * h, nz * h])
    at_bound = (np.abs(x) < 1e-9) | (np.abs(x - L) < 1e-9)
    nb = at_bound.sum(axis=1)
    face = nb == 1
    inner = nb == 0
    assert face.sum() > 20 and inner.sum() > 20
    nrm = o.node_normal.reshape(3, n).T
    # face nodes: the normal is the face's axis direction
    ax = np.argmax(at_bound[face], axis=1)
    assert np.allclose(np.abs(nrm[face][np.arange(face.sum()), ax]), 1.0, atol=1e-14)
    S, V = o.node_area, o.nodevol
    expect = -(D * S / V)[:, None] * np.cross(np.tile(mv, (n, 1)), nrm)
    scale = np.max(np.abs(expect[face]))
    assert scale > 0
    assert np.max(np.abs(H[face] - expect[face])) < 1e-12 * scale
    assert np.max(np.abs(H[inner])) < 1e-12 * scale
    # what two plausible mistakes would give, for the record: 3x (flipped surface sign), 2x (no surface term)
    assert np.max(np.abs(3 * expect[face] - expect[face])) > 1.0 * scale


def test_total_field_invjs_scaling(oracle_mod):
    o, mesh = _oracle(oracle_mod, 5, 4, 3, 2.0, 0.1, 5)
    n = o.n
    rng = np.random.default_rng(7)
    Js = MU0 * 8e5 * (0.5 + rng.random(n))
    Js[n // 3] = 0.0  # a non-magnetic node: invJs = 0
    A = np.full(n, 1.3e-11 / 1e-18)
    Ku = 5e4 * (1 + rng.random(n))
    kv = np.array([0.48, -0.6, 0.64])
    axis = _uniform(n, kv)
    mv = np.array([0.8, 0.0, 0.6])
    m = _uniform(n, mv)
    Hext = np.array([1.5e4, -2.0e4, 3.0e3])
    out = o.eval_rhs(m, Js=Js, A=A, Ku=Ku, axis=axis, Hext=Hext, alpha=0.5, demag=False)
    with np.errstate(divide="ignore"):
        inv = np.where(Js > 0, 1.0 / Js, 0.0)
    expect = (2 * Ku * (kv @ mv) * inv)[None, :] * kv[:, None] + Hext[:, None]
    got = out["Heff"].reshape(3, n)
    assert np.max(np.abs(got - expect)) < 1e-9 * np.max(np.abs(expect))
    assert np.allclose(got[:, n // 3], Hext, rtol=0, atol=1e-9 * np.max(np.abs(Hext)))


def test_total_field_spin_wave_scaling(oracle_mod):
    nx, ny, nz, h = 40, 4, 4, 1.0
    o, mesh = _oracle(oracle_mod, nx, ny, nz, h)
    n = o.n
    Js = np.full(n, 0.7)
    A = np.full(n, 1.3e-11 / 1e-18)
    k = 2 * np.pi / (nx * h)
    x = mesh.xyz[:, 0]
    m = np.concatenate([np.cos(k * x), np.sin(k * x), np.zeros(n)])
    out = o.eval_rhs(m, Js=Js, A=A, alpha=0.5, demag=False)
    interior = (x > 0) & (x < nx * h)
    expect = -2 * A[0] * k * k / Js[0]
    for c in range(2):
        got = out["Heff"][c * n:(c + 1) * n][interior]
        assert np.max(np.abs(got - expect * m[c * n:(c + 1) * n][interior])) < 0.01 * abs(expect)


def test_stt_uterm_linear_field_exact(oracle_mod):
    o, mesh = _oracle(oracle_mod, 5, 4, 3, 2.0, 0.15, 8)
    n = o.n
    rng = np.random.default_rng(9)
    a = rng.normal(size=3)
    Bm = rng.normal(size=(3, 3))  # m_c(r) = a_c + B[c] . r (not normalised: U is linear in m)
    m = (a[:, None] + Bm @ mesh.xyz.T).reshape(-1)
    eta = rng.normal(size=(3, n)) * 1e-3  # per node, per current direction
    U = o.stt_uterm(m, eta).reshape(3, n)
    expect = Bm @ eta  # U_c = sum_d eta_d B[c, d]
    assert np.max(np.abs(U - expect)) < 1e-12 * np.max(np.abs(expect))


def test_stt_torque_closed_forms(oracle_mod):
    o, mesh = _oracle(oracle_mod, 1, 1, 1, 1.0)
    n = o.n
    alpha, beta = 0.3, 0.05
    m = _uniform(n, [1, 0, 0])
    tau = o.stt_llg(m, _uniform(n, [0, 1, 0]), alpha, beta).reshape(3, n)
    assert np.allclose(tau[0], 0, atol=1e-16)
    assert np.allclose(tau[1], 1 + alpha * beta, rtol=1e-15)
    assert np.allclose(tau[2], -(beta - alpha), rtol=1e-15)
    assert np.max(np.abs(o.stt_llg(m, _uniform(n, [2.5, 0, 0]), alpha, beta))) == 0.0
    rng = np.random.default_rng(10)
    M = rng.normal(size=(n, 3))
    M /= np.linalg.norm(M, axis=1, keepdims=True)
    Uv = rng.normal(size=(n, 3))
    T = o.stt_llg(M.T.reshape(-1), Uv.T.reshape(-1), alpha, beta).reshape(3, n).T
    uperp = Uv - np.sum(Uv * M, axis=1, keepdims=True) * M
    assert np.max(np.abs(np.sum(T * M, axis=1))) < 1e-14
    want = np.sqrt((beta - alpha) ** 2 + (1 + alpha * beta) ** 2) * np.linalg.norm(uperp, axis=1)
    assert np.allclose(np.linalg.norm(T, axis=1), want, rtol=1e-13)


def test_bem_rows_equal_matrix_rows(oracle_mod):
    o, mesh = _oracle(oracle_mod, 6, 5, 3, 2.0, 0.1, 2)
    rows = np.array([0, 3, o.bnx // 2, o.bnx - 1])
    R = o.bem_rows(rows)
    assert np.array_equal(R, o.B[rows, :])
    assert np.max(np.abs(R.sum(axis=1) + 1.0)) < 1e-10


@pytest.mark.parametrize("which", ["neumann", "dirichlet"])
def test_amg_gmres_against_dense(oracle_mod, which):
    import pyoracle as O
    o, mesh = _oracle(oracle_mod, 10, 8, 4, 1.0, 0.1, 2)
    n = o.n
    rp, col, val = o.neumann_matrix() if which == "neumann" else o.dirichlet_matrix()
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    b = np.random.default_rng(0).normal(size=n)
    s = O.Solver(rp, col, val, precond="amg")
    assert len(s.levels()) >= 2 and s.levels()[-1] <= 100
    x = s.solve(b, tol=1e-14, warm=False)
    assert rel_l2(x, np.linalg.solve(A, b)) < 1e-10
    assert s.last_relres <= 1e-14
    # warm start from the solution: no iterations
    s.solve(b, tol=1e-13)
    assert s.last_iters == 0


def test_per_node_applied_field_is_additive(oracle_mod):
    o, mesh = _oracle(oracle_mod, 4, 3, 2, 2.0, 0.1, 3)
    n = o.n
    rng = np.random.default_rng(11)
    m = rng.normal(size=(3, n))
    m /= np.linalg.norm(m, axis=0)
    m = m.reshape(-1)
    kw = dict(Js=np.full(n, 1.0), A=np.full(n, 1e7), alpha=0.4, demag=False)
    Hn = rng.normal(size=3 * n) * 1e4
    a = o.eval_rhs(m, Hext=Hn, **kw)["Heff"]
    b = o.eval_rhs(m, **kw)["Heff"]
    assert np.max(np.abs((a - b) - Hn)) < 1e-9 * np.max(np.abs(Hn))
