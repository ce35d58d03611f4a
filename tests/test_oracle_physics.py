"""Pinning of the oracle's field, demag, LLG and solver functions by physics.

  * demag: volume-averaged field of a uniformly magnetised cuboid against
    Aharoni's closed-form demagnetising factors (J. Appl. Phys. 83, 3432 (1998));
    the formula itself is pinned by D_x + D_y + D_z = 1 and D = 1/3 for a cube;
  * demag gauge: H does not depend on which node the Neumann problem pins;
  * exchange: zero for uniform m, -2 A k^2 m for a spin wave (interior nodes);
  * bulk DMI: 2 D k m for a helix m = (0, cos kx, sin kx) (interior nodes);
  * uniaxial: 2 Ku k for m || k and 0 for m perpendicular to k;
  * LLG: the macrospin right-hand side in closed form, m . dm/dt = 0;
  * GMRES/ILU(0): against a dense direct solve of the row-replaced matrices.
"""
import numpy as np
import pytest

from conftest import rel_l2

MU0 = 1.256637061435917295376e-6


def aharoni_dz(a, b, c):
    """Demagnetising factor along z of the prism 2a x 2b x 2c."""
    r = np.sqrt(a * a + b * b + c * c)
    ab, bc, ac = np.sqrt(a * a + b * b), np.sqrt(b * b + c * c), np.sqrt(a * a + c * c)
    t = ((b * b - c * c) / (2 * b * c) * np.log((r - a) / (r + a))
         + (a * a - c * c) / (2 * a * c) * np.log((r - b) / (r + b))
         + b / (2 * c) * np.log((ab + a) / (ab - a))
         + a / (2 * c) * np.log((ab + b) / (ab - b))
         + c / (2 * a) * np.log((bc - b) / (bc + b))
         + c / (2 * b) * np.log((ac - a) / (ac + a))
         + 2 * np.arctan(a * b / (c * r))
         + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c)
         + (a * a + b * b - 2 * c * c) / (3 * a * b * c) * r
         + c / (a * b) * (ac + bc)
         - (ab ** 3 + bc ** 3 + ac ** 3) / (3 * a * b * c))
    return t / np.pi


def test_aharoni_formula_itself():
    assert abs(aharoni_dz(1, 1, 1) - 1 / 3) < 1e-14
    a, b, c = 5.0, 3.0, 1.0
    assert abs(aharoni_dz(a, b, c) + aharoni_dz(b, c, a) + aharoni_dz(c, a, b) - 1) < 1e-13
    assert aharoni_dz(10, 10, 1) > 0.8  # thin film, field along the normal


def _oracle(oracle_mod, nx, ny, nz, h, jitter=0.0, seed=1):
    from tetmag_amd import mesh as M
    m = M.kuhn_cuboid(nx, ny, nz, h, jitter, seed)
    return oracle_mod.Oracle(m.xyz, m.fel), m


def _avg_demag_factor(oracle_mod, nx, ny, nz, h, axis):
    o, mesh = _oracle(oracle_mod, nx, ny, nz, h)
    n = o.n
    Ms = 8e5
    Js = np.full(n, MU0 * Ms)
    m = np.zeros(3 * n)
    m[axis * n:(axis + 1) * n] = 1.0
    H, _, _ = o.demag(m, Js, tol=1e-13)
    V = o.nodevol
    Hc = H[axis * n:(axis + 1) * n]
    E = -0.5 * (Js * V * Hc).sum()
    return -(V * Hc).sum() / V.sum() / Ms, E / (0.5 * MU0 * Ms ** 2 * V.sum())


@pytest.mark.parametrize("cells,axis", [((1, 1, 1), 2), ((3, 2, 1), 2), ((3, 2, 1), 0)])
def test_uniform_cuboid_demag_factor(oracle_mod, cells, axis):
    """Volume-averaged field of a uniform cuboid converges to Aharoni's factor
    (P1 error ~h^1.5-2), and the demag energy equals 1/2 mu0 D Ms^2 Vol."""
    a, b, c = cells
    half = np.array([a, b, c], float) * 4.0 / 2
    D = aharoni_dz(*np.roll(half, 2 - axis))
    errs = []
    for k in (4, 8):
        Dh, Eh = _avg_demag_factor(oracle_mod, a * k, b * k, c * k, 4.0 / k, axis)
        assert abs(Eh - Dh) < 1e-9 * Dh  # E = -1/2 sum Js V H.m, consistent with the mean field
        errs.append(abs(Dh - D) / D)
    assert errs[1] < 0.025
    assert errs[1] < errs[0] / 2.2  # converging, not a constant bias


def test_demag_gauge_invariance(oracle_mod):
    """Pinning a different node of the Neumann problem shifts u1 by a constant
    and leaves H unchanged (B 1 = -1)."""
    o, mesh = _oracle(oracle_mod, 6, 5, 4, 2.0, 0.1, 3)
    import pyoracle as O
    from tetmag_amd import mesh as M
    n = o.n
    m = M.init_random_unit(n, 5)
    Js = np.full(n, 1.0)
    H1, u1a, _ = o.demag(m, Js, tol=1e-14, warm=False)
    fx = np.zeros(n, bool)
    fx[0] = True
    sN = O.Solver(*o.replaced(fx))
    sD = o.solvers()[1]
    b = o.div(m, Js)
    u1b = sN.solve(b, tol=1e-14, warm=False)
    d = u1b - u1a
    assert np.ptp(d) < 1e-9 * np.max(np.abs(u1a))  # a constant shift
    assert abs(d.mean()) > 1e-6 * np.max(np.abs(u1a))  # ... that is not zero
    u2b = sD.solve(o.boundary_integral(u1b), tol=1e-14, warm=False)
    H2 = o.grad_field(u1b + u2b)
    assert rel_l2(H2, H1) < 1e-9


def test_exchange_uniform_and_spin_wave(oracle_mod):
    nx, ny, nz, h = 40, 4, 4, 1.0
    o, mesh = _oracle(oracle_mod, nx, ny, nz, h)
    n = o.n
    A = np.full(n, 1.3e-11 / 1e-18)
    mu = np.concatenate([np.full(n, 0.6), np.full(n, 0.8), np.zeros(n)])
    H0 = o.exchange(mu, A)
    scale = 2 * A[0] * np.max(np.abs(o.K)) / o.nodevol.min()
    assert np.max(np.abs(H0)) < 1e-12 * scale
    k = 2 * np.pi / (nx * h)
    x = mesh.xyz[:, 0]
    m = np.concatenate([np.cos(k * x), np.sin(k * x), np.zeros(n)])
    H = o.exchange(m, A)
    interior = (x > 0) & (x < nx * h)
    idx = np.nonzero(interior)[0]
    expect = -2 * A[0] * k * k
    for c in range(2):
        got = H[c * n + idx]
        ref = expect * m[c * n + idx]
        assert np.max(np.abs(got - ref)) < 0.01 * abs(expect)


def test_bulk_dmi_helix(oracle_mod):
    nx, ny, nz, h = 40, 4, 4, 1.0
    o, mesh = _oracle(oracle_mod, nx, ny, nz, h)
    n = o.n
    D = np.full(n, 3e-3 / 1e-9)
    k = 2 * np.pi / (nx * h)
    x = mesh.xyz[:, 0]
    m = np.concatenate([np.zeros(n), np.cos(k * x), np.sin(k * x)])
    H = o.dmi(m, D)
    bn = np.zeros(n, bool)
    bn[o.bnodes] = True
    idx = np.nonzero(~bn)[0]
    expect = 2 * D[0] * k  # curl m = -k m  ->  H = -2 D curl m = 2 D k m
    for c in (1, 2):
        assert np.max(np.abs(H[c * n + idx] - expect * m[c * n + idx])) < 0.01 * expect
    assert np.max(np.abs(H[idx])) < 1e-6 * expect


def test_uniaxial(oracle_mod):
    o, mesh = _oracle(oracle_mod, 3, 3, 2, 1.0)
    n = o.n
    k = np.array([0.36, 0.48, 0.8])
    axis = np.concatenate([np.full(n, k[0]), np.full(n, k[1]), np.full(n, k[2])])
    Ku = np.full(n, 5e4)
    H = o.uniaxial(axis, Ku, axis)
    assert np.allclose(H, 2 * 5e4 * axis, rtol=1e-15)
    perp = np.array([0.8, 0.0, -0.36])
    perp /= np.linalg.norm(perp)
    mp = np.concatenate([np.full(n, perp[0]), np.full(n, perp[1]), np.full(n, perp[2])])
    assert np.max(np.abs(o.uniaxial(mp, Ku, axis))) < 1e-10


def test_llg_macrospin(oracle_mod):
    o, mesh = _oracle(oracle_mod, 1, 1, 1, 1.0)
    n = o.n
    hz, alpha = 3.0, 0.5
    m = np.concatenate([np.ones(n), np.zeros(n), np.zeros(n)])
    H = np.concatenate([np.zeros(n), np.zeros(n), np.full(n, hz)])
    dm = o.llg(m, H, alpha, variant=1).reshape(3, n)
    # -m x H - alpha m x (m x H) = (0, hz, alpha hz) for m = x, H = hz z
    assert np.allclose(dm[0], 0) and np.allclose(dm[1], hz) and np.allclose(dm[2], alpha * hz)
    dm0 = o.llg(m, H, alpha, variant=0).reshape(3, n)
    assert np.allclose(dm0[1], 0) and np.allclose(dm0[2], alpha * hz)
    assert o.max_torque(m, H) == pytest.approx(hz)
    # H || m: no torque
    assert np.max(np.abs(o.llg(m, 2.0 * m, alpha))) == 0.0


def test_gmres_against_dense(oracle_mod):
    o, mesh = _oracle(oracle_mod, 4, 4, 3, 1.0, 0.1, 2)
    import pyoracle as O
    n = o.n
    rng = np.random.default_rng(0)
    for rp, col, val in (o.neumann_matrix(), o.dirichlet_matrix()):
        A = np.zeros((n, n))
        for i in range(n):
            A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
        b = rng.normal(size=n)
        b -= b.mean()
        s = O.Solver(rp, col, val)
        x = s.solve(b, tol=1e-14)
        xd = np.linalg.solve(A, b)
        assert rel_l2(x, xd) < 1e-11
        assert s.last_relres <= 1e-14


def test_interfacial_dmi_cycloid(oracle_mod):
    """Neel cycloid m = (sin kx, 0, cos kx): interior field -2 D k m and
    energy density D k (within the P1 error)."""
    nx, ny, nz, h = 40, 4, 2, 1.0
    o, mesh = _oracle(oracle_mod, nx, ny, nz, h)
    n = o.n
    Dv = np.full(n, 3e-3 / 1e-9)
    k = 2 * np.pi / (nx * h)
    x = mesh.xyz[:, 0]
    m = np.concatenate([np.sin(k * x), np.zeros(n), np.cos(k * x)])
    H = o.idmi(m, Dv)
    bn = np.zeros(n, bool)
    bn[o.bnodes] = True
    inner = ~bn  # surface nodes carry the natural-boundary terms of the energy
    expect = -2 * Dv[0] * k
    for c in (0, 2):
        assert np.max(np.abs(H[c * n:(c + 1) * n][inner] - expect * m[c * n:(c + 1) * n][inner])) < 0.006 * abs(expect)
    e = -0.5 * np.sum(o.nodevol * (H.reshape(3, n) * m.reshape(3, n)).sum(axis=0)) / o.nodevol.sum()
    assert e == pytest.approx(Dv[0] * k, rel=0.03)


def test_interfacial_dmi_is_the_energy_gradient(oracle_mod):
    """H = -(1/V) dE_h/dm for E_h = sum_i D_i [m_z (Q_x m_x + Q_y m_y) - m_x Q_x m_z
    - m_y Q_y m_z]_i with Q_d = tGrad_d^T, written here independently."""
    import scipy.sparse as sp
    o, mesh = _oracle(oracle_mod, 5, 4, 3, 2.0, 0.1, 7)
    n = o.n
    rng = np.random.default_rng(1)
    Dv = 3e6 * (1 + 0.3 * rng.random(n))
    m = rng.normal(size=3 * n)
    Qx = sp.csr_matrix((o.tX, o.col, o.rowptr), shape=(n, n)).T.tocsr()
    Qy = sp.csr_matrix((o.tY, o.col, o.rowptr), shape=(n, n)).T.tocsr()

    def E(v):
        mx, my, mz = v.reshape(3, n)
        return np.sum(Dv * (mz * (Qx @ mx + Qy @ my) - mx * (Qx @ mz) - my * (Qy @ mz)))

    H = o.idmi(m, Dv)
    d = rng.normal(size=3 * n)
    dE = (E(m + d) - E(m - d)) / 2  # E is quadratic: exact
    assert -np.sum(np.tile(o.nodevol, 3) * H * d) == pytest.approx(dE, rel=1e-11)
