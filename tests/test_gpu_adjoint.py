"""Adjoints on the GPU (tm_eval_rhs_vjp, tm_demag_field_vjp,
tm_bem_apply_transposed, tm_solver_solve_transposed) against the FP64 CPU
adjoint of the oracle (itself pinned by finite differences), exact transpose
identities, and gradients through checkpointed RK4 trajectories against
central finite differences (bar 1e-4)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def base(gpu_available, oracle_mod):
    from tetmag_amd import problem as P
    from tetmag_amd.context import FieldContext
    prob = P.small_problem(nx=7, ny=6, nz=3, h=3.0, jitter=0.1, seed=13, Ku=4e4, D=3e-3, B_ext=(0.02, -0.01, 0.03))
    print(f"\n[case] adjoint base n={prob.n}", flush=True)
    o = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    fc = FieldContext(prob, tol=1e-13)
    o.set_B(fc.bem_matrix())
    lam = np.random.default_rng(17).normal(size=3 * prob.n)
    yield prob, o, fc, lam
    fc.close()


def test_bem_transposed(base):
    prob, o, fc, lam = base
    B = fc.bem_matrix()
    y = np.random.default_rng(2).normal(size=prob.bnd.bnx)
    assert rel_l2(fc.bem_apply(y, transposed=True), B.T @ y) < 1e-13
    assert rel_l2(fc.bem_apply(y), B @ y) < 1e-13


def test_solver_transposed(base):
    """A^T x = b for the Dirichlet (many identity rows) and Neumann matrices."""
    import scipy.sparse as sp
    from tetmag_amd import _native as N
    prob, o, fc, lam = base
    L = N.lib()
    for rp, col, val in (prob.dirichlet_matrix(), prob.neumann_matrix()):
        A = sp.csr_matrix((val, col, rp), shape=(prob.n, prob.n))
        h = C.c_void_p()
        N.check(L.tm_solver_create(prob.n, N.ptr(N.i32(rp)), N.ptr(N.i32(col)), N.ptr(N.f64(val)), N.TM_PRECOND_AMG,
                                   C.byref(h)), "tm_solver_create")
        b = np.random.default_rng(4).normal(size=prob.n)
        x = np.zeros(prob.n)
        it = C.c_int()
        rr = C.c_double()
        N.check(L.tm_solver_solve_transposed(h, N.ptr(b), N.TM_HOST, N.ptr(x), N.TM_HOST, 1e-13, 10000, 0,
                                             C.byref(it), C.byref(rr)), "tm_solver_solve_transposed")
        assert np.linalg.norm(A.T @ x - b) <= 1e-11 * np.linalg.norm(b)
        L.tm_solver_destroy(h)


def test_demag_vjp(base):
    prob, o, fc, lam = base
    g = fc.demag_vjp_host(lam)
    assert rel_l2(g, o.demag_vjp(lam, prob.Js)) < TOL
    m = prob.m0
    H, _, _ = o.demag(m, prob.Js, tol=1e-14, warm=False)
    assert g @ m == pytest.approx(lam @ H, rel=1e-10)  # Hdem is linear in m


@pytest.mark.parametrize("variant", [1, 0])
def test_rhs_vjp_matches_cpu_adjoint(base, variant):
    prob, o, fc, lam = base
    m = prob.m0
    r = fc.vjp_host(m, lam, variant=variant)
    ref = o.eval_rhs_vjp(m, lam, Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=prob.hext,
                         alpha=prob.alpha, variant=variant, tol=1e-14)
    assert rel_l2(r["gm"], ref["gm"]) < TOL
    assert rel_l2(r["gH"], ref["gH"]) < TOL
    assert rel_l2(r["g_hext"], ref["g_hext"]) < TOL
    assert rel_l2(r["g_ku"], ref["g_ku"]) < TOL
    assert rel_l2(r["g_d"], ref["g_d"]) < TOL


def test_rhs_vjp_finite_difference(base):
    prob, o, fc, lam = base
    m = prob.m0
    g = fc.vjp_host(m, lam)["gm"]
    d = np.random.default_rng(8).normal(size=m.size)
    eps = 1e-5
    fp = fc.eval_rhs_host(m + eps * d)[0]
    fm = fc.eval_rhs_host(m - eps * d)[0]
    fd = lam @ (fp - fm) / (2 * eps)
    assert g @ d == pytest.approx(fd, rel=1e-6)


def _objective(fc, prob, m0, target, amps, ku_e, d_e, e2n, nsteps, dtau, every):
    import torch
    from tetmag_amd import autograd as AG
    dirs = torch.tensor([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], dtype=torch.float64, device=fc.device)
    h0 = torch.as_tensor(prob.hext, device=fc.device)

    def hext_of_step(k):
        return h0 + amps[k % amps.numel()] * dirs[k % 2]

    mT = AG.integrate(fc, m0, dtau, nsteps, hext_of_step, e2n(ku_e), e2n(d_e), checkpoint_every=every)
    return torch.dot(mT, target)


def test_trajectory_gradients_vs_finite_differences(base):
    """J = <m(T), target> after 8 RK4 steps under 4 piecewise-constant field
    amplitudes; gradients w.r.t. the amplitudes and per-element Ku and D
    against central differences, and checkpointed == plain."""
    import torch
    from tetmag_amd import autograd as AG
    prob, o, fc, lam = base
    dev = fc.device
    e2n = AG.ElementToNode(prob, dev)
    ne = prob.mesh.ntet
    m0 = torch.as_tensor(prob.m0, device=dev)
    target = torch.as_tensor(np.roll(prob.m0, 7), device=dev)
    dtau = 20e-3 * prob.time_scale()
    amps = torch.tensor([3e3, -2e3, 1e3, 4e3], dtype=torch.float64, device=dev, requires_grad=True)
    ku_e = torch.full((ne,), 4e4, dtype=torch.float64, device=dev, requires_grad=True)
    d_e = torch.full((ne,), float(prob.D[0]), dtype=torch.float64, device=dev, requires_grad=True)
    J = _objective(fc, prob, m0, target, amps, ku_e, d_e, e2n, 8, dtau, 3)
    J.backward()
    ga, gk, gd = amps.grad.clone(), ku_e.grad.clone(), d_e.grad.clone()
    with torch.no_grad():
        J0 = _objective(fc, prob, m0, target, amps, ku_e, d_e, e2n, 8, dtau, 100)
        assert float(J0) == pytest.approx(float(J), rel=1e-12)

        def fd(fn, h):
            return (float(fn(h)) - float(fn(-h))) / 2

        for i in range(4):
            h = 1e-4 * 3e3
            e = torch.zeros(4, dtype=torch.float64, device=dev)
            e[i] = 1.0
            v = fd(lambda s: _objective(fc, prob, m0, target, amps + s * e, ku_e, d_e, e2n, 8, dtau, 100), h) / h
            assert float(ga[i]) == pytest.approx(v, rel=1e-4, abs=1e-4 * float(ga.abs().max()))
        rng = np.random.default_rng(3)
        dk = torch.as_tensor(rng.normal(size=ne), device=dev)
        v = fd(lambda s: _objective(fc, prob, m0, target, amps, ku_e + s * dk, d_e, e2n, 8, dtau, 100), 4.0) / 4.0
        assert float(gk @ dk) == pytest.approx(v, rel=1e-4)
        dd = torch.as_tensor(rng.normal(size=ne), device=dev) * float(prob.D[0])
        v = fd(lambda s: _objective(fc, prob, m0, target, amps, ku_e, d_e + s * dd, e2n, 8, dtau, 100), 1e-3) / 1e-3
        assert float(gd @ dd) == pytest.approx(v, rel=1e-4)
    fc.set_ku(prob.Ku)
    fc.set_d(prob.D)
    fc.set_hext(prob.hext)


def test_rhs_vjp_interfacial_dmi(gpu_available, oracle_mod):
    """Interfacial DMI film (bulk DMI off): VJP w.r.t. m and per-node Di."""
    from tetmag_amd import problem as P
    from tetmag_amd.context import FieldContext
    prob = P.small_problem(nx=8, ny=7, nz=2, h=2.0, jitter=0.1, seed=19, Ku=8e5, D=0.0, B_ext=(0.0, 0.01, 0.0))
    prob.Di = np.full(prob.n, 3e-3 / prob.mesh.scale) * (1 + 0.2 * np.random.default_rng(3).random(prob.n))
    print(f"\n[case] interfacial adjoint n={prob.n}", flush=True)
    o = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    fc = FieldContext(prob, tol=1e-13)
    o.set_B(fc.bem_matrix())
    lam = np.random.default_rng(23).normal(size=3 * prob.n)
    m = prob.m0
    r = fc.vjp_host(m, lam)
    ref = o.eval_rhs_vjp(m, lam, Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=prob.hext,
                         alpha=prob.alpha, tol=1e-14, Di=prob.Di)
    assert rel_l2(r["gm"], ref["gm"]) < TOL
    assert rel_l2(r["g_di"], ref["g_di"]) < TOL
    e = fc.energies(m)
    assert e["dmi"] != 0.0
    fc.close()
