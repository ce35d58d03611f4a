"""GPUSolver / SpMatCUDA through the C ABI: solutions of the reference's
row-replaced demag matrices against the oracle, preconditioner behaviour and
the error statuses."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mats(oracle_mod, gpu_available):
    from tetmag_amd import mesh as M
    m = M.kuhn_cuboid(30, 24, 6, 3.0, 0.1, 17)
    o = oracle_mod.Oracle(m.xyz, m.fel)
    return o, m


def _csr(rp, col, val, n):
    return sp.csr_matrix((val, col, rp), shape=(n, n))


@pytest.mark.parametrize("which", ["neumann", "dirichlet"])
@pytest.mark.parametrize("pre", ["amg", "jacobi"])
def test_solver_matches_oracle(mats, oracle_mod, which, pre):
    from tetmag_amd.gpu import GPUSolver
    from tetmag_amd import mesh as M
    o, m = mats
    rp, col, val = o.neumann_matrix() if which == "neumann" else o.dirichlet_matrix()
    rng = np.random.default_rng(3)
    b = o.div(M.init_random_unit(o.n, 9), np.ones(o.n)) if which == "neumann" else rng.normal(size=o.n)
    ref = oracle_mod.Solver(rp, col, val).solve(b, tol=1e-14)
    s = GPUSolver()
    s.setMatrix(_csr(rp, col, val, o.n))
    s.setCGTolerance(1e-13)
    s.definePreconditioner(pre)
    assert s.compute()
    s.setLoadVector(b)
    s.solve()
    print(which, pre, "iters", s.iters, "relres", s.error)
    assert s.wasSuccessful()
    assert rel_l2(s.result(), ref) < 1e-10
    if pre == "amg":
        print("levels", s.amgLevels())
        assert s.iters < 30
        assert len(s.amgLevels()) >= 2


def test_amg_warm_start_and_fewer_iterations(mats):
    from tetmag_amd.gpu import GPUSolver
    o, m = mats
    rp, col, val = o.neumann_matrix()
    rng = np.random.default_rng(1)
    b = rng.normal(size=o.n)
    it = {}
    for pre in ("jacobi", "amg"):
        s = GPUSolver()
        s.setMatrix(_csr(rp, col, val, o.n))
        s.setCGTolerance(1e-10)
        s.definePreconditioner(pre)
        assert s.compute()
        s.setLoadVector(b)
        s.solve()
        it[pre] = s.iters
        first = s.iters
        s.solve()  # warm start from the stored solution: already converged
        assert s.iters <= 1
        assert first > 0
    assert it["amg"] * 4 < it["jacobi"]


def test_successive_right_hand_sides_reuse_earlier_solutions(mats):
    """b_k = b0 + 0.1 k b1: once two solutions are kept they span every later
    right-hand side, so the later solves start at the answer; every result
    still meets the tolerance on its own residual."""
    from tetmag_amd.gpu import GPUSolver
    o, m = mats
    rp, col, val = o.neumann_matrix()
    A = _csr(rp, col, val, o.n)
    rng = np.random.default_rng(5)
    b0, b1 = rng.normal(size=o.n), rng.normal(size=o.n)
    b0[o.n // 2] = b1[o.n // 2] = 0.0
    s = GPUSolver()
    s.setMatrix(A)
    s.setCGTolerance(1e-10)
    s.definePreconditioner("amg")
    assert s.compute()
    its = []
    for k in range(6):
        b = b0 + 0.1 * k * b1
        s.setLoadVector(b)
        s.solve()
        assert s.wasSuccessful()
        x = s.result()
        assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
        its.append(s.iters)
    print("iterations", its)
    assert its[0] > 5 and max(its[3:]) <= 1


def test_nonsymmetric_matrix_is_rejected(gpu_available):
    from tetmag_amd.gpu import GPUSolver
    A = sp.csr_matrix(np.array([[4.0, 1.0, 0.0], [0.5, 4.0, 1.0], [0.0, 1.0, 4.0]]))
    s = GPUSolver()
    s.setMatrix(A)
    assert not s.compute()
    from tetmag_amd import _native as N
    assert s.status == N.TM_ERR_UNSUPPORTED


def test_maxiter_status(mats):
    from tetmag_amd.gpu import GPUSolver
    from tetmag_amd import _native as N
    o, m = mats
    rp, col, val = o.dirichlet_matrix()
    s = GPUSolver()
    s.setMatrix(_csr(rp, col, val, o.n))
    s.setCGTolerance(1e-14)
    s.definePreconditioner("jacobi")
    s.maxit = 3
    assert s.compute()
    s.setLoadVector(np.random.default_rng(0).normal(size=o.n))
    s.solve()
    assert s.status == N.TM_ERR_MAXITER and s.iters == 3 and not s.wasSuccessful()


def test_spmat_rectangular_and_empty_rows(gpu_available):
    from tetmag_amd.gpu import SpMatCUDA
    rng = np.random.default_rng(5)
    for shape in ((1, 1), (70, 130), (129, 64), (300, 300)):
        A = sp.random(*shape, density=0.05, random_state=7, format="csr")
        A = A.tolil()
        A[0, :] = 0  # an empty row
        A = A.tocsr()
        x = rng.normal(size=shape[1])
        y = SpMatCUDA(A).mvp(x)
        assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-13)


def test_bad_arguments(gpu_available):
    import ctypes as C
    from tetmag_amd import _native as N
    L = N.lib()
    h = C.c_void_p()
    assert L.tm_ctx_create(0, C.byref(h)) == N.TM_ERR_BAD_ARG
    assert L.tm_ctx_create(10, C.byref(h)) == N.TM_OK
    rp = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10], np.int32)
    col = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 11], np.int32)  # column out of range
    v = np.ones(10)
    assert L.tm_ctx_set_op(h, N.TM_OP_EXCHANGE, N.ptr(rp), N.ptr(col), N.ptr(v)) == N.TM_ERR_BAD_ARG
    out = np.empty(30)
    assert L.tm_exchange_field(h, N.ptr(out), N.TM_HOST) == N.TM_ERR_NOT_READY
    assert L.tm_ctx_set_op(h, 99, N.ptr(rp), N.ptr(col), N.ptr(v)) == N.TM_ERR_BAD_ARG
    L.tm_ctx_destroy(h)
