"""Unstructured meshes on the GPU path, against the oracle.

  * a seeded jittered Delaunay disk (80 nm x 10 nm, h = 2.5) whose matrix rows
    hold up to ~23 entries: the plain form against the oracle, and 64-sample
    batches (FP64 and mixed) against the single-sample FP64 path, which runs
    the batched V-cycle's multi-chunk level-0 smoother and long-row slot walks
    that Kuhn cuboids (at most 15 entries per row) never reach;
  * the reference's own example meshes (examples/example2/sp4_.vtk, muMAG
    standard problem 4, 16,599 of 16,602 nodes on the surface; and
    examples/example3/sp_stt.vtk, rows up to 32 entries), read from the
    compressed fixtures in tests/golden/: plain form and a small mixed batch
    against the oracle, and the spin-transfer-torque terms on sp_stt.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_eval(prob, o, m, tol=1e-13):
    return o.eval_rhs(m, Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=prob.hext,
                      alpha=prob.alpha, tol=tol, Di=prob.Di)


def _disk():
    from tetmag_amd import mesh as M, problem as P
    mesh = M.disk_delaunay(80.0, 10.0, 2.5, 0.15, 5)
    mat = P.Material(Ku=3e4, ku_axis=(0.6, 0.0, 0.8), D=2e-3, B_ext=(0.01, 0.0, -0.005))
    return P.mesh_problem("disk-80x10-h2.5", mesh, mat, "random", 5)


def _reference_mesh(name, **mat):
    from tetmag_amd import meshio as IO, problem as P
    mesh, _ = IO.read_vtk(os.path.join(GOLDEN, name))
    return P.mesh_problem(name.split(".")[0], mesh, P.Material(**mat), "random", 13)


CASES = {
    "disk": _disk,
    "sp4": lambda: _reference_mesh("sp4_.vtk.xz", Ms=8e5, A=1.3e-11, B_ext=(-0.0246, 0.0043, 0.0)),
    "sp_stt": lambda: _reference_mesh("sp_stt.vtk.xz", Ms=8e5, A=1.3e-11, Ku=1e4, ku_axis=(0.0, 0.0, 1.0)),
}


@pytest.fixture(scope="module", params=list(CASES))
def case(request, oracle_mod, gpu_available):
    from tetmag_amd.context import FieldContext
    name = request.param
    prob = CASES[name]()
    rl = np.diff(prob.ops.rowptr)
    print(f"\n[case] {name} n={prob.n} b={prob.bnd.bnx} max row {rl.max()} rows>15 {(rl > 15).sum()}", flush=True)
    orc = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    ctx = FieldContext(prob, tol=1e-13)
    yield name, prob, orc, ctx
    ctx.close()


def test_plain_against_oracle(case):
    name, prob, orc, ctx = case
    if name == "disk":
        assert np.diff(prob.ops.rowptr).max() > 16
    assert rel_l2(ctx.bem_matrix(), orc.B) < 1e-13
    dm, H, st = ctx.eval_rhs_host(prob.m0)
    ref = _oracle_eval(prob, orc, prob.m0)
    print(name, st, "oracle iterations", orc.last_iters, flush=True)
    assert rel_l2(H, ref["Heff"]) < 1e-10
    assert rel_l2(dm, ref["dmdt"]) < 1e-10
    u1, u2 = ctx.potentials()
    assert rel_l2(u1 + u2, ref["u1"] + ref["u2"]) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_against_single_sample(case, prec):
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    name, prob, orc, ctx = case
    S = 64 if name == "disk" else 8
    ms = [prob.m0] + P.smooth_random_fields(prob.mesh, S - 1, seed=21, lam_min=10.0, lam_max=80.0)
    ctx.set_solver_params(1e-13, 100000)  # both sides to 1e-13: results agree to cond(A) x 1e-13
    try:
        bc = BatchContext(ctx, S, N.TM_PREC_FP64 if prec == "fp64" else N.TM_PREC_MIXED)
        dm, H, st = bc.eval_rhs_host(to_batch(ms))
        bc.close()
        print(name, prec, st["outer_neumann"], st["inner_neumann"], st["outer_dirichlet"], st["inner_dirichlet"],
              flush=True)
        for s in sorted({0, S // 4, S // 2, S - 1}):
            d1, h1, _ = ctx.eval_rhs_host(ms[s])
            assert rel_l2(from_batch(H, s), h1) < 1e-8, s
            assert rel_l2(from_batch(dm, s), d1) < 1e-8, s
        assert st["max_relres_neumann"] <= 1e-12 and st["max_relres_dirichlet"] <= 1e-12
    finally:
        ctx.set_solver_params(1e-13, 100000)


def test_batch_sample_against_oracle(case):
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    name, prob, orc, ctx = case
    ms = [prob.m0] + P.smooth_random_fields(prob.mesh, 3, seed=22, lam_min=10.0, lam_max=80.0)
    bc = BatchContext(ctx, len(ms), N.TM_PREC_MIXED)
    dm, H, st = bc.eval_rhs_host(to_batch(ms))
    bc.close()
    for s in (1, 3):
        ref = _oracle_eval(prob, orc, ms[s])
        assert rel_l2(from_batch(H, s), ref["Heff"]) < 1e-9
        assert rel_l2(from_batch(dm, s), ref["dmdt"]) < 1e-9


def test_stt_on_reference_mesh(case):
    """Zhang-Li terms with a uniform current along x: eta = jc btimesMs invJs
    (ProgramSpecs.cpp:393-421), against the oracle's STT twin."""
    from tetmag_amd.gpu import EffFieldGPU
    name, prob, orc, ctx = case
    if name != "sp_stt":
        pytest.skip("STT runs on the reference's STT example mesh")
    n = prob.n
    mat = lambda v: sp.csr_matrix((v, orc.col, orc.rowptr), shape=(n, n))  # noqa: E731
    mu0, mub, gamma0, e = 1.256637061435917295376e-6, 9.27400968e-24, 221276.14886372426, 1.602176634e-19
    P_, beta, jc = 0.5, 0.04, 1.0
    btimesMs = P_ * mub * mu0 / (gamma0 * e * (1 + beta * beta)) / 1e-9
    eta = np.zeros((3, n))
    eta[0] = 1e12 * jc * btimesMs / prob.Js
    g = EffFieldGPU(n)
    g.setSTTDataOnDevice(mat(orc.gX), mat(orc.gY), mat(orc.gZ), eta[0], eta[1], eta[2])
    g.setMagDev(prob.m0.reshape(3, n).T)
    U = g.UTermSTT_GPU()
    Uo = orc.stt_uterm(prob.m0, eta)
    assert rel_l2(U.T.reshape(-1), Uo) < 1e-13
    assert rel_l2(g.STT_term_LLG_hst(U, prob.alpha, beta), orc.stt_llg(prob.m0, Uo, prob.alpha, beta)) < 1e-13
