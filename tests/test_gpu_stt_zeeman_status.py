"""Spin-transfer torque against the oracle's STT twin, per-node applied
fields (single-sample and batched) against the oracle, and the non-finite
status path, all through the C ABI."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(oracle_mod, gpu_available):
    from tetmag_amd import problem as P
    prob = P.small_problem(nx=9, ny=7, nz=4, h=3.0, jitter=0.1, seed=17, Ku=6e4, D=2.5e-3,
                           B_ext=(0.004, -0.01, 0.02))
    orc = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    return prob, orc


def _mat(o, vals):
    return sp.csr_matrix((vals, o.col, o.rowptr), shape=(o.n, o.n))


def test_stt_against_oracle_twin(setup):
    """UTermSTT_GPU / STT_term_LLG_hst (EffFieldGPU.cu:107-148, 434-482) against
    ref_stt_uterm / ref_stt_llg (TheLLG.cpp:201-220, 402-416)."""
    from tetmag_amd.gpu import EffFieldGPU
    prob, o = setup
    n = o.n
    print(f"\n[case] stt n={n}", flush=True)
    g = EffFieldGPU(n)
    rng = np.random.default_rng(5)
    eta = rng.normal(size=(3, n)) * 1e-2
    g.setSTTDataOnDevice(_mat(o, o.gX), _mat(o, o.gY), _mat(o, o.gZ), eta[0], eta[1], eta[2])
    m = prob.m0
    g.setMagDev(m.reshape(3, n).T)
    U = g.UTermSTT_GPU()
    Uo = o.stt_uterm(m, eta)
    assert rel_l2(U.T.reshape(-1), Uo) < 1e-13
    for alpha, beta in ((0.3, 0.04), (0.02, 0.5)):
        out = g.STT_term_LLG_hst(U, alpha, beta)
        assert rel_l2(out, o.stt_llg(m, Uo, alpha, beta)) < 1e-13


def _oracle_eval(prob, o, m, hext):
    return o.eval_rhs(m, Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=hext,
                      alpha=prob.alpha, tol=1e-14)


def _node_field(prob):
    """A smooth non-uniform applied field (A/m), e.g. a local field."""
    x = prob.mesh.xyz
    c = x.mean(axis=0)
    r2 = np.sum((x - c) ** 2, axis=1)
    bump = np.exp(-r2 / 40.0)
    return np.concatenate([3e4 * bump, -2e4 * bump * np.sin(0.3 * x[:, 0]), 1e4 * np.cos(0.2 * x[:, 1])])


def test_per_node_zeeman_against_oracle(setup):
    from tetmag_amd.context import FieldContext
    prob, o = setup
    n = prob.n
    print(f"\n[case] per-node zeeman n={n}", flush=True)
    Hn = _node_field(prob)
    ctx = FieldContext(prob, tol=1e-13)
    try:
        ctx.set_hext(prob.hext, Hn)
        dm, H, st = ctx.eval_rhs_host(prob.m0)
        tot = Hn + np.repeat(prob.hext, n)
        ref = _oracle_eval(prob, o, prob.m0, tot)
        assert rel_l2(H, ref["Heff"]) < 1e-10
        assert rel_l2(dm, ref["dmdt"]) < 1e-10
        e = ctx.energies(prob.m0)
        V = o.nodevol.sum()
        ez = o.energies(prob.m0, {"Hexc": ref["Hexc"], "Hext": tot}, Js=prob.Js)["zeeman"] / V
        assert e["zeeman"] == pytest.approx(ez, rel=1e-10)
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_per_node_zeeman_batched(setup, prec):
    """The context's per-node field is added to every sample of a batch, on
    top of each sample's uniform field."""
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    from tetmag_amd.context import FieldContext
    # a mesh with AMG levels above the dense coarse solve (the batched
    # solvers need at least one)
    prob = P.small_problem(nx=16, ny=12, nz=4, h=3.0, jitter=0.1, seed=17, Ku=6e4, D=2.5e-3,
                           B_ext=(0.004, -0.01, 0.02))
    print(f"\n[case] per-node zeeman batched {prec} n={prob.n}", flush=True)
    Hn = _node_field(prob)
    ms = [prob.m0] + P.smooth_random_fields(prob.mesh, 2, seed=12, lam_min=8.0, lam_max=40.0)
    Hs = np.array([[0.0, 2e4, -1e4], [0.0, 0.0, 5e3], [0.0, -1e4, 3e4]])  # (3, S)
    ctx = FieldContext(prob, tol=1e-12)
    try:
        bc = BatchContext(ctx, len(ms), N.TM_PREC_FP64 if prec == "fp64" else N.TM_PREC_MIXED)
        ctx.set_hext((0.0, 0.0, 0.0), Hn)  # after creation: read at every evaluation
        bc.set_zeeman(Hs)
        dm, H, _ = bc.eval_rhs_host(to_batch(ms))
        tol = 1e-9 if prec == "fp64" else 1e-6
        for s, m in enumerate(ms):
            ctx.set_hext(Hs[:, s], Hn)
            d1, h1, _ = ctx.eval_rhs_host(m)
            assert rel_l2(from_batch(H, s), h1) < tol
            assert rel_l2(from_batch(dm, s), d1) < tol
        bc.close()
    finally:
        ctx.close()


def test_nonfinite_magnetisation_returns_status_and_recovers(setup):
    """A NaN in m makes the Neumann CG see a NaN right-hand side: the device
    check stops it (every kernel still exits normally), tm_eval_rhs returns
    TM_ERR_NONFINITE, and the next finite call is unaffected."""
    from tetmag_amd import _native as N
    from tetmag_amd.context import FieldContext
    prob, o = setup
    n = prob.n
    print("\n[case] nonfinite status", flush=True)
    ctx = FieldContext(prob, tol=1e-13)
    try:
        ctx.eval_rhs_host(prob.m0)  # warm start in place
        bad = prob.m0.copy()
        bad[n // 2] = np.nan
        with pytest.raises(N.TetmagError) as ei:
            ctx.eval_rhs_host(bad)
        assert ei.value.status == N.TM_ERR_NONFINITE
        dm, H, st = ctx.eval_rhs_host(prob.m0)
        ref = _oracle_eval(prob, o, prob.m0, prob.hext)
        assert rel_l2(H, ref["Heff"]) < 1e-10
        assert np.all(np.isfinite(dm))
    finally:
        ctx.close()
