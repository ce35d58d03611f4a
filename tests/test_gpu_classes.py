"""The reference class interfaces (EffFieldGPU, GpuPotential, SetDevice) over
the C ABI against the oracle, called the way TheLLG / DemagField call them."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(oracle_mod, gpu_available):
    from tetmag_amd import problem as P
    prob = P.small_problem(nx=8, ny=7, nz=4, h=3.0, jitter=0.1, seed=31, Ku=6e4, D=2.5e-3)
    orc = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    return prob, orc


def _mat(o, vals):
    return sp.csr_matrix((vals, o.col, o.rowptr), shape=(o.n, o.n))


def test_set_device():
    from tetmag_amd.gpu import SetDevice
    d = SetDevice(123)  # out of range -> falls back to device 0 (SetDevice.cu:54-57)
    assert d.set() and d.mydev == 0


def test_efffieldgpu_fields_and_llg(setup):
    from tetmag_amd.gpu import EffFieldGPU
    prob, o = setup
    n = o.n
    g = EffFieldGPU(n)
    # TheLLG.cpp:88-108: XC = -2 diag(A/V) K; curl ops = -tGrad; DMI data; uniaxial
    rows = np.repeat(np.arange(n), np.diff(o.rowptr))
    XC = _mat(o, -2.0 * (prob.A / o.nodevol)[rows] * o.K)
    g.setExchangeMatOnDev(XC)
    g.setGradientMatsOnDev(_mat(o, -o.tX), _mat(o, -o.tY), _mat(o, -o.tZ))
    g.setDMIdata(prob.D, 1.0 / o.nodevol, o.node_normal.reshape(3, n).T, o.node_area)
    g.setUniaxialAnisotropy(prob.kaxis.reshape(3, n).T, prob.Ku)
    m = prob.m0
    g.setMagDev(m.reshape(3, n).T)
    Hx = g.ExchangeFieldGPU()
    assert Hx.shape == (n, 3)
    assert rel_l2(Hx.T.reshape(-1), o.exchange(m, prob.A)) < 1e-12
    assert rel_l2(g.UniaxialAnisotropyField().T.reshape(-1), o.uniaxial(m, prob.Ku, prob.kaxis)) < 1e-13
    assert rel_l2(g.DMIField().T.reshape(-1), o.dmi(m, prob.D)) < 1e-12
    assert np.all(g.CubicAnisotropyField() == 0.0)
    H = np.random.default_rng(2).normal(size=3 * n) * 1e5
    assert rel_l2(g.ClassicLLG_hst(H.reshape(3, n).T, 0.3), o.llg(m, H, 0.3, 1)) < 1e-14
    assert rel_l2(g.LLG_noPrec_hst(H.reshape(3, n).T, 0.3), o.llg(m, H, 0.3, 0)) < 1e-14
    assert g.MaxTorque(H.reshape(3, n).T) == pytest.approx(o.max_torque(m, H), rel=1e-14)
    dev = g.ClassicLLG_dev(H.reshape(3, n).T, 0.3)
    assert rel_l2(dev.cpu().numpy(), o.llg(m, H, 0.3, 1)) < 1e-14
    M2 = 2.5 * m.reshape(3, n).T.copy()
    g.NormalizeMag(M2, n)
    assert np.allclose(np.linalg.norm(M2, axis=1), 1.0, atol=1e-15)


def test_stt_terms(setup):
    """UTermSTT_GPU and STT_term_LLG (EffFieldGPU.cu:107-148, 434-482) against
    the CPU formulas TheLLG.cpp:201-220 and 402-416."""
    from tetmag_amd.gpu import EffFieldGPU
    prob, o = setup
    n = o.n
    g = EffFieldGPU(n)
    rng = np.random.default_rng(4)
    eta = rng.normal(size=(3, n))
    g.setSTTDataOnDevice(_mat(o, o.gX), _mat(o, o.gY), _mat(o, o.gZ), eta[0], eta[1], eta[2])
    m = prob.m0
    g.setMagDev(m.reshape(3, n).T)
    U = g.UTermSTT_GPU()
    G = [_mat(o, o.gX), _mat(o, o.gY), _mat(o, o.gZ)]
    mc = m.reshape(3, n)
    Uref = np.stack([eta[0] * (G[0] @ mc[c]) + eta[1] * (G[1] @ mc[c]) + eta[2] * (G[2] @ mc[c]) for c in range(3)])
    assert rel_l2(U.T, Uref) < 1e-13
    alpha, beta = 0.3, 0.04
    out = g.STT_term_LLG_hst(U, alpha, beta).reshape(3, n)
    mm = mc.T
    mxu = np.cross(mm, U)
    ref = -(beta - alpha) * mxu - (1 + alpha * beta) * np.cross(mm, mxu)
    assert rel_l2(out.T, ref) < 1e-13


def test_gpupotential(setup):
    from tetmag_amd.gpu import GpuPotential
    prob, o = setup
    n = o.n
    gp = GpuPotential()
    gp.setGradientMatrices(_mat(o, o.tX), _mat(o, o.tY), _mat(o, o.tZ), _mat(o, o.gX), _mat(o, o.gY), _mat(o, o.gZ),
                           prob.Js)
    m = prob.m0
    # GpuPotential::calcDivM does not divide by mu0 (DemagField.cpp:135 does)
    mu0 = 1.256637061435917295376e-6
    assert rel_l2(gp.calcDivM(m.reshape(3, n).T), o.div(m, prob.Js) * mu0) < 1e-13
    u = np.random.default_rng(8).normal(size=n)
    assert rel_l2(gp.calcGradientField(u).T.reshape(-1), o.grad_field(u)) < 1e-13


def test_temporaries_survive_the_call(setup):
    """Arrays converted on the fly (non-contiguous or non-float64 inputs) stay
    alive until the C call has read them."""
    import gc
    from tetmag_amd.gpu import EffFieldGPU
    prob, o = setup
    n = o.n
    g = EffFieldGPU(n)
    m = prob.m0
    g.setMagDev(m.reshape(3, n).T)
    H = np.random.default_rng(3).normal(size=(n, 3))
    for _ in range(5):
        out = g.ClassicLLG_hst(np.asfortranarray(H)[:, :], 0.2)  # converted copy inside the wrapper
        gc.collect()
        ref = o.llg(m, np.ascontiguousarray(H.T).reshape(-1), 0.2, variant=1)
        assert rel_l2(out, ref) < 1e-14
