"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Both sides run on the same seeded meshes and magnetisations.  The oracle
builds its own operators from (xyz, fel) with its restatement of the
reference setup and solves the reference's row-replaced systems with
GMRES/ILU(0) to 1e-14; the HIP path uses the product's setup, the GPU
Lindholm assembly and CG on the same systems.  Bar: fields within 1e-10
relative L2 (the FP64 parity tolerance of the path).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
GPU_CG_TOL = 1e-13
ORACLE_TOL = 1e-14


def _case(name):
    from tetmag_amd import problem as P
    if name == "cfg0":
        return P.cuboid_20x20x4()
    if name == "small_ku_dmi_zeeman":
        return P.small_problem(nx=7, ny=6, nz=3, h=4.0, jitter=0.1, seed=11, Ku=4e4, D=3e-3,
                               B_ext=(0.01, -0.02, 0.015))
    if name == "ragged":  # n not a multiple of 64, thin in z, jittered
        return P.small_problem(nx=9, ny=3, nz=1, h=2.5, jitter=0.15, seed=5, Ku=1e5, D=-2e-3)
    if name == "idmi":  # interfacial DMI film with bulk DMI off
        p = P.small_problem(nx=10, ny=8, nz=2, h=2.0, jitter=0.1, seed=31, Ku=8e5, D=0.0, B_ext=(0.0, 0.01, 0.0))
        p.mat.Di = 3e-3
        p.Di = np.full(p.n, 3e-3 / p.mesh.scale) * (1 + 0.2 * np.random.default_rng(3).random(p.n))
        return p
    if name == "mid":  # two AMG levels above the dense coarse grid
        return P.small_problem(nx=36, ny=30, nz=8, h=3.0, jitter=0.1, seed=21, Ku=5e4, D=0.0, init="vortex")
    raise KeyError(name)


CASES = [("cfg0", "amg"), ("cfg0", "jacobi"), ("small_ku_dmi_zeeman", "amg"), ("ragged", "amg"),
         ("idmi", "amg"), ("mid", "amg")]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[0]}-{c[1]}")
def pair(request, oracle_mod, gpu_available):
    from tetmag_amd.context import FieldContext
    from tetmag_amd import _native as N
    name, pre = request.param
    print(f"\n[case] {name} precond={pre}", flush=True)
    prob = _case(name)
    orc = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    ctx = FieldContext(prob, tol=GPU_CG_TOL,
                       precond=N.TM_PRECOND_AMG if pre == "amg" else N.TM_PRECOND_JACOBI)
    yield prob, orc, ctx
    ctx.close()


def _oracle_eval(prob, orc, m):
    return orc.eval_rhs(m, Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=prob.hext,
                        alpha=prob.alpha, variant=1, tol=ORACLE_TOL, Di=prob.Di)


def test_bem_matrix_matches_oracle(pair):
    """Reference form: same expression, same operation order as Lindholm.cpp."""
    prob, orc, ctx = pair
    Bg = ctx.bem_matrix()
    Bo = orc.B
    assert rel_l2(Bg, Bo) < 1e-13
    # B 1 = -1 row-wise, to the precision the reference's acos form allows
    assert np.max(np.abs(Bg.sum(axis=1) + 1.0)) < 1e-9


def test_bem_accurate_form(pair):
    """atan2 form: B 1 = -1 to rounding; differs from the reference form only by
    the acos rounding noise of distant triangles."""
    prob, orc, ctx = pair
    from tetmag_amd import _native as N
    ctx.assemble_bem(N.TM_BEM_ACCURATE)
    try:
        Ba = ctx.bem_matrix()
    finally:
        ctx.assemble_bem(N.TM_BEM_REFERENCE)
    assert np.max(np.abs(Ba.sum(axis=1) + 1.0)) < 1e-13
    assert rel_l2(Ba, orc.B) < 1e-8


def test_individual_fields(pair):
    prob, orc, ctx = pair
    m = prob.m0
    ctx.set_mag(m)
    assert rel_l2(ctx.exchange_field(), orc.exchange(m, prob.A)) < FIELD_TOL
    if prob.use_uniaxial:
        assert rel_l2(ctx.uniaxial_field(), orc.uniaxial(m, prob.Ku, prob.kaxis)) < FIELD_TOL
    if prob.use_dmi:
        assert rel_l2(ctx.dmi_field(), orc.dmi(m, prob.D)) < FIELD_TOL
    if prob.use_idmi:
        assert rel_l2(ctx.idmi_field(), orc.idmi(m, prob.Di)) < FIELD_TOL


def test_demag_field_and_potentials(pair):
    prob, orc, ctx = pair
    m = prob.m0
    Hg, st = ctx.demag_field(m)
    Ho, u1o, u2o = orc.demag(m, prob.Js, tol=ORACLE_TOL, warm=False)
    print("iters", st, "oracle", orc.last_iters)
    assert st["relres_neumann"] <= GPU_CG_TOL and st["relres_dirichlet"] <= GPU_CG_TOL
    assert rel_l2(Hg, Ho) < FIELD_TOL
    u1g, u2g = ctx.potentials()
    # identity-row elimination keeps the reference gauge u1(n/2) = b(n/2)
    assert rel_l2(u1g, u1o) < FIELD_TOL
    assert rel_l2(u2g, u2o) < FIELD_TOL


def test_eval_rhs_host(pair):
    prob, orc, ctx = pair
    m = prob.m0
    dm, H, st = ctx.eval_rhs_host(m)
    ref = _oracle_eval(prob, orc, m)
    assert rel_l2(H, ref["Heff"]) < FIELD_TOL
    assert rel_l2(dm, ref["dmdt"]) < FIELD_TOL
    assert abs(st["max_torque"] - ref["max_torque"]) <= 1e-10 * ref["max_torque"]


def test_eval_rhs_device_matches_host(pair):
    import torch
    prob, orc, ctx = pair
    m = prob.m0
    dm_h, H_h, _ = ctx.eval_rhs_host(m)
    with torch.cuda.stream(ctx.stream):
        md = torch.as_tensor(m).to(ctx.device)
    ctx.stream.synchronize()
    dm_d = ctx.eval_rhs(md, want_heff=True)
    ctx.stream.synchronize()
    assert rel_l2(dm_d.cpu().numpy(), dm_h) < 1e-12
    assert rel_l2(ctx.heff.cpu().numpy(), H_h) < 1e-12


def test_noprec_variant(pair):
    prob, orc, ctx = pair
    m = prob.m0
    dm, H, _ = ctx.eval_rhs_host(m, variant=0)
    ref = orc.llg(m, H, prob.alpha, variant=0)
    assert rel_l2(dm, ref) < 1e-13


def test_llg_invariants(pair):
    """m . dm/dtau = 0 and dm/dtau = 0 for H || m."""
    prob, orc, ctx = pair
    n = prob.n
    m = prob.m0
    dm, H, _ = ctx.eval_rhs_host(m)
    dots = dm[:n] * m[:n] + dm[n:2 * n] * m[n:2 * n] + dm[2 * n:] * m[2 * n:]
    assert np.max(np.abs(dots)) <= 1e-12 * np.max(np.abs(dm))


def test_energies(pair):
    """tm_energies_eval against the oracle's energy functions on the oracle's
    fields (densities, J/m^3), plus volume averages of m and of the applied field."""
    prob, orc, ctx = pair
    n = prob.n
    m = prob.m0
    ref = _oracle_eval(prob, orc, m)
    fields = {"Hexc": ref["Hexc"], "Hdem": ref["Hdem"]}
    if "Hdmi" in ref:
        fields["Hdmi"] = ref["Hdmi"]
    fields["Hext"] = np.concatenate([np.full(n, c) for c in prob.hext])
    eo = orc.energies(m, fields, Js=prob.Js, Ku=prob.Ku, axis=prob.kaxis)
    V = orc.nodevol.sum()
    eg = ctx.energies(m)
    assert eg["volume"] == pytest.approx(V, rel=1e-13)
    scale = max(abs(v) for v in eo.values()) / V
    for k, v in eo.items():
        assert abs(eg[k] - v / V) <= 1e-10 * scale, (k, eg[k], v / V)
    tot = sum(eo.values()) / V
    assert abs(eg["total"] - tot) <= 1e-10 * scale
    w = orc.nodevol / V
    for d in range(3):
        assert eg["mean_m"][d] == pytest.approx(float(w @ m[d * n:(d + 1) * n]), abs=1e-13)
        assert eg["mean_h"][d] == pytest.approx(prob.hext[d], abs=1e-9 * (1 + abs(prob.hext[d])))
    # frozen potentials: reusing the last solve gives the same demag energy
    eg2 = ctx.energies(None, recompute_demag=False)
    assert eg2["demag"] == pytest.approx(eg["demag"], rel=1e-12)
