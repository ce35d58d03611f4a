"""The differentiation workload at its stated mesh and material: a disk of
diameter 120 nm and thickness 1 nm (layered prism mesh at 0.6 nm: 95,238
nodes, 377,166 tetrahedra, 64,121 boundary nodes, a 32.9 GB dense B), interfacial DMI
D = 3e-3 J/m^2, Ku = 8e5 J/m^3 along z, Ms = 1.1e6 A/m, A = 1.6e-11 J/m,
alpha = 0.3, a seeded Neel skyrmion of radius 15 nm; 40 piecewise-constant
in-plane pulse amplitudes (seeded, |B| <= 20 mT); objective: the
volume-weighted overlap of the final m with the skyrmion shifted by 20 nm.

The horizon is 1 ps (40 pulses of 25 fs, RK4 at 2.5 fs, 400 steps,
checkpoints every 100 steps) instead of the workload's 2 ns: the stiffest
exchange mode of the 0.6 nm cells (omega ~ 5.5e14 rad/s) limits RK4 to
~5 fs, so 2 ns is ~8e5 steps and hours per forward pass, and central
differences need many passes.  The mesh, material, controls and objective
are the stated ones.

  * gradients by the adjoint (transposed gradient/divergence, B^T, adjoint
    CG solves, LLG vector-Jacobian product) with respect to the 40 pulse
    amplitudes (three of them) and the per-element Ku and D (directional
    derivatives) against central differences with h = 1e-4 (relative),
    agreement <= 1e-4;
  * one right-hand-side vector-Jacobian product at this size against the
    FP64 CPU adjoint of the oracle (its own operators, its own B, SA-AMG
    solves of the transposed systems).
"""
import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

NSTEPS = 400
STEPS_PER_PULSE = 10
MU0 = 1.256637061435917295376e-6


def log(msg):
    print(f"[config4 {time.strftime('%H:%M:%S')}] {msg}", flush=True)


@pytest.fixture(scope="module")
def cfg4(gpu_available):
    from tetmag_amd import problem as P
    from tetmag_amd.context import FieldContext
    prob = P.disk_skyrmion()
    log(f"mesh {prob.summary()}")
    fc = FieldContext(prob, tol=1e-12)
    log("GPU context ready")
    yield prob, fc
    fc.close()


def _target(prob, shift=20.0, radius=15.0):
    """The initial Neel skyrmion (theta = 2 atan(R / r), core down, in-plane
    part along r_hat) with its centre moved by +20 nm along x."""
    xyz = prob.mesh.xyz
    cx = 0.5 * (xyz[:, 0].min() + xyz[:, 0].max()) + shift
    cy = 0.5 * (xyz[:, 1].min() + xyz[:, 1].max())
    x, y = xyz[:, 0] - cx, xyz[:, 1] - cy
    r = np.hypot(x, y)
    th = 2.0 * np.arctan2(radius, r)
    with np.errstate(invalid="ignore", divide="ignore"):
        ux, uy = np.where(r > 0, x / r, 0.0), np.where(r > 0, y / r, 0.0)
    m = np.stack([np.sin(th) * ux, np.sin(th) * uy, np.cos(th)], axis=1)
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    return np.ascontiguousarray(m.T.reshape(-1))


def _objective(fc, prob, m0, wtarget, amps, ku_e, d_e, e2n, every):
    import torch
    from tetmag_amd import autograd as AG
    dtau = 2.5e-3 * prob.time_scale()
    ex = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64, device=fc.device)

    def hext_of_step(k):
        return amps[k // STEPS_PER_PULSE] / MU0 * ex

    mT = AG.integrate(fc, m0, dtau, NSTEPS, hext_of_step, ku=e2n(ku_e), di=e2n(d_e), checkpoint_every=every)
    return torch.dot(mT, wtarget)


def test_gradients_vs_central_differences(cfg4):
    import torch
    from tetmag_amd import autograd as AG
    prob, fc = cfg4
    dev = fc.device
    n = prob.n
    e2n = AG.ElementToNode(prob, dev)
    ne = prob.mesh.ntet
    m0 = torch.as_tensor(prob.m0, device=dev)
    V = np.tile(prob.ops.nodevol / prob.ops.nodevol.sum(), 3)
    wtarget = torch.as_tensor(V * _target(prob), device=dev)
    rng = np.random.default_rng(2028)
    amps = torch.as_tensor(0.02 * (2 * rng.random(40) - 1), device=dev).requires_grad_(True)
    ku0, d0 = 8e5, 3e-3 / prob.mesh.scale
    ku_e = torch.full((ne,), ku0, dtype=torch.float64, device=dev, requires_grad=True)
    d_e = torch.full((ne,), d0, dtype=torch.float64, device=dev, requires_grad=True)
    t0 = time.time()
    J = _objective(fc, prob, m0, wtarget, amps, ku_e, d_e, e2n, 100)
    log(f"forward {time.time() - t0:.1f}s, J = {float(J):.12e}")
    t0 = time.time()
    J.backward()
    log(f"backward {time.time() - t0:.1f}s")
    ga, gk, gd = amps.grad.clone(), ku_e.grad.clone(), d_e.grad.clone()
    with torch.no_grad():
        def fd(fn, h):
            return (float(fn(h)) - float(fn(-h))) / (2 * h)

        t0 = time.time()
        for i in (0, 17, 39):
            e = torch.zeros(40, dtype=torch.float64, device=dev)
            e[i] = 1.0
            h = 1e-4 * 0.02
            v = fd(lambda s: _objective(fc, prob, m0, wtarget, amps + s * e, ku_e, d_e, e2n, NSTEPS), h)
            log(f"dJ/da[{i}]: adjoint {float(ga[i]):.10e}  central difference {v:.10e}")
            assert float(ga[i]) == pytest.approx(v, rel=1e-4, abs=1e-4 * float(ga.abs().max()))
        dk = torch.as_tensor(rng.normal(size=ne), device=dev) * ku0
        v = fd(lambda s: _objective(fc, prob, m0, wtarget, amps, ku_e + s * dk, d_e, e2n, NSTEPS), 1e-4)
        log(f"dJ along a per-element Ku direction: adjoint {float(gk @ dk):.10e}  central difference {v:.10e}")
        assert float(gk @ dk) == pytest.approx(v, rel=1e-4)
        dd = torch.as_tensor(rng.normal(size=ne), device=dev) * d0
        v = fd(lambda s: _objective(fc, prob, m0, wtarget, amps, ku_e, d_e + s * dd, e2n, NSTEPS), 1e-4)
        log(f"dJ along a per-element D direction: adjoint {float(gd @ dd):.10e}  central difference {v:.10e}")
        assert float(gd @ dd) == pytest.approx(v, rel=1e-4)
        log(f"central differences {time.time() - t0:.1f}s")
    fc.set_ku(prob.Ku)
    fc.set_di(prob.Di)
    fc.set_hext(prob.hext)


def test_rhs_vjp_against_cpu_adjoint(cfg4, oracle_mod):
    prob, fc = cfg4
    n = prob.n
    o = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    # the oracle's own 64,121 x 64,121 matrix, assembled in blocks of rows
    t0 = time.time()
    bnx = o.bnx
    B = np.empty((bnx, bnx), order="F")
    for r0 in range(0, bnx, 8192):
        rows = np.arange(r0, min(r0 + 8192, bnx))
        B[r0:r0 + rows.shape[0], :] = o.bem_rows(rows)
        log(f"oracle B rows {rows[-1] + 1} of {bnx} ({time.time() - t0:.1f}s)")
    o.set_B(B)
    del B
    rng = np.random.default_rng(7)
    lam = rng.normal(size=3 * n)
    m = prob.m0
    t0 = time.time()
    r = fc.vjp_host(m, lam, want_ku=True)
    ref = o.eval_rhs_vjp(m, lam, Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=prob.hext,
                         alpha=prob.alpha, tol=1e-12, Di=prob.Di)
    log(f"VJP on both sides {time.time() - t0:.1f}s")
    errs = {"gm": rel_l2(r["gm"], ref["gm"]), "g_ku": rel_l2(r["g_ku"], ref["g_ku"]),
            "g_di": rel_l2(r["g_di"], ref["g_di"])}
    log("relative L2 vs the CPU adjoint: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    for k, v in errs.items():
        assert v < 1e-8, k
