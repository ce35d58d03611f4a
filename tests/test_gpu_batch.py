"""Batched evaluation (tm_batch_*): every sample of a batch matches the
single-sample FP64 path and the oracle; the mixed-precision form (FP32 inner
CG + FP64 refinement) reaches the FP64 answer; per-sample applied fields."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def base(gpu_available):
    from tetmag_amd import problem as P
    from tetmag_amd.context import FieldContext
    prob = P.small_problem(nx=24, ny=20, nz=6, h=3.0, jitter=0.1, seed=41, Ku=5e4, D=2e-3,
                           B_ext=(0.01, 0.0, -0.02))
    ms = [prob.m0] + P.smooth_random_fields(prob.mesh, 4, seed=3, lam_min=10.0, lam_max=60.0)
    # solves to 1e-13: two independently converged solutions differ by up to
    # cond(A) times the tolerance, so batch and single-sample results are
    # compared at 1e-8
    ctx = FieldContext(prob, tol=1e-13)
    yield prob, ctx, ms
    ctx.close()


PREC = {"fp64": "TM_PREC_FP64", "mixed": "TM_PREC_MIXED", "bem32": "TM_PREC_MIXED_BEM32",
        "bemi8": "TM_PREC_MIXED_BEMI8"}


def prec_id(name):
    from tetmag_amd import _native as N
    return getattr(N, PREC[name])


@pytest.mark.parametrize("prec", ["fp64", "mixed", "bemi8"])
def test_batch_matches_single_sample(base, prec):
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    prob, ctx, ms = base
    bc = BatchContext(ctx, len(ms), prec_id(prec))
    dm, H, st = bc.eval_rhs_host(to_batch(ms))
    print(prec, st, bc.gemm_info())
    # mixed: FP32 inner CG; bemi8: the boundary product from int8 digits as well
    tol = 1e-8 if prec == "fp64" else 1e-6
    for s, m in enumerate(ms):
        d1, h1, st1 = ctx.eval_rhs_host(m)
        assert rel_l2(from_batch(H, s), h1) < tol
        assert rel_l2(from_batch(dm, s), d1) < tol
        assert st["max_torque"][s] == pytest.approx(st1["max_torque"], rel=10 * tol)
    assert st["max_relres_neumann"] <= 1e-12 and st["max_relres_dirichlet"] <= 1e-12
    if prec != "fp64":
        assert st["outer_neumann"] >= 2  # refinement steps took place
    bc.close()


def test_batch_against_oracle(base, oracle_mod):
    from tetmag_amd import _native as N
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    prob, ctx, ms = base
    orc = oracle_mod.Oracle(prob.mesh.xyz, prob.mesh.fel)
    bc = BatchContext(ctx, 2, N.TM_PREC_FP64)
    dm, H, st = bc.eval_rhs_host(to_batch(ms[:2]))
    for s in range(2):
        ref = orc.eval_rhs(ms[s], Js=prob.Js, A=prob.A, Ku=prob.Ku, axis=prob.kaxis, D=prob.D, Hext=prob.hext,
                           alpha=prob.alpha, tol=1e-14)
        assert rel_l2(from_batch(H, s), ref["Heff"]) < 1e-9
        assert rel_l2(from_batch(dm, s), ref["dmdt"]) < 1e-9
    bc.close()


def test_per_sample_zeeman(base):
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    from tetmag_amd.context import FieldContext
    prob, ctx, ms = base
    S = 3
    bc = BatchContext(ctx, S, N.TM_PREC_MIXED)
    Hs = np.array([[1e4, -2e4, 0.0], [0.0, 5e3, 3e4], [-1e4, 0.0, 2e4]])  # (3, S)
    bc.set_zeeman(Hs)
    dm, H, st = bc.eval_rhs_host(to_batch([ms[0]] * S))
    for s in range(S):
        # the applied field enters Heff additively
        d = from_batch(H, s) - from_batch(H, 0)
        n = prob.n
        expect = np.concatenate([np.full(n, Hs[c, s] - Hs[c, 0]) for c in range(3)])
        assert np.max(np.abs(d - expect)) < 1e-6 * np.max(np.abs(from_batch(H, 0)))
    bc.close()


def test_full_batch_and_euler(base):
    """64 samples (one wavefront of samples) advance together."""
    import torch
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    prob, ctx, ms = base
    ms64 = P.smooth_random_fields(prob.mesh, 64, seed=9, lam_min=10.0, lam_max=60.0)
    bc = BatchContext(ctx, 64, N.TM_PREC_MIXED)
    bc.set_mag(to_batch(ms64))
    for _ in range(3):
        bc.eval_rhs(None)
        bc.step_euler(1e-3)
    ctx.synchronize()
    st = bc.last_stats.as_dict()
    assert st["max_relres_neumann"] <= 1e-12
    m = bc.get_mag()
    assert np.allclose(np.linalg.norm(m, axis=0), 1.0, atol=1e-12)
    # sample 17 after the batch steps equals three single-sample steps
    d1 = None
    m17 = ms64[17]
    for _ in range(3):
        d1, _, _ = ctx.eval_rhs_host(m17)
        m17 = m17 + 1e-3 * d1
        n = prob.n
        l = np.sqrt(m17[:n] ** 2 + m17[n:2 * n] ** 2 + m17[2 * n:] ** 2)
        m17 = m17 / np.concatenate([l, l, l])
    assert rel_l2(from_batch(m, 17), m17) < 1e-8
    bc.close()


def test_batch_rk4_matches_single_samples(base):
    """Batched RK4 (FP64, per-sample applied fields) equals the single-sample
    on-device RK4 of each sample."""
    from tetmag_amd import _native as N
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    from tetmag_amd.context import FieldContext
    prob, ctx, ms = base
    S = 3
    dtau = 20e-3 * prob.time_scale()
    Hs = np.array([[2e4, -1e4, 0.0], [0.0, 3e4, -2e4], [1e4, 0.0, 5e3]])
    bc = BatchContext(ctx, S, N.TM_PREC_FP64)
    bc.set_zeeman(Hs)
    bc.set_mag(to_batch(ms[1:1 + S]))
    st, rk = bc.run_rk4(dtau, 4)
    mb = bc.get_mag()
    bc.close()
    assert st["max_relres_neumann"] <= 1e-12 and rk["sync_steps"] == 4
    for s in range(S):
        fc = FieldContext(prob, tol=1e-13)
        fc.set_hext(Hs[:, s])  # the per-sample field replaces the context's uniform one
        fc.set_mag(ms[1 + s])
        fc.run_rk4(dtau, 4, use_graph=False)
        assert rel_l2(from_batch(mb, s), fc.get_mag()) < 1e-8
        fc.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed", "bemi8"])
def test_batch_rk4_graph_matches_eager(base, prec):
    """One batched RK4 step captured in a HIP graph (fixed solver budgets,
    kernels gated past convergence) and replayed: every replay is checked
    against the eager host-polled steps of an identical batch, with the
    program's own eager work between the replays."""
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch
    prob, ctx, ms = base
    S = 64
    pid = prec_id(prec)
    fields = P.smooth_random_fields(prob.mesh, S, seed=31, lam_min=10.0, lam_max=60.0)
    dtau = 20e-3 * prob.time_scale()
    Hs = np.stack([np.linspace(-3e4, 3e4, S), np.zeros(S), np.full(S, 1e4)])
    ga, gb = BatchContext(ctx, S, pid), BatchContext(ctx, S, pid)
    for b in (ga, gb):
        b.set_zeeman(Hs)
        b.set_mag(to_batch(fields))
    info = ga.gemm_info()
    print(prec, info, flush=True)
    total = {"graph_steps": 0, "captures": 0, "redone_steps": 0}
    for rep in range(3):  # replays separated by eager work and synchronisations
        _, ra = ga.run_rk4(dtau, 3, use_graph=True)
        _, rb = gb.run_rk4(dtau, 3, use_graph=False)
        for k in total:
            total[k] += ra[k]
        print(rep, ra, flush=True)
        assert rb["graph_steps"] == 0
        ctx.synchronize()
        d = rel_l2(ga.get_mag(), gb.get_mag())
        assert d < (1e-11 if prec == "fp64" else 1e-9), (rep, d)
        ga.eval_rhs_host(ga.get_mag())  # eager work between replays
    assert total["graph_steps"] >= 4 and total["captures"] >= 1
    ga.close()
    gb.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_trajectory_with_kept_solutions(base, prec):
    """Euler steps of a batch and of the single-sample path, both starting
    their solves from kept earlier solutions: the states agree, and the
    solves after the first few need fewer iterations than the first."""
    from tetmag_amd import _native as N
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    prob, ctx, ms = base
    S = len(ms)
    dtau = 20e-3 * prob.time_scale()
    bc = BatchContext(ctx, S, N.TM_PREC_FP64 if prec == "fp64" else N.TM_PREC_MIXED)
    bc.set_mag(to_batch(ms))
    iters = []
    for _ in range(8):
        bc.eval_rhs(None, stats=True)
        st = bc.last_stats.as_dict()
        iters.append(st["inner_neumann"] + st["inner_dirichlet"])
        assert st["max_relres_neumann"] <= 1e-12 and st["max_relres_dirichlet"] <= 1e-12
        bc.step_euler(dtau)
    mb = bc.get_mag()
    bc.close()
    print(prec, "inner iterations per evaluation", iters)
    assert max(iters[5:]) < iters[0]
    for s in (0, S - 1):
        ctx.set_mag(ms[s])
        for _ in range(8):
            ctx.eval_rhs(None)
            ctx.step_euler(dtau)
        assert rel_l2(from_batch(mb, s), ctx.get_mag()) < 1e-8


@pytest.mark.parametrize("prec,S", [("fp64", 64), ("mixed", 64), ("fp64", 3), ("bem32", 64), ("bemi8", 64),
                                    ("bemi8", 3)])
def test_batch_bem_apply(base, prec, S):
    """tm_batch_bem_apply against the dense FP64 product B X on the host: the
    boundary GEMM the batched evaluation runs (tuned MFMA kernel or rocBLAS for
    S = 64, rocBLAS otherwise; the int8 digit-slice kernel), host and device
    operands."""
    import torch
    from tetmag_amd.batch import BatchContext
    prob, ctx, _ = base
    bc = BatchContext(ctx, S, prec_id(prec))
    print(prec, S, bc.gemm_info())
    B = ctx.bem_matrix()
    rng = np.random.default_rng(5)
    X = rng.standard_normal((B.shape[0], S))
    ref = B @ X
    Y = bc.bem_apply(X)
    # bem32: FP32 products summed in FP64 per 512-column chunk; bemi8: five
    # base-2^7 digits per value (~2^-35 of the row / column maxima per
    # entry) and the digit products with p + q <= 4
    tol = {"bem32": 1e-5, "bemi8": 2e-7}.get(prec, 1e-13)
    print("rel err", rel_l2(Y, ref))
    assert rel_l2(Y, ref) < tol
    Xd = torch.from_numpy(X).to(ctx.device)
    torch.cuda.synchronize()
    Yd = bc.bem_apply(Xd)
    ctx.synchronize()
    assert np.array_equal(Yd.cpu().numpy(), Y)  # same kernels, same order
    Y2 = bc.bem_apply(X[::-1].copy())  # a second call on new values
    assert rel_l2(Y2, B @ X[::-1]) < tol
    bc.close()


def test_wide_batch_medium_mesh():
    """64 samples in the mixed form on a 22k-node mesh (the 16-byte-lane
    kernels, deeper AMG hierarchy, DMI): every checked sample matches the
    single-sample FP64 path."""
    from tetmag_amd import _native as N, problem as P
    from tetmag_amd.batch import BatchContext, to_batch, from_batch
    from tetmag_amd.context import FieldContext
    prob = P.small_problem(nx=48, ny=40, nz=10, h=3.0, jitter=0.1, seed=52, Ku=5e4, D=2e-3,
                           B_ext=(0.0, 0.02, 0.0))
    ctx = FieldContext(prob, tol=1e-13)
    print("AMG levels", ctx.amg_levels(0), ctx.amg_levels(1))
    ms = P.smooth_random_fields(prob.mesh, 64, seed=8, lam_min=10.0, lam_max=80.0)
    bc = BatchContext(ctx, 64, N.TM_PREC_MIXED)
    dm, H, st = bc.eval_rhs_host(to_batch(ms))
    assert st["max_relres_neumann"] <= 1e-12 and st["max_relres_dirichlet"] <= 1e-12
    for s in (0, 21, 42, 63):
        d1, h1, _ = ctx.eval_rhs_host(ms[s])
        assert rel_l2(from_batch(H, s), h1) < 1e-8
        assert rel_l2(from_batch(dm, s), d1) < 1e-8
    # the kernel timing hook: every hot kernel timed, and the next evaluation
    # is unaffected by the state the timing runs leave behind
    kt = bc.time_kernels(3)
    names = [k["kernel"] for k in kt]
    print("kernel times", kt)
    assert len(kt) >= 6 and names[0] == "k_b_div" and names[1].startswith("k_b_eval")
    assert all(k["ms"] > 0 and k["bytes"] > 0 for k in kt)
    dm2, H2, st2 = bc.eval_rhs_host(to_batch(ms))
    assert st2["max_relres_neumann"] <= 1e-12
    assert rel_l2(H2, H) < 1e-10
    bc.close()
    ctx.close()


@pytest.mark.parametrize("prec", ["mixed", "bemi8"])
def test_fused_euler_step_matches_eval_then_step(base, prec):
    """tm_batch_eval_euler (the Euler step written by the field kernel into a
    second buffer) gives the numbers of tm_batch_eval_rhs followed by
    tm_batch_step_euler, step after step."""
    from tetmag_amd import problem as P
    from tetmag_amd.batch import BatchContext, to_batch
    prob, ctx, _ = base
    S = 64
    fields = P.smooth_random_fields(prob.mesh, S, seed=17, lam_min=10.0, lam_max=60.0)
    dtau = 20e-3 * prob.time_scale()
    a, b = BatchContext(ctx, S, prec_id(prec)), BatchContext(ctx, S, prec_id(prec))
    for bc in (a, b):
        bc.set_mag(to_batch(fields))
    for _ in range(4):
        a.eval_rhs(None, stats=False)
        a.step_euler(dtau)
        b.eval_euler(dtau, stats=False)
    ctx.synchronize()
    ma, mb = a.get_mag(), b.get_mag()
    da, db = a.dmdt.cpu().numpy(), b.dmdt.cpu().numpy()
    print(prec, "max |m diff|", np.max(np.abs(ma - mb)), "max |dm/dt diff|", np.max(np.abs(da - db)))
    assert np.max(np.abs(ma - mb)) <= 1e-14  # |m| = 1: the same arithmetic on the same dm/dt
    assert np.max(np.abs(da - db)) <= 1e-14 * np.max(np.abs(da))
    assert np.all(np.isfinite(mb))
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "bemi8"])
def test_gemm_live_events(base, prec):
    """tm_batch_gemm_live records one event pair per evaluation around the
    boundary GEMM and reports their mean; recording changes no result."""
    from tetmag_amd import problem as P
    from tetmag_amd.batch import BatchContext, to_batch
    prob, ctx, _ = base
    S = 8
    fields = P.smooth_random_fields(prob.mesh, S, seed=29, lam_min=10.0, lam_max=60.0)
    dtau = 20e-3 * prob.time_scale()
    a, b = BatchContext(ctx, S, prec_id(prec)), BatchContext(ctx, S, prec_id(prec))
    for bc in (a, b):
        bc.set_mag(to_batch(fields))
    a.gemm_live(True)
    for _ in range(3):
        a.eval_euler(dtau, stats=False)
        b.eval_euler(dtau, stats=False)
    ms, cnt = a.gemm_live(False)
    print(prec, "live GEMM ms", ms, "launches", cnt)
    assert cnt == 3 and ms > 0.0
    a.eval_euler(dtau, stats=False)  # recording stopped: no more pairs
    assert a.gemm_live(False)[1] == 3
    b.eval_euler(dtau, stats=False)
    ctx.synchronize()
    assert np.array_equal(a.get_mag(), b.get_mag())
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "bemi8"])
def test_bem_apply_leaves_evaluations_unchanged(base, prec):
    """A boundary product on host operands between two evaluations changes
    neither: the Dirichlet right-hand side stays zero off the boundary (its
    window dots read only the boundary rows) and the batch's buffers are
    rewritten before they are read."""
    from tetmag_amd import problem as P
    from tetmag_amd.batch import BatchContext, to_batch
    prob, ctx, _ = base
    S = 64
    fields = P.smooth_random_fields(prob.mesh, S, seed=31, lam_min=10.0, lam_max=60.0)
    dtau = 20e-3 * prob.time_scale()
    a, b = BatchContext(ctx, S, prec_id(prec)), BatchContext(ctx, S, prec_id(prec))
    for bc in (a, b):
        bc.set_mag(to_batch(fields))
    X = np.random.default_rng(7).standard_normal((ctx.bem_matrix().shape[0], S))
    for k in range(3):
        a.eval_euler(dtau, stats=False)
        b.eval_euler(dtau, stats=False)
        b.bem_apply(X)
    ctx.synchronize()
    ma, mb = a.get_mag(), b.get_mag()
    da, db = a.dmdt.cpu().numpy(), b.dmdt.cpu().numpy()
    print(prec, "rel dm/dt diff", rel_l2(db, da), "max |m diff|", np.max(np.abs(ma - mb)))
    # the two batches tune their boundary GEMMs separately (the summation
    # order of the split K ranges may differ)
    assert rel_l2(db, da) < 1e-10
    assert np.max(np.abs(ma - mb)) < 1e-12
    a.close()
    b.close()
