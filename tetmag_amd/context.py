"""Device-resident field context: the composition TheLLG + DemagField run on the GPU.

:class:`FieldContext` owns one ``tm_ctx`` (include/tetmag_hip.h) holding the
operators, material data, the dense boundary-element matrix, both demag
solvers and the magnetisation, and exposes the hot path:

* :meth:`FieldContext.eval_rhs` -- one effective-field + LLG right-hand-side
  evaluation with the demag field recomputed (TheLLG::rhs_d ->
  classicVersion_GPU -> evaluateAllEffectiveFields -> totalEffectiveField ->
  ClassicLLG_dev, LLGWrapper.cu:78-82, 213-228, TheLLG.cpp:238-262).

Device buffers are torch tensors on ``cuda:0``; the context runs on a torch
stream so ordering with torch work is by stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .problem import Problem


def _torch():
    import torch
    return torch


class FieldContext:
    def __init__(self, prob: Problem, *, tol: float = 1e-10, maxit: int = 100000, device: int = 0,
                 B: np.ndarray | None = None, stream=None, bem_formula: int = N.TM_BEM_REFERENCE,
                 precond: int = N.TM_PRECOND_AMG, uniaxial: bool | None = None, dmi: bool | None = None,
                 idmi: bool | None = None):
        torch = _torch()
        L = N.lib()
        self.L = L
        self.prob = prob
        self.n = prob.n
        chosen = C.c_int()
        N.check(L.tm_set_device(device, C.byref(chosen)), "tm_set_device")
        self.device = torch.device("cuda", chosen.value)
        self.stream = stream if stream is not None else torch.cuda.Stream(device=self.device)
        h = C.c_void_p()
        N.check(L.tm_ctx_create(self.n, C.byref(h)), "tm_ctx_create")
        self.h = h
        N.check(L.tm_ctx_set_stream(self.h, C.c_void_p(self.stream.cuda_stream)), "tm_ctx_set_stream")
        ops = prob.ops
        rp, col = N.i32(ops.rowptr), N.i32(ops.col)
        self._set_op(N.TM_OP_EXCHANGE, rp, col, prob.exchange_values())
        for o, v in ((N.TM_OP_TGRAD_X, ops.tgx), (N.TM_OP_TGRAD_Y, ops.tgy), (N.TM_OP_TGRAD_Z, ops.tgz),
                     (N.TM_OP_GRAD_X, ops.gx), (N.TM_OP_GRAD_Y, ops.gy), (N.TM_OP_GRAD_Z, ops.gz)):
            self._set_op(o, rp, col, v)
        N.check(L.tm_ctx_set_js(self.h, N.ptr(N.f64(prob.Js))), "tm_ctx_set_js")
        N.check(L.tm_ctx_set_node_volume(self.h, N.ptr(N.f64(prob.ops.nodevol))), "tm_ctx_set_node_volume")
        self.has_uniaxial = prob.use_uniaxial if uniaxial is None else bool(uniaxial)
        self.has_dmi = prob.use_dmi if dmi is None else bool(dmi)
        if self.has_uniaxial:
            N.check(L.tm_ctx_set_uniaxial(self.h, N.ptr(N.f64(prob.kaxis)), N.ptr(N.f64(prob.Ku))),
                    "tm_ctx_set_uniaxial")
        if self.has_dmi:
            # TheLLG.cpp:94-98: curl operators are -tGrad
            for o, v in ((N.TM_OP_CURL_X, ops.tgx), (N.TM_OP_CURL_Y, ops.tgy), (N.TM_OP_CURL_Z, ops.tgz)):
                self._set_op(o, rp, col, -v)
            inv_v = 1.0 / ops.nodevol
            N.check(L.tm_ctx_set_dmi(self.h, N.ptr(N.f64(prob.D)), N.ptr(N.f64(inv_v)),
                                     N.ptr(N.f64(prob.bnd.node_normal)), N.ptr(N.f64(prob.bnd.node_area))),
                    "tm_ctx_set_dmi")
        self.has_idmi = prob.use_idmi if idmi is None else bool(idmi)
        if self.has_idmi:
            N.check(L.tm_ctx_set_dmi_interfacial(self.h, N.ptr(N.f64(prob.Di))), "tm_ctx_set_dmi_interfacial")
        hx, hy, hz = (float(v) for v in prob.hext)
        N.check(L.tm_ctx_set_zeeman(self.h, None, hx, hy, hz), "tm_ctx_set_zeeman")
        bnd = prob.bnd
        N.check(L.tm_ctx_set_boundary(self.h, bnd.bnx, N.ptr(N.i32(bnd.bnodes))), "tm_ctx_set_boundary")
        if B is not None:
            Bf = np.asfortranarray(B, dtype=np.float64)
            N.check(L.tm_ctx_set_bem_dense(self.h, Bf.ctypes.data, bnd.bnx, N.TM_HOST), "tm_ctx_set_bem_dense")
        else:
            self.assemble_bem(bem_formula)
        rpN, colN, valN = prob.neumann_matrix()
        rpD, colD, valD = prob.dirichlet_matrix()
        N.check(L.tm_ctx_set_demag_solvers(self.h, N.ptr(rpN), N.ptr(colN), N.ptr(valN), N.ptr(rpD), N.ptr(colD),
                                           N.ptr(valD), int(precond)), "tm_ctx_set_demag_solvers")
        self.set_solver_params(tol, maxit)
        with torch.cuda.stream(self.stream):
            self.m = torch.as_tensor(N.f64(prob.m0)).to(self.device)
            self.dmdt = torch.empty(3 * self.n, dtype=torch.float64, device=self.device)
            self.heff = torch.empty(3 * self.n, dtype=torch.float64, device=self.device)
        self.stream.synchronize()
        N.check(L.tm_ctx_set_mag(self.h, N.ptr(N.f64(prob.m0)), N.TM_HOST), "tm_ctx_set_mag")
        self.last_stats = N.EvalStats()
        # parameters last applied from torch tensors (autograd), by key; the
        # setters below drop their key
        self._param_cache = {}

    # -- setup -------------------------------------------------------------------
    def _set_op(self, op, rp, col, vals):
        N.check(self.L.tm_ctx_set_op(self.h, op, N.ptr(rp), N.ptr(col), N.ptr(N.f64(vals))), f"tm_ctx_set_op({op})")

    def assemble_bem(self, formula: int = N.TM_BEM_REFERENCE):
        """Dense double-layer matrix assembled on the GPU (BEMprocessing.cpp:358-411)."""
        bnd = self.prob.bnd
        bxyz = N.f64(self.prob.mesh.xyz[bnd.bnodes])
        faces_b = N.i32(bnd.bindex[bnd.faces])
        N.check(self.L.tm_ctx_assemble_bem(self.h, bnd.nbel, N.ptr(bxyz), N.ptr(faces_b), N.ptr(N.f64(bnd.face_area)),
                                           N.ptr(N.f64(bnd.face_normal)), N.ptr(N.f64(bnd.solid_angle)),
                                           int(formula)),
                "tm_ctx_assemble_bem")

    def bem_matrix_device(self):
        """Copy of the dense matrix as a (b, b) column-major device tensor (B[:, j] contiguous)."""
        torch = _torch()
        b = self.prob.bnd.bnx
        with torch.cuda.stream(self.stream):
            T = torch.empty((b, b), dtype=torch.float64, device=self.device)
        self.stream.synchronize()
        N.check(self.L.tm_ctx_get_bem_dense(self.h, N.ptr(T), N.TM_DEVICE), "tm_ctx_get_bem_dense")
        return T.t()

    def bem_matrix(self) -> np.ndarray:
        b = self.prob.bnd.bnx
        B = np.empty((b, b), order="F")
        N.check(self.L.tm_ctx_get_bem_dense(self.h, B.ctypes.data, N.TM_HOST), "tm_ctx_get_bem_dense")
        return B

    def set_solver_params(self, tol: float, maxit: int):
        N.check(self.L.tm_ctx_set_solver_params(self.h, float(tol), int(maxit)), "tm_ctx_set_solver_params")

    def set_mag(self, m):
        """EffFieldGPU::setMagDev: m is a (3n,) numpy array or device tensor."""
        torch = _torch()
        if isinstance(m, np.ndarray):
            N.check(self.L.tm_ctx_set_mag(self.h, N.ptr(N.f64(m)), N.TM_HOST), "tm_ctx_set_mag")
            with torch.cuda.stream(self.stream):
                self.m.copy_(torch.as_tensor(N.f64(m)), non_blocking=False)
        else:
            with torch.cuda.stream(self.stream):
                self.m.copy_(m)
            N.check(self.L.tm_ctx_set_mag(self.h, N.ptr(self.m), N.TM_DEVICE), "tm_ctx_set_mag")

    # -- hot path ----------------------------------------------------------------
    def eval_rhs(self, m=None, *, alpha=None, variant=N.TM_LLG_CLASSIC, want_heff=False, stats=True):
        """One Heff + LLG RHS evaluation on the device.  m: (3n,) device tensor
        (None = current context magnetisation).  Returns the dm/dtau tensor
        (owned by the context, overwritten by the next call)."""
        alpha = self.prob.alpha if alpha is None else alpha
        mptr = N.ptr(m) if m is not None else None
        st = self.last_stats if stats else None
        status = self.L.tm_eval_rhs(self.h, mptr, N.TM_DEVICE, float(alpha), int(variant), N.ptr(self.dmdt),
                                    N.TM_DEVICE, N.ptr(self.heff) if want_heff else None, N.TM_DEVICE,
                                    C.byref(st) if st is not None else None)
        N.check(status, "tm_eval_rhs")
        return self.dmdt

    def eval_rhs_host(self, m: np.ndarray, alpha=None, variant=N.TM_LLG_CLASSIC):
        """Same evaluation with host arrays in and out (synchronous)."""
        alpha = self.prob.alpha if alpha is None else alpha
        n3 = 3 * self.n
        dm = np.empty(n3)
        H = np.empty(n3)
        st = N.EvalStats()
        N.check(self.L.tm_eval_rhs(self.h, N.ptr(N.f64(m)), N.TM_HOST, float(alpha), int(variant), N.ptr(dm),
                                   N.TM_HOST, N.ptr(H), N.TM_HOST, C.byref(st)), "tm_eval_rhs")
        self.last_stats = st
        return dm, H, st.as_dict()

    # -- individual pieces (host arrays) -----------------------------------------
    def _field(self, fn, name):
        out = np.empty(3 * self.n)
        N.check(fn(self.h, N.ptr(out), N.TM_HOST), name)
        return out

    def exchange_field(self):
        return self._field(self.L.tm_exchange_field, "tm_exchange_field")

    def uniaxial_field(self):
        return self._field(self.L.tm_uniaxial_field, "tm_uniaxial_field")

    def dmi_field(self):
        return self._field(self.L.tm_dmi_field, "tm_dmi_field")

    def demag_field(self, m: np.ndarray | None = None):
        H = np.empty(3 * self.n)
        st = N.EvalStats()
        mp = N.ptr(N.f64(m)) if m is not None else None
        N.check(self.L.tm_demag_field(self.h, mp, N.TM_HOST, N.ptr(H), N.TM_HOST, C.byref(st)), "tm_demag_field")
        return H, st.as_dict()

    def potentials(self):
        u1 = np.empty(self.n)
        u2 = np.empty(self.n)
        N.check(self.L.tm_demag_last_potentials(self.h, N.ptr(u1), N.ptr(u2), N.TM_HOST), "tm_demag_last_potentials")
        return u1, u2

    def step_euler(self, dtau: float, dmdt=None):
        """m <- normalize(m + dtau dm/dtau) on the device (dmdt: device tensor,
        default the last eval_rhs result)."""
        d = self.dmdt if dmdt is None else dmdt
        N.check(self.L.tm_ctx_step_euler(self.h, N.ptr(d), N.TM_DEVICE, float(dtau)), "tm_ctx_step_euler")

    def run_rk4(self, dtau: float, nsteps: int, use_graph: bool = True, alpha=None,
                variant=N.TM_LLG_CLASSIC) -> dict:
        """nsteps classical RK4 steps of the context magnetisation on the device
        (tm_ctx_run_rk4); returns the step statistics."""
        alpha = self.prob.alpha if alpha is None else alpha
        st = N.RK4Stats()
        N.check(self.L.tm_ctx_run_rk4(self.h, float(dtau), int(nsteps), float(alpha), int(variant), int(use_graph),
                                      C.byref(st)), "tm_ctx_run_rk4")
        return st.as_dict()

    def integrate(self, tau_span: float, rtol: float = 1e-6, atol: float = 1e-6, h0: float = 0.0,
                  max_steps: int = 1000000, use_graph: bool = True, normalize: bool = True, alpha=None,
                  variant=N.TM_LLG_CLASSIC) -> dict:
        """Adaptive Dormand-Prince 5(4) integration of the context magnetisation
        over tau_span (tm_ctx_integrate; the reference runs CVODE with
        rtol = atol = 1e-6).  h0 <= 0 picks the first step; the returned
        statistics carry h_next for the next window."""
        alpha = self.prob.alpha if alpha is None else alpha
        st = N.OdeStats()
        h = C.c_double(h0)
        N.check(self.L.tm_ctx_integrate(self.h, float(tau_span), float(rtol), float(atol), C.byref(h), int(max_steps),
                                        float(alpha), int(variant), int(use_graph), int(normalize), C.byref(st)),
                "tm_ctx_integrate")
        return st.as_dict()

    def enable_timers(self, on: bool = True):
        N.check(self.L.tm_ctx_enable_timers(self.h, int(on)), "tm_ctx_enable_timers")

    def timers(self, reset: bool = False) -> dict:
        t = N.PhaseTimes()
        N.check(self.L.tm_ctx_get_timers(self.h, C.byref(t), int(reset)), "tm_ctx_get_timers")
        return t.as_dict()

    def amg_levels(self, which: int) -> list:
        """Unknowns per AMG level of the Neumann (0) or Dirichlet (1) solver."""
        nl = C.c_int()
        sizes = np.zeros(16, np.int32)
        N.check(self.L.tm_ctx_solver_amg_info(self.h, which, C.byref(nl), N.ptr(sizes), 16), "tm_ctx_solver_amg_info")
        return sizes[:nl.value].tolist()

    def energies(self, m: np.ndarray | None = None, recompute_demag: bool = True) -> dict:
        """Energy densities (J/m^3) and volume averages of m (None = current m)."""
        e = N.Energies()
        mp = N.ptr(N.f64(m)) if m is not None else None
        N.check(self.L.tm_energies_eval(self.h, mp, N.TM_HOST, int(recompute_demag), C.byref(e)), "tm_energies_eval")
        return e.as_dict()

    def set_freeze_demag(self, on: bool):
        N.check(self.L.tm_ctx_set_freeze_demag(self.h, int(on)), "tm_ctx_set_freeze_demag")

    def refresh_demag(self, m=None):
        st = N.EvalStats()
        if m is None:
            N.check(self.L.tm_ctx_refresh_demag(self.h, None, N.TM_DEVICE, C.byref(st)), "tm_ctx_refresh_demag")
        elif isinstance(m, np.ndarray):
            N.check(self.L.tm_ctx_refresh_demag(self.h, N.ptr(N.f64(m)), N.TM_HOST, C.byref(st)),
                    "tm_ctx_refresh_demag")
        else:
            N.check(self.L.tm_ctx_refresh_demag(self.h, N.ptr(m), N.TM_DEVICE, C.byref(st)), "tm_ctx_refresh_demag")
        return st.as_dict()

    # -- material / field parameters (host arrays) --------------------------------
    def set_hext(self, h3, field3n=None):
        """Applied field (A/m): a uniform (3,) vector plus, optionally, a
        per-node (3n,) field added to it (TheLLG.cpp:320-332 carries the
        pulse, sweep and local fields this way)."""
        self._param_cache.pop("hext", None)
        hx, hy, hz = (float(v) for v in h3)
        fp = N.ptr(N.f64(field3n)) if field3n is not None else None
        N.check(self.L.tm_ctx_set_zeeman(self.h, fp, hx, hy, hz), "tm_ctx_set_zeeman")

    def set_ku(self, Ku):
        """Per-node uniaxial constant (J/m^3) along the problem's axes."""
        self._param_cache.pop("ku", None)
        if not self.has_uniaxial:
            raise RuntimeError("context built without uniaxial anisotropy (pass uniaxial=True)")
        Kd = N.f64(Ku)
        N.check(self.L.tm_ctx_set_uniaxial(self.h, N.ptr(N.f64(self.prob.kaxis)), N.ptr(Kd)), "tm_ctx_set_uniaxial")

    def set_d(self, D):
        """Per-node bulk DMI constant (J/m^2 divided by the mesh scale)."""
        self._param_cache.pop("d", None)
        if not self.has_dmi:
            raise RuntimeError("context built without DMI (pass dmi=True)")
        ops, bnd = self.prob.ops, self.prob.bnd
        N.check(self.L.tm_ctx_set_dmi(self.h, N.ptr(N.f64(D)), N.ptr(N.f64(1.0 / ops.nodevol)),
                                      N.ptr(N.f64(bnd.node_normal)), N.ptr(N.f64(bnd.node_area))), "tm_ctx_set_dmi")

    def set_di(self, Di):
        """Per-node interfacial DMI constant (J/m^2 divided by the mesh scale)."""
        self._param_cache.pop("di", None)
        if not self.has_idmi:
            raise RuntimeError("context built without interfacial DMI (pass idmi=True)")
        N.check(self.L.tm_ctx_set_dmi_interfacial(self.h, N.ptr(N.f64(Di))), "tm_ctx_set_dmi_interfacial")

    def idmi_field(self):
        return self._field(self.L.tm_idmi_field, "tm_idmi_field")

    # -- adjoint ------------------------------------------------------------------
    def vjp_host(self, m: np.ndarray, lam: np.ndarray, alpha=None, variant=N.TM_LLG_CLASSIC, want_ku=None,
                 want_d=None, want_di=None) -> dict:
        """lam^T d(dm/dtau)/d(m, Hext, Ku, D) at m (host arrays, synchronous)."""
        alpha = self.prob.alpha if alpha is None else alpha
        n = self.n
        gm = np.empty(3 * n)
        gh = np.empty(3)
        gf = np.empty(3 * n)
        want_ku = self.has_uniaxial if want_ku is None else want_ku
        want_d = self.has_dmi if want_d is None else want_d
        gku = np.empty(n) if want_ku else None
        gd = np.empty(n) if want_d else None
        want_di = self.has_idmi if want_di is None else want_di
        gdi = np.empty(n) if want_di else None
        st = N.EvalStats()
        N.check(self.L.tm_eval_rhs_vjp(self.h, N.ptr(N.f64(m)), N.TM_HOST, float(alpha), int(variant),
                                       N.ptr(N.f64(lam)), N.TM_HOST, N.ptr(gm), N.TM_HOST, N.ptr(gh), N.ptr(gf),
                                       N.ptr(gku), N.ptr(gd), N.ptr(gdi), C.byref(st)), "tm_eval_rhs_vjp")
        out = {"gm": gm, "g_hext": gh, "gH": gf, "stats": st.as_dict()}
        if want_ku:
            out["g_ku"] = gku
        if want_d:
            out["g_d"] = gd
        if want_di:
            out["g_di"] = gdi
        return out

    def vjp(self, m, lam, alpha=None, variant=N.TM_LLG_CLASSIC, want_ku=False, want_d=False):
        """Device form: m, lam are (3n,) float64 tensors on the context's device.
        Returns (g_m tensor, g_hext numpy (3,), g_ku tensor | None, g_d tensor | None)."""
        torch = _torch()
        alpha = self.prob.alpha if alpha is None else alpha
        n = self.n
        with torch.cuda.stream(self.stream):
            gm = torch.empty(3 * n, dtype=torch.float64, device=self.device)
            gku = torch.empty(n, dtype=torch.float64, device=self.device) if want_ku else None
            gd = torch.empty(n, dtype=torch.float64, device=self.device) if want_d else None
        gh = np.empty(3)
        N.check(self.L.tm_eval_rhs_vjp(self.h, N.ptr(m), N.TM_DEVICE, float(alpha), int(variant), N.ptr(lam),
                                       N.TM_DEVICE, N.ptr(gm), N.TM_DEVICE, N.ptr(gh), None, N.ptr(gku), N.ptr(gd),
                                       None, None), "tm_eval_rhs_vjp")
        return gm, gh, gku, gd

    def demag_vjp_host(self, v: np.ndarray) -> np.ndarray:
        g = np.empty(3 * self.n)
        N.check(self.L.tm_demag_field_vjp(self.h, N.ptr(N.f64(v)), N.TM_HOST, N.ptr(g), N.TM_HOST, None),
                "tm_demag_field_vjp")
        return g

    def bem_apply(self, x: np.ndarray, transposed: bool = False) -> np.ndarray:
        y = np.empty(self.prob.bnd.bnx)
        fn = self.L.tm_bem_apply_transposed if transposed else self.L.tm_bem_apply
        N.check(fn(self.h, N.ptr(N.f64(x)), N.TM_HOST, N.ptr(y), N.TM_HOST), "tm_bem_apply")
        return y

    def get_mag(self) -> np.ndarray:
        out = np.empty(3 * self.n)
        N.check(self.L.tm_ctx_get_mag(self.h, N.ptr(out), N.TM_HOST), "tm_ctx_get_mag")
        return out

    def synchronize(self):
        N.check(self.L.tm_ctx_synchronize(self.h), "tm_ctx_synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.L.tm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
