"""Batched evaluation of many magnetisation samples on one mesh (tm_batch_*).

Samples share the operators, the dense boundary matrix and the AMG
hierarchies of a :class:`~tetmag_amd.context.FieldContext`; fields are
``(3, n, S)`` arrays (component, node, sample), potentials ``(n, S)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .context import FieldContext


class BatchContext:
    def __init__(self, ctx: FieldContext, nsamples: int, precision: int = N.TM_PREC_MIXED):
        import torch
        self.ctx = ctx
        self.n = ctx.n
        self.S = int(nsamples)
        self.precision = precision
        h = C.c_void_p()
        N.check(N.lib().tm_batch_create(ctx.h, self.S, int(precision), C.byref(h)), "tm_batch_create")
        self.h = h
        with torch.cuda.stream(ctx.stream):
            self.dmdt = torch.empty((3, self.n, self.S), dtype=torch.float64, device=ctx.device)
            self.heff = torch.empty((3, self.n, self.S), dtype=torch.float64, device=ctx.device)
        ctx.stream.synchronize()
        self.last_stats = N.BatchStats()

    def set_zeeman(self, H3s):
        """Per-sample uniform applied field in A/m, shape (3, S)."""
        h = N.f64(np.asarray(H3s, np.float64).reshape(3, self.S))
        N.check(N.lib().tm_batch_set_zeeman(self.h, N.ptr(h)), "tm_batch_set_zeeman")

    def set_mag(self, m):
        if isinstance(m, np.ndarray):
            a = N.f64(m.reshape(3, self.n, self.S))
            N.check(N.lib().tm_batch_set_mag(self.h, N.ptr(a), N.TM_HOST), "tm_batch_set_mag")
        else:
            N.check(N.lib().tm_batch_set_mag(self.h, N.ptr(m), N.TM_DEVICE), "tm_batch_set_mag")
            self.ctx.synchronize()

    def get_mag(self) -> np.ndarray:
        out = np.empty((3, self.n, self.S))
        N.check(N.lib().tm_batch_get_mag(self.h, N.ptr(out), N.TM_HOST), "tm_batch_get_mag")
        return out

    def eval_rhs(self, m=None, *, alpha=None, variant=N.TM_LLG_CLASSIC, want_heff=False, stats=True):
        alpha = self.ctx.prob.alpha if alpha is None else alpha
        st = self.last_stats if stats else None
        N.check(N.lib().tm_batch_eval_rhs(self.h, N.ptr(m) if m is not None else None, N.TM_DEVICE, float(alpha),
                                          int(variant), N.ptr(self.dmdt), N.TM_DEVICE,
                                          N.ptr(self.heff) if want_heff else None, N.TM_DEVICE,
                                          C.byref(st) if st is not None else None), "tm_batch_eval_rhs")
        return self.dmdt

    def eval_rhs_host(self, m: np.ndarray, alpha=None, variant=N.TM_LLG_CLASSIC):
        alpha = self.ctx.prob.alpha if alpha is None else alpha
        a = N.f64(np.asarray(m).reshape(3, self.n, self.S))
        dm = np.empty((3, self.n, self.S))
        H = np.empty((3, self.n, self.S))
        st = N.BatchStats()
        N.check(N.lib().tm_batch_eval_rhs(self.h, N.ptr(a), N.TM_HOST, float(alpha), int(variant), N.ptr(dm),
                                          N.TM_HOST, N.ptr(H), N.TM_HOST, C.byref(st)), "tm_batch_eval_rhs")
        self.last_stats = st
        return dm, H, st.as_dict()

    def potentials(self):
        u1 = np.empty((self.n, self.S))
        u2 = np.empty((self.n, self.S))
        N.check(N.lib().tm_batch_get_potentials(self.h, N.ptr(u1), N.ptr(u2), N.TM_HOST), "tm_batch_get_potentials")
        return u1, u2

    def step_euler(self, dtau: float, dmdt=None):
        d = self.dmdt if dmdt is None else dmdt
        N.check(N.lib().tm_batch_step_euler(self.h, N.ptr(d), N.TM_DEVICE, float(dtau)), "tm_batch_step_euler")

    def eval_euler(self, dtau: float, *, alpha=None, variant=N.TM_LLG_CLASSIC, stats=True):
        """dm/dt at the batch magnetisation into self.dmdt, then the Euler step
        of m, fused into the field kernel (tm_batch_eval_euler)."""
        alpha = self.ctx.prob.alpha if alpha is None else alpha
        st = self.last_stats if stats else None
        N.check(N.lib().tm_batch_eval_euler(self.h, float(alpha), int(variant), float(dtau), N.ptr(self.dmdt),
                                            N.TM_DEVICE, C.byref(st) if st is not None else None),
                "tm_batch_eval_euler")
        return self.dmdt

    def run_rk4(self, dtau: float, nsteps: int, alpha=None, variant=N.TM_LLG_CLASSIC, use_graph: bool = False):
        """nsteps RK4 steps of every sample on the device (tm_batch_run_rk4);
        use_graph replays one captured step with fixed solver budgets.
        Returns (statistics of the last evaluation, step statistics)."""
        alpha = self.ctx.prob.alpha if alpha is None else alpha
        st = N.BatchStats()
        rk = N.RK4Stats()
        N.check(N.lib().tm_batch_run_rk4(self.h, float(dtau), int(nsteps), float(alpha), int(variant),
                                         int(use_graph), C.byref(st), C.byref(rk)), "tm_batch_run_rk4")
        return st.as_dict(), rk.as_dict()

    def time_kernels(self, reps: int = 20) -> list:
        """Mean ms and algorithmic bytes per launch of the batch's hot kernels,
        each launched alone (tm_batch_time_kernels)."""
        arr = (N.KernelTime * 16)()
        cnt = C.c_int()
        N.check(N.lib().tm_batch_time_kernels(self.h, int(reps), arr, 16, C.byref(cnt)), "tm_batch_time_kernels")
        return [{"kernel": arr[k].name.decode(), "ms": arr[k].ms, "bytes": arr[k].bytes} for k in range(cnt.value)]

    def gemm_info(self) -> dict:
        """The tuned boundary GEMM: own_kernel True means the library's FP64
        MFMA kernel with tile shape `solution`, else rocBLAS solution index
        `solution` (0 = library default; rocBLAS indices may be negative)."""
        sol, own, cand, ks = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        d, b = C.c_double(), C.c_double()
        N.check(N.lib().tm_batch_gemm_info(self.h, C.byref(sol), C.byref(own), C.byref(cand), C.byref(d),
                                           C.byref(b), C.byref(ks)), "tm_batch_gemm_info")
        return {"solution": sol.value, "own_kernel": bool(own.value), "candidates": cand.value,
                "ms_default": d.value, "ms_tuned": b.value, "ksplit": ks.value}

    def gemm_live(self, on: bool):
        """Event pairs around the boundary GEMM of each following evaluation
        (tm_batch_gemm_live).  gemm_live(True) starts recording; gemm_live(False)
        stops and returns (mean ms per launch, launches recorded)."""
        ms, cnt = C.c_double(), C.c_int()
        N.check(N.lib().tm_batch_gemm_live(self.h, int(bool(on)), C.byref(ms), C.byref(cnt)), "tm_batch_gemm_live")
        return ms.value, cnt.value

    def bem_apply(self, x):
        """Boundary product of every sample, y[j][s] = sum_k B[j][k] x[k][s]
        (tm_batch_bem_apply), in the batch's precision.  x: (b, S) numpy
        array (returns numpy) or device tensor (returns a device tensor on the
        context's stream)."""
        b = self.ctx.prob.bnd.bnx
        if isinstance(x, np.ndarray):
            a = N.f64(np.asarray(x).reshape(b, self.S))
            y = np.empty((b, self.S))
            N.check(N.lib().tm_batch_bem_apply(self.h, N.ptr(a), N.TM_HOST, N.ptr(y), N.TM_HOST),
                    "tm_batch_bem_apply")
            return y
        import torch
        with torch.cuda.stream(self.ctx.stream):
            y = torch.empty((b, self.S), dtype=torch.float64, device=self.ctx.device)
        N.check(N.lib().tm_batch_bem_apply(self.h, N.ptr(x), N.TM_DEVICE, N.ptr(y), N.TM_DEVICE),
                "tm_batch_bem_apply")
        return y

    def enable_timers(self, on: bool = True):
        N.check(N.lib().tm_batch_enable_timers(self.h, int(on)), "tm_batch_enable_timers")

    def timers(self, reset: bool = False) -> dict:
        t = N.PhaseTimes()
        N.check(N.lib().tm_batch_get_timers(self.h, C.byref(t), int(reset)), "tm_batch_get_timers")
        return t.as_dict()

    def close(self):
        if getattr(self, "h", None):
            N.lib().tm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_batch(ms) -> np.ndarray:
    """Stack per-sample 3n magnetisations into a (3, n, S) array."""
    ms = [np.asarray(m, np.float64) for m in ms]
    n = ms[0].shape[0] // 3
    return np.ascontiguousarray(np.stack([m.reshape(3, n) for m in ms], axis=-1))


def from_batch(a: np.ndarray, s: int) -> np.ndarray:
    return np.ascontiguousarray(a[:, :, s].reshape(-1))
