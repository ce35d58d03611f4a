"""torch.autograd bindings of the right-hand side and its adjoint.

``rhs(fc, m, hext, ku, d, di=...)`` evaluates dm/dtau = LLG(m, Heff(m; hext,
Ku, D, Di)) through ``tm_eval_rhs`` and differentiates through
``tm_eval_rhs_vjp``: gradients flow to m, to a uniform applied field ``hext``
(3,), and to per-node ``ku`` / ``d`` (bulk DMI) / ``di`` (interfacial DMI)
(n,).  ``ElementToNode`` maps per-element Ku or D to nodes
(volume-weighted, the lumped form of an element-wise material), so gradients
with respect to element values follow by autograd.  ``integrate`` advances m
with classical RK4 steps (renormalised) in checkpointed segments: the forward
keeps only segment boundaries and recomputes each segment in the backward.

Streams: the library works on the field context's stream; every call here
orders it after torch's current stream and torch's current stream after it.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np
import torch

from . import _native as N
from .context import FieldContext


def _sync_in(fc: FieldContext):
    fc.stream.wait_stream(torch.cuda.current_stream(fc.device))


def _sync_out(fc: FieldContext):
    torch.cuda.current_stream(fc.device).wait_stream(fc.stream)


def _upload(fc: FieldContext, key: str, t, setter):
    """setter(t) unless t (the same tensor object, unmodified since) is what
    the context holds already: a trajectory passes the same per-node Ku / D
    tensors to every stage, so they cross the host once, not at every stage.
    The context's own setters drop the record (FieldContext._param_cache)."""
    ent = fc._param_cache.get(key)
    if ent is not None and ent[0]() is t and ent[1] == t._version:
        return
    setter(t.detach().double().cpu().numpy())
    fc._param_cache[key] = (weakref.ref(t), t._version)


def _apply_params(fc: FieldContext, hext, ku, d, di=None):
    if di is not None:
        _upload(fc, "di", di, fc.set_di)
    if hext is not None:
        _upload(fc, "hext", hext, fc.set_hext)
    if ku is not None:
        _upload(fc, "ku", ku, fc.set_ku)
    if d is not None:
        _upload(fc, "d", d, fc.set_d)


class LLGRhs(torch.autograd.Function):
    """dm/dtau of the field context ``fc`` (see module docstring)."""

    @staticmethod
    def forward(ctx, m, hext, ku, d, di, fc, alpha, variant):
        _apply_params(fc, hext, ku, d, di)
        m = m.detach().contiguous()
        out = torch.empty_like(m)
        _sync_in(fc)
        N.check(fc.L.tm_eval_rhs(fc.h, N.ptr(m), N.TM_DEVICE, float(alpha), int(variant), N.ptr(out), N.TM_DEVICE,
                                 None, N.TM_DEVICE, None), "tm_eval_rhs")
        _sync_out(fc)
        ctx.fc, ctx.alpha, ctx.variant = fc, alpha, variant
        ctx.has = (hext is not None, ku is not None, d is not None, di is not None)
        ctx.save_for_backward(m, *(t for t in (hext, ku, d, di) if t is not None))
        return out

    @staticmethod
    def backward(ctx, g):
        fc = ctx.fc
        saved = list(ctx.saved_tensors)
        m = saved.pop(0)
        hext = saved.pop(0) if ctx.has[0] else None
        ku = saved.pop(0) if ctx.has[1] else None
        d = saved.pop(0) if ctx.has[2] else None
        di = saved.pop(0) if ctx.has[3] else None
        _apply_params(fc, hext, ku, d, di)
        g = g.detach().contiguous()
        gm = torch.empty_like(m)
        want_ku = ku is not None and ctx.needs_input_grad[2]
        want_d = d is not None and ctx.needs_input_grad[3]
        gku = torch.empty(fc.n, dtype=torch.float64, device=m.device) if want_ku else None
        gd = torch.empty(fc.n, dtype=torch.float64, device=m.device) if want_d else None
        want_di = di is not None and ctx.needs_input_grad[4]
        gdi = torch.empty(fc.n, dtype=torch.float64, device=m.device) if want_di else None
        gh = np.empty(3)
        _sync_in(fc)
        N.check(fc.L.tm_eval_rhs_vjp(fc.h, N.ptr(m), N.TM_DEVICE, float(ctx.alpha), int(ctx.variant), N.ptr(g),
                                     N.TM_DEVICE, N.ptr(gm), N.TM_DEVICE, N.ptr(gh), None, N.ptr(gku), N.ptr(gd),
                                     N.ptr(gdi), None), "tm_eval_rhs_vjp")
        _sync_out(fc)
        g_h = None
        if hext is not None and ctx.needs_input_grad[1]:
            g_h = torch.as_tensor(gh, dtype=hext.dtype).to(hext.device)
        return gm, g_h, gku, gd, gdi, None, None, None


def rhs(fc: FieldContext, m, hext=None, ku=None, d=None, alpha=None, variant=N.TM_LLG_CLASSIC, di=None):
    alpha = fc.prob.alpha if alpha is None else alpha
    return LLGRhs.apply(m, hext, ku, d, di, fc, alpha, variant)


# -- per-element material --------------------------------------------------------
def element_volumes(prob) -> np.ndarray:
    x = prob.mesh.xyz
    f = prob.mesh.fel
    a = x[f[:, 1]] - x[f[:, 0]]
    b = x[f[:, 2]] - x[f[:, 0]]
    c = x[f[:, 3]] - x[f[:, 0]]
    return np.abs(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0


class ElementToNode:
    """value_i = sum_{e containing i} (V_e / 4) value_e / V_i, with V_i the
    node volume (sum_e V_e / 4): the lumped-mass form of an element-wise
    coefficient, differentiable in the element values."""

    def __init__(self, prob, device):
        ve = element_volumes(prob)
        f = prob.mesh.fel.astype(np.int64)
        self.node = torch.as_tensor(f.reshape(-1), device=device)
        w = np.repeat(ve / 4.0, 4)
        vn = np.zeros(prob.n)
        np.add.at(vn, f.reshape(-1), w)
        self.w = torch.as_tensor(w / vn[f.reshape(-1)], device=device)
        self.n = prob.n

    def __call__(self, values_e):
        v = torch.repeat_interleave(values_e, 4) * self.w
        return torch.zeros(self.n, dtype=values_e.dtype, device=values_e.device).index_add(0, self.node, v)


# -- trajectory -----------------------------------------------------------------
def normalize(m, n):
    M = m.view(3, n)
    return (M / torch.linalg.vector_norm(M, dim=0, keepdim=True)).reshape(-1)


def rk4_step(fc, m, dtau, hext=None, ku=None, d=None, alpha=None, variant=N.TM_LLG_CLASSIC, di=None):
    f = lambda x: rhs(fc, x, hext, ku, d, alpha, variant, di)
    k1 = f(m)
    k2 = f(m + 0.5 * dtau * k1)
    k3 = f(m + 0.5 * dtau * k2)
    k4 = f(m + dtau * k3)
    return normalize(m + (dtau / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4), fc.n)


def integrate(fc, m0, dtau, nsteps, hext_of_step=None, ku=None, d=None, alpha=None, variant=N.TM_LLG_CLASSIC,
              checkpoint_every: int = 100, di=None):
    """RK4 trajectory from m0; ``hext_of_step(k)`` gives the (3,) applied field
    of step k (a tensor, possibly depending on control parameters).  Segments
    of ``checkpoint_every`` steps are recomputed in the backward pass."""
    from torch.utils.checkpoint import checkpoint

    def segment(m, k0, k1, ku_, d_, di_, *hs):
        for j, k in enumerate(range(k0, k1)):
            m = rk4_step(fc, m, dtau, hs[j] if hs else None, ku_, d_, alpha, variant, di_)
        return m

    m = m0
    k = 0
    while k < nsteps:
        k1 = min(nsteps, k + checkpoint_every)
        hs = [hext_of_step(j) for j in range(k, k1)] if hext_of_step is not None else []
        if torch.is_grad_enabled():
            m = checkpoint(segment, m, k, k1, ku, d, di, *hs, use_reentrant=False)
        else:
            m = segment(m, k, k1, ku, d, di, *hs)
        k = k1
    return m
