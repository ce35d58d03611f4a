"""ctypes binding of libtetmag_hip.so (the C ABI declared in include/tetmag_hip.h).

The library is the only compute path: if it cannot be loaded every call
raises :class:`NativeLibraryError`; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TETMAG_HIP_LIB", os.path.join(_HERE, "libtetmag_hip.so"))

TM_OK = 0
TM_ERR_NO_DEVICE = 1
TM_ERR_BAD_ARG = 2
TM_ERR_NONFINITE = 3
TM_ERR_MAXITER = 4
TM_ERR_HIP = 5
TM_ERR_NOT_READY = 6
TM_ERR_UNSUPPORTED = 7
STATUS_NAMES = {
    0: "TM_OK", 1: "TM_ERR_NO_DEVICE", 2: "TM_ERR_BAD_ARG", 3: "TM_ERR_NONFINITE", 4: "TM_ERR_MAXITER",
    5: "TM_ERR_HIP", 6: "TM_ERR_NOT_READY", 7: "TM_ERR_UNSUPPORTED",
}

TM_HOST = 0
TM_DEVICE = 1

TM_OP_EXCHANGE = 0
TM_OP_CURL_X, TM_OP_CURL_Y, TM_OP_CURL_Z = 1, 2, 3
TM_OP_TGRAD_X, TM_OP_TGRAD_Y, TM_OP_TGRAD_Z = 4, 5, 6
TM_OP_GRAD_X, TM_OP_GRAD_Y, TM_OP_GRAD_Z = 7, 8, 9
TM_OP_STT_GRAD_X, TM_OP_STT_GRAD_Y, TM_OP_STT_GRAD_Z = 10, 11, 12

TM_LLG_NOPREC = 0
TM_LLG_CLASSIC = 1
TM_PRECOND_JACOBI = 0
TM_PRECOND_AMG = 1
TM_PREC_FP64 = 0
TM_PREC_MIXED = 1
TM_PREC_MIXED_BEM32 = 2
TM_PREC_MIXED_BEMI8 = 3
TM_BATCH_MAX = 64
TM_BEM_REFERENCE = 0
TM_BEM_ACCURATE = 1


class NativeLibraryError(RuntimeError):
    pass


class TetmagError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what} failed: {STATUS_NAMES.get(status, status)}")
        self.status = status


class EvalStats(C.Structure):
    _fields_ = [
        ("iters_neumann", C.c_int),
        ("iters_dirichlet", C.c_int),
        ("relres_neumann", C.c_double),
        ("relres_dirichlet", C.c_double),
        ("max_torque", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PhaseTimes(C.Structure):
    _fields_ = [
        ("div_ms", C.c_double),
        ("neumann_ms", C.c_double),
        ("bem_ms", C.c_double),
        ("dirichlet_ms", C.c_double),
        ("fields_ms", C.c_double),
        ("iters_neumann", C.c_longlong),
        ("iters_dirichlet", C.c_longlong),
        ("count", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Energies(C.Structure):
    _fields_ = [("demag", C.c_double), ("exchange", C.c_double), ("uniaxial", C.c_double),
                ("zeeman", C.c_double), ("cubic", C.c_double), ("surface", C.c_double), ("dmi", C.c_double),
                ("total", C.c_double), ("mean_m", C.c_double * 3), ("mean_h", C.c_double * 3),
                ("volume", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("mean_m", "mean_h")}
        d["mean_m"] = list(self.mean_m)
        d["mean_h"] = list(self.mean_h)
        return d


class RK4Stats(C.Structure):
    _fields_ = [("sync_steps", C.c_int), ("graph_steps", C.c_int), ("redone_steps", C.c_int), ("captures", C.c_int),
                ("budget_neumann", C.c_int), ("budget_dirichlet", C.c_int), ("max_iters_neumann", C.c_int),
                ("max_iters_dirichlet", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OdeStats(C.Structure):
    _fields_ = [("accepted", C.c_int), ("rejected", C.c_int), ("evaluations", C.c_int), ("sync_steps", C.c_int),
                ("graph_steps", C.c_int), ("redone_steps", C.c_int), ("captures", C.c_int),
                ("budget_neumann", C.c_int), ("budget_dirichlet", C.c_int), ("max_iters_neumann", C.c_int),
                ("max_iters_dirichlet", C.c_int), ("tau", C.c_double), ("h_next", C.c_double),
                ("last_err", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("ms", C.c_double), ("bytes", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [
        ("nsamples", C.c_int),
        ("outer_neumann", C.c_int),
        ("outer_dirichlet", C.c_int),
        ("inner_neumann", C.c_int),
        ("inner_dirichlet", C.c_int),
        ("max_relres_neumann", C.c_double),
        ("max_relres_dirichlet", C.c_double),
        ("max_torque", C.c_double * 64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "max_torque"}
        d["max_torque"] = list(self.max_torque[:self.nsamples])
        return d


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_U64 = C.c_uint64
_PP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int unless listed in _RESTYPE)
_SIGNATURES = {
    "tm_device_count": [_P],
    "tm_set_device": [_I, _P],
    "tm_spmat_create": [_I, _I, _P, _P, _P, _PP],
    "tm_spmat_destroy": [_P],
    "tm_spmat_mvp": [_P, _P, _I, _P, _I],
    "tm_solver_create": [_I, _P, _P, _P, _I, _PP],
    "tm_solver_destroy": [_P],
    "tm_solver_solve": [_P, _P, _I, _P, _I, _D, _I, _I, _P, _P],
    "tm_ctx_create": [_I, _PP],
    "tm_ctx_destroy": [_P],
    "tm_ctx_set_stream": [_P, _P],
    "tm_ctx_synchronize": [_P],
    "tm_ctx_set_op": [_P, _I, _P, _P, _P],
    "tm_ctx_set_js": [_P, _P],
    "tm_ctx_set_uniaxial": [_P, _P, _P],
    "tm_ctx_set_dmi": [_P, _P, _P, _P, _P],
    "tm_ctx_set_stt_eta": [_P, _P, _P, _P],
    "tm_ctx_set_zeeman": [_P, _P, _D, _D, _D],
    "tm_ctx_set_mag": [_P, _P, _I],
    "tm_ctx_get_mag": [_P, _P, _I],
    "tm_ctx_normalize_mag": [_P, _P, _I],
    "tm_exchange_field": [_P, _P, _I],
    "tm_uniaxial_field": [_P, _P, _I],
    "tm_cubic_field": [_P, _P, _I],
    "tm_dmi_field": [_P, _P, _I],
    "tm_stt_uterm": [_P, _P, _I],
    "tm_llg_rhs": [_P, _P, _I, _D, _I, _P, _I],
    "tm_stt_llg": [_P, _P, _I, _D, _D, _P, _I],
    "tm_max_torque": [_P, _P, _I, _P],
    "tm_div_m": [_P, _P, _I, _P, _I],
    "tm_grad_field": [_P, _P, _I, _P, _I],
    "tm_ctx_set_boundary": [_P, _I, _P],
    "tm_ctx_set_bem_dense": [_P, _P, _I, _I],
    "tm_ctx_assemble_bem": [_P, _I, _P, _P, _P, _P, _P, _I],
    "tm_ctx_get_bem_dense": [_P, _P, _I],
    "tm_bem_apply_noalloc": [_I, _I, _P, _P, _P, _P],
    "tm_bem_apply": [_P, _P, _I, _P, _I],
    "tm_ctx_set_demag_solvers": [_P, _P, _P, _P, _P, _P, _P, _I],
    "tm_ctx_set_solver_params": [_P, _D, _I],
    "tm_demag_potential": [_P, _P, _I, _P, _I, _P],
    "tm_demag_field": [_P, _P, _I, _P, _I, _P],
    "tm_demag_last_potentials": [_P, _P, _P, _I],
    "tm_eval_rhs": [_P, _P, _I, _D, _I, _P, _I, _P, _I, _P],
    "tm_mesh_kuhn_cuboid_size": [_I, _I, _I, _P, _P],
    "tm_mesh_locality_order": [_I, _P, _P],
    "tm_mesh_kuhn_cuboid": [_I, _I, _I, _D, _D, _U64, _P, _P],
    "tm_fem_pattern": [_I, _I, _P, _P],
    "tm_fem_assemble": [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "tm_boundary_faces": [_I, _P, _P, _P, _P],
    "tm_boundary_props": [_I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "tm_boundary_solid_angles": [_I, _I, _P, _P, _P, _I, _P],
    "tm_init_random_unit": [_I, _U64, _P],
    "tm_uniform": [_U64, _I, _P],
    "tm_ctx_step_euler": [_P, _P, _I, _D],
    "tm_ctx_enable_timers": [_P, _I],
    "tm_ctx_get_timers": [_P, _P, _I],
    "tm_solver_amg_info": [_P, _P, _P, _I],
    "tm_ctx_solver_amg_info": [_P, _I, _P, _P, _I],
    "tm_ctx_set_node_volume": [_P, _P],
    "tm_ctx_set_dmi_interfacial": [_P, _P],
    "tm_idmi_field": [_P, _P, _I],
    "tm_energies_eval": [_P, _P, _I, _I, _P],
    "tm_ctx_set_freeze_demag": [_P, _I],
    "tm_ctx_refresh_demag": [_P, _P, _I, _P],
    "tm_eval_rhs_vjp": [_P, _P, _I, _D, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P],
    "tm_demag_field_vjp": [_P, _P, _I, _P, _I, _P],
    "tm_bem_apply_transposed": [_P, _P, _I, _P, _I],
    "tm_solver_solve_transposed": [_P, _P, _I, _P, _I, _D, _I, _I, _P, _P],
    "tm_ctx_run_rk4": [_P, _D, _I, _D, _I, _I, _P],
    "tm_ctx_integrate": [_P, _D, _D, _D, _P, _I, _D, _I, _I, _I, _P],
    "tm_batch_create": [_P, _I, _I, _PP],
    "tm_batch_destroy": [_P],
    "tm_batch_set_zeeman": [_P, _P],
    "tm_batch_set_mag": [_P, _P, _I],
    "tm_batch_get_mag": [_P, _P, _I],
    "tm_batch_eval_rhs": [_P, _P, _I, _D, _I, _P, _I, _P, _I, _P],
    "tm_batch_get_potentials": [_P, _P, _P, _I],
    "tm_batch_step_euler": [_P, _P, _I, _D],
    "tm_batch_eval_euler": [_P, _D, _I, _D, _P, _I, _P],
    "tm_batch_run_rk4": [_P, _D, _I, _D, _I, _I, _P, _P],
    "tm_batch_time_kernels": [_P, _I, _P, _I, _P],
    "tm_batch_gemm_info": [_P, _P, _P, _P, _P, _P, _P],
    "tm_batch_gemm_live": [_P, _I, _P, _P],
    "tm_batch_bem_apply": [_P, _P, _I, _P, _I],
    "tm_batch_enable_timers": [_P, _I],
    "tm_batch_get_timers": [_P, _P, _I],
    "tm_version": [],
}
_RESTYPE = {"tm_version": C.c_char_p}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def lib():
    """Loads libtetmag_hip.so once; raises NativeLibraryError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    # torch bundles its own libamdhip64.so under the same soname as the ROCm
    # runtime this library links; whichever loads first serves the whole
    # process.  Loading torch first keeps one runtime that both can use.
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is optional for pure C-ABI use
        pass
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} not found: build it with `python -m tetmag_amd.build` (no CPU fallback exists)")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the environment
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    for name, args in _SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int)
    _lib = L
    return L


def check(status: int, what: str) -> int:
    if status != TM_OK:
        raise TetmagError(status, what)
    return status


class _Buf:
    """A buffer address for a ctypes call that keeps its owner alive.

    ``ptr(f64(x))`` may be handed a temporary array; returning a bare integer
    would let Python free the array before the C function reads it.  ctypes
    converts this object through ``_as_parameter_`` while the argument (and so
    the owner) is alive for the duration of the call."""
    __slots__ = ("_as_parameter_", "owner")

    def __init__(self, addr: int, owner):
        self._as_parameter_ = C.c_void_p(addr)
        self.owner = owner

    def __int__(self):
        return self._as_parameter_.value or 0


def ptr(a):
    """Address of a C-contiguous numpy array or a torch tensor (host or device),
    as an object that keeps the array or tensor alive while it is referenced."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return _Buf(a.ctypes.data, a)
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return _Buf(a.data_ptr(), a)
    raise TypeError(f"unsupported buffer type {type(a)}")


def where(a) -> int:
    """TM_DEVICE for torch CUDA/HIP tensors, TM_HOST otherwise."""
    if hasattr(a, "is_cuda") and a.is_cuda:
        return TM_DEVICE
    return TM_HOST


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)
