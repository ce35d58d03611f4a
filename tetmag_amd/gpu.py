"""Host-side mirror of the reference GPU classes over the C ABI.

Same class and method names, argument meaning and return conventions as
tetmag's gpu/ directory, with numpy / scipy.sparse standing in for Eigen:

* ``SpMat`` (Eigen RowMajor SparseMatrix<double>) -> ``scipy.sparse.csr_matrix``
* ``MatrixXd`` n x 3 (column-major)               -> ``numpy`` (n, 3) array
* ``std::vector<double>`` of 3n                   -> ``numpy`` (3n,) array
* ``thrust::device_vector<double>``               -> ``torch`` tensor on cuda:0

Every method calls libtetmag_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _csr(A):
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.shape, N.i32(A.indptr), N.i32(A.indices), N.f64(A.data)


def _to3n(M, n):
    """(n, 3) matrix or 3n vector -> 3n SoA vector (x-block, y-block, z-block)."""
    a = np.asarray(M, np.float64)
    if a.ndim == 2:
        if a.shape != (n, 3):
            raise ValueError(f"expected ({n}, 3), got {a.shape}")
        return np.ascontiguousarray(a.T.reshape(-1))
    if a.shape != (3 * n,):
        raise ValueError(f"expected ({3 * n},), got {a.shape}")
    return N.f64(a)


def _from3n(v, n):
    return np.ascontiguousarray(v.reshape(3, n).T)


class SetDevice:
    """SetDevice (SetDevice.h:36-44)."""

    def __init__(self, mydev: int):
        self.mydev = mydev
        self.ndev = 0
        self.success = False

    def set(self) -> bool:
        chosen = C.c_int()
        st = N.lib().tm_set_device(self.mydev, C.byref(chosen))
        if st == N.TM_ERR_NO_DEVICE:
            self.success = False  # the caller falls back to the CPU path (tetmag.cpp:82-90)
            return False
        N.check(st, "tm_set_device")
        self.mydev = chosen.value
        self.success = True
        return True


class SpMatCUDA:
    """SpMatCUDA (SpMatCUDA.h:49-79): y = A x on the device."""

    def __init__(self, A):
        (self.rows, self.cols), rp, ci, v = _csr(A)
        h = C.c_void_p()
        N.check(N.lib().tm_spmat_create(self.rows, self.cols, N.ptr(rp), N.ptr(ci), N.ptr(v), C.byref(h)),
                "tm_spmat_create")
        self.h = h
        self._x = None

    def setX(self, x):
        self._x = x

    def mvp(self, x=None):
        """mvp(devVecD) returns a device tensor; mvp() multiplies the vector set by setX."""
        x = self._x if x is None else x
        if hasattr(x, "is_cuda") and x.is_cuda:
            import torch
            y = torch.empty(self.rows, dtype=torch.float64, device=x.device)
            N.check(N.lib().tm_spmat_mvp(self.h, N.ptr(x), N.TM_DEVICE, N.ptr(y), N.TM_DEVICE), "tm_spmat_mvp")
            return y
        xv = N.f64(x)
        y = np.empty(self.rows)
        N.check(N.lib().tm_spmat_mvp(self.h, N.ptr(xv), N.TM_HOST, N.ptr(y), N.TM_HOST), "tm_spmat_mvp")
        self._y = y
        return y

    def mvpResEig(self):
        return self._y

    def __del__(self):
        try:
            if getattr(self, "h", None):
                N.lib().tm_spmat_destroy(self.h)
                self.h = None
        except Exception:
            pass


class GPUSolver:
    """GPUSolver / SolverFactory (GPUSolver.h:65-87, SolverFactory.h:42-65).

    Accepts the reference's row-replaced stiffness matrices and solves the
    symmetric free block with preconditioned CG ("amg" = smoothed
    aggregation, "jacobi" = diagonal)."""

    maxit = 100000

    def __init__(self):
        self.A = None
        self.b = None
        self.x = None
        self.cgTol = 1e-7  # ProgramSpecs.cpp:70 default
        self.preconditionerType = "amg"
        self.h = None
        self.iters = 0
        self.error = float("nan")

    def setMatrix(self, A):
        self.A = A

    def setLoadVector(self, b):
        self.b = N.f64(b)

    def setCGTolerance(self, tol: float):
        self.cgTol = float(tol)

    def definePreconditioner(self, p: str):
        self.preconditionerType = p

    def compute(self) -> bool:
        (n, m), rp, ci, v = _csr(self.A)
        pre = N.TM_PRECOND_JACOBI if str(self.preconditionerType).lower() in ("jacobi", "diagonal") else N.TM_PRECOND_AMG
        h = C.c_void_p()
        st = N.lib().tm_solver_create(n, N.ptr(rp), N.ptr(ci), N.ptr(v), pre, C.byref(h))
        if st != N.TM_OK:  # GPUSolver::setup catches and returns false (GPUSolver.cu:57-79)
            self.h = None
            self.status = st
            return False
        self.h = h
        self.n = n
        self.x = np.zeros(n)
        return True

    def solve(self):
        it = C.c_int()
        rr = C.c_double()
        x = self.x
        st = N.lib().tm_solver_solve(self.h, N.ptr(self.b), N.TM_HOST, N.ptr(x), N.TM_HOST, self.cgTol, self.maxit, 2,
                                     C.byref(it), C.byref(rr))
        self.iters, self.error = it.value, rr.value
        self.status = st
        if st not in (N.TM_OK, N.TM_ERR_MAXITER, N.TM_ERR_NONFINITE):
            N.check(st, "tm_solver_solve")

    def result(self):
        return self.x.copy()

    def amgLevels(self) -> list:
        nl = C.c_int()
        sizes = np.zeros(16, np.int32)
        N.check(N.lib().tm_solver_amg_info(self.h, C.byref(nl), N.ptr(sizes), 16), "tm_solver_amg_info")
        return sizes[:nl.value].tolist()

    def wasSuccessful(self) -> bool:
        return bool(np.isfinite(self.error)) and self.status == N.TM_OK

    def report(self):
        print(f"iterations: {self.iters}, relative residual: {self.error:.3e}")

    def __del__(self):
        try:
            if self.h:
                N.lib().tm_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass


class EffFieldGPU:
    """EffFieldGPU (EffFieldGPU.h:54-127): device magnetisation and the field
    terms computed from it.  Field getters return (n, 3) host matrices."""

    def __init__(self, nx: int):
        self.nx = nx
        h = C.c_void_p()
        N.check(N.lib().tm_ctx_create(nx, C.byref(h)), "tm_ctx_create")
        self.h = h
        self.L = N.lib()

    def init(self, nx: int):
        if nx != self.nx:
            raise ValueError("EffFieldGPU.init: size differs from the constructor's")

    def _set(self, op, A):
        (n, m), rp, ci, v = _csr(A)
        N.check(self.L.tm_ctx_set_op(self.h, op, N.ptr(rp), N.ptr(ci), N.ptr(v)), "tm_ctx_set_op")

    def setExchangeMatOnDev(self, XC):
        self._set(N.TM_OP_EXCHANGE, XC)

    def setGradientMatsOnDev(self, tGx, tGy, tGz):
        for op, A in ((N.TM_OP_CURL_X, tGx), (N.TM_OP_CURL_Y, tGy), (N.TM_OP_CURL_Z, tGz)):
            self._set(op, A)

    def setUniaxialAnisotropy(self, KuAxis, Ku1):
        N.check(self.L.tm_ctx_set_uniaxial(self.h, N.ptr(_to3n(KuAxis, self.nx)), N.ptr(N.f64(Ku1))),
                "tm_ctx_set_uniaxial")

    def setCubicAnisotropy(self, cubicAxes, Kc1, Kc2):
        pass  # the reference leaves this unimplemented on the GPU (EffFieldGPU.cu:66-68)

    def setDMIdata(self, D, invNodeVol, nv, nodeArea):
        N.check(self.L.tm_ctx_set_dmi(self.h, N.ptr(N.f64(D)), N.ptr(N.f64(invNodeVol)), N.ptr(_to3n(nv, self.nx)),
                                      N.ptr(N.f64(nodeArea))), "tm_ctx_set_dmi")

    def setSTTDataOnDevice(self, Gx, Gy, Gz, eta_x, eta_y, eta_z):
        for op, A in ((N.TM_OP_STT_GRAD_X, Gx), (N.TM_OP_STT_GRAD_Y, Gy), (N.TM_OP_STT_GRAD_Z, Gz)):
            self._set(op, A)
        N.check(self.L.tm_ctx_set_stt_eta(self.h, N.ptr(N.f64(eta_x)), N.ptr(N.f64(eta_y)), N.ptr(N.f64(eta_z))),
                "tm_ctx_set_stt_eta")

    def setMagDev(self, M):
        if hasattr(M, "is_cuda") and M.is_cuda:
            N.check(self.L.tm_ctx_set_mag(self.h, N.ptr(M), N.TM_DEVICE), "tm_ctx_set_mag")
            N.check(self.L.tm_ctx_synchronize(self.h), "tm_ctx_synchronize")
        else:
            N.check(self.L.tm_ctx_set_mag(self.h, N.ptr(_to3n(M, self.nx)), N.TM_HOST), "tm_ctx_set_mag")

    def _field(self, fn, name):
        out = np.empty(3 * self.nx)
        N.check(fn(self.h, N.ptr(out), N.TM_HOST), name)
        return _from3n(out, self.nx)

    def ExchangeFieldGPU(self):
        return self._field(self.L.tm_exchange_field, "tm_exchange_field")

    def UniaxialAnisotropyField(self):
        return self._field(self.L.tm_uniaxial_field, "tm_uniaxial_field")

    def CubicAnisotropyField(self):
        return self._field(self.L.tm_cubic_field, "tm_cubic_field")

    def DMIField(self):
        return self._field(self.L.tm_dmi_field, "tm_dmi_field")

    def UTermSTT_GPU(self):
        return self._field(self.L.tm_stt_uterm, "tm_stt_uterm")

    def _llg(self, Heff, alpha, variant):
        out = np.empty(3 * self.nx)
        N.check(self.L.tm_llg_rhs(self.h, N.ptr(_to3n(Heff, self.nx)), N.TM_HOST, float(alpha), variant, N.ptr(out),
                                  N.TM_HOST), "tm_llg_rhs")
        return out

    def ClassicLLG_hst(self, Heff, alpha):
        return self._llg(Heff, alpha, N.TM_LLG_CLASSIC)

    def LLG_noPrec_hst(self, Heff, alpha):
        return self._llg(Heff, alpha, N.TM_LLG_NOPREC)

    def STT_term_LLG_hst(self, U, alpha, beta):
        out = np.empty(3 * self.nx)
        N.check(self.L.tm_stt_llg(self.h, N.ptr(_to3n(U, self.nx)), N.TM_HOST, float(alpha), float(beta), N.ptr(out),
                                  N.TM_HOST), "tm_stt_llg")
        return out

    def ClassicLLG_dev(self, Heff, alpha):
        import torch
        out = torch.empty(3 * self.nx, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        N.check(self.L.tm_llg_rhs(self.h, N.ptr(_to3n(Heff, self.nx)), N.TM_HOST, float(alpha), N.TM_LLG_CLASSIC,
                                  N.ptr(out), N.TM_DEVICE), "tm_llg_rhs")
        N.check(self.L.tm_ctx_synchronize(self.h), "tm_ctx_synchronize")
        return out

    def MaxTorque(self, Heff):
        out = C.c_double()
        N.check(self.L.tm_max_torque(self.h, N.ptr(_to3n(Heff, self.nx)), N.TM_HOST, C.byref(out)), "tm_max_torque")
        return out.value

    def NormalizeMag(self, M, nx):
        v = _to3n(M, nx)
        N.check(self.L.tm_ctx_normalize_mag(self.h, N.ptr(v), N.TM_HOST), "tm_ctx_normalize_mag")
        M[...] = _from3n(v, nx) if np.asarray(M).ndim == 2 else v

    def displayTimer(self):
        pass

    def __del__(self):
        try:
            if getattr(self, "h", None):
                N.lib().tm_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass


class GpuPotential:
    """GpuPotential (GpuPotential.h:50-75)."""

    def __init__(self):
        self.h = None

    def setGradientMatrices(self, tGx, tGy, tGz, gx, gy, gz, Js):
        import scipy.sparse as sp
        n = sp.csr_matrix(gx).shape[0]
        self.nx = n
        h = C.c_void_p()
        N.check(N.lib().tm_ctx_create(n, C.byref(h)), "tm_ctx_create")
        self.h = h
        for op, A in ((N.TM_OP_TGRAD_X, tGx), (N.TM_OP_TGRAD_Y, tGy), (N.TM_OP_TGRAD_Z, tGz),
                      (N.TM_OP_GRAD_X, gx), (N.TM_OP_GRAD_Y, gy), (N.TM_OP_GRAD_Z, gz)):
            (r, c), rp, ci, v = _csr(A)
            N.check(N.lib().tm_ctx_set_op(self.h, op, N.ptr(rp), N.ptr(ci), N.ptr(v)), "tm_ctx_set_op")
        N.check(N.lib().tm_ctx_set_js(self.h, N.ptr(N.f64(Js))), "tm_ctx_set_js")

    def calcDivM(self, mag):
        b = np.empty(self.nx)
        N.check(N.lib().tm_div_m(self.h, N.ptr(_to3n(mag, self.nx)), N.TM_HOST, N.ptr(b), N.TM_HOST), "tm_div_m")
        return b

    def calcGradientField(self, u):
        H = np.empty(3 * self.nx)
        N.check(N.lib().tm_grad_field(self.h, N.ptr(N.f64(u)), N.TM_HOST, N.ptr(H), N.TM_HOST), "tm_grad_field")
        return _from3n(H, self.nx)

    def __del__(self):
        try:
            if self.h:
                N.lib().tm_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LLGWrapper:
    """LLGWrapper (LLGWrapper.h:57-63): device scratch of the RHS callback.
    Scratch lives in the field context here, so init only records the size."""

    def init(self, nx: int):
        self.nx3 = 3 * nx
