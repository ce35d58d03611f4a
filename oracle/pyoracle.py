"""Python driver of the FP64 CPU oracle (liboracle.so built from refcpu.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  Never imported by
the product package.

Given a mesh (xyz, fel) and material data, :class:`Oracle` builds every
operator with the oracle's own restatement of the reference's setup
(FEMprocessing.cpp, Boundary.cpp, BEMprocessing.cpp, Lindholm.cpp) and
evaluates the effective field and the LLG right-hand side the way the
reference CPU path does (EffFieldCalc.cpp, DemagField.cpp, TheLLG.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MU0 = 1.256637061435917295376e-6

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        sig = {
            "ref_volumes": ([I, I, P, P, P, P], None),
            "ref_shape_grads": ([I, P, P, P, P, P], None),
            "ref_fem_pattern": ([I, I, P, P], I),
            "ref_fem_assemble": ([I, I, P, P, P, P, P, P, P, P, P, P, P, P], None),
            "ref_boundary_faces": ([I, P, P, P], I),
            "ref_boundary_props": ([I, P, P, I, P, P, P, P, P, P, P, P], I),
            "ref_solid_angle": ([P, P, P, P], D),
            "ref_lindholm_weights": ([P, P, P, P, D, P, P], None),
            "ref_solid_angles": ([I, P, P, P, I, P], None),
            "ref_bem_matrix": ([I, I, P, P, P, P, P, P, P, P], None),
            "ref_spmv": ([I, P, P, P, P, P], None),
            "ref_dense_gemv": ([I, P, P, P], None),
            "ref_replace_rows_identity": ([I, P, P, P, P, P, P, P], I),
            "ref_solver_create": ([I, P, P, P], P),
            "ref_solver_create_amg": ([I, P, P, P, D, D, I], P),
            "ref_solver_levels": ([P, P, I], I),
            "ref_bem_rows": ([I, I, P, P, P, P, P, P, P, I, P, P], None),
            "ref_stt_uterm": ([I, P, P, P, P, P, P, P, P, P, P], None),
            "ref_stt_llg": ([I, D, D, P, P, P], None),
            "ref_solver_destroy": ([P], None),
            "ref_solver_solve": ([P, P, P, D, I, I, P], I),
            "ref_exchange_field": ([I, P, P, P, P, P, P, P], None),
            "ref_uniaxial_field": ([I, P, P, P, P], None),
            "ref_dmi_field": ([I, P, P, P, P, P, P, P, P, P, P, P], None),
            "ref_idmi_field": ([I, P, P, P, P, P, P, P, P, P, P], None),
            "ref_zeeman_field": ([I, D, D, D, P], None),
            "ref_total_field": ([I, P, P, P, P, P, P, P], None),
            "ref_llg_rhs": ([I, D, I, P, P, P], None),
            "ref_max_torque": ([I, P, P], D),
            "ref_normalize": ([I, P], None),
            "ref_div_m": ([I, P, P, P, P, P, P, P, P], None),
            "ref_grad_field": ([I, P, P, P, P, P, P, P], None),
            "ref_boundary_integral": ([I, I, P, P, P, P], None),
            "ref_demag_energy": ([I, P, P, P, P], D),
            "ref_field_energy": ([I, P, P, P], D),
            "ref_uniaxial_energy": ([I, P, P, P, P], D),
            "ref_zeeman_energy": ([I, P, P, P, P], D),
            "ref_num_threads": ([], I),
            "ref_set_num_threads": ([I], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def p(a):
    # ndarray.ctypes carries _as_parameter_ and a reference to the array, so a
    # temporary (p(f64(x))) stays alive until the ctypes call has returned.
    return None if a is None else a.ctypes


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Solver:
    """SolverFactory-like handle: matrix set once, warm-started solves.

    precond "amg": smoothed-aggregation AMG with ILU(0) smoothing inside
    GMRES(30), the reference's configuration (AMGCLSolver.h:89-101,
    AMGCLSolver.cpp:39-44); "ilu0": ILU(0)-preconditioned GMRES(60)."""

    def __init__(self, rp, col, val, precond="amg"):
        self.rp, self.col, self.val = i32(rp), i32(col), f64(val)
        self.n = self.rp.shape[0] - 1
        self.precond = precond
        if precond == "amg":
            self.h = lib().ref_solver_create_amg(self.n, p(self.rp), p(self.col), p(self.val), 0.0, 0.0, 0)
            self.restart = 30
        elif precond == "ilu0":
            self.h = lib().ref_solver_create(self.n, p(self.rp), p(self.col), p(self.val))
            self.restart = 60
        else:
            raise ValueError(precond)
        if not self.h:
            raise RuntimeError(f"{precond} preconditioner set-up failed")
        self.x = np.zeros(self.n)
        self.last_iters = 0
        self.last_relres = 0.0

    def levels(self):
        sizes = np.zeros(32, np.int32)
        k = lib().ref_solver_levels(self.h, p(sizes), 32)
        return [int(v) for v in sizes[:k]]

    def solve(self, b, tol=1e-14, maxit=20000, restart=None, warm=True):
        restart = self.restart if restart is None else restart
        b = f64(b)
        if not warm:
            self.x[:] = 0.0
        rr = C.c_double()
        it = lib().ref_solver_solve(self.h, p(b), p(self.x), tol, maxit, restart, C.byref(rr))
        self.last_iters, self.last_relres = it, rr.value
        if it < 0:
            raise RuntimeError(f"oracle GMRES did not converge: relres {rr.value:.3e}")
        return self.x.copy()

    def __del__(self):
        try:
            if self.h:
                lib().ref_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass


def csr_matvec(rp, col, val, x):
    y = np.empty(rp.shape[0] - 1)
    lib().ref_spmv(y.shape[0], p(rp), p(col), p(val), p(f64(x)), p(y))
    return y


class Oracle:
    """FP64 restatement of tetmag's CPU effective-field + LLG path on one mesh."""

    def __init__(self, xyz, fel):
        L = lib()
        self.xyz = f64(xyz)
        self.fel = i32(fel)
        n, nt = self.xyz.shape[0], self.fel.shape[0]
        self.n, self.ntet = n, nt
        # FEM operators (FEMprocessing.cpp)
        self.rowptr = np.empty(n + 1, np.int32)
        L.ref_fem_pattern(n, nt, p(self.fel), p(self.rowptr))
        nnz = int(self.rowptr[-1])
        self.col = np.empty(nnz, np.int32)
        self.K, self.gX, self.gY, self.gZ, self.tX, self.tY, self.tZ = [np.empty(nnz) for _ in range(7)]
        self.nodevol = np.empty(n)
        L.ref_fem_assemble(n, nt, p(self.xyz), p(self.fel), p(self.rowptr), p(self.col), p(self.K), p(self.gX),
                           p(self.gY), p(self.gZ), p(self.tX), p(self.tY), p(self.tZ), p(self.nodevol))
        self.elvol = np.empty(nt)
        nv2 = np.empty(n)
        L.ref_volumes(n, nt, p(self.xyz), p(self.fel), p(self.elvol), p(nv2))
        # boundary (Boundary.cpp)
        faces = np.empty((4 * nt, 3), np.int32)
        ftet = np.empty(4 * nt, np.int32)
        nbel = L.ref_boundary_faces(nt, p(self.fel), p(faces), p(ftet))
        self.faces = np.ascontiguousarray(faces[:nbel])
        self.face_tet = np.ascontiguousarray(ftet[:nbel])
        self.nbel = nbel
        bn = np.empty(n, np.int32)
        self.bindex = np.empty(n, np.int32)
        self.face_area = np.empty(nbel)
        self.face_normal = np.empty((nbel, 3))
        self.node_area = np.empty(n)
        nn = np.empty((n, 3))
        bnx = L.ref_boundary_props(n, p(self.xyz), p(self.fel), nbel, p(self.faces), p(self.face_tet), p(bn),
                                   p(self.bindex), p(self.face_area), p(self.face_normal), p(self.node_area), p(nn))
        self.bnodes = np.ascontiguousarray(bn[:bnx])
        self.bnx = bnx
        self.node_normal = np.ascontiguousarray(nn.T.reshape(-1))  # 3n SoA
        self.sub = np.empty(bnx)
        L.ref_solid_angles(nt, p(self.xyz), p(self.fel), p(self.bindex), bnx, p(self.sub))
        self._B = None
        self._solvers = None

    # -- BEM (BEMprocessing.cpp:375-411) ------------------------------------
    @property
    def B(self):
        if self._B is None:
            self._B = np.empty((self.bnx, self.bnx), order="F")
            lib().ref_bem_matrix(self.bnx, self.nbel, p(self.xyz), p(self.bnodes), p(self.bindex), p(self.faces),
                                 p(self.face_area), p(self.face_normal), p(self.sub), self._B.ctypes.data)
        return self._B

    def set_B(self, B):
        self._B = np.asfortranarray(B, dtype=np.float64)

    def bem_rows(self, rows):
        """Rows `rows` of B (BEMprocessing.cpp:375-411) without forming B:
        (len(rows), bnx), row k = B[rows[k], :]."""
        rows = i32(rows)
        out = np.empty((rows.shape[0], self.bnx))
        lib().ref_bem_rows(self.bnx, self.nbel, p(self.xyz), p(self.bnodes), p(self.bindex), p(self.faces),
                           p(self.face_area), p(self.face_normal), p(self.sub), rows.shape[0], p(rows), p(out))
        return out

    # -- demag solvers (DemagField.cpp:204-220, FEMprocessing.cpp:329-369) ----
    def replaced(self, fixed_mask):
        L = lib()
        nnz = int(self.rowptr[-1])
        rp = np.empty(self.n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz)
        fx = np.ascontiguousarray(fixed_mask.astype(np.int8))
        k = L.ref_replace_rows_identity(self.n, p(self.rowptr), p(self.col), p(self.K), p(fx), p(rp), p(col),
                                        p(val))
        return rp, col[:k].copy(), val[:k].copy()

    def neumann_matrix(self):
        fx = np.zeros(self.n, bool)
        fx[self.n // 2] = True
        return self.replaced(fx)

    def dirichlet_matrix(self):
        fx = np.zeros(self.n, bool)
        fx[self.bnodes] = True
        return self.replaced(fx)

    def solvers(self, precond=None):
        """(Neumann, Dirichlet) solvers; precond "amg" (default, the
        reference's SA-AMG) or "ilu0", fixed by the first call."""
        if self._solvers is None:
            pc = precond or "amg"
            self._solvers = (Solver(*self.neumann_matrix(), precond=pc), Solver(*self.dirichlet_matrix(), precond=pc))
        elif precond is not None and precond != self._solvers[0].precond:
            raise ValueError(f"solvers already built with {self._solvers[0].precond}")
        return self._solvers

    # -- fields ---------------------------------------------------------------
    def exchange(self, m, A):
        H = np.empty(3 * self.n)
        lib().ref_exchange_field(self.n, p(self.rowptr), p(self.col), p(self.K), p(f64(A)), p(self.nodevol),
                                 p(f64(m)), p(H))
        return H

    def uniaxial(self, m, Ku, axis):
        H = np.empty(3 * self.n)
        lib().ref_uniaxial_field(self.n, p(f64(Ku)), p(f64(axis)), p(f64(m)), p(H))
        return H

    def dmi(self, m, D):
        H = np.empty(3 * self.n)
        lib().ref_dmi_field(self.n, p(self.rowptr), p(self.col), p(self.tX), p(self.tY), p(self.tZ), p(f64(D)),
                            p(self.nodevol), p(self.node_normal), p(self.node_area), p(f64(m)), p(H))
        return H

    def idmi(self, m, D):
        """Interfacial DMI field (J/m^3), interface normal z (no reference counterpart)."""
        H = np.empty(3 * self.n)
        lib().ref_idmi_field(self.n, p(self.rowptr), p(self.col), p(self.gX), p(self.gY), p(self.tX), p(self.tY),
                             p(f64(D)), p(self.nodevol), p(f64(m)), p(H))
        return H

    def stt_uterm(self, m, eta):
        """Zhang-Li u term U_c = sum_d eta_d .* (grad_d m_c) (TheLLG.cpp:402-416);
        eta: (3, n) or (3n,) per-node current factors."""
        e = f64(eta).reshape(3, self.n)
        U = np.empty(3 * self.n)
        lib().ref_stt_uterm(self.n, p(self.rowptr), p(self.col), p(self.gX), p(self.gY), p(self.gZ), p(f64(e[0])),
                            p(f64(e[1])), p(f64(e[2])), p(f64(m)), p(U))
        return U

    def stt_llg(self, m, U, alpha, beta):
        """-(beta - alpha) m x U - (1 + alpha beta) m x (m x U) (TheLLG.cpp:209-212)."""
        out = np.empty(3 * self.n)
        lib().ref_stt_llg(self.n, float(alpha), float(beta), p(f64(m)), p(f64(U)), p(out))
        return out

    def div(self, m, Js):
        b = np.empty(self.n)
        lib().ref_div_m(self.n, p(self.rowptr), p(self.col), p(self.tX), p(self.tY), p(self.tZ), p(f64(Js)),
                        p(f64(m)), p(b))
        return b

    def grad_field(self, u):
        H = np.empty(3 * self.n)
        lib().ref_grad_field(self.n, p(self.rowptr), p(self.col), p(self.gX), p(self.gY), p(self.gZ), p(f64(u)),
                             p(H))
        return H

    def boundary_integral(self, u1):
        g = np.empty(self.n)
        lib().ref_boundary_integral(self.n, self.bnx, p(self.bnodes), self.B.ctypes.data, p(f64(u1)), p(g))
        return g

    def demag(self, m, Js, tol=1e-14, warm=True):
        """DemagField::calcField (DemagField.cpp:92-94, 163-181): returns Hdem, u1, u2."""
        sN, sD = self.solvers()
        b = self.div(m, Js)  # already divided by mu0 (DemagField.cpp:135)
        u1 = sN.solve(b, tol=tol, warm=warm)
        g = self.boundary_integral(u1)
        u2 = sD.solve(g, tol=tol, warm=warm)
        H = self.grad_field(u1 + u2)
        self.last_iters = (sN.last_iters, sD.last_iters)
        return H, u1, u2

    def total(self, Js, Hexc=None, Hku=None, Hdmi=None, Hext=None, Hdem=None):
        Heff = np.empty(3 * self.n)
        lib().ref_total_field(self.n, p(f64(Js)), p(Hexc), p(Hku), p(Hdmi), p(Hext), p(Hdem), p(Heff))
        return Heff

    def llg(self, m, H, alpha, variant=1):
        out = np.empty(3 * self.n)
        lib().ref_llg_rhs(self.n, alpha, variant, p(f64(m)), p(f64(H)), p(out))
        return out

    def max_torque(self, m, H):
        return lib().ref_max_torque(self.n, p(f64(m)), p(f64(H)))

    def eval_rhs(self, m, *, Js, A, Ku=None, axis=None, D=None, Hext=None, alpha=0.5, variant=1, demag=True,
                 tol=1e-14, warm=True, Di=None):
        """TheLLG::classicVersion with every term of the composition
        (TheLLG.cpp:150-178, 238-254); Hext is a uniform (3,) field or a
        per-node (3n,) field in A/m.  Material arrays are per node and
        already scaled (A / scale^2, D / scale)."""
        n = self.n
        m = f64(m)
        out = {"Hexc": self.exchange(m, A)}
        if Ku is not None and np.any(Ku != 0):
            out["Hku"] = self.uniaxial(m, Ku, axis)
        if D is not None and np.any(D != 0):
            out["Hdmi"] = self.dmi(m, D)
        if Di is not None and np.any(Di != 0):  # interfacial DMI, summed with the bulk term
            out["Hidmi"] = self.idmi(m, Di)
            out["Hdmi"] = out["Hidmi"] + out.get("Hdmi", 0.0)
        if Hext is not None and np.any(np.asarray(Hext) != 0):
            He = f64(Hext).ravel()
            if He.shape[0] == 3 * n:
                out["Hext"] = He
            else:
                out["Hext"] = np.concatenate([np.full(n, He[0]), np.full(n, He[1]), np.full(n, He[2])])
        if demag:
            out["Hdem"], out["u1"], out["u2"] = self.demag(m, Js, tol=tol, warm=warm)
        out["Heff"] = self.total(Js, out.get("Hexc"), out.get("Hku"), out.get("Hdmi"), out.get("Hext"),
                                 out.get("Hdem"))
        out["dmdt"] = self.llg(m, out["Heff"], alpha, variant)
        out["max_torque"] = self.max_torque(m, out["Heff"])
        return out

    # -- FP64 CPU adjoint (no reference counterpart) ---------------------------
    # Restated from the same operators with scipy: transposed CSR products and
    # GMRES/ILU(0) on the transposed row-replaced matrices.  Pinned by central
    # finite differences of eval_rhs (tests/test_oracle_adjoint.py).
    def _csr(self, vals):
        import scipy.sparse as sp
        return sp.csr_matrix((vals, self.col, self.rowptr), shape=(self.n, self.n))

    def transposed_solvers(self):
        if getattr(self, "_tsolvers", None) is None:
            import scipy.sparse as sp
            pc = self.solvers()[0].precond  # the forward solvers' preconditioner
            out = []
            for rp, col, val in (self.neumann_matrix(), self.dirichlet_matrix()):
                T = sp.csr_matrix((val, col, rp), shape=(self.n, self.n)).T.tocsr()
                T.sort_indices()
                out.append(Solver(i32(T.indptr), i32(T.indices), f64(T.data), precond=pc))
            self._tsolvers = tuple(out)
        return self._tsolvers

    def demag_vjp(self, v, Js, tol=1e-14):
        """(dHdem/dm)^T v for Hdem = -grad(u1 + u2) (DemagField.cpp:92-240)."""
        n = self.n
        v = f64(v).reshape(3, n)
        G = [self._csr(g) for g in (self.gX, self.gY, self.gZ)]
        T = [self._csr(t) for t in (self.tX, self.tY, self.tZ)]
        w0 = -(G[0].T @ v[0] + G[1].T @ v[1] + G[2].T @ v[2])
        sNt, sDt = self.transposed_solvers()
        y2 = sDt.solve(w0, tol=tol, warm=False)
        t = w0.copy()
        t[self.bnodes] += self.B.T @ y2[self.bnodes]
        y1 = sNt.solve(t, tol=tol, warm=False)
        return np.concatenate([f64(Js) / MU0 * (T[d].T @ y1) for d in range(3)])

    def eval_rhs_vjp(self, m, lam, *, Js, A, Ku=None, axis=None, D=None, Hext=None, alpha=0.5, variant=1,
                     demag=True, tol=1e-14, Di=None):
        """lam^T d f / d(m, Hext, Ku, D, Di) for f = eval_rhs(m)["dmdt"]."""
        n = self.n
        m = f64(m)
        fw = self.eval_rhs(m, Js=Js, A=A, Ku=Ku, axis=axis, D=D, Hext=Hext, alpha=alpha, variant=variant,
                           demag=demag, tol=tol, warm=False, Di=Di)
        M = m.reshape(3, n).T
        H = fw["Heff"].reshape(3, n).T
        L = f64(lam).reshape(3, n).T
        ml = np.cross(M, L)
        gH = -alpha * np.cross(M, ml)
        gm = alpha * (np.cross(L, np.cross(M, H)) - np.cross(H, np.cross(L, M)))
        if variant == 1:
            gH += ml
            gm += np.cross(L, H)
        Jsv = f64(Js)
        with np.errstate(divide="ignore"):
            invjs = np.where(Jsv != 0, 1.0 / Jsv, 0.0)
        gp = gH * invjs[:, None]
        V = self.nodevol
        Kop = self._csr(self.K)
        XC = Kop.multiply((-2.0 * f64(A) / V)[:, None]).tocsr()
        gm += np.stack([XC.T @ gp[:, c] for c in range(3)], axis=1)
        out = {"gH": np.ascontiguousarray(gH.T.reshape(-1)), "g_hext": gH.sum(axis=0)}
        if Ku is not None and np.any(Ku != 0):
            k = f64(axis).reshape(3, n).T
            kg = np.sum(k * gp, axis=1)
            gm += (2.0 * f64(Ku) * kg)[:, None] * k
            out["g_ku"] = 2.0 * np.sum(k * M, axis=1) * kg
        if D is not None and np.any(D != 0):
            f = f64(D) / V
            S = (self.node_normal.reshape(3, n) * self.node_area).T
            C = [-self._csr(t) for t in (self.tX, self.tY, self.tZ)]
            gm += f[:, None] * np.cross(S, gp)
            w = -2.0 * f[:, None] * gp
            gm[:, 0] += C[2].T @ w[:, 1] - C[1].T @ w[:, 2]
            gm[:, 1] += C[0].T @ w[:, 2] - C[2].T @ w[:, 0]
            gm[:, 2] += C[1].T @ w[:, 0] - C[0].T @ w[:, 1]
            curl = np.stack([C[1] @ M[:, 2] - C[2] @ M[:, 1], C[2] @ M[:, 0] - C[0] @ M[:, 2],
                             C[0] @ M[:, 1] - C[1] @ M[:, 0]], axis=1)
            out["g_d"] = np.sum(gp * (np.cross(M, S) - 2.0 * curl), axis=1) / V
        if Di is not None and np.any(Di != 0):
            Dv = f64(Di)
            Gx, Gy = self._csr(self.gX), self._csr(self.gY)
            Tx, Ty = self._csr(self.tX), self._csr(self.tY)
            w = gp
            gm[:, 0] += -(Gx.T @ (Dv * w[:, 2])) + Dv * (Tx.T @ (w[:, 2] / V))
            gm[:, 1] += -(Gy.T @ (Dv * w[:, 2])) + Dv * (Ty.T @ (w[:, 2] / V))
            gm[:, 2] += (Gx.T @ (Dv * w[:, 0]) + Gy.T @ (Dv * w[:, 1])
                         - Dv * (Tx.T @ (w[:, 0] / V) + Ty.T @ (w[:, 1] / V)))
            out["g_di"] = (w[:, 0] * (Gx @ M[:, 2]) + w[:, 1] * (Gy @ M[:, 2])
                           - w[:, 2] * (Gx @ M[:, 0] + Gy @ M[:, 1])
                           - M[:, 2] * (Tx.T @ (w[:, 0] / V) + Ty.T @ (w[:, 1] / V))
                           + M[:, 0] * (Tx.T @ (w[:, 2] / V)) + M[:, 1] * (Ty.T @ (w[:, 2] / V)))
        g = np.ascontiguousarray(gm.T.reshape(-1))
        if demag:
            g += self.demag_vjp(out["gH"], Js, tol=tol)
        out["gm"] = g
        out["dmdt"] = fw["dmdt"]
        return out

    # -- energies (EffFieldCalc.cpp, DemagField.cpp:97-99, TheLLG.cpp:345-348) --
    def energies(self, m, fields, *, Js, Ku=None, axis=None):
        L = lib()
        n = self.n
        e = {"exchange": L.ref_field_energy(n, p(self.nodevol), p(fields["Hexc"]), p(f64(m)))}
        if "Hdem" in fields:
            e["demag"] = L.ref_demag_energy(n, p(f64(Js)), p(self.nodevol), p(fields["Hdem"]), p(f64(m)))
        if "Hdmi" in fields:
            e["dmi"] = L.ref_field_energy(n, p(self.nodevol), p(fields["Hdmi"]), p(f64(m)))
        if Ku is not None and np.any(Ku != 0):
            e["uniaxial"] = L.ref_uniaxial_energy(n, p(f64(Ku)), p(f64(axis)), p(self.nodevol), p(f64(m)))
        if "Hext" in fields:
            e["zeeman"] = L.ref_zeeman_energy(n, p(f64(Js)), p(self.nodevol), p(fields["Hext"]), p(f64(m)))
        return e


def solid_angle(x, t1, t2, t3):
    return lib().ref_solid_angle(p(f64(x)), p(f64(t1)), p(f64(t2)), p(f64(t3)))


def lindholm(x0, p1, p2, p3, area, nv):
    L = np.empty(3)
    lib().ref_lindholm_weights(p(f64(x0)), p(f64(p1)), p(f64(p2)), p(f64(p3)), float(area), p(f64(nv)), p(L))
    return L


def set_threads(t):
    lib().ref_set_num_threads(int(t))


def num_threads():
    return lib().ref_num_threads()
