"""Problem definition: material data at nodes and the operators handed to the GPU.

Mirrors what the reference host prepares before it calls the GPU classes:
SimulationData::setMaterialParametersAtNodes / scaleToRealSize
(SimulationData.cpp:45-49, 125-163), the exchange operator and the invJs /
DMI data of the TheLLG constructor (TheLLG.cpp:88-110), the Zeeman field
(TheLLG.cpp:442-449, 477-483) and the two row-replaced demag matrices
(DemagField.cpp:204-220, FEMprocessing.cpp:329-369).

Also defines the seeded workloads the tests and the benchmark run.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M

MU0 = 1.256637061435917295376e-6          # PhysicalConstants.h:39
GAMMA0 = 221276.14886372426114160400      # PhysicalConstants.h:43


@dataclass
class Material:
    Ms: float = 8.0e5          # A/m
    A: float = 1.3e-11         # J/m
    Ku: float = 0.0            # J/m^3
    ku_axis: tuple = (0.0, 0.0, 1.0)
    D: float = 0.0             # J/m^2, bulk DMI (EffFieldCalc.cpp:159-170)
    Di: float = 0.0            # J/m^2, interfacial DMI, interface normal z (no reference counterpart)
    alpha: float = 0.5
    gamma: float = 2.211e5     # m/(A s)
    B_ext: tuple = (0.0, 0.0, 0.0)  # applied field in T

    @property
    def Js(self) -> float:     # MaterialReader.cpp:94
        return MU0 * self.Ms


@dataclass
class Problem:
    name: str
    mesh: M.Mesh
    ops: M.Operators
    bnd: M.Boundary
    mat: Material
    Js: np.ndarray = field(default=None)       # (n,)  T
    A: np.ndarray = field(default=None)        # (n,)  J/m / scale^2
    Ku: np.ndarray = field(default=None)       # (n,)
    kaxis: np.ndarray = field(default=None)    # (3n,) SoA
    D: np.ndarray = field(default=None)        # (n,)  J/m^2 / scale
    Di: np.ndarray = field(default=None)       # (n,)  J/m^2 / scale (interfacial)
    hext: np.ndarray = field(default=None)     # (3,)  A/m
    m0: np.ndarray = field(default=None)       # (3n,) initial magnetisation

    def __post_init__(self):
        n = self.mesh.n
        s = self.mesh.scale
        mat = self.mat
        self.Js = np.full(n, mat.Js)
        self.A = np.full(n, mat.A / (s * s))           # SimulationData.cpp:46
        self.Ku = np.full(n, mat.Ku)
        ax = np.asarray(mat.ku_axis, np.float64)
        ax = ax / np.linalg.norm(ax)
        self.kaxis = np.concatenate([np.full(n, ax[0]), np.full(n, ax[1]), np.full(n, ax[2])])
        self.D = np.full(n, mat.D / s)                  # SimulationData.cpp:48
        self.Di = np.full(n, mat.Di / s)
        self.hext = np.asarray(mat.B_ext, np.float64) / MU0

    @property
    def n(self) -> int:
        return self.mesh.n

    @property
    def alpha(self) -> float:
        return self.mat.alpha

    def time_scale(self) -> float:
        """tau per ps: gamma / (1 + alpha^2) * 1e-12 (SimulationData.cpp:52-57)."""
        return self.mat.gamma / (1.0 + self.mat.alpha ** 2) * 1e-12

    # -- operators as the reference host builds them ------------------------
    def row_of(self) -> np.ndarray:
        return np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.ops.rowptr))

    def exchange_values(self) -> np.ndarray:
        """XC_field_OP = -2 diag(A / V) K (TheLLG.cpp:88-93)."""
        f = -2.0 * (self.A * (1.0 / self.ops.nodevol))
        return f[self.row_of()] * self.ops.K

    def neumann_matrix(self):
        """Stiffness with row nx/2 replaced by the identity (DemagField.cpp:79-80, 204-220)."""
        fixed = np.zeros(self.n, bool)
        fixed[self.n // 2] = True
        return M.replace_rows_identity(self.ops.rowptr, self.ops.col, self.ops.K, fixed)

    def dirichlet_matrix(self):
        """Stiffness with every boundary row replaced by the identity (FEMprocessing.cpp:329-369)."""
        fixed = np.zeros(self.n, bool)
        fixed[self.bnd.bnodes] = True
        return M.replace_rows_identity(self.ops.rowptr, self.ops.col, self.ops.K, fixed)

    @property
    def use_uniaxial(self) -> bool:
        return bool(np.any(self.Ku != 0))

    @property
    def use_dmi(self) -> bool:       # TheLLG.cpp:84
        return bool(np.any(self.D != 0))

    @property
    def use_idmi(self) -> bool:
        return bool(np.any(self.Di != 0))

    def summary(self) -> dict:
        return {"n": self.n, "tets": self.mesh.ntet, "bnx": self.bnd.bnx, "nbel": self.bnd.nbel,
                "nnz": self.ops.nnz}


def _unit_axis(seed: int) -> tuple:
    u = M.uniform(seed, 2)
    phi = 2.0 * math.pi * u[0]
    ct = 2.0 * u[1] - 1.0
    st = math.sqrt(max(0.0, 1.0 - ct * ct))
    return (st * math.cos(phi), st * math.sin(phi), ct)


def vortex(mesh: M.Mesh, core_radius: float = 10.0) -> np.ndarray:
    """m = (-y, x, c)/|.| about the xy-centre; c = 1 for r < core_radius, else 0."""
    xyz = mesh.xyz
    cx = 0.5 * (xyz[:, 0].min() + xyz[:, 0].max())
    cy = 0.5 * (xyz[:, 1].min() + xyz[:, 1].max())
    x = xyz[:, 0] - cx
    y = xyz[:, 1] - cy
    r = np.hypot(x, y)
    c = np.where(r < core_radius, 1.0, 0.0)
    v = np.stack([-y, x, c], axis=1)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return np.ascontiguousarray(v.T.reshape(-1))


def mesh_problem(name: str, mesh: M.Mesh, mat: Material | None = None, init="random", seed: int = 0) -> Problem:
    """Problem on any tetrahedral mesh (e.g. one read by meshio.read_vtk);
    init is "random", "vortex" or a (3n,) array."""
    ops = M.fem_operators(mesh)
    bnd = M.boundary(mesh)
    p = Problem(name, mesh, ops, bnd, mat or Material())
    if isinstance(init, np.ndarray):
        p.m0 = np.ascontiguousarray(init, dtype=np.float64)
    elif init == "random":
        p.m0 = M.init_random_unit(mesh.n, seed + 1)
    elif init == "vortex":
        p.m0 = vortex(mesh)
    else:
        raise ValueError(init)
    return p


def cuboid_problem(name: str, nx: int, ny: int, nz: int, h: float, jitter: float = 0.0, seed: int = 0,
                   mat: Material | None = None, init: str = "random", locality: bool = False) -> Problem:
    """Kuhn cuboid; its nodes numbered row by row (x fastest), or along a
    Z-order curve (mesh.locality_order) with locality=True."""
    mesh = M.kuhn_cuboid(nx, ny, nz, h, jitter, seed)
    if locality:
        mesh = M.renumber(mesh, M.locality_order(mesh))
    return mesh_problem(name, mesh, mat, init, seed)


def cuboid_20x20x4(seed: int = 2024) -> Problem:
    """Kuhn cuboid 20x20x4 cells at 5 nm: 2,205 nodes, 9,600 tets, 1,122 boundary
    nodes; Ms 8e5 A/m, A 1.3e-11 J/m, K = D = 0, alpha 0.5; seeded random m."""
    return cuboid_problem("cuboid-20x20x4-h5", 20, 20, 4, 5.0, 0.0, seed, Material(), "random")


def cuboid_128x128x16(seed: int = 2025, locality: bool = False) -> Problem:
    """Kuhn cuboid 128x128x16 cells at 3 nm with +-0.1h interior jitter:
    282,897 nodes, 1,572,864 tets, 40,962 boundary nodes; the 20x20x4 cuboid's material plus
    Ku = 5e4 J/m^3 along a seeded unit axis; vortex initial m.  locality:
    nodes numbered along a Z-order curve instead of row by row."""
    mat = Material(Ku=5e4, ku_axis=_unit_axis(seed + 7))
    return cuboid_problem("cuboid-128x128x16-h3", 128, 128, 16, 3.0, 0.1, seed, mat, "vortex", locality)


def smooth_random_fields(mesh: M.Mesh, count: int, seed: int, modes: int = 8, lam_min: float = 20.0,
                         lam_max: float = 200.0) -> list:
    """Seeded smooth fields m = normalise(sum_k a_k cos(q_k . r + phi_k)), |q_k| =
    2 pi / lambda_k with lambda_k ~ U[lam_min, lam_max] (mesh units), random
    directions, amplitudes a_k (3-vectors) and phases."""
    out = []
    xyz = mesh.xyz
    for s in range(count):
        u = M.uniform(seed * 1000003 + s, modes * 9)
        acc = np.zeros((mesh.n, 3))
        for k in range(modes):
            v = u[9 * k:9 * k + 9]
            lam = lam_min + (lam_max - lam_min) * v[0]
            phi_d = 2 * math.pi * v[1]
            ct = 2 * v[2] - 1
            st = math.sqrt(max(0.0, 1 - ct * ct))
            q = (2 * math.pi / lam) * np.array([st * math.cos(phi_d), st * math.sin(phi_d), ct])
            ph = 2 * math.pi * v[3]
            a = 2 * v[4:7] - 1
            acc += np.cos(xyz @ q + ph)[:, None] * a[None, :]
        nrm = np.linalg.norm(acc, axis=1, keepdims=True)
        nrm[nrm == 0] = 1.0
        out.append(np.ascontiguousarray((acc / nrm).T.reshape(-1)))
    return out


def neel_skyrmion(mesh: M.Mesh, radius: float, chirality: float = 1.0) -> np.ndarray:
    """Neel skyrmion about the xy-centre: core down in an up background,
    theta(r) = 2 atan(radius / r) (theta = pi/2 at r = radius), in-plane part
    along +-r_hat (chirality +1 or -1)."""
    xyz = mesh.xyz
    cx = 0.5 * (xyz[:, 0].min() + xyz[:, 0].max())
    cy = 0.5 * (xyz[:, 1].min() + xyz[:, 1].max())
    x, y = xyz[:, 0] - cx, xyz[:, 1] - cy
    r = np.hypot(x, y)
    th = 2.0 * np.arctan2(radius, r)
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(r > 0, x / r, 0.0)
        uy = np.where(r > 0, y / r, 0.0)
    m = np.stack([chirality * np.sin(th) * ux, chirality * np.sin(th) * uy, np.cos(th)], axis=1)
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    return np.ascontiguousarray(m.T.reshape(-1))


def disk_hysteresis(seed: int = 2027, h: float = 2.5, nsamples: int = 64):
    """Disk 200 nm x 10 nm, layered prism mesh at edge h (mesh.disk_prism); the 20x20x4
    cuboid's material; `nsamples` states, each under B_ext = mu0 H_k (cos t_i, sin t_i, 0)
    with mu0 H_k spaced from -80 to +80 mT and t_i seeded in [0, 5 deg],
    starting from a slightly perturbed saturated state along x.
    Returns (problem, fields (3, S) in A/m, m0 (3, n, S))."""
    mesh = M.disk_prism(200.0, 10.0, h, 0.15, seed)
    prob = mesh_problem(f"disk-200x10-h{h}", mesh, Material(), "random", seed)
    u = M.uniform(seed + 11, nsamples)
    th = np.deg2rad(5.0) * u
    Bk = np.linspace(-0.08, 0.08, nsamples)
    H = np.stack([Bk * np.cos(th), Bk * np.sin(th), np.zeros(nsamples)]) / MU0
    n = prob.n
    m0 = np.empty((3, n, nsamples))
    for s in range(nsamples):
        p = M.init_random_unit(n, seed + 100 + s).reshape(3, n)
        v = np.array([1.0, 0.0, 0.0])[:, None] + 0.05 * p
        m0[:, :, s] = v / np.linalg.norm(v, axis=0, keepdims=True)
    return prob, H, m0


def disk_skyrmion(seed: int = 2028, h: float = 0.6) -> Problem:
    """Disk 120 nm x 1 nm (layered prism mesh at edge h, mesh.disk_prism) with interfacial
    DMI 3e-3 J/m^2, Ku 8e5 J/m^3 along z, Ms 1.1e6 A/m, A 1.6e-11 J/m,
    alpha 0.3; Neel skyrmion of radius 15 nm."""
    mesh = M.disk_prism(120.0, 1.0, h, 0.15, seed)
    mat = Material(Ms=1.1e6, A=1.6e-11, Ku=8e5, ku_axis=(0.0, 0.0, 1.0), alpha=0.3, Di=3e-3)
    return mesh_problem(f"disk-120x1-h{h}", mesh, mat, neel_skyrmion(mesh, 15.0), seed)


def small_problem(nx=6, ny=5, nz=3, h=4.0, jitter=0.1, seed=7, Ku=3e4, D=0.0, B_ext=(0.0, 0.0, 0.0),
                  init="random") -> Problem:
    """Small seeded problem for parity tests (all field terms switchable)."""
    mat = Material(Ku=Ku, ku_axis=_unit_axis(seed + 3), D=D, B_ext=B_ext)
    return cuboid_problem(f"cuboid-{nx}x{ny}x{nz}-h{h}", nx, ny, nz, h, jitter, seed, mat, init)
