"""Mesh generation and P1/boundary operator preparation (host, C++ via the C ABI).

This is the setup stage of the reference (preproc/FEMprocessing.cpp,
preproc/Boundary.cpp, preproc/BEMprocessing.cpp): it runs once per mesh on
the host and produces the operators the GPU path consumes.  The dense
boundary-element matrix itself is assembled on the GPU by
``FieldContext.assemble_bem``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class Mesh:
    xyz: np.ndarray  # (n, 3) float64, mesh units
    fel: np.ndarray  # (ntet, 4) int32
    scale: float = 1e-9  # metres per mesh unit (ProgramSpecs "scale")

    @property
    def n(self) -> int:
        return self.xyz.shape[0]

    @property
    def ntet(self) -> int:
        return self.fel.shape[0]


@dataclass
class Operators:
    """P1 operators on one shared CSR pattern (sorted columns)."""
    rowptr: np.ndarray
    col: np.ndarray
    K: np.ndarray     # stiffness  (FEMprocessing.cpp:266-286)
    gx: np.ndarray    # grad_d     (FEMprocessing.cpp:243-245)
    gy: np.ndarray
    gz: np.ndarray
    tgx: np.ndarray   # tGrad_d    (FEMprocessing.cpp:247-249)
    tgy: np.ndarray
    tgz: np.ndarray
    nodevol: np.ndarray
    elvol: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])


@dataclass
class Boundary:
    faces: np.ndarray        # (nbel, 3) global node ids, sorted per face
    face_tet: np.ndarray     # (nbel,)
    bnodes: np.ndarray       # (bnx,) ascending
    bindex: np.ndarray       # (n,) boundary index or -1
    face_area: np.ndarray    # (nbel,)
    face_normal: np.ndarray  # (nbel, 3) outward unit normals
    node_area: np.ndarray    # (n,)
    node_normal: np.ndarray  # (3n,) SoA, zero off the boundary
    solid_angle: np.ndarray  # (bnx,) interior solid angle

    @property
    def bnx(self) -> int:
        return self.bnodes.shape[0]

    @property
    def nbel(self) -> int:
        return self.faces.shape[0]


def kuhn_cuboid(nx: int, ny: int, nz: int, h: float, jitter: float = 0.0, seed: int = 0,
                scale: float = 1e-9) -> Mesh:
    L = N.lib()
    n = C.c_int()
    t = C.c_int()
    N.check(L.tm_mesh_kuhn_cuboid_size(nx, ny, nz, C.byref(n), C.byref(t)), "tm_mesh_kuhn_cuboid_size")
    xyz = np.empty((n.value, 3), np.float64)
    fel = np.empty((t.value, 4), np.int32)
    N.check(L.tm_mesh_kuhn_cuboid(nx, ny, nz, float(h), float(jitter), int(seed), N.ptr(xyz), N.ptr(fel)),
            "tm_mesh_kuhn_cuboid")
    return Mesh(xyz, fel, scale)


def locality_order(mesh: Mesh) -> np.ndarray:
    """perm[k] = the old number of new node k along a Z-order curve of the
    coordinates (tm_mesh_locality_order): spatially close nodes get close
    numbers, which keeps the batched gather kernels' rows in cache."""
    perm = np.empty(mesh.n, np.int32)
    N.check(N.lib().tm_mesh_locality_order(mesh.n, N.ptr(N.f64(mesh.xyz)), N.ptr(perm)), "tm_mesh_locality_order")
    return perm


def renumber(mesh: Mesh, perm: np.ndarray) -> Mesh:
    """The same mesh with node k = old node perm[k]."""
    perm = np.asarray(perm, np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.shape[0])
    return Mesh(np.ascontiguousarray(mesh.xyz[perm]), np.ascontiguousarray(inv[mesh.fel].astype(np.int32)),
                mesh.scale)


def disk_delaunay(diameter: float, thickness: float, h: float, jitter: float = 0.15, seed: int = 0,
                  scale: float = 1e-9) -> Mesh:
    """Seeded jittered Delaunay tetrahedralisation of a disk (cylinder of the
    given diameter and thickness, mesh units): layers of points every ~h in z,
    a square lattice of spacing h inside each layer plus a ring of points on
    the rim, interior points jittered by +-jitter*h (surface points stay on
    their surface).  Tetrahedra are oriented positively; flat ones (volume
    below 1e-9 h^3, from coplanar surface points) are dropped, then unused
    points.  Points are ordered layer by layer for locality."""
    from scipy.spatial import Delaunay
    R = 0.5 * diameter
    nz = max(1, int(round(thickness / h)))
    zs = np.linspace(0.0, thickness, nz + 1)
    g = np.arange(-R, R + 1e-9 * h, h)
    X, Y = np.meshgrid(g, g, indexing="xy")
    X, Y = X.reshape(-1), Y.reshape(-1)
    inside = X * X + Y * Y < (R - 0.5 * h) ** 2
    X, Y = X[inside], Y[inside]
    nr = max(8, int(np.ceil(2 * np.pi * R / h)))
    ang = 2 * np.pi * np.arange(nr) / nr
    RX, RY = R * np.cos(ang), R * np.sin(ang)
    u = uniform(seed, 3 * (X.size + nr) * (nz + 1))
    pts = []
    k = 0
    for iz, z in enumerate(zs):
        top = iz == 0 or iz == nz
        jx = jitter * h * (2 * u[k:k + X.size] - 1)
        jy = jitter * h * (2 * u[k + X.size:k + 2 * X.size] - 1)
        jz = np.zeros(X.size) if top else jitter * h * (2 * u[k + 2 * X.size:k + 3 * X.size] - 1)
        k += 3 * X.size
        px, py = X + jx, Y + jy
        r = np.hypot(px, py)
        f = np.where(r > R - 0.25 * h, (R - 0.25 * h) / np.maximum(r, 1e-30), 1.0)  # keep off the rim ring
        pts.append(np.stack([px * f, py * f, np.full(X.size, z) + jz], axis=1))
        pts.append(np.stack([RX, RY, np.full(nr, z)], axis=1))
    P = np.concatenate(pts)
    P = P[np.lexsort((P[:, 0], P[:, 1], np.round(P[:, 2] / h).astype(np.int64)))]
    tri = Delaunay(P, qhull_options="QJ Qbb Q12")  # joggled input: no degenerate cospherical sets
    T = tri.simplices.astype(np.int64)
    a = P[T[:, 1]] - P[T[:, 0]]
    b = P[T[:, 2]] - P[T[:, 0]]
    c = P[T[:, 3]] - P[T[:, 0]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    keep = np.abs(vol) > 1e-9 * h ** 3
    T, vol = T[keep], vol[keep]
    neg = vol < 0
    T[neg] = T[neg][:, [0, 2, 1, 3]]
    used = np.unique(T)
    remap = -np.ones(P.shape[0], np.int64)
    remap[used] = np.arange(used.size)
    return Mesh(np.ascontiguousarray(P[used]), np.ascontiguousarray(remap[T].astype(np.int32)), scale)


def disk_prism(diameter: float, thickness: float, h: float, jitter: float = 0.15, seed: int = 0,
               scale: float = 1e-9) -> Mesh:
    """Layered tetrahedral mesh of a disk (cylinder of the given diameter and
    thickness, mesh units), the usual mesh of a thin film: a 2D Delaunay
    triangulation of the cross-section (a square lattice of spacing h,
    jittered by +-jitter*h, plus a ring of points on the rim) extruded into
    round(thickness / h) layers of prisms, each prism split into three
    tetrahedra.  A 2D Delaunay triangulation maximises the smallest angle, so
    the mesh has no slivers (a 3D Delaunay of a jittered lattice has many,
    whose stiffness entries make the exchange operator, and so the explicit
    time step, hopeless).  The split of a prism over the triangle a < b < c
    (node numbers) into (a,b,c,c'), (a,b,b',c'), (a,a',b',c') takes each
    side quad's diagonal from its lowest-numbered node, so neighbouring
    prisms share faces.  Nodes are numbered column by column (z fastest),
    columns in raster order of the cross-section: a node's neighbours lie
    within a few rows of the raster."""
    from scipy.spatial import Delaunay
    R = 0.5 * diameter
    nz = max(1, int(round(thickness / h)))
    zs = np.linspace(0.0, thickness, nz + 1)
    g = np.arange(-R, R + 1e-9 * h, h)
    X, Y = np.meshgrid(g, g, indexing="xy")
    X, Y = X.reshape(-1), Y.reshape(-1)
    inside = X * X + Y * Y < (R - 0.5 * h) ** 2
    X, Y = X[inside], Y[inside]
    u = uniform(seed, 2 * X.size)
    X = X + jitter * h * (2 * u[:X.size] - 1)
    Y = Y + jitter * h * (2 * u[X.size:] - 1)
    nr = max(8, int(np.ceil(2 * np.pi * R / h)))
    ang = 2 * np.pi * np.arange(nr) / nr
    P2 = np.concatenate([np.stack([X, Y], axis=1), np.stack([R * np.cos(ang), R * np.sin(ang)], axis=1)])
    P2 = P2[np.lexsort((P2[:, 0], np.round(P2[:, 1] / h).astype(np.int64)))]  # raster order
    T2 = Delaunay(P2).simplices.astype(np.int64)
    e1, e2 = P2[T2[:, 1]] - P2[T2[:, 0]], P2[T2[:, 2]] - P2[T2[:, 0]]
    T2 = T2[np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]) > 1e-9 * h * h]
    T2 = np.sort(T2, axis=1)
    n2, L = P2.shape[0], nz + 1
    xyz = np.empty((n2 * L, 3))
    xyz[:, 0] = np.repeat(P2[:, 0], L)
    xyz[:, 1] = np.repeat(P2[:, 1], L)
    xyz[:, 2] = np.tile(zs, n2)
    tets = []
    for k in range(nz):
        a, b, c = T2[:, 0] * L + k, T2[:, 1] * L + k, T2[:, 2] * L + k
        tets += [np.stack([a, b, c, c + 1], 1), np.stack([a, b, b + 1, c + 1], 1), np.stack([a, a + 1, b + 1, c + 1], 1)]
    T = np.concatenate(tets)
    p = xyz[T]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    T[vol < 0] = T[vol < 0][:, [0, 2, 1, 3]]
    return Mesh(np.ascontiguousarray(xyz), np.ascontiguousarray(T.astype(np.int32)), scale)


def fem_operators(mesh: Mesh) -> Operators:
    L = N.lib()
    n, nt = mesh.n, mesh.ntet
    xyz = N.f64(mesh.xyz)
    fel = N.i32(mesh.fel)
    rowptr = np.empty(n + 1, np.int32)
    N.check(L.tm_fem_pattern(n, nt, N.ptr(fel), N.ptr(rowptr)), "tm_fem_pattern")
    nnz = int(rowptr[-1])
    col = np.empty(nnz, np.int32)
    arrs = [np.empty(nnz, np.float64) for _ in range(7)]
    nodevol = np.empty(n, np.float64)
    elvol = np.empty(nt, np.float64)
    N.check(L.tm_fem_assemble(n, nt, N.ptr(xyz), N.ptr(fel), N.ptr(rowptr), N.ptr(col),
                              *[N.ptr(a) for a in arrs], N.ptr(nodevol), N.ptr(elvol)), "tm_fem_assemble")
    K, gx, gy, gz, tgx, tgy, tgz = arrs
    return Operators(rowptr, col, K, gx, gy, gz, tgx, tgy, tgz, nodevol, elvol)


def boundary(mesh: Mesh) -> Boundary:
    L = N.lib()
    n, nt = mesh.n, mesh.ntet
    xyz = N.f64(mesh.xyz)
    fel = N.i32(mesh.fel)
    faces = np.empty((4 * nt, 3), np.int32)
    face_tet = np.empty(4 * nt, np.int32)
    nb = C.c_int()
    N.check(L.tm_boundary_faces(nt, N.ptr(fel), N.ptr(faces), N.ptr(face_tet), C.byref(nb)), "tm_boundary_faces")
    nbel = nb.value
    faces = np.ascontiguousarray(faces[:nbel])
    face_tet = np.ascontiguousarray(face_tet[:nbel])
    bnodes = np.empty(n, np.int32)
    bindex = np.empty(n, np.int32)
    farea = np.empty(nbel, np.float64)
    fnorm = np.empty((nbel, 3), np.float64)
    narea = np.empty(n, np.float64)
    nnorm = np.empty(3 * n, np.float64)
    bnx = C.c_int()
    N.check(L.tm_boundary_props(n, N.ptr(xyz), N.ptr(fel), nbel, N.ptr(faces), N.ptr(face_tet), N.ptr(bnodes),
                                N.ptr(bindex), N.ptr(farea), N.ptr(fnorm), N.ptr(narea), N.ptr(nnorm),
                                C.byref(bnx)), "tm_boundary_props")
    bnodes = np.ascontiguousarray(bnodes[:bnx.value])
    omega = np.empty(bnx.value, np.float64)
    N.check(L.tm_boundary_solid_angles(n, nt, N.ptr(xyz), N.ptr(fel), N.ptr(bindex), bnx.value, N.ptr(omega)),
            "tm_boundary_solid_angles")
    return Boundary(faces, face_tet, bnodes, bindex, farea, fnorm, narea, nnorm, omega)


def replace_rows_identity(rowptr, col, val, fixed_mask):
    """Rows flagged in fixed_mask become identity rows; explicit zeros are
    dropped (the row replacement + prune(0,0) of FEMprocessing.cpp:329-369 and
    DemagField.cpp:204-220)."""
    rowptr = np.asarray(rowptr)
    n = rowptr.shape[0] - 1
    row_of = np.repeat(np.arange(n, dtype=np.int32), np.diff(rowptr))
    keep = (~fixed_mask[row_of]) & (val != 0.0)
    rows = np.concatenate([row_of[keep], np.nonzero(fixed_mask)[0].astype(np.int32)])
    cols = np.concatenate([col[keep], np.nonzero(fixed_mask)[0].astype(np.int32)])
    vals = np.concatenate([val[keep], np.ones(int(fixed_mask.sum()))])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    return rp, N.i32(cols), N.f64(vals)


def init_random_unit(n: int, seed: int) -> np.ndarray:
    m = np.empty(3 * n, np.float64)
    N.check(N.lib().tm_init_random_unit(n, int(seed), N.ptr(m)), "tm_init_random_unit")
    return m


def uniform(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float64)
    N.check(N.lib().tm_uniform(int(seed), count, N.ptr(out)), "tm_uniform")
    return out
