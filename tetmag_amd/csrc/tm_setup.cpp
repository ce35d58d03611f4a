// tm_setup.cpp -- host-side mesh and operator preparation for the HIP path.
//
// The reference assembles its operators on the host (preproc/FEMprocessing.cpp,
// preproc/Boundary.cpp, preproc/BEMprocessing.cpp) and hands them to the GPU
// classes.  This file is the MI355X build's own host implementation of that
// setup stage, written independently of the FP64 oracle:
//   * element gradients come from the inverse Jacobian (not from cofactor
//     determinants as in FEMprocessing.cpp:153-195),
//   * boundary faces are matched with a hash table (not a priority queue as in
//     Boundary.cpp:159-197),
//   * solid angles use the Van Oosterom-Strackee atan2 form (not the acos form
//     of Lindholm.cpp:41-60).
// The three give the same operators to rounding; the oracle tests check that.
//
// Layouts: xyz n x 3 row-major (mesh units), fel ntet x 4 row-major, CSR with
// sorted column indices, fields 3n SoA.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "tetmag_hip.h"

namespace {

// splitmix64: seeded, portable 64-bit generator used for all synthetic inputs.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // uniform in [0,1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Grads {
    double vol;
    double g[4][3];
};

// P1 gradients of a tetrahedron from the inverse Jacobian J = [p1-p0 | p2-p0 | p3-p0]:
// grad N_k (k=1..3) are the rows of J^{-1}; grad N_0 = -(sum of the others).
inline Grads tet_grads(const double* p0, const double* p1, const double* p2, const double* p3) {
    double J[3][3];
    for (int d = 0; d < 3; ++d) {
        J[d][0] = p1[d] - p0[d];
        J[d][1] = p2[d] - p0[d];
        J[d][2] = p3[d] - p0[d];
    }
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double inv[3][3];
    double id = 1.0 / det;
    inv[0][0] = c00 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][0] = c01 * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][0] = c02 * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    Grads G;
    G.vol = std::fabs(det) / 6.0;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) G.g[k + 1][d] = inv[k][d];
    for (int d = 0; d < 3; ++d) G.g[0][d] = -(G.g[1][d] + G.g[2][d] + G.g[3][d]);
    return G;
}

void node_incidence(int n, int ntet, const int* fel, std::vector<int>& ptr, std::vector<int>& idx) {
    ptr.assign((size_t)n + 1, 0);
    for (int e = 0; e < ntet; ++e)
        for (int j = 0; j < 4; ++j) ptr[fel[4 * e + j] + 1]++;
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    idx.resize(ptr[n]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < ntet; ++e)
        for (int j = 0; j < 4; ++j) idx[fill[fel[4 * e + j]]++] = e;
}

void row_neighbours(int i, const std::vector<int>& ptr, const std::vector<int>& idx, const int* fel,
                    std::vector<int>& out) {
    out.clear();
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
        for (int j = 0; j < 4; ++j) out.push_back(fel[4 * idx[q] + j]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

inline void sub3(const double* a, const double* b, double* r) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

extern "C" {

int tm_mesh_kuhn_cuboid_size(int nx, int ny, int nz, int* n_nodes, int* n_tets) {
    if (nx < 1 || ny < 1 || nz < 1 || !n_nodes || !n_tets) return TM_ERR_BAD_ARG;
    *n_nodes = (nx + 1) * (ny + 1) * (nz + 1);
    *n_tets = 6 * nx * ny * nz;
    return TM_OK;
}

int tm_mesh_kuhn_cuboid(int nx, int ny, int nz, double h, double jitter, uint64_t seed, double* xyz, int* fel) {
    if (nx < 1 || ny < 1 || nz < 1 || !(h > 0) || !xyz || !fel || jitter < 0 || jitter >= 0.25)
        return TM_ERR_BAD_ARG;
    const int sx = nx + 1, sy = ny + 1;
    auto id = [&](int i, int j, int k) { return i + sx * (j + sy * k); };
    SplitMix64 rng(seed);
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                int v = id(i, j, k);
                double p[3] = {i * h, j * h, k * h};
                bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                for (int d = 0; d < 3; ++d) {
                    double r = rng.uniform();  // drawn for every node so the stream is size-stable
                    if (interior && jitter > 0) p[d] += (2.0 * r - 1.0) * jitter * h;
                }
                xyz[3 * v + 0] = p[0];
                xyz[3 * v + 1] = p[1];
                xyz[3 * v + 2] = p[2];
            }
    // Kuhn (Freudenthal) split: six tets per cube along the (0,0,0)-(1,1,1)
    // diagonal, one per ordering of the axes; conforming across cubes.
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int t = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int q = 0; q < 6; ++q) {
                    int c[3] = {i, j, k};
                    fel[4 * t + 0] = id(c[0], c[1], c[2]);
                    for (int s = 0; s < 3; ++s) {
                        c[perm[q][s]] += 1;
                        fel[4 * t + 1 + s] = id(c[0], c[1], c[2]);
                    }
                    ++t;
                }
    return TM_OK;
}

// Node numbering for gather locality: nodes sorted along a Z-order (Morton)
// curve through a cubic grid of 1024 cells over the largest extent of the
// bounding box, the axis of smallest extent varying fastest.  Nodes close in
// space get close numbers, so the rows a kernel walks in sequence gather
// nearby rows that are still in cache.
int tm_mesh_locality_order(int n, const double* xyz, int* perm) {
    if (n < 0 || (n > 0 && (!xyz || !perm))) return TM_ERR_BAD_ARG;
    if (n == 0) return TM_OK;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = xyz[d];
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            const double v = xyz[3 * (size_t)i + d];
            if (!std::isfinite(v)) return TM_ERR_BAD_ARG;
            lo[d] = std::min(lo[d], v);
            hi[d] = std::max(hi[d], v);
        }
    double ext = 0.0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    const double cell = ext > 0.0 ? ext / 1024.0 : 1.0;
    int axis[3] = {0, 1, 2};  // bit order: smallest extent first (fastest)
    std::sort(axis, axis + 3, [&](int a, int b) { return hi[a] - lo[a] < hi[b] - lo[b]; });
    std::vector<std::pair<uint64_t, int>> key((size_t)n);
    for (int i = 0; i < n; ++i) {
        uint64_t q[3];
        for (int d = 0; d < 3; ++d)
            q[d] = (uint64_t)std::min(1023.0, std::floor((xyz[3 * (size_t)i + d] - lo[d]) / cell));
        uint64_t code = 0;
        for (int bit = 0; bit < 10; ++bit)
            for (int k = 0; k < 3; ++k) code |= ((q[axis[k]] >> bit) & 1u) << (3 * bit + k);
        key[(size_t)i] = {code, i};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int>& a, const std::pair<uint64_t, int>& b) { return a.first < b.first; });
    for (int k = 0; k < n; ++k) perm[k] = key[(size_t)k].second;
    return TM_OK;
}

int tm_fem_pattern(int n, int ntet, const int* fel, int* rowptr) {
    if (n <= 0 || ntet <= 0 || !fel || !rowptr) return TM_ERR_BAD_ARG;
    std::vector<int> ptr, idx;
    node_incidence(n, ntet, fel, ptr, idx);
    rowptr[0] = 0;
#pragma omp parallel
    {
        std::vector<int> nb;
#pragma omp for schedule(dynamic, 512)
        for (int i = 0; i < n; ++i) {
            row_neighbours(i, ptr, idx, fel, nb);
            rowptr[i + 1] = (int)nb.size();
        }
    }
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    return TM_OK;
}

int tm_fem_assemble(int n, int ntet, const double* xyz, const int* fel, const int* rowptr, int* col, double* K,
                    double* gx, double* gy, double* gz, double* tgx, double* tgy, double* tgz, double* nodevol,
                    double* elvol) {
    if (n <= 0 || ntet <= 0 || !xyz || !fel || !rowptr || !col || !K || !nodevol) return TM_ERR_BAD_ARG;
    std::vector<Grads> G((size_t)ntet);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ntet; ++e) {
        const int* t = fel + 4 * e;
        G[e] = tet_grads(xyz + 3 * t[0], xyz + 3 * t[1], xyz + 3 * t[2], xyz + 3 * t[3]);
    }
    std::vector<int> ptr, idx;
    node_incidence(n, ntet, fel, ptr, idx);
    for (int i = 0; i < n; ++i) {
        double v = 0.0;
        for (int q = ptr[i]; q < ptr[i + 1]; ++q) v += G[idx[q]].vol * 0.25;
        nodevol[i] = v;
    }
    if (elvol)
        for (int e = 0; e < ntet; ++e) elvol[e] = G[e].vol;
#pragma omp parallel
    {
        std::vector<int> nb;
#pragma omp for schedule(dynamic, 512)
        for (int i = 0; i < n; ++i) {
            row_neighbours(i, ptr, idx, fel, nb);
            const int lo = rowptr[i];
            const int len = rowptr[i + 1] - lo;
            for (int q = 0; q < len; ++q) {
                col[lo + q] = nb[q];
                K[lo + q] = 0.0;
                if (gx) gx[lo + q] = gy[lo + q] = gz[lo + q] = 0.0;
                if (tgx) tgx[lo + q] = tgy[lo + q] = tgz[lo + q] = 0.0;
            }
            const double iv = 1.0 / nodevol[i];
            for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
                const int e = idx[q];
                const int* t = fel + 4 * e;
                int a = 0;
                while (t[a] != i) ++a;
                const Grads& g = G[e];
                const double w = 0.25 * g.vol;
                for (int b = 0; b < 4; ++b) {
                    const int p = lo + (int)(std::lower_bound(col + lo, col + lo + len, t[b]) - (col + lo));
                    K[p] += g.vol * dot3(g.g[a], g.g[b]);
                    if (gx) {
                        // grad_d(i,j) = sum_e dN_j/dx_d V_e/4 / V_i   (FEMprocessing.cpp:243-245)
                        gx[p] += g.g[b][0] * w * iv;
                        gy[p] += g.g[b][1] * w * iv;
                        gz[p] += g.g[b][2] * w * iv;
                    }
                    if (tgx) {
                        // tGrad_d(i,j) = sum_e dN_i/dx_d V_e/4          (FEMprocessing.cpp:247-249)
                        tgx[p] += g.g[a][0] * w;
                        tgy[p] += g.g[a][1] * w;
                        tgz[p] += g.g[a][2] * w;
                    }
                }
            }
        }
    }
    return TM_OK;
}

// Boundary faces: faces that belong to exactly one tet.  Output ordered by
// ascending sorted node triple.  faces needs 4*ntet*3 ints, face_tet 4*ntet.
int tm_boundary_faces(int ntet, const int* fel, int* faces, int* face_tet, int* nbel_out) {
    if (ntet <= 0 || !fel || !faces || !face_tet || !nbel_out) return TM_ERR_BAD_ARG;
    struct Key {
        int a, b, c;
        bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
    };
    struct Hash {
        size_t operator()(const Key& k) const {
            uint64_t h = (uint64_t)(uint32_t)k.a * 0x9E3779B97F4A7C15ULL;
            h ^= (uint64_t)(uint32_t)k.b * 0xC2B2AE3D27D4EB4FULL + (h << 6) + (h >> 2);
            h ^= (uint64_t)(uint32_t)k.c * 0x165667B19E3779F9ULL + (h << 6) + (h >> 2);
            return (size_t)h;
        }
    };
    std::unordered_map<Key, std::pair<int, int>, Hash> seen;  // face -> (count, tet)
    seen.reserve((size_t)ntet * 3);
    for (int e = 0; e < ntet; ++e) {
        for (int skip = 0; skip < 4; ++skip) {
            int v[3], k = 0;
            for (int q = 0; q < 4; ++q)
                if (q != skip) v[k++] = fel[4 * e + q];
            std::sort(v, v + 3);
            auto it = seen.find(Key{v[0], v[1], v[2]});
            if (it == seen.end())
                seen.emplace(Key{v[0], v[1], v[2]}, std::make_pair(1, e));
            else
                it->second.first++;
        }
    }
    std::vector<std::pair<Key, int>> bf;
    for (auto& kv : seen)
        if (kv.second.first == 1) bf.push_back({kv.first, kv.second.second});
    std::sort(bf.begin(), bf.end(), [](const std::pair<Key, int>& x, const std::pair<Key, int>& y) {
        if (x.first.a != y.first.a) return x.first.a < y.first.a;
        if (x.first.b != y.first.b) return x.first.b < y.first.b;
        return x.first.c < y.first.c;
    });
    for (size_t f = 0; f < bf.size(); ++f) {
        faces[3 * f + 0] = bf[f].first.a;
        faces[3 * f + 1] = bf[f].first.b;
        faces[3 * f + 2] = bf[f].first.c;
        face_tet[f] = bf[f].second;
    }
    *nbel_out = (int)bf.size();
    return TM_OK;
}

// Boundary node list (ascending), index map, face areas, outward unit face
// normals, node areas (sum A_f/3) and node normals (normalised sum of
// A_f/3-weighted face normals; zero away from the boundary).  node_normal is
// 3n SoA.  bnodes needs n ints.
int tm_boundary_props(int n, const double* xyz, const int* fel, int nbel, const int* faces, const int* face_tet,
                      int* bnodes, int* bindex, double* face_area, double* face_normal, double* node_area,
                      double* node_normal, int* bnx_out) {
    if (n <= 0 || nbel <= 0 || !xyz || !fel || !faces || !face_tet || !bnodes || !bindex || !face_area ||
        !face_normal || !node_area || !node_normal || !bnx_out)
        return TM_ERR_BAD_ARG;
    std::vector<char> isb((size_t)n, 0);
    for (int f = 0; f < 3 * nbel; ++f) isb[faces[f]] = 1;
    int b = 0;
    for (int i = 0; i < n; ++i) {
        bindex[i] = isb[i] ? b : -1;
        if (isb[i]) bnodes[b++] = i;
    }
    std::vector<double> acc((size_t)3 * n, 0.0);
    for (int i = 0; i < n; ++i) node_area[i] = 0.0;
    for (int f = 0; f < nbel; ++f) {
        const double* p0 = xyz + 3 * faces[3 * f];
        const double* p1 = xyz + 3 * faces[3 * f + 1];
        const double* p2 = xyz + 3 * faces[3 * f + 2];
        double e1[3], e2[3], c[3];
        sub3(p1, p0, e1);
        sub3(p2, p0, e2);
        cross3(e1, e2, c);
        double len = std::sqrt(dot3(c, c));
        face_area[f] = 0.5 * len;
        // orient away from the centroid of the owning tet
        const int* t = fel + 4 * face_tet[f];
        double opp[3] = {0, 0, 0};
        for (int q = 0; q < 4; ++q)
            for (int d = 0; d < 3; ++d) opp[d] += 0.25 * xyz[3 * t[q] + d];
        double w[3];
        sub3(p0, opp, w);
        double s = dot3(c, w) >= 0 ? 1.0 : -1.0;
        for (int d = 0; d < 3; ++d) face_normal[3 * f + d] = s * c[d] / len;
        for (int j = 0; j < 3; ++j) {
            int v = faces[3 * f + j];
            node_area[v] += face_area[f] / 3.0;
            for (int d = 0; d < 3; ++d) acc[3 * (size_t)v + d] += face_normal[3 * f + d] * face_area[f] / 3.0;
        }
    }
    for (int i = 0; i < n; ++i) {
        double l = std::sqrt(acc[3 * (size_t)i] * acc[3 * (size_t)i] + acc[3 * (size_t)i + 1] * acc[3 * (size_t)i + 1] +
                             acc[3 * (size_t)i + 2] * acc[3 * (size_t)i + 2]);
        for (int d = 0; d < 3; ++d) node_normal[(size_t)d * n + i] = l > 0 ? acc[3 * (size_t)i + d] / l : 0.0;
    }
    *bnx_out = b;
    return TM_OK;
}

// Interior solid angle at every boundary node (sum over incident tets of the
// angle subtended by the opposite face), Van Oosterom-Strackee form:
// tan(Omega/2) = |r1.(r2 x r3)| / (r1 r2 r3 + (r1.r2) r3 + (r2.r3) r1 + (r3.r1) r2).
int tm_boundary_solid_angles(int n, int ntet, const double* xyz, const int* fel, const int* bindex, int bnx,
                             double* omega) {
    if (n <= 0 || ntet <= 0 || !xyz || !fel || !bindex || bnx <= 0 || !omega) return TM_ERR_BAD_ARG;
    for (int i = 0; i < bnx; ++i) omega[i] = 0.0;
    for (int e = 0; e < ntet; ++e) {
        const int* t = fel + 4 * e;
        for (int a = 0; a < 4; ++a) {
            int bi = bindex[t[a]];
            if (bi < 0) continue;
            const double* x = xyz + 3 * t[a];
            double r1[3], r2[3], r3[3], c[3];
            sub3(xyz + 3 * t[(a + 1) % 4], x, r1);
            sub3(xyz + 3 * t[(a + 2) % 4], x, r2);
            sub3(xyz + 3 * t[(a + 3) % 4], x, r3);
            double l1 = std::sqrt(dot3(r1, r1)), l2 = std::sqrt(dot3(r2, r2)), l3 = std::sqrt(dot3(r3, r3));
            cross3(r2, r3, c);
            double num = std::fabs(dot3(r1, c));
            double den = l1 * l2 * l3 + dot3(r1, r2) * l3 + dot3(r2, r3) * l1 + dot3(r3, r1) * l2;
            double half = std::atan2(num, den);
            if (half < 0) half += M_PI;
            omega[bi] += 2.0 * half;
        }
    }
    return TM_OK;
}

// Seeded i.i.d. unit vectors (normalised Gaussian triples, Box-Muller), 3n SoA.
int tm_init_random_unit(int n, uint64_t seed, double* m) {
    if (n <= 0 || !m) return TM_ERR_BAD_ARG;
    SplitMix64 rng(seed);
    for (int i = 0; i < n; ++i) {
        double g[3];
        double l = 0.0;
        do {
            for (int d = 0; d < 3; ++d) {
                double u1 = rng.uniform(), u2 = rng.uniform();
                if (u1 < 1e-300) u1 = 1e-300;
                g[d] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            }
            l = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        } while (l < 1e-12);
        for (int d = 0; d < 3; ++d) m[(size_t)d * n + i] = g[d] / l;
    }
    return TM_OK;
}

// Seeded uniform 64-bit-derived doubles in [0,1): helper for workload seeds.
int tm_uniform(uint64_t seed, int count, double* out) {
    if (count < 0 || (count > 0 && !out)) return TM_ERR_BAD_ARG;
    SplitMix64 rng(seed);
    for (int i = 0; i < count; ++i) out[i] = rng.uniform();
    return TM_OK;
}

}  // extern "C"
