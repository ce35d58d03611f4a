/*
 * refcpu.c -- FP64 CPU restatement of tetmag's effective-field + LLG-RHS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product (tetmag_amd/) never links or calls it.
 *
 * Every function restates one piece of the reference (tetmag v1.7.7) and cites
 * the file:line it follows.  The reference cannot be compiled in this
 * environment (Eigen3/Boost/VTK/AMGCL/SUNDIALS/H2Lib absent), so these are
 * hand restatements of the formulas, written in plain C with OpenMP.
 *
 * Conventions (all FP64, int32 indices):
 *   xyz   : n x 3 row-major node coordinates (mesh units)
 *   fel   : ntet x 4 row-major tetrahedron connectivity
 *   fields/magnetisation : "3n SoA" = x-block, y-block, z-block, i.e. the
 *           column-major n x 3 layout of the reference state vector
 *           (TheLLG.cpp:156, typedefs.h:39-41)
 *   sparse: CSR with sorted column indices (Eigen RowMajor SparseMatrix,
 *           typedefs.h:35)
 *   B     : dense bnx x bnx column-major (BEMprocessing.cpp:375-411)
 *
 * Pinning status of each function is recorded in the test files; see
 * tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_PI 3.14159265358979323846264338327950288
/* PhysicalConstants.h:39 */
#define REF_MU0 (1.256637061435917295376 * 1.e-6)

/* ------------------------------------------------------------------------ */
/* small 3-vector helpers                                                    */
/* ------------------------------------------------------------------------ */
static inline void v3sub(const double* a, const double* b, double* r) {
    r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline double v3dot(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void v3cross(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double v3norm(const double* a) { return sqrt(v3dot(a, a)); }

/* auxiliaries.cpp:42-47 -- areEqual(x, y, ulp=100) */
static int ref_are_equal(double x, double y, int ulp) {
    const double eps = DBL_EPSILON;
    int almost = fabs(x - y) < (eps * fabs(x + y) * ulp);
    if (fabs(x) < eps * ulp && fabs(y) < eps * ulp) almost = 1;
    return almost;
}

/* 3x3 determinant, rows r0,r1,r2 */
static inline double det3(const double* r0, const double* r1, const double* r2) {
    return r0[0] * (r1[1] * r2[2] - r1[2] * r2[1])
         - r0[1] * (r1[0] * r2[2] - r1[2] * r2[0])
         + r0[2] * (r1[0] * r2[1] - r1[1] * r2[0]);
}

/* FEMprocessing.cpp:115-123 -- det of [1 x y z] rows = 6 * signed volume.
 * Expanded along the column of ones. */
static double signed_vol6(const double* p1, const double* p2, const double* p3, const double* p4) {
    /* det [[1,p1],[1,p2],[1,p3],[1,p4]]: Laplace expansion along column 0 */
    return det3(p2, p3, p4) - det3(p1, p3, p4) + det3(p1, p2, p4) - det3(p1, p2, p3);
}

/* ------------------------------------------------------------------------ */
/* FEM setup                                                                 */
/* ------------------------------------------------------------------------ */

/* FEMprocessing.cpp:104-113 -- element volumes |det|/6, node volume = sum V_e/4 */
void ref_volumes(int n, int ntet, const double* xyz, const int* fel, double* elvol, double* nodevol) {
    for (int i = 0; i < n; ++i) nodevol[i] = 0.0;
    for (int e = 0; e < ntet; ++e) {
        const int* t = fel + 4 * e;
        double v = fabs(signed_vol6(xyz + 3 * t[0], xyz + 3 * t[1], xyz + 3 * t[2], xyz + 3 * t[3])) / 6.0;
        elvol[e] = v;
        for (int j = 0; j < 4; ++j) nodevol[t[j]] += v / 4.0;
    }
}

/* FEMprocessing.cpp:153-195 -- shape-function gradients by cofactors.
 * For local node a, node_mat rows are the coordinates of nodes (a+1)%4,
 * (a+2)%4, (a+3)%4; column d is replaced by ones; grad_d N_a =
 * (-1)^(a+1) det(.) / V6. */
void ref_shape_grads(int ntet, const double* xyz, const int* fel, double* dNx, double* dNy, double* dNz) {
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ntet; ++e) {
        const int* t = fel + 4 * e;
        const double* P[4] = {xyz + 3 * t[0], xyz + 3 * t[1], xyz + 3 * t[2], xyz + 3 * t[3]};
        double V6 = signed_vol6(P[0], P[1], P[2], P[3]);
        for (int a = 0; a < 4; ++a) {
            double rows[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) rows[r][c] = P[(a + 1 + r) % 4][c];
            double sgn = ((a + 1) % 2 == 0) ? 1.0 : -1.0;
            double out[3];
            for (int d = 0; d < 3; ++d) {
                double m[3][3];
                memcpy(m, rows, sizeof(m));
                for (int r = 0; r < 3; ++r) m[r][d] = 1.0;
                out[d] = sgn * det3(m[0], m[1], m[2]) / V6;
            }
            dNx[4 * e + a] = out[0];
            dNy[4 * e + a] = out[1];
            dNz[4 * e + a] = out[2];
        }
    }
}

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* node -> incident tets incidence (CSR) */
static void node_tets(int n, int ntet, const int* fel, int** ptr_out, int** idx_out) {
    int* ptr = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int e = 0; e < ntet; ++e)
        for (int j = 0; j < 4; ++j) ptr[fel[4 * e + j] + 1]++;
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    int* idx = (int*)malloc(sizeof(int) * (size_t)ptr[n]);
    int* fill = (int*)malloc(sizeof(int) * (size_t)n);
    memcpy(fill, ptr, sizeof(int) * (size_t)n);
    for (int e = 0; e < ntet; ++e)
        for (int j = 0; j < 4; ++j) idx[fill[fel[4 * e + j]]++] = e;
    free(fill);
    *ptr_out = ptr;
    *idx_out = idx;
}

/* Sparsity of the P1 operators: (i,j) is present iff i and j share a tet.
 * This is the set of entries touched by coeffRef in FEMprocessing.cpp:237-252
 * and 256-264.  Returns row lengths into rowptr (size n+1, prefix-summed). */
int ref_fem_pattern(int n, int ntet, const int* fel, int* rowptr) {
    int *tp, *ti;
    node_tets(n, ntet, fel, &tp, &ti);
    rowptr[0] = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        int cnt = tp[i + 1] - tp[i];
        int* buf = (int*)malloc(sizeof(int) * (size_t)(4 * cnt + 1));
        int k = 0;
        for (int q = tp[i]; q < tp[i + 1]; ++q)
            for (int j = 0; j < 4; ++j) buf[k++] = fel[4 * ti[q] + j];
        qsort(buf, (size_t)k, sizeof(int), cmp_int);
        int u = 0;
        for (int q = 0; q < k; ++q)
            if (q == 0 || buf[q] != buf[q - 1]) ++u;
        rowptr[i + 1] = u;
        free(buf);
    }
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    free(tp);
    free(ti);
    return rowptr[n];
}

static inline int find_col(const int* col, int lo, int hi, int j) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (col[mid] < j) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* FEMprocessing.cpp:228-253 (gradient operators) and 266-286 (stiffness).
 *   grad_d(i,j)  += dN_d[j] * (V_e/4) / NodeVolume(i)
 *   tGrad_d(i,j) += dN_d[i] * (V_e/4)
 *   K(i,j)       += V_e * sum_d dN_d[i] dN_d[j]
 * Output on the pattern from ref_fem_pattern (col filled here). */
void ref_fem_assemble(int n, int ntet, const double* xyz, const int* fel, const int* rowptr, int* col,
                      double* K, double* gX, double* gY, double* gZ, double* tX, double* tY, double* tZ,
                      double* nodevol) {
    int *tp, *ti;
    node_tets(n, ntet, fel, &tp, &ti);
    double* elvol = (double*)malloc(sizeof(double) * (size_t)ntet);
    double* dN = (double*)malloc(sizeof(double) * (size_t)ntet * 12);
    double *dNx = dN, *dNy = dN + 4 * (size_t)ntet, *dNz = dN + 8 * (size_t)ntet;
    ref_volumes(n, ntet, xyz, fel, elvol, nodevol);
    ref_shape_grads(ntet, xyz, fel, dNx, dNy, dNz);
    /* column pattern */
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        int cnt = tp[i + 1] - tp[i];
        int* buf = (int*)malloc(sizeof(int) * (size_t)(4 * cnt + 1));
        int k = 0;
        for (int q = tp[i]; q < tp[i + 1]; ++q)
            for (int j = 0; j < 4; ++j) buf[k++] = fel[4 * ti[q] + j];
        qsort(buf, (size_t)k, sizeof(int), cmp_int);
        int u = rowptr[i];
        for (int q = 0; q < k; ++q)
            if (q == 0 || buf[q] != buf[q - 1]) col[u++] = buf[q];
        free(buf);
    }
    /* row-wise accumulation: row i gathers contributions of its incident tets,
     * in increasing tet order (the reference adds tets in increasing order). */
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        int lo = rowptr[i], hi = rowptr[i + 1];
        for (int p = lo; p < hi; ++p) {
            K[p] = gX[p] = gY[p] = gZ[p] = tX[p] = tY[p] = tZ[p] = 0.0;
        }
        for (int q = tp[i]; q < tp[i + 1]; ++q) {
            int e = ti[q];
            const int* t = fel + 4 * e;
            int li = 0;
            while (t[li] != i) ++li;
            double w = elvol[e] / 4.0;
            for (int lj = 0; lj < 4; ++lj) {
                int j = t[lj];
                int p = find_col(col, lo, hi, j);
                gX[p] += dNx[4 * e + lj] * w / nodevol[i];
                gY[p] += dNy[4 * e + lj] * w / nodevol[i];
                gZ[p] += dNz[4 * e + lj] * w / nodevol[i];
                tX[p] += dNx[4 * e + li] * w;
                tY[p] += dNy[4 * e + li] * w;
                tZ[p] += dNz[4 * e + li] * w;
                K[p] += elvol[e] * (dNx[4 * e + li] * dNx[4 * e + lj] + dNy[4 * e + li] * dNy[4 * e + lj] +
                                    dNz[4 * e + li] * dNz[4 * e + lj]);
            }
        }
    }
    free(elvol);
    free(dN);
    free(tp);
    free(ti);
}

/* ------------------------------------------------------------------------ */
/* Boundary                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct { int a, b, c, tet; } face_t;

static int cmp_face(const void* x, const void* y) {
    const face_t* f = (const face_t*)x;
    const face_t* g = (const face_t*)y;
    if (f->a != g->a) return (f->a > g->a) - (f->a < g->a);
    if (f->b != g->b) return (f->b > g->b) - (f->b < g->b);
    return (f->c > g->c) - (f->c < g->c);
}

/* Boundary.cpp:159-197 -- every tet contributes its 4 faces (node ids
 * sorted); faces shared by two tets are interior, faces seen once form the
 * boundary.  faces_out must hold 4*ntet*3 ints, face_tet_out 4*ntet ints.
 * The reference pops a max-priority-queue, i.e. it lists the boundary faces
 * in decreasing lexicographic order; this restatement lists them in the same
 * decreasing order.  Returns nbel. */
int ref_boundary_faces(int ntet, const int* fel, int* faces_out, int* face_tet_out) {
    size_t nf = 4 * (size_t)ntet;
    face_t* F = (face_t*)malloc(sizeof(face_t) * nf);
    for (int e = 0; e < ntet; ++e) {
        for (int j = 0; j < 4; ++j) {
            int v[3], k = 0;
            for (int q = 0; q < 4; ++q)
                if (q != j) v[k++] = fel[4 * e + q];
            qsort(v, 3, sizeof(int), cmp_int);
            face_t f = {v[0], v[1], v[2], e};
            F[4 * (size_t)e + j] = f;
        }
    }
    qsort(F, nf, sizeof(face_t), cmp_face);
    int nb = 0;
    /* walk from the largest face down, as the priority queue does */
    long i = (long)nf - 1;
    while (i >= 0) {
        if (i > 0 && cmp_face(&F[i], &F[i - 1]) == 0) {
            i -= 2;
        } else {
            faces_out[3 * nb + 0] = F[i].a;
            faces_out[3 * nb + 1] = F[i].b;
            faces_out[3 * nb + 2] = F[i].c;
            face_tet_out[nb] = F[i].tet;
            ++nb;
            i -= 1;
        }
    }
    free(F);
    return nb;
}

/* auxiliaries.cpp:58-72 */
static double tri_area(const double* p1, const double* p2, const double* p3) {
    double v12[3], v13[3], c[3];
    v3sub(p1, p2, v12);
    v3sub(p1, p3, v13);
    v3cross(v12, v13, c);
    return 0.5 * v3norm(c);
}
static void tri_unit_normal(const double* p1, const double* p2, const double* p3, double* nrm) {
    double v12[3], v13[3], c[3];
    v3sub(p1, p2, v12);
    v3sub(p1, p3, v13);
    v3cross(v12, v13, c);
    double l = v3norm(c);
    nrm[0] = c[0] / l; nrm[1] = c[1] / l; nrm[2] = c[2] / l;
}

/* Boundary.cpp:200-211 (sorted unique boundary nodes), 58-69 (index map),
 * 149-156 (face areas), 214-224 (node areas = sum A_f/3), 109-125 (outward
 * unit face normals via the owning tet's centre), 78-89 (node normals =
 * normalised area-weighted face normals, non-finite -> 0).
 * bnodes_out needs n ints; returns bnx. */
int ref_boundary_props(int n, const double* xyz, const int* fel, int nbel, const int* faces, const int* face_tet,
                       int* bnodes_out, int* bindex_out, double* face_area, double* face_normal,
                       double* node_area, double* node_normal) {
    char* isb = (char*)calloc((size_t)n, 1);
    for (int f = 0; f < nbel; ++f)
        for (int j = 0; j < 3; ++j) isb[faces[3 * f + j]] = 1;
    int bnx = 0;
    for (int i = 0; i < n; ++i) {
        bindex_out[i] = -1;
        if (isb[i]) { bindex_out[i] = bnx; bnodes_out[bnx++] = i; }
    }
    free(isb);
    for (int i = 0; i < n; ++i) {
        node_area[i] = 0.0;
        node_normal[3 * i] = node_normal[3 * i + 1] = node_normal[3 * i + 2] = 0.0;
    }
    for (int f = 0; f < nbel; ++f) {
        const double* p0 = xyz + 3 * faces[3 * f];
        const double* p1 = xyz + 3 * faces[3 * f + 1];
        const double* p2 = xyz + 3 * faces[3 * f + 2];
        double A = tri_area(p0, p1, p2);
        face_area[f] = A;
        double nv[3];
        tri_unit_normal(p0, p1, p2, nv);
        const int* t = fel + 4 * face_tet[f];
        double ctr[3] = {0, 0, 0};
        for (int q = 0; q < 4; ++q)
            for (int d = 0; d < 3; ++d) ctr[d] += xyz[3 * t[q] + d] / 4.0;
        double toc[3];
        v3sub(ctr, p0, toc);
        double s = v3dot(nv, toc);
        double sg = (s > 0) ? 1.0 : ((s < 0) ? -1.0 : 0.0);
        for (int d = 0; d < 3; ++d) face_normal[3 * f + d] = -nv[d] * sg;
    }
    for (int f = 0; f < nbel; ++f) {
        for (int j = 0; j < 3; ++j) {
            int v = faces[3 * f + j];
            node_area[v] += face_area[f] / 3.0;
            for (int d = 0; d < 3; ++d) node_normal[3 * v + d] += face_normal[3 * f + d] * face_area[f] / 3.0;
        }
    }
    for (int i = 0; i < n; ++i) {
        double l = v3norm(node_normal + 3 * i);
        for (int d = 0; d < 3; ++d) {
            double v = node_normal[3 * i + d] / l;
            node_normal[3 * i + d] = isfinite(v) ? v : 0.0;
        }
    }
    return bnx;
}

/* ------------------------------------------------------------------------ */
/* BEM                                                                       */
/* ------------------------------------------------------------------------ */

/* Lindholm.cpp:41-60 -- solid angle subtended by triangle (T1,T2,T3) at x,
 * 2*acos(numer/denom), with the areEqual(arg, 1) clamp. */
double ref_solid_angle(const double* x, const double* T1, const double* T2, const double* T3) {
    double r1[3], r2[3], r3[3];
    v3sub(T1, x, r1);
    v3sub(T2, x, r2);
    v3sub(T3, x, r3);
    double r1n = v3norm(r1), r2n = v3norm(r2), r3n = v3norm(r3);
    double r123 = r1n * r2n * r3n;
    double r23 = v3dot(r2, r3), r12 = v3dot(r1, r2), r31 = v3dot(r3, r1);
    double numer = r123 + r1n * r23 + r2n * r31 + r3n * r12;
    double denom = sqrt(2.0 * (r2n * r3n + r23) * (r3n * r1n + r31) * (r1n * r2n + r12));
    double arg = numer / denom;
    if (ref_are_equal(arg, 1.0, 100)) arg = 1.0;
    return 2.0 * acos(arg);
}

/* Lindholm.cpp:63-128 (isOrdered = false branch, as called from
 * BEMprocessing.cpp:401) -- double-layer weights of the three vertices of a
 * flat triangle seen from x0, D.A. Lindholm, IEEE Trans. Mag. 20, 2025 (1984). */
void ref_lindholm_weights(const double* x0, const double* p1, const double* p2, const double* p3, double A,
                          const double* nv, double* L) {
    const double* p[3] = {p1, p2, p3};
    for (int i = 0; i < 3; ++i) {
        double d[3];
        v3sub(p[i], x0, d);
        /* Eigen isMuchSmallerThan(1.): |d|^2 <= (1e-12)^2 */
        if (v3dot(d, d) <= 1e-12 * 1e-12) { L[0] = L[1] = L[2] = 0.0; return; }
    }
    double chi[3];
    tri_unit_normal(p1, p2, p3, chi);
    double rho_vec[3][3], xi[3][3], s[3], eta[3][3], eta0[3], rho[3];
    for (int i = 0; i < 3; ++i) v3sub(p[i], x0, rho_vec[i]);
    for (int i = 0; i < 3; ++i) v3sub(rho_vec[(i + 1) % 3], rho_vec[i], xi[i]);
    for (int i = 0; i < 3; ++i) {
        s[i] = v3norm(xi[i]);
        for (int d = 0; d < 3; ++d) xi[i][d] /= s[i];
        v3cross(chi, xi[i], eta[i]);
        eta0[i] = v3dot(eta[i], rho_vec[i]);
        rho[i] = v3norm(rho_vec[i]);
    }
    double chi0 = v3dot(chi, rho_vec[0]);
    double gamma[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) gamma[i][j] = v3dot(xi[(i + 1) % 3], xi[j]);
    double P[3];
    for (int i = 0; i < 3; ++i) {
        int i1 = (i + 1) % 3;
        P[i] = log((rho[i] + rho[i1] + s[i]) / (rho[i] + rho[i1] - s[i]));
    }
    double Omega = ref_solid_angle(x0, p1, p2, p3);
    if (chi0 < 0.) Omega = -Omega;
    for (int i = 0; i < 3; ++i) {
        int i1 = (i + 1) % 3;
        double gP = gamma[i][0] * P[0] + gamma[i][1] * P[1] + gamma[i][2] * P[2];
        L[i] = s[i1] / (8. * REF_PI * A) * (eta0[i1] * Omega - chi0 * gP);
    }
    if (v3dot(nv, chi) < 0) { L[0] = -L[0]; L[1] = -L[1]; L[2] = -L[2]; }
}

/* BEMprocessing.cpp:358-372 -- interior solid angle at each boundary node:
 * sum over tets touching the node of the angle subtended by the opposite face. */
void ref_solid_angles(int ntet, const double* xyz, const int* fel, const int* bindex, int bnx, double* sub) {
    for (int i = 0; i < bnx; ++i) sub[i] = 0.0;
    for (int e = 0; e < ntet; ++e) {
        const int* t = fel + 4 * e;
        for (int j = 0; j < 4; ++j) {
            int bi = bindex[t[j]];
            if (bi >= 0) {
                sub[bi] += ref_solid_angle(xyz + 3 * t[j], xyz + 3 * t[(j + 1) % 4], xyz + 3 * t[(j + 2) % 4],
                                           xyz + 3 * t[(j + 3) % 4]);
            }
        }
    }
}

/* BEMprocessing.cpp:375-411 -- dense double-layer matrix, column-major
 * bnx x bnx: B(j,j) = sub_j/(4 pi) - 1; B(n, m) += Lindholm weight of vertex
 * m of every face not containing node n. */
static void bem_row(int nrow, int bnx, int nbel, const double* xyz, const int* bnodes, const int* bindex,
                    const int* faces, const double* face_area, const double* face_normal, const double* sub,
                    double* row, size_t stride) {
    for (int m = 0; m < bnx; ++m) row[(size_t)m * stride] = 0.0;
    row[(size_t)nrow * stride] = sub[nrow] / (4. * REF_PI) - 1.;
    const double* x0 = xyz + 3 * bnodes[nrow];
    for (int f = 0; f < nbel; ++f) {
        const int* v = faces + 3 * f;
        int b0 = bindex[v[0]], b1 = bindex[v[1]], b2 = bindex[v[2]];
        if (nrow == b0 || nrow == b1 || nrow == b2) continue;
        double L[3];
        ref_lindholm_weights(x0, xyz + 3 * v[0], xyz + 3 * v[1], xyz + 3 * v[2], face_area[f], face_normal + 3 * f,
                             L);
        row[(size_t)b0 * stride] += L[0];
        row[(size_t)b1 * stride] += L[1];
        row[(size_t)b2 * stride] += L[2];
    }
}

void ref_bem_matrix(int bnx, int nbel, const double* xyz, const int* bnodes, const int* bindex, const int* faces,
                    const double* face_area, const double* face_normal, const double* sub, double* B) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int nrow = 0; nrow < bnx; ++nrow)
        bem_row(nrow, bnx, nbel, xyz, bnodes, bindex, faces, face_area, face_normal, sub, B + nrow, (size_t)bnx);
}

/* The same matrix, rows rows[0..nrows) only (BEMprocessing.cpp:375-411 visit
 * every face for each row, so a row costs nbel Lindholm evaluations): out is
 * nrows x bnx row-major, out[k][m] = B(rows[k], m). */
void ref_bem_rows(int bnx, int nbel, const double* xyz, const int* bnodes, const int* bindex, const int* faces,
                  const double* face_area, const double* face_normal, const double* sub, int nrows, const int* rows,
                  double* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nrows; ++k)
        bem_row(rows[k], bnx, nbel, xyz, bnodes, bindex, faces, face_area, face_normal, sub, out + (size_t)k * bnx,
                1);
}

/* ------------------------------------------------------------------------ */
/* Sparse helpers                                                            */
/* ------------------------------------------------------------------------ */
void ref_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) s += val[p] * x[col[p]];
        y[i] = s;
    }
}

/* dense column-major GEMV y = B x (DemagField.cpp:145-147, denseMVP) */
void ref_dense_gemv(int b, const double* B, const double* x, double* y) {
#pragma omp parallel
    {
#ifdef _OPENMP
        int nt = omp_get_num_threads(), id = omp_get_thread_num();
#else
        int nt = 1, id = 0;
#endif
        int chunk = (b + nt - 1) / nt;
        int lo = id * chunk, hi = lo + chunk;
        if (hi > b) hi = b;
        for (int r = lo; r < hi; ++r) y[r] = 0.0;
        for (int c = 0; c < b; ++c) {
            const double* colp = B + (size_t)c * b;
            double xc = x[c];
            for (int r = lo; r < hi; ++r) y[r] += colp[r] * xc;
        }
    }
}

/* Replace the listed rows of a CSR matrix by identity rows and drop explicit
 * zeros, on the same column pattern (prune(0,0) after the replacement).
 * FEMprocessing.cpp:329-369 (Dirichlet: all boundary rows) and
 * DemagField.cpp:204-220 (Neumann: row nx/2).  Output arrays must hold nnz
 * entries; returns the new nnz. */
int ref_replace_rows_identity(int n, const int* rowptr, const int* col, const double* val, const char* fixed,
                              int* rp_out, int* col_out, double* val_out) {
    int k = 0;
    rp_out[0] = 0;
    for (int i = 0; i < n; ++i) {
        if (fixed[i]) {
            col_out[k] = i;
            val_out[k] = 1.0;
            ++k;
        } else {
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                if (val[p] != 0.0) {
                    col_out[k] = col[p];
                    val_out[k] = val[p];
                    ++k;
                }
            }
        }
        rp_out[i + 1] = k;
    }
    return k;
}

/* ------------------------------------------------------------------------ */
/* Krylov solver: right-preconditioned restarted GMRES with ILU(0).          */
/* The reference solves both row-replaced (non-symmetric) systems with a     */
/* Krylov method: AMGCL smoothed-aggregation + ILU0 smoothing + GMRES         */
/* (GPUSolver.h:53-63, GPUSolver.cu:43-48; AMGCLSolver.cpp:39-103), warm-     */
/* started from the previous solution (AMGCLSolver.cpp:97-103).  AMGCL is     */
/* not vendored; this restatement keeps the Krylov method (GMRES) and the     */
/* ILU0 smoother as the preconditioner, without the AMG hierarchy.            */
/* Convergence: ||b - A x|| <= tol * ||b||.                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    const int* rowptr;
    const int* col;
    const double* val;
    double* lu;  /* ILU(0) factors on the pattern */
    int* diag;   /* position of the diagonal entry per row */
} ilu0_t;

static int ilu0_factor(ilu0_t* f) {
    int n = f->n;
    f->lu = (double*)malloc(sizeof(double) * (size_t)f->rowptr[n]);
    f->diag = (int*)malloc(sizeof(int) * (size_t)n);
    memcpy(f->lu, f->val, sizeof(double) * (size_t)f->rowptr[n]);
    for (int i = 0; i < n; ++i) {
        int lo = f->rowptr[i], hi = f->rowptr[i + 1];
        int d = find_col(f->col, lo, hi, i);
        if (d >= hi || f->col[d] != i) return -1;
        f->diag[i] = d;
    }
    int* pos = (int*)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) pos[i] = -1;
    for (int i = 0; i < n; ++i) {
        int lo = f->rowptr[i], hi = f->rowptr[i + 1];
        for (int p = lo; p < hi; ++p) pos[f->col[p]] = p;
        for (int p = lo; p < hi && f->col[p] < i; ++p) {
            int k = f->col[p];
            double lik = f->lu[p] / f->lu[f->diag[k]];
            f->lu[p] = lik;
            for (int q = f->diag[k] + 1; q < f->rowptr[k + 1]; ++q) {
                int j = f->col[q];
                if (pos[j] >= 0) f->lu[pos[j]] -= lik * f->lu[q];
            }
        }
        for (int p = lo; p < hi; ++p) pos[f->col[p]] = -1;
        if (f->lu[f->diag[i]] == 0.0) { free(pos); return -2; }
    }
    free(pos);
    return 0;
}

static void ilu0_apply(const ilu0_t* f, const double* r, double* z) {
    int n = f->n;
    for (int i = 0; i < n; ++i) {
        double s = r[i];
        for (int p = f->rowptr[i]; p < f->diag[i]; ++p) s -= f->lu[p] * z[f->col[p]];
        z[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = z[i];
        for (int p = f->diag[i] + 1; p < f->rowptr[i + 1]; ++p) s -= f->lu[p] * z[f->col[p]];
        z[i] = s / f->lu[f->diag[i]];
    }
}

static double dot_n(int n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* ------------------------------------------------------------------------ */
/* Smoothed-aggregation AMG with ILU(0) smoothing, as the reference          */
/* configures AMGCL (ddemidov/amgcl, fetched at configure time with          */
/* GIT_TAG master, i.e. unpinned, external/CMakeLists.txt; not vendored):     */
/* amg<smoothed_aggregation, ilu0> inside gmres (AMGCLSolver.h:89-101), with  */
/* coarsening.relax = 0.75 (AMGCLSolver.cpp:40, GPUSolver.cu:45),             */
/* coarse_enough = 100 and direct_coarse = false (AMGCLSolver.cpp:43-44), and */
/* AMGCL's defaults otherwise: strength eps_strong = 0.08, plain greedy       */
/* aggregation, Gershgorin bound for rho(D^-1 A_F), one pre- and one          */
/* post-smoothing sweep, V-cycle, GMRES(30).  Restated from the published     */
/* method (Vanek, Mandel, Brezina, Computing 56 (1996) 179) and AMGCL's       */
/* documented parameters; iterates are not claimed to match AMGCL's, only     */
/* the converged solution (||b - A x|| <= tol ||b||) is relied upon.          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, m; /* rows, columns */
    int* rp;
    int* col;
    double* val;
} csr_t;

static void csr_free(csr_t* a) {
    free(a->rp); free(a->col); free(a->val);
    a->rp = NULL; a->col = NULL; a->val = NULL;
}

static int csr_alloc(csr_t* a, int n, int m, size_t nnz) {
    a->n = n; a->m = m;
    a->rp = (int*)calloc((size_t)n + 1, sizeof(int));
    a->col = (int*)malloc(sizeof(int) * (nnz ? nnz : 1));
    a->val = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    return a->rp && a->col && a->val ? 0 : -1;
}

static void csr_spmv(const csr_t* a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < a->n; ++i) {
        double s = 0.0;
        for (int p = a->rp[i]; p < a->rp[i + 1]; ++p) s += a->val[p] * x[a->col[p]];
        y[i] = s;
    }
}

/* C = A^T (columns sorted because rows are visited in order) */
static int csr_transpose(const csr_t* a, csr_t* t) {
    size_t nnz = (size_t)a->rp[a->n];
    if (csr_alloc(t, a->m, a->n, nnz)) return -1;
    for (size_t p = 0; p < nnz; ++p) t->rp[a->col[p] + 1]++;
    for (int i = 0; i < a->m; ++i) t->rp[i + 1] += t->rp[i];
    int* pos = (int*)malloc(sizeof(int) * ((size_t)a->m + 1));
    memcpy(pos, t->rp, sizeof(int) * ((size_t)a->m + 1));
    for (int i = 0; i < a->n; ++i)
        for (int p = a->rp[i]; p < a->rp[i + 1]; ++p) {
            int q = pos[a->col[p]]++;
            t->col[q] = i;
            t->val[q] = a->val[p];
        }
    free(pos);
    return 0;
}

static int cmp_int_asc(const void* x, const void* y) {
    int a = *(const int*)x, b = *(const int*)y;
    return (a > b) - (a < b);
}

/* C = A B (Gustavson row-by-row product, sorted columns) */
static int csr_spgemm(const csr_t* A, const csr_t* B, csr_t* C) {
    const int n = A->n, m = B->m;
    int* mark = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    double* acc = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    int* rp = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int j = 0; j < m; ++j) mark[j] = -1;
    size_t nnz = 0;
    for (int i = 0; i < n; ++i) { /* symbolic */
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p) {
            int k = A->col[p];
            for (int q = B->rp[k]; q < B->rp[k + 1]; ++q)
                if (mark[B->col[q]] != i) { mark[B->col[q]] = i; ++nnz; }
        }
        rp[i + 1] = (int)nnz;
    }
    if (nnz > 0x7fffffffULL) { free(mark); free(acc); free(rp); return -1; }
    C->n = n; C->m = m; C->rp = rp;
    C->col = (int*)malloc(sizeof(int) * (nnz ? nnz : 1));
    C->val = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    for (int j = 0; j < m; ++j) mark[j] = -1;
    for (int i = 0; i < n; ++i) { /* numeric */
        int q0 = rp[i], cnt = 0;
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p) {
            int k = A->col[p];
            double a = A->val[p];
            for (int q = B->rp[k]; q < B->rp[k + 1]; ++q) {
                int j = B->col[q];
                if (mark[j] != i) { mark[j] = i; C->col[q0 + cnt++] = j; acc[j] = 0.0; }
                acc[j] += a * B->val[q];
            }
        }
        qsort(C->col + q0, (size_t)cnt, sizeof(int), cmp_int_asc);
        for (int c = 0; c < cnt; ++c) C->val[q0 + c] = acc[C->col[q0 + c]];
    }
    free(mark); free(acc);
    return 0;
}

/* Plain greedy aggregation on the strength graph (AMGCL plain_aggregates):
 * j != i is a strong neighbour of i when a_ij^2 > eps^2 |a_ii a_jj|.  Nodes
 * without strong neighbours are left out (zero rows of P).  Nodes are
 * visited in order; an unassigned node starts an aggregate, takes its strong
 * neighbours (also ones tentatively given to an earlier aggregate), and
 * tentatively claims the unassigned strong neighbours of those; aggregates
 * left empty are dropped and the rest renumbered.  strong[p] flags entry p.
 * Returns the number of aggregates. */
static int sa_aggregates(const csr_t* A, double eps, const double* dia, char* strong, int* id) {
    const int n = A->n;
    const int UNDEF = -1, REMOVED = -2;
    for (int i = 0; i < n; ++i) {
        int any = 0;
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p) {
            int j = A->col[p];
            double v = A->val[p];
            strong[p] = (char)(j != i && v * v > eps * eps * fabs(dia[i] * dia[j]));
            any |= strong[p];
        }
        id[i] = any ? UNDEF : REMOVED;
    }
    int count = 0;
    int* neib = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        if (id[i] != UNDEF) continue;
        int cur = count++;
        id[i] = cur;
        int nn = 0;
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p) {
            int c = A->col[p];
            if (strong[p] && id[c] != REMOVED) { id[c] = cur; neib[nn++] = c; }
        }
        for (int k = 0; k < nn; ++k) {
            int c = neib[k];
            for (int p = A->rp[c]; p < A->rp[c + 1]; ++p) {
                int cc = A->col[p];
                if (strong[p] && id[cc] == UNDEF) id[cc] = cur;
            }
        }
    }
    free(neib);
    int* renum = (int*)malloc(sizeof(int) * (size_t)(count > 0 ? count : 1));
    for (int k = 0; k < count; ++k) renum[k] = 0;
    for (int i = 0; i < n; ++i) if (id[i] >= 0) renum[id[i]] = 1;
    int nagg = 0;
    for (int k = 0; k < count; ++k) renum[k] = renum[k] ? nagg++ : -1;
    for (int i = 0; i < n; ++i) if (id[i] >= 0) id[i] = renum[id[i]];
    free(renum);
    return nagg;
}

/* Smoothed prolongator P = (I - omega D_F^-1 A_F) P_tent with the filtered
 * matrix A_F (strong off-diagonal entries kept, weak ones lumped into the
 * diagonal), omega = relax (4/3) / rho, rho = max_i sum_j |A_F,ij| / |D_F,i|
 * (Gershgorin), P_tent(i, id[i]) = 1. */
static int sa_prolongator(const csr_t* A, const char* strong, const int* id, int nagg, double relax, csr_t* P) {
    const int n = A->n;
    double* dF = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double rho = 0.0;
    for (int i = 0; i < n; ++i) {
        double d = 0.0, s = 0.0;
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p)
            if (A->col[p] == i || !strong[p]) d += A->val[p];
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p)
            if (A->col[p] != i && strong[p]) s += fabs(A->val[p]);
        dF[i] = d;
        double r = d != 0.0 ? (s + fabs(d)) / fabs(d) : 0.0;
        if (r > rho) rho = r;
    }
    const double omega = rho > 0 ? relax * (4.0 / 3.0) / rho : 0.0;
    /* S = I - omega D_F^-1 A_F on A's pattern, then P = S P_tent */
    csr_t S, T;
    csr_alloc(&S, n, n, (size_t)A->rp[n]);
    size_t k = 0;
    for (int i = 0; i < n; ++i) {
        for (int p = A->rp[i]; p < A->rp[i + 1]; ++p) {
            int j = A->col[p];
            double v;
            if (j == i) v = 1.0 - omega * (dF[i] != 0.0 ? dF[i] / dF[i] : 0.0);
            else if (strong[p]) v = -omega * A->val[p] / dF[i];
            else continue;
            S.col[k] = j;
            S.val[k] = v;
            ++k;
        }
        S.rp[i + 1] = (int)k;
    }
    csr_alloc(&T, n, nagg, (size_t)n);
    k = 0;
    for (int i = 0; i < n; ++i) {
        if (id[i] >= 0) { T.col[k] = id[i]; T.val[k] = 1.0; ++k; }
        T.rp[i + 1] = (int)k;
    }
    int r = csr_spgemm(&S, &T, P);
    csr_free(&S); csr_free(&T); free(dF);
    return r;
}

typedef struct {
    csr_t A, P, R; /* P, R: to and from the next coarser level */
    ilu0_t f;
    double *x, *b, *r;
} amg_level_t;

typedef struct {
    int nlev;
    amg_level_t lev[32];
} amg_t;

static void amg_free(amg_t* a) {
    if (!a) return;
    for (int l = 0; l < a->nlev; ++l) {
        amg_level_t* L = &a->lev[l];
        csr_free(&L->A); csr_free(&L->P); csr_free(&L->R);
        free(L->f.lu); free(L->f.diag); free(L->x); free(L->b); free(L->r);
    }
    free(a);
}

static amg_t* amg_setup(int n, const int* rowptr, const int* col, const double* val, double relax, double eps,
                        int coarse_enough) {
    amg_t* a = (amg_t*)calloc(1, sizeof(amg_t));
    csr_t* A0 = &a->lev[0].A;
    csr_alloc(A0, n, n, (size_t)rowptr[n]);
    memcpy(A0->rp, rowptr, sizeof(int) * ((size_t)n + 1));
    memcpy(A0->col, col, sizeof(int) * (size_t)rowptr[n]);
    memcpy(A0->val, val, sizeof(double) * (size_t)rowptr[n]);
    a->nlev = 1;
    while (a->nlev < 32 && a->lev[a->nlev - 1].A.n > coarse_enough) {
        amg_level_t* L = &a->lev[a->nlev - 1];
        const csr_t* A = &L->A;
        double* dia = (double*)calloc((size_t)A->n, sizeof(double));
        for (int i = 0; i < A->n; ++i)
            for (int p = A->rp[i]; p < A->rp[i + 1]; ++p) if (A->col[p] == i) dia[i] = A->val[p];
        char* strong = (char*)malloc((size_t)A->rp[A->n] + 1);
        int* id = (int*)malloc(sizeof(int) * (size_t)A->n);
        int nagg = sa_aggregates(A, eps, dia, strong, id);
        int ok = nagg > 0 && nagg < A->n && sa_prolongator(A, strong, id, nagg, relax, &L->P) == 0;
        free(dia); free(strong); free(id);
        if (!ok) break;
        csr_transpose(&L->P, &L->R);
        csr_t AP;
        if (csr_spgemm(A, &L->P, &AP)) break;
        csr_t* Ac = &a->lev[a->nlev].A;
        int r = csr_spgemm(&L->R, &AP, Ac);
        csr_free(&AP);
        if (r) break;
        a->nlev++;
    }
    for (int l = 0; l < a->nlev; ++l) {
        amg_level_t* L = &a->lev[l];
        L->f.n = L->A.n;
        L->f.rowptr = L->A.rp;
        L->f.col = L->A.col;
        L->f.val = L->A.val;
        if (ilu0_factor(&L->f) != 0) { a->nlev = l; amg_free(a); return NULL; }
        L->x = (double*)calloc((size_t)L->A.n, sizeof(double));
        L->b = (double*)calloc((size_t)L->A.n, sizeof(double));
        L->r = (double*)calloc((size_t)L->A.n, sizeof(double));
    }
    return a;
}

/* x += ILU0^-1 (b - A x) (AMGCL relaxation::ilu0 with damping 1) */
static void amg_smooth(amg_level_t* L, const double* b, double* x) {
    csr_spmv(&L->A, x, L->r);
    for (int i = 0; i < L->A.n; ++i) L->r[i] = b[i] - L->r[i];
    double* z = L->r; /* ilu0_apply reads r[i] before writing z[i] */
    ilu0_apply(&L->f, L->r, z);
    for (int i = 0; i < L->A.n; ++i) x[i] += z[i];
}

/* One V-cycle from x = 0 on level l (AMGCL amg::cycle, npre = npost = 1; the
 * coarsest level, direct_coarse = false, is only smoothed). */
static void amg_cycle(amg_t* a, int l, const double* b, double* x) {
    amg_level_t* L = &a->lev[l];
    const int n = L->A.n;
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    amg_smooth(L, b, x);
    if (l + 1 < a->nlev) {
        amg_level_t* C = &a->lev[l + 1];
        csr_spmv(&L->A, x, L->r);
        for (int i = 0; i < n; ++i) L->r[i] = b[i] - L->r[i];
        csr_spmv(&L->R, L->r, C->b);
        amg_cycle(a, l + 1, C->b, C->x);
        csr_spmv(&L->P, C->x, L->r);
        for (int i = 0; i < n; ++i) x[i] += L->r[i];
    }
    amg_smooth(L, b, x);
}

typedef struct {
    ilu0_t f;
    int* rowptr;
    int* col;
    double* val;
    amg_t* amg; /* non-null: AMG V-cycle preconditioner instead of ILU(0) */
} ref_solver_t;

static void precond_apply(ref_solver_t* s, const double* r, double* z) {
    if (s->amg) amg_cycle(s->amg, 0, r, z);
    else ilu0_apply(&s->f, r, z);
}

/* Number of levels and level sizes of an AMG solver (0 for ILU(0)). */
int ref_solver_levels(void* h, int* sizes, int cap) {
    ref_solver_t* s = (ref_solver_t*)h;
    if (!s || !s->amg) return 0;
    for (int l = 0; l < s->amg->nlev && l < cap; ++l) sizes[l] = s->amg->lev[l].A.n;
    return s->amg->nlev;
}

/* SolverFactory::setMatrix + compute (SolverFactory.cpp:45-48;
 * AMGCLSolver.cpp:39-95) with the AMG preconditioner described above.
 * relax <= 0 selects the reference values (0.75, 0.08, 100). */
void* ref_solver_create_amg(int n, const int* rowptr, const int* col, const double* val, double relax, double eps,
                            int coarse_enough) {
    if (relax <= 0) { relax = 0.75; eps = 0.08; coarse_enough = 100; }
    ref_solver_t* s = (ref_solver_t*)calloc(1, sizeof(ref_solver_t));
    int nnz = rowptr[n];
    s->rowptr = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    s->col = (int*)malloc(sizeof(int) * (size_t)nnz);
    s->val = (double*)malloc(sizeof(double) * (size_t)nnz);
    memcpy(s->rowptr, rowptr, sizeof(int) * (size_t)(n + 1));
    memcpy(s->col, col, sizeof(int) * (size_t)nnz);
    memcpy(s->val, val, sizeof(double) * (size_t)nnz);
    s->f.n = n;
    s->amg = amg_setup(n, rowptr, col, val, relax, eps, coarse_enough);
    if (!s->amg) { free(s->rowptr); free(s->col); free(s->val); free(s); return NULL; }
    return s;
}

/* SolverFactory::setMatrix + compute (SolverFactory.cpp:45-48;
 * AMGCLSolver.cpp:59-95): copy the matrix and build the preconditioner once. */
void* ref_solver_create(int n, const int* rowptr, const int* col, const double* val) {
    ref_solver_t* s = (ref_solver_t*)calloc(1, sizeof(ref_solver_t));
    int nnz = rowptr[n];
    s->rowptr = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    s->col = (int*)malloc(sizeof(int) * (size_t)nnz);
    s->val = (double*)malloc(sizeof(double) * (size_t)nnz);
    memcpy(s->rowptr, rowptr, sizeof(int) * (size_t)(n + 1));
    memcpy(s->col, col, sizeof(int) * (size_t)nnz);
    memcpy(s->val, val, sizeof(double) * (size_t)nnz);
    s->f.n = n;
    s->f.rowptr = s->rowptr;
    s->f.col = s->col;
    s->f.val = s->val;
    if (ilu0_factor(&s->f) != 0) {
        free(s->f.lu); free(s->f.diag); free(s->rowptr); free(s->col); free(s->val); free(s);
        return NULL;
    }
    return s;
}

void ref_solver_destroy(void* h) {
    ref_solver_t* s = (ref_solver_t*)h;
    if (!s) return;
    amg_free(s->amg);
    free(s->f.lu); free(s->f.diag); free(s->rowptr); free(s->col); free(s->val); free(s);
}

/* SolverFactory::solve for the CPU solvers (AMGCLSolver.cpp:97-103):
 * x holds the initial guess on entry (warm start) and the solution on exit.
 * Returns the number of inner iterations (>=0) when ||b - A x|| <= tol ||b||
 * was reached, -1 when maxit was exhausted first.  *relres receives the
 * final true relative residual. */
int ref_solver_solve(void* h, const double* b, double* x, double tol, int maxit, int restart, double* relres) {
    ref_solver_t* s = (ref_solver_t*)h;
    int n = s->f.n;
    const int *rowptr = s->rowptr, *col = s->col;
    const double* val = s->val;
    int m = restart;
    double* V = (double*)malloc(sizeof(double) * (size_t)n * (m + 1));
    double* Z = (double*)malloc(sizeof(double) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double*)malloc(sizeof(double) * m), *sn = (double*)malloc(sizeof(double) * m);
    double* g = (double*)malloc(sizeof(double) * (m + 1));
    double* yv = (double*)malloc(sizeof(double) * m);
    double bnorm = sqrt(dot_n(n, b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    int it = 0, status = -1;
    double res = 0.0;
    for (;;) {
        /* true residual r = b - A x */
        ref_spmv(n, rowptr, col, val, x, w);
        for (int i = 0; i < n; ++i) V[i] = b[i] - w[i];
        double beta = sqrt(dot_n(n, V, V));
        res = beta / bnorm;
        if (res <= tol) { status = it; break; }
        if (it >= maxit) break;
        for (int i = 0; i < n; ++i) V[i] /= beta;
        for (int i = 0; i <= m; ++i) g[i] = 0.0;
        g[0] = beta;
        int kdone = 0;
        for (int j = 0; j < m && it < maxit; ++j) {
            precond_apply(s, V + (size_t)j * n, Z);
            ref_spmv(n, rowptr, col, val, Z, w);
            for (int i = 0; i <= j; ++i) { /* modified Gram-Schmidt */
                const double* vi = V + (size_t)i * n;
                double hij = dot_n(n, w, vi);
                H[i * m + j] = hij;
#pragma omp parallel for schedule(static)
                for (int q = 0; q < n; ++q) w[q] -= hij * vi[q];
            }
            double hn = sqrt(dot_n(n, w, w));
            H[(j + 1) * m + j] = hn;
            if (hn != 0.0) {
                double* vn = V + (size_t)(j + 1) * n;
                for (int q = 0; q < n; ++q) vn[q] = w[q] / hn;
            }
            for (int i = 0; i < j; ++i) { /* apply previous Givens rotations */
                double t = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
                H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
                H[i * m + j] = t;
            }
            double a = H[j * m + j], bb = H[(j + 1) * m + j];
            double r = hypot(a, bb);
            cs[j] = a / r;
            sn[j] = bb / r;
            H[j * m + j] = r;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++it;
            kdone = j + 1;
            if (fabs(g[j + 1]) / bnorm <= 0.5 * tol || hn == 0.0) break;
        }
        for (int i = kdone - 1; i >= 0; --i) { /* back substitution */
            double sacc = g[i];
            for (int q = i + 1; q < kdone; ++q) sacc -= H[i * m + q] * yv[q];
            yv[i] = sacc / H[i * m + i];
        }
        for (int q = 0; q < n; ++q) w[q] = 0.0;
        for (int i = 0; i < kdone; ++i) {
            const double* vi = V + (size_t)i * n;
            for (int q = 0; q < n; ++q) w[q] += yv[i] * vi[q];
        }
        precond_apply(s, w, Z); /* right preconditioning: x += M^-1 V y */
        for (int q = 0; q < n; ++q) x[q] += Z[q];
    }
    if (relres) *relres = res;
    free(V); free(Z); free(w); free(H); free(cs); free(sn); free(g); free(yv);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Effective fields                                                           */
/* ------------------------------------------------------------------------ */

/* EffFieldCalc.cpp:59-63, 97-100 (and TheLLG.cpp:88-93 for the GPU copy):
 * XC = -2 * diag(A .* invNodeVol) * K;  Hexc = XC * m  (J/m^3 units before
 * the invJs scaling of TheLLG.cpp:245). */
void ref_exchange_field(int n, const int* rowptr, const int* col, const double* K, const double* A,
                        const double* nodevol, const double* m, double* H) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double f = -2.0 * (A[i] * (1.0 / nodevol[i]));
        double sx = 0, sy = 0, sz = 0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            double v = f * K[p];
            int j = col[p];
            sx += v * m[j];
            sy += v * m[n + j];
            sz += v * m[2 * n + j];
        }
        H[i] = sx; H[n + i] = sy; H[2 * n + i] = sz;
    }
}

/* EffFieldCalc.cpp:71-84 -- Hani_i = 2 Ku_i (k_i . m_i) k_i; axis is 3n SoA. */
void ref_uniaxial_field(int n, const double* Ku, const double* axis, const double* m, double* H) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double kx = axis[i], ky = axis[n + i], kz = axis[2 * n + i];
        double d = kx * m[i] + ky * m[n + i] + kz * m[2 * n + i];
        double f = 2.0 * Ku[i] * d;
        H[i] = f * kx; H[n + i] = f * ky; H[2 * n + i] = f * kz;
    }
}

/* EffFieldCalc.cpp:159-170 and 280-298 -- bulk DMI:
 *   curl_x = tGy m_z - tGz m_y (cyclic) with tG = -tGrad,
 *   Hdmi = (S_i (m_i x nv_i) - 2 curl_i) * invNodeVol_i * D_i.
 * tX,tY,tZ are the tGrad operators as assembled (without the minus sign). */
void ref_dmi_field(int n, const int* rowptr, const int* col, const double* tX, const double* tY, const double* tZ,
                   const double* D, const double* nodevol, const double* node_normal_soa,
                   const double* node_area, const double* m, double* H) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double gxz = 0, gyz = 0, gzy = 0, gxy = 0, gzx = 0, gyx = 0; /* (tG_d m_c) */
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int j = col[p];
            double ax = -tX[p], ay = -tY[p], az = -tZ[p];
            gyz += ay * m[2 * n + j];
            gzy += az * m[n + j];
            gzx += az * m[j];
            gxz += ax * m[2 * n + j];
            gxy += ax * m[n + j];
            gyx += ay * m[j];
        }
        double curl[3] = {gyz - gzy, gzx - gxz, gxy - gyx};
        double mi[3] = {m[i], m[n + i], m[2 * n + i]};
        double nvi[3] = {node_normal_soa[i], node_normal_soa[n + i], node_normal_soa[2 * n + i]};
        double c[3];
        v3cross(mi, nvi, c);
        double f = (1.0 / nodevol[i]) * D[i];
        for (int d = 0; d < 3; ++d) H[d * n + i] = (node_area[i] * c[d] - 2. * curl[d]) * f;
    }
}

/* Interfacial DMI (interface normal z), no reference counterpart.  Energy
 * density D [m_z div m - (m . grad) m_z] with P1 m and the lumped (nodal)
 * non-derivative factor:
 *   E_h = sum_i D_i [m_z,i (Q_x m_x + Q_y m_y)_i - m_x,i (Q_x m_z)_i - m_y,i (Q_y m_z)_i],
 *   Q_d(i,j) = int N_i d_d N_j = tGrad_d(j,i) = V_i grad_d(i,j),
 * and H = -(1/V) dE_h/dm (J/m^3, before invJs):
 *   H_x = D grad_x m_z - (1/V) tGrad_x (D m_z)
 *   H_y = D grad_y m_z - (1/V) tGrad_y (D m_z)
 *   H_z = -D (grad_x m_x + grad_y m_y) + (1/V) (tGrad_x (D m_x) + tGrad_y (D m_y)). */
void ref_idmi_field(int n, const int* rowptr, const int* col, const double* gX, const double* gY, const double* tX,
                    const double* tY, const double* D, const double* nodevol, const double* m, double* H) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double gxz = 0, gyz = 0, gxx = 0, gyy = 0, txz = 0, tyz = 0, txx = 0, tyy = 0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int j = col[p];
            const double mx = m[j], my = m[n + j], mz = m[2 * n + j];
            gxz += gX[p] * mz;
            gyz += gY[p] * mz;
            gxx += gX[p] * mx;
            gyy += gY[p] * my;
            txz += tX[p] * D[j] * mz;
            tyz += tY[p] * D[j] * mz;
            txx += tX[p] * D[j] * mx;
            tyy += tY[p] * D[j] * my;
        }
        const double iv = 1.0 / nodevol[i];
        H[i] = D[i] * gxz - iv * txz;
        H[n + i] = D[i] * gyz - iv * tyz;
        H[2 * n + i] = -D[i] * (gxx + gyy) + iv * (txx + tyy);
    }
}

/* TheLLG.cpp:442-449, 477-483 -- uniform Zeeman field |B|/mu0 along (theta, phi) in radians */
void ref_zeeman_field(int n, double Babs, double theta, double phi, double* H) {
    double hx = sin(theta) * cos(phi), hy = sin(theta) * sin(phi), hz = cos(theta);
    for (int i = 0; i < n; ++i) {
        H[i] = hx * (Babs / REF_MU0);
        H[n + i] = hy * (Babs / REF_MU0);
        H[2 * n + i] = hz * (Babs / REF_MU0);
    }
}

/* TheLLG.cpp:238-254 -- Heff = (Hexc + Hku + Hdmi) * invJs + Hext + Hdem,
 * invJs = 1/Js with non-finite -> 0 (TheLLG.cpp:110).  Null pointers are
 * treated as zero fields. */
void ref_total_field(int n, const double* Js, const double* Hexc, const double* Hku, const double* Hdmi,
                     const double* Hext, const double* Hdem, double* Heff) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double inv = 1.0 / Js[i];
        if (!isfinite(inv)) inv = 0.0;
        for (int d = 0; d < 3; ++d) {
            size_t q = (size_t)d * n + i;
            double a = (Hexc ? Hexc[q] : 0.0) + (Hku ? Hku[q] : 0.0) + (Hdmi ? Hdmi[q] : 0.0);
            Heff[q] = a * inv + (Hext ? Hext[q] : 0.0) + (Hdem ? Hdem[q] : 0.0);
        }
    }
}

/* TheLLG.cpp:165-175 (classic, variant 1) and 191-196 (no precession,
 * variant 0): dm/dtau = -m x H - alpha m x (m x H)  or  -alpha m x (m x H). */
void ref_llg_rhs(int n, double alpha, int variant, const double* m, const double* H, double* dmdt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double mi[3] = {m[i], m[n + i], m[2 * n + i]};
        double h[3] = {H[i], H[n + i], H[2 * n + i]};
        double mxh[3], mxmxh[3];
        v3cross(mi, h, mxh);
        v3cross(mi, mxh, mxmxh);
        for (int d = 0; d < 3; ++d)
            dmdt[d * n + i] = (variant == 0) ? -alpha * mxmxh[d] : -mxh[d] - alpha * mxmxh[d];
    }
}

/* TheLLG.cpp:426-439 -- max_i |m_i x H_i| (the caller multiplies by mu0). */
double ref_max_torque(int n, const double* m, const double* H) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        double mi[3] = {m[i], m[n + i], m[2 * n + i]};
        double h[3] = {H[i], H[n + i], H[2 * n + i]};
        double c[3];
        v3cross(mi, h, c);
        double t = v3norm(c);
        if (t > mx) mx = t;
    }
    return mx;
}

/* Zhang-Li spin-transfer torque, u term: TheLLG.cpp:402-416 (calcUtermSTT,
 * CPU branch; GPU twin EffFieldGPU.cu:107-148):
 *   U_c = eta_x .* (gradX m_c) + eta_y .* (gradY m_c) + eta_z .* (gradZ m_c)
 * for c = x, y, z, with the grad operators of the mesh (TheSimulation.cpp:170)
 * and eta_d = jc_d * btimesMs * invJs (ProgramSpecs.cpp:393-421). */
void ref_stt_uterm(int n, const int* rowptr, const int* col, const double* gX, const double* gY, const double* gZ,
                   const double* eta_x, const double* eta_y, const double* eta_z, const double* m, double* U) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            const double* mc = m + (size_t)c * n;
            double sx = 0, sy = 0, sz = 0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                double v = mc[col[p]];
                sx += gX[p] * v;
                sy += gY[p] * v;
                sz += gZ[p] * v;
            }
            U[(size_t)c * n + i] = eta_x[i] * sx + eta_y[i] * sy + eta_z[i] * sz;
        }
    }
}

/* TheLLG.cpp:209-212 (sttDynamics, CPU branch; GPU twin EffFieldGPU.cu:434-475):
 *   tau = -(beta - alpha) m x U - (1 + alpha beta) m x (m x U). */
void ref_stt_llg(int n, double alpha, double beta, const double* m, const double* U, double* out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double mi[3] = {m[i], m[n + i], m[2 * n + i]};
        double u[3] = {U[i], U[n + i], U[2 * n + i]};
        double mxu[3], mxmxu[3];
        v3cross(mi, u, mxu);
        v3cross(mi, mxu, mxmxu);
        for (int d = 0; d < 3; ++d) out[(size_t)d * n + i] = -(beta - alpha) * mxu[d] - (1. + alpha * beta) * mxmxu[d];
    }
}

/* EffFieldGPU.cu:625-663 -- m_i <- m_i / |m_i| */
void ref_normalize(int n, double* m) {
    for (int i = 0; i < n; ++i) {
        double l = sqrt(m[i] * m[i] + m[n + i] * m[n + i] + m[2 * n + i] * m[2 * n + i]);
        m[i] /= l; m[n + i] /= l; m[2 * n + i] /= l;
    }
}

/* ------------------------------------------------------------------------ */
/* Demag (FEM-BEM)                                                           */
/* ------------------------------------------------------------------------ */

/* DemagField.cpp:124-136 -- b = (tGx (mx.*Js) + tGy (my.*Js) + tGz (mz.*Js)) / mu0 */
void ref_div_m(int n, const int* rowptr, const int* col, const double* tX, const double* tY, const double* tZ,
               const double* Js, const double* m, double* b) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double sx = 0, sy = 0, sz = 0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int j = col[p];
            sx += tX[p] * (m[j] * Js[j]);
            sy += tY[p] * (m[n + j] * Js[j]);
            sz += tZ[p] * (m[2 * n + j] * Js[j]);
        }
        b[i] = (sx + sy + sz) / REF_MU0;
    }
}

/* DemagField.cpp:102-113 -- Hdem = -(gradX u, gradY u, gradZ u) */
void ref_grad_field(int n, const int* rowptr, const int* col, const double* gX, const double* gY,
                    const double* gZ, const double* u, double* H) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double sx = 0, sy = 0, sz = 0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            double uj = u[col[p]];
            sx += gX[p] * uj;
            sy += gY[p] * uj;
            sz += gZ[p] * uj;
        }
        H[i] = -sx; H[n + i] = -sy; H[2 * n + i] = -sz;
    }
}

/* DemagField.cpp:184-201 -- gather u1 on the boundary, dense B product,
 * scatter into an n-vector that is zero elsewhere. */
void ref_boundary_integral(int n, int bnx, const int* bnodes, const double* B, const double* u1, double* g) {
    double* ub = (double*)calloc((size_t)bnx, sizeof(double));
    double* gb = (double*)malloc(sizeof(double) * (size_t)bnx);
    for (int i = 0; i < bnx; ++i) ub[i] = u1[bnodes[i]];
    ref_dense_gemv(bnx, B, ub, gb);
    for (int i = 0; i < n; ++i) g[i] = 0.0;
    for (int i = 0; i < bnx; ++i) g[bnodes[i]] = gb[i];
    free(ub);
    free(gb);
}

/* DemagField.cpp:97-99 -- E = -1/2 sum Js V Hdem . m */
double ref_demag_energy(int n, const double* Js, const double* nodevol, const double* H, const double* m) {
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s += Js[i] * nodevol[i] * (H[i] * m[i] + H[n + i] * m[n + i] + H[2 * n + i] * m[2 * n + i]);
    return -0.5 * s;
}

/* EffFieldCalc.cpp:65-69 (exchange) and 172-175 (DMI): -1/2 sum V H . m */
double ref_field_energy(int n, const double* nodevol, const double* H, const double* m) {
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s += nodevol[i] * (H[i] * m[i] + H[n + i] * m[n + i] + H[2 * n + i] * m[2 * n + i]);
    return -0.5 * s;
}

/* EffFieldCalc.cpp:86-95 -- sum V Ku - sum Ku V (k.m)^2 */
double ref_uniaxial_energy(int n, const double* Ku, const double* axis, const double* nodevol, const double* m) {
    double off = 0.0, s = 0.0;
    for (int i = 0; i < n; ++i) {
        off += nodevol[i] * Ku[i];
        double d = axis[i] * m[i] + axis[n + i] * m[n + i] + axis[2 * n + i] * m[2 * n + i];
        s += Ku[i] * nodevol[i] * d * d;
    }
    return off - s;
}

/* TheLLG.cpp:345-348 -- E = -sum Js V Hext . m */
double ref_zeeman_energy(int n, const double* Js, const double* nodevol, const double* Hext, const double* m) {
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s += Js[i] * nodevol[i] * (Hext[i] * m[i] + Hext[n + i] * m[n + i] + Hext[2 * n + i] * m[2 * n + i]);
    return -s;
}

int ref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_set_num_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
