// tm_cg.hip -- preconditioned conjugate gradients for tetmag's demag systems.
//
// Replaces GPUSolver (GPUSolver.h:53-87, GPUSolver.cu:43-98), which runs
// AMGCL smoothed-aggregation + ILU0 + GMRES on the row-replaced stiffness
// matrices.  Those matrices are symmetric except for identity rows
// (Neumann: DemagField.cpp:204-220; Dirichlet: FEMprocessing.cpp:329-369).
// With x preset to b on the identity rows, every residual and search
// direction vanishes on those rows and the iteration only sees the free
// block, which is symmetric positive definite -- so CG applies to exactly the
// reference's systems and returns the reference's solution (same gauge for
// the Neumann problem).
//
// Each iteration is two kernels with no host round trip:
//   k_cg_spmv   : p = z + beta p_old (fused into the gathers), q = A p, (p,q)
//   k_cg_update : x += alpha p, r -= alpha q, z = D^-1 r, (r,z), (r,r)
// Dot products finish inside the launch (last-block reduction) and the
// scalars stay in device memory.  The host checks convergence once per
// chunk of iterations.
#include <algorithm>
#include <cmath>

#include "tm_internal.hpp"

namespace tmh {

int spmv(const SellPattern& P, const double* val, const double* x, double* y, hipStream_t s);

namespace {

__global__ __launch_bounds__(kBlock) void k_cg_init_fixed(int n, const unsigned char* __restrict__ fixed,
                                                          const double* __restrict__ b, double* __restrict__ x,
                                                          double tol, int maxit, double* partials, CgState* st) {
    __shared__ double smem[kBlock / kWave];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const double bi = b[i];
        if (fixed[i]) x[i] = bi;
        v[0] = bi * bi;
    }
    block_sum<1>(v, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
    if (last_block_arrive(&st->counter)) {
        double tot[1];
        block_reduce_partials<1>(partials, 0, gridDim.x, tot, smem);
        if (threadIdx.x == 0) {
            st->bb = tot[0];
            st->tol2bb = tol * tol * tot[0];
            st->iters = 0;
            st->done = 0;
            st->beta = 0.0;
            st->alpha = 0.0;
            st->pad = maxit;
            atomicExch(&st->counter, 0u);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_init_res(SellView A, const double* __restrict__ val,
                                                        const double* __restrict__ dinv,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ r, double* __restrict__ z,
                                                        double* partials, CgState* st) {
    __shared__ double smem[2 * kBlock / kWave];
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    double v[2] = {0.0, 0.0};
    if (s < (A.n + kWave - 1) / kWave) {
        const int64_t base = A.off[s] + lane;
        const int w = A.w[s];
        double acc = 0.0;
        for (int k = 0; k < w; ++k) {
            const int64_t q = base + (int64_t)k * kWave;
            acc += val[q] * x[A.col[q]];
        }
        if (row < A.n) {
            const double ri = b[row] - acc;
            const double zi = dinv[row] * ri;
            r[row] = ri;
            z[row] = zi;
            v[0] = ri * zi;
            v[1] = ri * ri;
        }
    }
    block_sum<2>(v, smem);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v[0];
        partials[gridDim.x + blockIdx.x] = v[1];
    }
    if (last_block_arrive(&st->counter)) {
        double tot[2];
        block_reduce_partials<2>(partials, gridDim.x, gridDim.x, tot, smem);
        if (threadIdx.x == 0) {
            st->gamma = tot[0];
            st->rr = tot[1];
            if (!isfinite(tot[0]) || !isfinite(tot[1])) st->done = 2;
            else if (tot[1] <= st->tol2bb) st->done = 1;
            atomicExch(&st->counter, 0u);
        }
    }
}

// p_new = z + beta p_old; q = A p_new; partial (p_new, q).
__global__ __launch_bounds__(kBlock) void k_cg_spmv(SellView A, const double* __restrict__ val,
                                                    const double* __restrict__ z, const double* __restrict__ p_old,
                                                    double* __restrict__ p_new, double* __restrict__ q,
                                                    int first, double* partials, CgState* st) {
    __shared__ double smem[kBlock / kWave];
    if (st->done) return;
    const double beta = first ? 0.0 : st->beta;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    double v[1] = {0.0};
    if (s < (A.n + kWave - 1) / kWave) {
        const int64_t base = A.off[s] + lane;
        const int w = A.w[s];
        double acc = 0.0;
        if (first) {
            for (int k = 0; k < w; ++k) {
                const int64_t e = base + (int64_t)k * kWave;
                acc += val[e] * z[A.col[e]];
            }
        } else {
            for (int k = 0; k < w; ++k) {
                const int64_t e = base + (int64_t)k * kWave;
                const int c = A.col[e];
                acc += val[e] * (z[c] + beta * p_old[c]);
            }
        }
        if (row < A.n) {
            const double pi = first ? z[row] : z[row] + beta * p_old[row];
            p_new[row] = pi;
            q[row] = acc;
            v[0] = pi * acc;
        }
    }
    block_sum<1>(v, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
    if (last_block_arrive(&st->counter)) {
        double tot[1];
        block_reduce_partials<1>(partials, 0, gridDim.x, tot, smem);
        if (threadIdx.x == 0) {
            const double pq = tot[0];
            if (!(pq > 0.0) || !isfinite(pq)) st->done = 2;
            else st->alpha = st->gamma / pq;
            atomicExch(&st->counter, 0u);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_update(int n, const double* __restrict__ dinv,
                                                      const double* __restrict__ p, const double* __restrict__ q,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ z, double* partials, CgState* st) {
    __shared__ double smem[2 * kBlock / kWave];
    if (st->done) return;
    const double alpha = st->alpha;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
    }
    block_sum<2>(v, smem);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = v[0];
        partials[gridDim.x + blockIdx.x] = v[1];
    }
    if (last_block_arrive(&st->counter)) {
        double tot[2];
        block_reduce_partials<2>(partials, gridDim.x, gridDim.x, tot, smem);
        if (threadIdx.x == 0) {
            const double g = tot[0];
            st->beta = g / st->gamma;
            st->gamma = g;
            st->rr = tot[1];
            st->iters += 1;
            if (!isfinite(g) || !isfinite(tot[1])) st->done = 2;
            else if (tot[1] <= st->tol2bb) st->done = 1;
            else if (st->iters >= st->pad) st->done = 3;
            atomicExch(&st->counter, 0u);
        }
    }
}

}  // namespace

CgSolver::~CgSolver() {
    if (h_st) (void)hipHostFree(h_st);
    h_st = nullptr;
}

int CgSolver::setup(int n_, const int* rowptr, const int* colidx, const double* vals, int precond,
                    hipStream_t s) {
    if (precond != TM_PRECOND_JACOBI) return TM_ERR_BAD_ARG;
    n = n_;
    // identity rows: a single stored entry on the diagonal equal to 1
    std::vector<unsigned char> fx((size_t)n, 0);
    std::vector<double> di((size_t)n, 0.0);
    nfixed = 0;
    double amax = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) amax = std::max(amax, std::fabs(vals[q]));
        int nz = 0;
        bool diag1 = false;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            if (vals[q] != 0.0) {
                ++nz;
                if (colidx[q] == i && vals[q] == 1.0) diag1 = true;
            }
        }
        if (nz == 1 && diag1) {
            fx[i] = 1;
            ++nfixed;
        }
    }
    // diagonal, symmetry of the free block
    const double stol = 1e-12 * amax;
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int j = colidx[q];
            if (j == i) d = vals[q];
            if (fx[i] || fx[j] || j == i) continue;
            const int* lo = colidx + rowptr[j];
            const int* hi = colidx + rowptr[j + 1];
            const int* it = std::lower_bound(lo, hi, i);
            const double aji = (it != hi && *it == i) ? vals[it - colidx] : 0.0;
            if (std::fabs(aji - vals[q]) > stol) return TM_ERR_UNSUPPORTED;
        }
        if (fx[i]) di[i] = 1.0;
        else if (d > 0.0) di[i] = 1.0 / d;
        else return TM_ERR_UNSUPPORTED;  // free block must be positive definite
    }
    pat.build(n, n, rowptr, colidx);
    std::vector<double> sv;
    TM_CHECK(pat.scatter(rowptr, colidx, vals, sv));
    TM_CHECK(pat.upload(s));
    TM_CHECK(val.upload(sv.data(), sv.size(), s));
    TM_CHECK(dinv.upload(di.data(), di.size(), s));
    TM_CHECK(fixed.upload(fx.data(), fx.size(), s));
    TM_CHECK(r.alloc(n));
    TM_CHECK(z.alloc(n));
    TM_CHECK(q.alloc(n));
    TM_CHECK(p0.alloc(n));
    TM_CHECK(p1.alloc(n));
    TM_CHECK(xw.alloc(n));
    TM_HIP(hipMemsetAsync(xw.p, 0, sizeof(double) * n, s));
    const int nb = grid_for((int64_t)pat.nslices * kWave);
    TM_CHECK(partials.alloc((size_t)2 * std::max(nb, grid_for(n))));
    TM_CHECK(st.alloc(1));
    TM_HIP(hipMemsetAsync(st.p, 0, sizeof(CgState), s));
    if (!h_st) TM_HIP(hipHostMalloc((void**)&h_st, sizeof(CgState), hipHostMallocDefault));
    TM_HIP(hipStreamSynchronize(s));
    last_iters = 0;
    has_solution = false;
    return TM_OK;
}

int CgSolver::solve(const double* b, double* x, double tol, int maxit, hipStream_t s, int* iters, double* relres) {
    if (n == 0) return TM_ERR_NOT_READY;
    if (!(tol > 0.0) || maxit < 1) return TM_ERR_BAD_ARG;
    const int gv = grid_for(n);
    const int gs = grid_for((int64_t)pat.nslices * kWave);
    const SellView A = view(pat);
    hipLaunchKernelGGL(k_cg_init_fixed, dim3(gv), dim3(kBlock), 0, s, n, fixed.p, b, x, tol, maxit, partials.p,
                       st.p);
    hipLaunchKernelGGL(k_cg_init_res, dim3(gs), dim3(kBlock), 0, s, A, val.p, dinv.p, b, x, r.p, z.p, partials.p,
                       st.p);
    TM_HIP(hipGetLastError());
    // First chunk sized from the previous solve (warm-started solves along a
    // trajectory need similar counts); then small chunks.
    int enq = 0;
    int chunk = std::max(4, std::min(maxit, last_iters > 8 ? last_iters - 2 : 8));
    double* pcur = p0.p;
    double* pold = p1.p;
    for (;;) {
        for (int k = 0; k < chunk && enq < maxit; ++k, ++enq) {
            hipLaunchKernelGGL(k_cg_spmv, dim3(gs), dim3(kBlock), 0, s, A, val.p, z.p, pold, pcur, q.p,
                               enq == 0 ? 1 : 0, partials.p, st.p);
            hipLaunchKernelGGL(k_cg_update, dim3(gv), dim3(kBlock), 0, s, n, dinv.p, pcur, q.p, x, r.p, z.p,
                               partials.p, st.p);
            std::swap(pcur, pold);
        }
        TM_HIP(hipGetLastError());
        TM_HIP(hipMemcpyAsync(h_st, st.p, sizeof(CgState), hipMemcpyDeviceToHost, s));
        TM_HIP(hipStreamSynchronize(s));
        if (h_st->done || enq >= maxit) break;
        chunk = 4;
    }
    const CgState hs = *h_st;
    if (hs.bb == 0.0) TM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    last_iters = hs.iters;
    has_solution = true;
    if (iters) *iters = hs.iters;
    if (relres) *relres = hs.bb > 0 ? std::sqrt(hs.rr / hs.bb) : 0.0;
    if (hs.done == 2) return TM_ERR_NONFINITE;
    if (hs.done == 1 || hs.bb == 0.0) return TM_OK;
    return TM_ERR_MAXITER;
}

}  // namespace tmh

using namespace tmh;

extern "C" {

int tm_solver_create(int n, const int* rowptr, const int* colidx, const double* vals, int precond, tm_solver** out) {
    if (!out || n <= 0 || !rowptr || !colidx || !vals) return TM_ERR_BAD_ARG;
    int nd = 0;
    TM_CHECK(tm_device_count(&nd));
    tm_solver* S = new tm_solver();
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return TM_ERR_HIP;
    }
    int st = S->cg.setup(n, rowptr, colidx, vals, precond, S->stream);
    if (st == TM_OK) st = S->b.alloc(n);
    if (st == TM_OK) st = S->x.alloc(n);
    if (st == TM_OK && hipMemsetAsync(S->x.p, 0, sizeof(double) * n, S->stream) != hipSuccess) st = TM_ERR_HIP;
    if (st == TM_OK && hipStreamSynchronize(S->stream) != hipSuccess) st = TM_ERR_HIP;
    if (st != TM_OK) {
        tm_solver_destroy(S);
        return st;
    }
    *out = S;
    return TM_OK;
}

int tm_solver_destroy(tm_solver* S) {
    if (!S) return TM_OK;
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return TM_OK;
}

int tm_solver_solve(tm_solver* S, const double* b, int bwhere, double* x, int xwhere, double tol, int maxit,
                    int warm, int* iters, double* relres) {
    if (!S || !b || !x) return TM_ERR_BAD_ARG;
    if ((bwhere != TM_HOST && bwhere != TM_DEVICE) || (xwhere != TM_HOST && xwhere != TM_DEVICE))
        return TM_ERR_BAD_ARG;
    const int n = S->cg.n;
    const double* bd = b;
    if (bwhere == TM_HOST) {
        TM_CHECK(copy_in(S->b.p, b, TM_HOST, n, S->stream));
        bd = S->b.p;
    }
    double* xd = S->x.p;  // the solver keeps its solution between calls (GPUSolver::x_d)
    if (warm == 1) TM_CHECK(copy_in(xd, x, xwhere, n, S->stream));
    else if (warm == 0) TM_HIP(hipMemsetAsync(xd, 0, sizeof(double) * n, S->stream));
    int st = S->cg.solve(bd, xd, tol, maxit, S->stream, iters, relres);
    int st2 = copy_out(x, xwhere, xd, n, S->stream);
    if (xwhere == TM_DEVICE) TM_HIP(hipStreamSynchronize(S->stream));
    return st != TM_OK ? st : st2;
}

}  // extern "C"
