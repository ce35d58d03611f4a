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
// Each iteration is two kernels with no host round trip and no device-scope
// fence:
//   k_cg_spmv(k)   : gamma_k, (r,r) from the update's partials; beta; p = z +
//                    beta p_old fused into the gathers; q = A p; partial (p,q)
//   k_cg_update(k) : alpha = gamma_k / (p,q); x += alpha p; r -= alpha q;
//                    z = D^-1 r; partials (r,z), (r,r)
// Every block reduces the producer's partials itself (tm_internal.hpp), so all
// blocks take the same convergence decision; block 0 publishes the scalars.
// Grids are at most kMaxCgBlocks blocks (grid-stride), which keeps the
// partial arrays short.  The host checks the state once per chunk.
#include <algorithm>
#include <cmath>

#include "tm_internal.hpp"

namespace tmh {

namespace {

// Slices handled by this block's wavefronts: s = wave_global, += total waves.
#define CG_SLICE_LOOP(A)                                                                             \
    const int lane = threadIdx.x & (kWave - 1);                                                       \
    const int nsl = ((A).n + kWave - 1) / kWave;                                                      \
    const int spb_ = ((nsl + (int)gridDim.x - 1) / (int)gridDim.x + kBlock / kWave - 1) / (kBlock / kWave) * \
                     (kBlock / kWave);                                                                \
    const int s0_ = xcd_block() * spb_, s1_ = min(nsl, s0_ + spb_);                                   \
    for (int s = s0_ + (int)threadIdx.x / kWave; s < s1_; s += kBlock / kWave)

__device__ inline void block_partials(double* part, int stride, double v0, double v1, int nv) {
    __shared__ double sp[2][kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v0 = wave_sum(v0);
    if (nv > 1) v1 = wave_sum(v1);
    if (lane == 0) {
        sp[0][wid] = v0;
        sp[1][wid] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) {
            a += sp[0][k];
            b += sp[1][k];
        }
        part[blockIdx.x] = a;
        if (nv > 1) part[stride + blockIdx.x] = b;
    }
}

// ---- initial guesses from earlier solves ---------------------------------
// x0 = x_0 + sum_j c_j dX_j over the kept pairs (backward differences at the
// newest, least squares in the residual norm; the method is described with
// kSolveWindow in tm_internal.hpp).  Slots are ordered newest first.

// per-block partials wpart[q * G + block] of the kWinDots dots.  The new b
// goes into the head slot as it is read: when the window is full that slot
// holds the oldest pair, whose value this thread has already loaded.
__global__ __launch_bounds__(kBlock) void k_cg_win_dots(int n, const CgWindow* w, double* __restrict__ Bw,
                                                        const double* __restrict__ b, double* wpart) {
    const int cnt = w->cnt, head = w->head;
    double acc[kWinDots] = {};
    const double* src[kSolveWindow];
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) src[i] = Bw + (size_t)win_slot(head, i) * n;
    double* dst = Bw + (size_t)head * n;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
        double v[kSolveWindow];
#pragma unroll
        for (int i = 0; i < kSolveWindow; ++i) v[i] = i < cnt ? src[i][e] : 0.0;
        const double bi = b[e];
        const double r = bi - v[0];
        win_differences(v, cnt);
        win_accumulate(acc, v, r);
        dst[e] = bi;
    }
    __shared__ double sp[kWinDots][kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < kWinDots; ++q) {
        const double t = wave_sum(acc[q]);
        if (lane == 0) sp[q][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < kWinDots) {
        double a = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) a += sp[threadIdx.x][k];
        wpart[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
    }
}

// one block: the dots, then the coefficients (thread 0)
// (one pass over all partials with every load independent; the solve on
// register copies, so the unrolled factorisation makes no memory accesses)
__global__ __launch_bounds__(kBlock) void k_cg_win_solve(int G, CgWindow* w, const double* wpart) {
    double v[kWinDots] = {};
    for (int g = threadIdx.x; g < G; g += kBlock)
#pragma unroll
        for (int q = 0; q < kWinDots; ++q) v[q] += wpart[(size_t)q * G + g];
    __shared__ double sr[kWinDots][kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < kWinDots; ++q) {
        const double s = wave_sum(v[q]);
        if (lane == 0) sr[q][wid] = s;
    }
    __syncthreads();
    const int cnt = w->cnt;
    if (threadIdx.x != 0 || cnt == 0) return;
    double t[kWinDots];
#pragma unroll
    for (int q = 0; q < kWinDots; ++q) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kBlock / kWave; ++k) s += sr[q][k];
        t[q] = s;
    }
    double a[kSolveWindow];
    win_coefficients(cnt, t, a);
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) w->a[i] = a[i];
}

// x0 = sum_i a_i x_slot(i) (nothing when the window is empty)
__global__ __launch_bounds__(kBlock) void k_cg_win_combine(int n, const CgWindow* w, const double* __restrict__ Xw,
                                                           double* __restrict__ x) {
    const int cnt = w->cnt;
    if (cnt == 0) return;
    double a[kSolveWindow];
    const double* src[kSolveWindow];
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) {
        a[i] = w->a[i];
        src[i] = Xw + (size_t)win_slot(w->head, i) * n;
    }
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < kSolveWindow; ++i)
            if (i < cnt) acc += a[i] * src[i][e];
        x[e] = acc;
    }
}

// the solution into the head slot (its b is there already)
__global__ __launch_bounds__(kBlock) void k_cg_win_store(int n, const CgWindow* w, const double* __restrict__ x,
                                                         double* __restrict__ Xw) {
    const size_t off = (size_t)w->head * n;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < n; e += gridDim.x * kBlock) Xw[off + e] = x[e];
}

// keep the pair and advance the head; a failed solve empties the window
__global__ void k_cg_win_advance(CgWindow* w, const CgState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (st->done != 1) {
        w->cnt = 0;
        w->head = 0;
        return;
    }
    w->head = (w->head + 1) % kSolveWindow;
    w->cnt = min(w->cnt + 1, kSolveWindow);
}

// part layout: [0, G) bb | [G, 2G) rz | [2G, 3G) rr | [3G, 4G) pq
__global__ __launch_bounds__(kBlock) void k_cg_init_fixed(int n, const unsigned char* __restrict__ fixed,
                                                          const double* __restrict__ b, double* __restrict__ x,
                                                          double* part) {
    double v = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double bi = b[i];
        if (fixed[i]) x[i] = bi;
        v += bi * bi;
    }
    block_partials(part, 0, v, 0.0, 1);
}

template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_cg_init_res(SellView A, const double* __restrict__ val,
                                                        const double* __restrict__ dinv,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ r, double* __restrict__ z,
                                                        double* part, double tol, CgState* st) {
    const int G = gridDim.x;
    double bb[1];
    reduce_partials_all<1>(part, G, G, bb);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->bb = bb[0];
        st->tol2bb = tol * tol * bb[0];
        st->iters = 0;
        st->done = 0;
        st->rr = 0.0;
    }
    double vrz = 0.0, vrr = 0.0;
    CG_SLICE_LOOP(A) {
        double acc = 0.0;
        sell_row(A, s, lane, [&](int64_t e, int c) { acc += val[e] * x[c]; });
        const int row = s * kWave + lane;
        if (row < A.n) {
            const double ri = b[row] - acc;
            r[row] = ri;
            vrr += ri * ri;
            if (JAC) {
                const double zi = dinv[row] * ri;
                z[row] = zi;
                vrz += ri * zi;
            }
        }
    }
    if (JAC) block_partials(part + G, G, vrz, vrr, 2);
    else block_partials(part + 2 * G, G, vrr, 0.0, 1);
}

__global__ __launch_bounds__(kBlock) void k_cg_spmv(SellView A, const double* __restrict__ val,
                                                    const double* __restrict__ z, const double* __restrict__ p_old,
                                                    double* __restrict__ p_new, double* __restrict__ q, int k,
                                                    double* part, CgState* st) {
    if (st->done) return;
    const int G = gridDim.x;
    double t[2];
    reduce_partials_all<2>(part + G, G, G, t);  // (r,z), (r,r)
    const double gam = t[0], rr = t[1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!isfinite(gam) || !isfinite(rr)) {
        if (lead) st->done = 2;
        return;
    }
    if (rr <= st->tol2bb) {
        if (lead) {
            st->done = 1;
            st->rr = rr;
            st->iters = k;
        }
        return;
    }
    const double beta = k > 0 ? gam / st->gam[(k - 1) & 1] : 0.0;
    if (lead) {
        st->gam[k & 1] = gam;
        st->rr = rr;
    }
    double v = 0.0;
    CG_SLICE_LOOP(A) {
        double acc = 0.0;
        if (k == 0) sell_row(A, s, lane, [&](int64_t e, int c) { acc += val[e] * z[c]; });
        else sell_row(A, s, lane, [&](int64_t e, int c) { acc += val[e] * (z[c] + beta * p_old[c]); });
        const int row = s * kWave + lane;
        if (row < A.n) {
            const double pi = k == 0 ? z[row] : z[row] + beta * p_old[row];
            p_new[row] = pi;
            q[row] = acc;
            v += pi * acc;
        }
    }
    block_partials(part + 3 * G, G, v, 0.0, 1);
}

template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_cg_update(int n, const double* __restrict__ dinv,
                                                      const double* __restrict__ p, const double* __restrict__ q,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ z, int k, double* part, CgState* st) {
    if (st->done) return;
    const int G = gridDim.x;
    double t[1];
    reduce_partials_all<1>(part + 3 * G, G, G, t);
    const double pq = t[0];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(pq > 0.0) || !isfinite(pq)) {
        if (lead) st->done = 2;
        return;
    }
    const double alpha = st->gam[k & 1] / pq;
    double vrz = 0.0, vrr = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += G * kBlock) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        vrr += ri * ri;
        if (JAC) {
            const double zi = dinv[i] * ri;
            z[i] = zi;
            vrz += ri * zi;
        }
    }
    if (lead) st->iters = k + 1;
    if (JAC) block_partials(part + G, G, vrz, vrr, 2);
    else block_partials(part + 2 * G, G, vrr, 0.0, 1);
}

// (r, r) after the last update when the loop stopped on maxit.
__global__ __launch_bounds__(kBlock) void k_cg_final(int G, double* part, CgState* st) {
    double t[2];
    reduce_partials_all<2>(part + G, G, G, t);
    if (threadIdx.x == 0 && st->done == 0) st->rr = t[1];
}

// Fixed-budget solves (CgSolver::enqueue): b = 0 gives x = 0 at once.
__global__ __launch_bounds__(kBlock) void k_cg_trivial(int n, double* __restrict__ x, CgState* st) {
    if (st->bb != 0.0) return;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->done = 1;
        st->rr = 0.0;
    }
}

// End of a fixed-budget solve: a residual that met the tolerance on the last
// update counts as converged; anything else increments *fails.  *imax keeps
// the largest iteration count seen.
__global__ __launch_bounds__(kBlock) void k_cg_check(int G, int budget, const double* part, CgState* st,
                                                     unsigned* fails, int* imax) {
    double t[2];
    reduce_partials_all<2>(part + G, G, G, t);
    if (threadIdx.x != 0) return;
    if (st->done == 0) {
        st->rr = t[1];
        if (t[1] <= st->tol2bb) {
            st->done = 1;
            st->iters = budget;
        }
    }
    if (st->done != 1) fails[0] += 1u;
    if (st->iters > imax[0]) imax[0] = st->iters;
}

}  // namespace

// Transposed solve helpers: zero the fixed entries of the right-hand side,
// then y_j = w_j - sum_i A_ij y_i over the free rows i coupled to fixed j.
__global__ __launch_bounds__(kBlock) void k_adj_zero_fixed(int n, const unsigned char* __restrict__ fixed,
                                                           const double* __restrict__ w, double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = fixed[i] ? 0.0 : w[i];
}

__global__ __launch_bounds__(kBlock) void k_adj_fix_rows(int nf, const int* __restrict__ idx,
                                                         const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const double* __restrict__ v, const double* __restrict__ w,
                                                         double* __restrict__ y) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= nf) return;
    const int j = idx[r];
    double acc = w[j];
    for (int q = rp[r]; q < rp[r + 1]; ++q) acc -= v[q] * y[ci[q]];
    y[j] = acc;
}

CgSolver::~CgSolver() {
    if (h_st) (void)hipHostFree(h_st);
    h_st = nullptr;
}

int CgSolver::setup(int n_, const int* rowptr, const int* colidx, const double* vals, int precond_,
                    hipStream_t s) {
    if (precond_ != TM_PRECOND_JACOBI && precond_ != TM_PRECOND_AMG) return TM_ERR_BAD_ARG;
    precond = precond_;
    n = n_;
    // identity rows: a single stored nonzero on the diagonal equal to 1
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
    // fixed-row coupling for the transposed solve
    {
        std::vector<int> fidx, frp(1, 0), fci, pos((size_t)n, -1);
        for (int i = 0; i < n; ++i)
            if (fx[i]) {
                pos[i] = (int)fidx.size();
                fidx.push_back(i);
            }
        std::vector<std::vector<std::pair<int, double>>> rows(fidx.size());
        for (int i = 0; i < n; ++i) {
            if (fx[i]) continue;
            for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
                const int j = colidx[q];
                if (fx[j] && vals[q] != 0.0) rows[pos[j]].push_back({i, vals[q]});
            }
        }
        std::vector<double> fv;
        for (auto& r : rows) {
            for (auto& e : r) {
                fci.push_back(e.first);
                fv.push_back(e.second);
            }
            frp.push_back((int)fci.size());
        }
        if (!fidx.empty()) {
            TM_CHECK(fix_idx.upload(fidx.data(), fidx.size(), s));
            TM_CHECK(fix_rp.upload(frp.data(), frp.size(), s));
            if (!fci.empty()) {
                TM_CHECK(fix_ci.upload(fci.data(), fci.size(), s));
                TM_CHECK(fix_v.upload(fv.data(), fv.size(), s));
            }
            TM_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
        }
    }
    pat.build(n, n, rowptr, colidx);
    h_val.assign(vals, vals + rowptr[n]);
    h_dinv = di;
    h_fixed = fx;
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
    grid = xcd_grid(std::min(kMaxCgBlocks, (pat.nslices + kBlock / kWave - 1) / (kBlock / kWave)));
    TM_CHECK(part.alloc((size_t)4 * grid));
    TM_CHECK(winX.alloc((size_t)kSolveWindow * n));
    TM_CHECK(winB.alloc((size_t)kSolveWindow * n));
    TM_CHECK(wpart.alloc((size_t)kWinDots * grid));
    TM_CHECK(win.alloc(1));
    TM_HIP(hipMemsetAsync(win.p, 0, sizeof(CgWindow), s));
    TM_CHECK(st.alloc(1));
    TM_HIP(hipMemsetAsync(st.p, 0, sizeof(CgState), s));
    if (!h_st) TM_HIP(hipHostMalloc((void**)&h_st, sizeof(CgState), hipHostMallocDefault));
    TM_HIP(hipStreamSynchronize(s));
    if (precond == TM_PRECOND_AMG) TM_CHECK(amg.setup(n, rowptr, colidx, vals, fx, &pat, val.p, dinv.p, s));
    last_iters = 0;
    return TM_OK;
}

void CgSolver::win_project(const double* b, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_win_dots, dim3(grid), dim3(kBlock), 0, s, n, (const CgWindow*)win.p, winB.p, b, wpart.p);
    hipLaunchKernelGGL(k_cg_win_solve, dim3(1), dim3(kBlock), 0, s, grid, win.p, (const double*)wpart.p);
    hipLaunchKernelGGL(k_cg_win_combine, dim3(grid), dim3(kBlock), 0, s, n, (const CgWindow*)win.p,
                       (const double*)winX.p, x);
}

// after win_project with the same b: the head slot holds b already
void CgSolver::win_record(const double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_win_store, dim3(grid), dim3(kBlock), 0, s, n, (const CgWindow*)win.p, x, winX.p);
    hipLaunchKernelGGL(k_cg_win_advance, dim3(1), dim3(1), 0, s, win.p, (const CgState*)st.p);
}

int CgSolver::solve(const double* b, double* x, double tol, int maxit, hipStream_t s, int* iters, double* relres,
                    bool project) {
    if (n == 0) return TM_ERR_NOT_READY;
    if (!(tol > 0.0) || maxit < 1) return TM_ERR_BAD_ARG;
    const int G = grid;
    const SellView A = view(pat);
    if (project) win_project(b, x, s);
    hipLaunchKernelGGL(k_cg_init_fixed, dim3(G), dim3(kBlock), 0, s, n, fixed.p, b, x, part.p);
    const bool jac = precond == TM_PRECOND_JACOBI;
    if (jac)
        hipLaunchKernelGGL(k_cg_init_res<true>, dim3(G), dim3(kBlock), 0, s, A, val.p, dinv.p, b, x, r.p, z.p, part.p,
                           tol, st.p);
    else
        hipLaunchKernelGGL(k_cg_init_res<false>, dim3(G), dim3(kBlock), 0, s, A, val.p, dinv.p, b, x, r.p, z.p,
                           part.p, tol, st.p);
    TM_HIP(hipGetLastError());
    if (!jac) TM_CHECK(amg.apply(r.p, z.p, part.p + G, G, s, st.p, part.p + 2 * G));
    // First chunk sized from the previous solve (warm-started solves along a
    // trajectory need similar counts), then small chunks.
    int enq = 0;
    int chunk = std::max(4, std::min(maxit, last_iters > 8 ? last_iters - 2 : 8));
    double* pcur = p0.p;
    double* pold = p1.p;
    for (;;) {
        for (int c = 0; c < chunk && enq < maxit; ++c, ++enq) {
            hipLaunchKernelGGL(k_cg_spmv, dim3(G), dim3(kBlock), 0, s, A, val.p, z.p, pold, pcur, q.p, enq, part.p,
                               st.p);
            if (jac) {
                hipLaunchKernelGGL(k_cg_update<true>, dim3(G), dim3(kBlock), 0, s, n, dinv.p, pcur, q.p, x, r.p, z.p,
                                   enq, part.p, st.p);
            } else {
                hipLaunchKernelGGL(k_cg_update<false>, dim3(G), dim3(kBlock), 0, s, n, dinv.p, pcur, q.p, x, r.p,
                                   z.p, enq, part.p, st.p);
                TM_CHECK(amg.apply(r.p, z.p, part.p + G, G, s, st.p, part.p + 2 * G));
            }
            std::swap(pcur, pold);
        }
        TM_HIP(hipGetLastError());
        TM_HIP(hipMemcpyAsync(h_st, st.p, sizeof(CgState), hipMemcpyDeviceToHost, s));
        TM_HIP(hipStreamSynchronize(s));
        if (h_st->done || enq >= maxit) break;
        chunk = 4;
    }
    if (!h_st->done) {
        // one more convergence test on the last residual
        hipLaunchKernelGGL(k_cg_final, dim3(1), dim3(kBlock), 0, s, G, part.p, st.p);
        TM_HIP(hipMemcpyAsync(h_st, st.p, sizeof(CgState), hipMemcpyDeviceToHost, s));
        TM_HIP(hipStreamSynchronize(s));
    }
    const CgState hs = *h_st;
    if (hs.bb == 0.0) TM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    if (project) win_record(x, s);
    last_iters = hs.iters;
    if (iters) *iters = hs.iters;
    if (relres) *relres = hs.bb > 0 ? std::sqrt(hs.rr / hs.bb) : 0.0;
    if (hs.done == 2) return TM_ERR_NONFINITE;
    if (hs.done == 1 || hs.bb == 0.0 || hs.rr <= hs.tol2bb) return TM_OK;
    return TM_ERR_MAXITER;
}

int CgSolver::enqueue(const double* b, double* x, double tol, int budget, hipStream_t s, unsigned* fails,
                      int* imax, bool project) {
    if (n == 0) return TM_ERR_NOT_READY;
    if (!(tol > 0.0) || budget < 1 || !fails || !imax) return TM_ERR_BAD_ARG;
    const int G = grid;
    const SellView A = view(pat);
    const bool jac = precond == TM_PRECOND_JACOBI;
    if (project) win_project(b, x, s);
    hipLaunchKernelGGL(k_cg_init_fixed, dim3(G), dim3(kBlock), 0, s, n, fixed.p, b, x, part.p);
    if (jac)
        hipLaunchKernelGGL(k_cg_init_res<true>, dim3(G), dim3(kBlock), 0, s, A, val.p, dinv.p, b, x, r.p, z.p, part.p,
                           tol, st.p);
    else
        hipLaunchKernelGGL(k_cg_init_res<false>, dim3(G), dim3(kBlock), 0, s, A, val.p, dinv.p, b, x, r.p, z.p,
                           part.p, tol, st.p);
    hipLaunchKernelGGL(k_cg_trivial, dim3(G), dim3(kBlock), 0, s, n, x, st.p);
    if (!jac) TM_CHECK(amg.apply(r.p, z.p, part.p + G, G, s, st.p, part.p + 2 * G));
    double* pcur = p0.p;
    double* pold = p1.p;
    for (int k = 0; k < budget; ++k) {
        hipLaunchKernelGGL(k_cg_spmv, dim3(G), dim3(kBlock), 0, s, A, val.p, z.p, pold, pcur, q.p, k, part.p, st.p);
        if (jac) {
            hipLaunchKernelGGL(k_cg_update<true>, dim3(G), dim3(kBlock), 0, s, n, dinv.p, pcur, q.p, x, r.p, z.p, k,
                               part.p, st.p);
        } else {
            hipLaunchKernelGGL(k_cg_update<false>, dim3(G), dim3(kBlock), 0, s, n, dinv.p, pcur, q.p, x, r.p, z.p, k,
                               part.p, st.p);
            if (k + 1 < budget)  // the z of the last update would go unused
                TM_CHECK(amg.apply(r.p, z.p, part.p + G, G, s, st.p, part.p + 2 * G));
        }
        std::swap(pcur, pold);
    }
    hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(kBlock), 0, s, G, budget, (const double*)part.p, st.p, fails, imax);
    if (project) win_record(x, s);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

int CgSolver::solve_transposed(const double* w, double* y, double tol, int maxit, hipStream_t s, int* iters,
                               double* relres) {
    if (n == 0) return TM_ERR_NOT_READY;
    TM_CHECK(tb.alloc(n));
    hipLaunchKernelGGL(k_adj_zero_fixed, dim3(grid_for(n)), dim3(kBlock), 0, s, n, fixed.p, w, tb.p);
    TM_HIP(hipGetLastError());
    const int st = solve(tb.p, y, tol, maxit, s, iters, relres);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    if (nfixed > 0) {
        hipLaunchKernelGGL(k_adj_fix_rows, dim3(grid_for(nfixed)), dim3(kBlock), 0, s, nfixed, fix_idx.p, fix_rp.p,
                           fix_ci.p, fix_v.p, w, y);
        TM_HIP(hipGetLastError());
    }
    return st;
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
    if (st == TM_OK) st = S->xt.alloc(n);
    if (st == TM_OK && hipMemsetAsync(S->x.p, 0, sizeof(double) * n, S->stream) != hipSuccess) st = TM_ERR_HIP;
    if (st == TM_OK && hipMemsetAsync(S->xt.p, 0, sizeof(double) * n, S->stream) != hipSuccess) st = TM_ERR_HIP;
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
    // warm == 2 (the solver's own previous solution): start from the best
    // combination of its last solutions instead
    int st = S->cg.solve(bd, xd, tol, maxit, S->stream, iters, relres, warm == 2);
    int st2 = copy_out(x, xwhere, xd, n, S->stream);
    if (xwhere == TM_DEVICE) TM_HIP(hipStreamSynchronize(S->stream));
    return st != TM_OK ? st : st2;
}

int tm_solver_solve_transposed(tm_solver* S, const double* b, int bwhere, double* x, int xwhere, double tol,
                               int maxit, int warm, int* iters, double* relres) {
    if (!S || !b || !x) return TM_ERR_BAD_ARG;
    if ((bwhere != TM_HOST && bwhere != TM_DEVICE) || (xwhere != TM_HOST && xwhere != TM_DEVICE))
        return TM_ERR_BAD_ARG;
    const int n = S->cg.n;
    const double* bd = b;
    if (bwhere == TM_HOST) {
        TM_CHECK(copy_in(S->b.p, b, TM_HOST, n, S->stream));
        bd = S->b.p;
    }
    double* xd = S->xt.p;
    if (warm == 1) TM_CHECK(copy_in(xd, x, xwhere, n, S->stream));
    else if (warm == 0) TM_HIP(hipMemsetAsync(xd, 0, sizeof(double) * n, S->stream));
    int st = S->cg.solve_transposed(bd, xd, tol, maxit, S->stream, iters, relres);
    int st2 = copy_out(x, xwhere, xd, n, S->stream);
    if (xwhere == TM_DEVICE) TM_HIP(hipStreamSynchronize(S->stream));
    return st != TM_OK ? st : st2;
}

}  // extern "C"

namespace tmh {
int amg_info(const CgSolver& cg, int* nlevels, int* sizes, int cap) {
    if (!nlevels || (cap > 0 && !sizes)) return TM_ERR_BAD_ARG;
    if (cg.precond != TM_PRECOND_AMG) {
        *nlevels = 0;
        return TM_OK;
    }
    const int L = (int)cg.amg.levels.size();
    *nlevels = L + 1;
    for (int l = 0; l < L + 1 && l < cap; ++l) sizes[l] = l < L ? cg.amg.levels[l].n : cg.amg.nc;
    return TM_OK;
}
}  // namespace tmh

extern "C" int tm_solver_amg_info(tm_solver* s, int* nlevels, int* sizes, int cap) {
    if (!s) return TM_ERR_BAD_ARG;
    return tmh::amg_info(s->cg, nlevels, sizes, cap);
}
