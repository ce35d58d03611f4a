// tm_internal.hpp -- shared internals of the HIP library (not part of the ABI).
#pragma once

#include <algorithm>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "tetmag_hip.h"

namespace tmh {

constexpr int kWave = 64;      // CDNA wavefront
constexpr int kBlock = 256;    // 4 wavefronts per workgroup
constexpr double kMu0 = 1.256637061435917295376 * 1.e-6;  // PhysicalConstants.h:39
constexpr double kPi = 3.14159265358979323846264338327950288;

#define TM_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "tetmag_hip: %s failed: %s (%s:%d)\n", #call,                \
                         hipGetErrorString(e_), __FILE__, __LINE__);                          \
            return TM_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)

#define TM_CHECK(st)                 \
    do {                             \
        int s_ = (st);               \
        if (s_ != TM_OK) return s_;  \
    } while (0)

// Device buffer owned by RAII.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        if (count == n && p) return TM_OK;
        release();
        if (count == 0) return TM_OK;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return TM_ERR_HIP;
        }
        n = count;
        return TM_OK;
    }
    int upload(const T* host, size_t count, hipStream_t s) {
        TM_CHECK(alloc(count));
        if (count) TM_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
        return TM_OK;
    }
};

// SELL-64 sparse storage: rows are grouped into slices of one wavefront (64
// rows); slice s stores its entries column-major (entry k of lane l at
// off[s] + 64 k + l), padded to the slice's longest row with col = own row
// (or 0 past n) and value 0.  Lane l of a wavefront owns row 64 s + l, so
// each load instruction of the SpMV reads 64 consecutive entries.
struct SellPattern {
    int n = 0;        // rows
    int ncols = 0;    // columns
    int nslices = 0;
    int64_t nnz = 0;        // true CSR nonzeros
    int64_t stored = 0;     // padded entries
    std::vector<int64_t> h_off;  // nslices+1
    std::vector<int> h_w;        // nslices
    std::vector<int> h_col;      // stored
    std::vector<int> h_rowptr;   // CSR of the pattern (n+1), sorted columns
    std::vector<int> h_csrcol;   // CSR columns (nnz)
    DevBuf<int64_t> off;
    DevBuf<int> w;
    DevBuf<int> col;
    // Builds from a CSR pattern with sorted columns.
    void build(int rows, int cols, const int* rowptr, const int* colidx);
    int upload(hipStream_t s);
    // Scatter CSR values (on this or a sub-pattern) into SELL order.  Entries
    // of the pattern absent from the given CSR get 0.  Returns TM_ERR_BAD_ARG
    // if the CSR has an entry outside the pattern.
    int scatter(const int* rowptr, const int* colidx, const double* vals, std::vector<double>& out) const;
};

struct SellView {
    int n;
    const int64_t* __restrict__ off;
    const int* __restrict__ w;
    const int* __restrict__ col;
};

inline SellView view(const SellPattern& P) { return SellView{P.n, P.off.p, P.w.p, P.col.p}; }

// Calls f(entry, column) for every stored entry of the row owned by `lane` in
// slice s.  Unrolled by four so the column loads of four entries (and the
// dependent value/gather loads) are in flight together.
template <class F>
__device__ __forceinline__ void sell_row(const SellView& A, int s, int lane, F&& f) {
    const int64_t base = A.off[s] + lane;
    const int w = A.w[s];
    int k = 0;
    for (; k + 4 <= w; k += 4) {
        const int64_t e0 = base + (int64_t)k * kWave;
        const int64_t e1 = e0 + kWave, e2 = e1 + kWave, e3 = e2 + kWave;
        const int c0 = A.col[e0], c1 = A.col[e1], c2 = A.col[e2], c3 = A.col[e3];
        f(e0, c0);
        f(e1, c1);
        f(e2, c2);
        f(e3, c3);
    }
    for (; k < w; ++k) {
        const int64_t e = base + (int64_t)k * kWave;
        f(e, A.col[e]);
    }
}

// Block-wide reductions for kBlock threads.
__device__ inline double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, kWave));
    return v;
}
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* smem /* NV * kBlock/kWave */) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wave_sum(v[q]);
        if (lane == 0) smem[q * (kBlock / kWave) + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int k = 0; k < kBlock / kWave; ++k) s += smem[q * (kBlock / kWave) + k];
            v[q] = s;
        }
    }
}

// Every block of a consumer kernel reduces the partials written by the
// producer kernel (one per producer block, at most a few hundred) itself; the
// sum is then known to every thread of the block.  Kernel boundaries order
// producer and consumer, so no device-scope fence is needed, and every block
// computes bit-identical totals (same data, same order), so decisions taken on
// them are uniform across the grid.
template <int NV>
__device__ inline void reduce_partials_all(const double* partials, int stride, int count, double (&out)[NV]) {
    __shared__ double sred[NV][kBlock / kWave];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int i = threadIdx.x; i < count; i += kBlock) {
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += partials[(size_t)q * stride + i];
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) sred[q][wid] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double v = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) v += sred[q][k];
        out[q] = v;
    }
    __syncthreads();
}

// Atomic max of a non-negative double (the IEEE bit patterns of non-negative
// doubles order like unsigned integers).
__device__ inline void atomic_max_nonneg(double* addr, double v) {
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

inline int grid_for(int64_t n, int block = kBlock) { return (int)((n + block - 1) / block); }

// XCD-aware block order.  The 8 XCDs of the chip each have their own L2, and
// consecutive blocks are dispatched to the XCDs in turn.  Kernels whose rows
// gather from shared vectors map physical block b to the logical block
// (b % 8) * (G / 8) + b / 8, so each XCD walks one contiguous eighth of the
// rows and neighbouring rows (which share columns) hit the same L2.  The
// grid must be a multiple of 8 (xcd_grid); otherwise the order is unchanged.
constexpr int kXcd = 8;
inline int xcd_grid(int g) { return (std::max(g, 1) + kXcd - 1) / kXcd * kXcd; }
__device__ inline int xcd_block() {
    const int G = gridDim.x, b = blockIdx.x;
    if (G % kXcd) return b;
    return (b % kXcd) * (G / kXcd) + b / kXcd;
}

// Copies between a caller pointer (host or device) and a device buffer.
int copy_in(double* dst_dev, const double* src, int where, size_t count, hipStream_t s);
int copy_out(double* dst, int where, const double* src_dev, size_t count, hipStream_t s);

// ---------------------------------------------------------------------------
// Preconditioned CG on "symmetric matrix with identity rows" systems.
// ---------------------------------------------------------------------------
// Initial guesses from earlier solves (successive right-hand sides).
// A solver keeps its last kSolveWindow solved pairs (x_i, b_i), A x_i = b_i,
// i = 0 the newest.  A solve with right-hand side b starts from
//   x0 = x_0 + sum_j c_j dX_j,
// dB_j, dX_j the j-th backward differences of the kept b, x at the newest
// pair (dB_0 = b_0, dB_1 = b_0 - b_1, dB_2 = b_0 - 2 b_1 + b_2, ...) and c
// minimising ||(b - b_0) - sum_j c_j dB_j||: least squares over the span of
// the kept right-hand sides (the projection of P. F. Fischer, Comput. Methods
// Appl. Mech. Engrg. 163 (1998) 193, in the residual norm), where x_0 alone
// (c = 0) is the plain warm start.  Along a smooth trajectory the
// differences fall off geometrically; forming them element by element
// before any dot product keeps their Gram matrix well conditioned after
// diagonal scaling.  The Gram matrix of the raw b_i, whose condition number
// is the square of the b_i's, loses every direction below ~1e-8 of |b| to
// rounding.
constexpr int kSolveWindow = 6;
// dots per solve: the Gram matrix (packed upper triangle), dB_j . (b - b_0), |b - b_0|^2
constexpr int kWinGram = kSolveWindow * (kSolveWindow + 1) / 2;
constexpr int kWinDots = kWinGram + kSolveWindow + 1;

struct CgWindow {
    int head, cnt;             // slot the next pair goes to; pairs held
    double a[kSolveWindow];    // x0 = sum_i a_i x_i (newest first), from the last projection
};

// slot of the i-th newest pair in a ring of kSolveWindow slots
__host__ __device__ inline int win_slot(int head, int i) { return (head - 1 - i + 2 * kSolveWindow) % kSolveWindow; }

// in place: v[j] <- j-th backward difference of v[0..cnt) at v[0]
__host__ __device__ inline void win_differences(double (&v)[kSolveWindow], int cnt) {
#pragma unroll
    for (int j = 1; j < kSolveWindow; ++j)
#pragma unroll
        for (int i = kSolveWindow - 1; i >= j; --i)
            if (i < cnt) v[i] = v[i - 1] - v[i];
}

// acc += the products of one element: dB (differences, zero past cnt) and r = b - b_0
__host__ __device__ inline void win_accumulate(double (&acc)[kWinDots], const double (&dv)[kSolveWindow], double r) {
    int q = 0;
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i)
#pragma unroll
        for (int j = i; j < kSolveWindow; ++j) acc[q++] += dv[i] * dv[j];
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) acc[kWinGram + i] += dv[i] * r;
    acc[kWinDots - 1] += r * r;
}

__host__ __device__ inline int win_q(int i, int j) {  // packed index of G_ij, i <= j
    return i * kSolveWindow - i * (i - 1) / 2 + (j - i);
}

// From the summed dots t[kWinDots] of cnt >= 1 kept pairs: the coefficients
// a over the kept x (newest first).  The Gram system is scaled to a unit
// diagonal and factorised by Cholesky, dropping a direction whose pivot
// falls below 1e-12 (a difference nearly in the span of the lower ones, or
// zero).  Non-finite dots, or a residual estimate above |b - b_0|^2, give
// the plain warm start a = e_0.
__host__ __device__ inline void win_coefficients(int cnt, const double* t, double* a) {
    // every loop runs to the compile-time window size with the count as a
    // condition, so the small arrays stay in registers
    double sc[kSolveWindow], c[kSolveWindow], y[kSolveWindow];
    double L[kSolveWindow][kSolveWindow];
    bool keep[kSolveWindow];
    bool fin = true;
#pragma unroll
    for (int q = 0; q < kWinDots; ++q) fin = fin && std::isfinite(t[q]);
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) {
        const double g = t[win_q(i, i)];
        sc[i] = (i < cnt && g > 0.0) ? 1.0 / sqrt(g) : 0.0;
        c[i] = y[i] = 0.0;
        keep[i] = false;
#pragma unroll
        for (int j = 0; j < kSolveWindow; ++j) L[i][j] = 0.0;
    }
    if (fin) {
#pragma unroll
        for (int i = 0; i < kSolveWindow; ++i) {
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                if (i >= cnt || (j < i && !keep[j])) continue;
                double v = t[win_q(j, i)] * sc[i] * sc[j];
#pragma unroll
                for (int k = 0; k < j; ++k)
                    if (keep[k]) v -= L[i][k] * L[j][k];
                if (j < i) {
                    L[i][j] = v / L[j][j];
                } else {
                    keep[i] = sc[i] > 0.0 && v > 1e-12;
                    L[i][i] = keep[i] ? sqrt(v) : 1.0;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kSolveWindow; ++i) {
            if (!keep[i]) continue;
            double v = t[kWinGram + i] * sc[i];
#pragma unroll
            for (int k = 0; k < i; ++k)
                if (keep[k]) v -= L[i][k] * y[k];
            y[i] = v / L[i][i];
        }
#pragma unroll
        for (int i = kSolveWindow - 1; i >= 0; --i) {
            if (!keep[i]) continue;
            double v = y[i];
#pragma unroll
            for (int k = i + 1; k < kSolveWindow; ++k)
                if (keep[k]) v -= L[k][i] * c[k];
            c[i] = v / L[i][i];
        }
        double est = t[kWinDots - 1];
#pragma unroll
        for (int i = 0; i < kSolveWindow; ++i) {
            c[i] *= sc[i];
            fin = fin && std::isfinite(c[i]);
        }
#pragma unroll
        for (int i = 0; i < kSolveWindow; ++i) {
            est -= 2.0 * c[i] * t[kWinGram + i];
#pragma unroll
            for (int j = 0; j < kSolveWindow; ++j) est += c[i] * c[j] * t[i < j ? win_q(i, j) : win_q(j, i)];
        }
        fin = fin && est <= t[kWinDots - 1];
    }
    // dX_j = sum_{i <= j} (-1)^i binom(j, i) x_i
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) a[i] = i == 0 ? 1.0 : 0.0;
    if (!fin) return;
#pragma unroll
    for (int j = 0; j < kSolveWindow; ++j) {
        double bin = 1.0;  // binom(j, i)
#pragma unroll
        for (int i = 0; i <= j; ++i) {
            a[i] += ((i & 1) ? -bin : bin) * c[j];
            bin = bin * (j - i) / (i + 1);
        }
    }
}

struct CgState {
    double gam[2];   // (r, z) of iterations k (slot k & 1)
    double bb;       // (b, b)
    double tol2bb;   // tol^2 (b, b)
    double rr;       // (r, r) at exit
    int iters;       // completed iterations
    int done;        // 0 running, 1 converged, 2 breakdown / non-finite
};

// Smoothed-aggregation AMG V-cycle used as CG preconditioner (tm_amg.hip).
struct AmgPrecond {
    static constexpr int kCoarseMax = 512;
    static constexpr int kMaxLevels = 8;
    // CSR on the device, processed with `lpr` lanes per row.
    struct CsrDev {
        int n = 0;
        int lpr = 4;
        DevBuf<int> rp, ci;
        DevBuf<double> v;
    };
    struct Level {
        int n = 0;
        double omega = 0.0;
        int64_t nnz_a = 0, nnz_p = 0;
        // level 0 reuses the solver's SELL matrix (not owned)
        const SellPattern* A_ext = nullptr;
        const double* Aval_ext = nullptr;
        const double* dinv_ext = nullptr;
        CsrDev A, P, R;  // A used on levels >= 1
        DevBuf<double> dinv;
        DevBuf<double> x, x2, r, b;  // b: restricted right-hand side of the next level
    };
    std::vector<Level> levels;
    int nc = 0;
    DevBuf<double> Cinv, xc, bc;
    int setup(int n, const int* rowptr, const int* colidx, const double* vals, const std::vector<unsigned char>& fixed,
              const SellPattern* A0, const double* A0val, const double* A0dinv, hipStream_t s);
    // z = M^-1 r (one V-cycle); every kernel returns at once when st != NULL
    // and st->done is set (the CG it serves has stopped).  With rr_part (the
    // grid partials of (r, r) from the CG update) the first kernel ends the
    // CG when ||r||^2 <= st->tol2bb, so a converged CG skips the V-cycle.
    int apply(const double* r, double* z, double* part, int grid, hipStream_t s, CgState* st = nullptr,
              const double* rr_part = nullptr);
};

struct CgSolver {
    int n = 0;
    int grid = 0;            // blocks of every CG kernel (<= kMaxCgBlocks)
    SellPattern pat;
    DevBuf<double> val;      // matrix values in SELL order
    std::vector<double> h_val;  // matrix values in CSR order (pat.h_rowptr / h_csrcol)
    std::vector<double> h_dinv;
    std::vector<unsigned char> h_fixed;
    DevBuf<double> dinv;     // inverse diagonal (1 on fixed rows)
    DevBuf<unsigned char> fixed;  // 1 on identity rows
    DevBuf<double> r, z, q, p0, p1;
    DevBuf<double> part;     // partial sums: bb | rz | rr | pq, grid entries each
    DevBuf<CgState> st;
    CgState* h_st = nullptr;  // pinned mirror for the convergence check
    // the last kSolveWindow solved pairs (x, b) and the initial-guess state
    DevBuf<double> winX, winB, wpart;
    DevBuf<CgWindow> win;
    int precond = TM_PRECOND_JACOBI;
    AmgPrecond amg;
    int nfixed = 0;
    int last_iters = 0;
    // Coupling of the identity rows to the free unknowns, for the transposed
    // solve: row r (node fix_idx[r]) holds A_ij over free rows i that have a
    // column j = fix_idx[r].
    DevBuf<int> fix_idx, fix_rp, fix_ci;
    DevBuf<double> fix_v, tb;
    ~CgSolver();
    int setup(int n, const int* rowptr, const int* colidx, const double* vals, int precond, hipStream_t s);
    // Solves A x = b for device vectors b, x (x is the warm start).  tol is
    // relative to ||b||.  Returns TM_OK, TM_ERR_MAXITER or TM_ERR_NONFINITE.
    // project: start from the combination of the kept solutions whose
    // right-hand sides best match b (least-squares residual) instead of x,
    // and keep the new pair.
    int solve(const double* b, double* x, double tol, int maxit, hipStream_t s, int* iters, double* relres,
              bool project = false);
    // Enqueues a solve of exactly `budget` iterations with no host
    // synchronisation (capturable in a HIP graph); every kernel returns at
    // once after convergence.  At the end *fails is incremented on the device
    // if the tolerance was not met, and *imax keeps the largest count.
    int enqueue(const double* b, double* x, double tol, int budget, hipStream_t s, unsigned* fails, int* imax,
                bool project = false);
    void win_project(const double* b, double* x, hipStream_t s);
    void win_record(const double* x, hipStream_t s);  // after win_project with the same b
    // Solves A^T y = w for the same matrix (identity rows on the fixed set F):
    // y_free = A_ff^-1 w_free (A_ff symmetric), y_F = w_F - A_{free,F}^T y_free.
    int solve_transposed(const double* w, double* y, double tol, int maxit, hipStream_t s, int* iters,
                         double* relres);
};

constexpr int kMaxCgBlocks = 512;

// Read-only view of a field context's device data (tm_field.hip), used by
// the batched evaluator.
struct CtxView {
    int n;
    hipStream_t stream;
    const SellPattern* pat;
    const std::vector<int>* hop_rp[TM_NUM_OPS];
    const std::vector<int>* hop_col[TM_NUM_OPS];
    const std::vector<double>* hop_val[TM_NUM_OPS];
    bool hop_set[TM_NUM_OPS];
    const double *js, *invjs, *ku, *kaxis, *dmi_fac, *nvs;
    bool has_ku, has_dmi, has_idmi;
    int bnx, bld;
    const int* bnodes;
    const double* B;
    const CgSolver* cgN;
    const CgSolver* cgD;
    double tol;
    int maxit;
    double hext[3];
};
int ctx_view(tm_ctx* c, CtxView* v);
// The context's per-node applied field (3n, A/m) set by tm_ctx_set_zeeman,
// or null when only a uniform field is set.
const double* ctx_hext_field(tm_ctx* c);
// Looks up, once per device, the device address of the no-allocation
// boundary product's scratch array on the current device (tm_bem.hip).
int noalloc_scratch_init();

}  // namespace tmh

// Opaque ABI types.
struct tm_spmat {
    tmh::SellPattern pat;
    tmh::DevBuf<double> val;
    tmh::DevBuf<double> x, y;
    hipStream_t stream = nullptr;
    int rows = 0, cols = 0;
};

struct tm_solver {
    tmh::CgSolver cg;
    tmh::DevBuf<double> b, x, xt;  // xt: warm start of the transposed solves
    hipStream_t stream = nullptr;
};
