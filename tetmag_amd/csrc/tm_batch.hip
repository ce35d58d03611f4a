// tm_batch.hip -- many magnetisation samples on one mesh in one pass.
//
// The reference evaluates one magnetisation at a time (TheLLG, DemagField)
// and its hysteresis driver runs field values one after another
// (Hysteresis.cpp:43-169).  Here up to 64 samples share every operator read:
// vectors are stored node-major, sample-fastest ([node][sample]); each sparse
// product runs one wavefront per row with one lane per sample, so the row's
// column indices and values are wavefront-uniform loads read once for all
// samples and every gather is one contiguous access.  The boundary-element
// product becomes a GEMM (B times the b x S boundary potentials).
//
// Both demag solves use the same scheme per sample:
//   outer (FP64):  x_fixed = b_fixed;  r = b - A x  (FP64 operator and vectors)
//   inner (T):     A d = r by AMG-preconditioned CG in precision T, from d = 0
//   x += d, repeat until ||b - A x|| <= tol ||b|| for every sample.
// T = double solves to the final tolerance in one outer step (the FP64
// batched form); T = float runs the inner iterations in FP32 to a relative
// 1e-5 and recovers FP64 accuracy by iterative refinement (the mixed form).
// Converged samples are frozen (their step lengths are zero).
#define ROCBLAS_BETA_FEATURES_API  // rocblas_gemm_ex_get_solutions (the GEMM autotuner)
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "tm_internal.hpp"

namespace tmh {

namespace {

constexpr int kMaxS = 64;
constexpr int kRowsPerBlock = kBlock / kWave;  // one row per wavefront

struct DCsr {
    int n = 0;
    DevBuf<int> rp, ci;
    DevBuf<double> v;
    int upload(int rows, const int* rowptr, const int* col, const double* val, hipStream_t s) {
        n = rows;
        TM_CHECK(rp.upload(rowptr, (size_t)rows + 1, s));
        TM_CHECK(ci.upload(col, (size_t)rowptr[rows], s));
        TM_CHECK(v.upload(val, (size_t)rowptr[rows], s));
        TM_HIP(hipStreamSynchronize(s));
        return TM_OK;
    }
};

struct CsrArgs {
    int n;
    const int* __restrict__ rp;
    const int* __restrict__ ci;
    const double* __restrict__ v;
};
inline CsrArgs args(const DCsr& M) { return CsrArgs{M.n, M.rp.p, M.ci.p, M.v.p}; }
inline CsrArgs args(const AmgPrecond::CsrDev& M) { return CsrArgs{M.n, M.rp.p, M.ci.p, M.v.p}; }

// Per-sample CG state of one inner solve.
struct BState {
    double gam[kMaxS];    // (r, z)
    double alpha[kMaxS];
    double beta[kMaxS];
    double itol2[kMaxS];  // inner tolerance^2 (absolute, on ||r||^2)
    double rr[kMaxS];
    int active[kMaxS];
    int iters[kMaxS];
};

// Per-sample outer state.
struct OState {
    double bb[kMaxS];
    double rr[kMaxS];
    int done[kMaxS];
};

// Wavefront w of block b walks rows b*4 + w, then + 4*G, ...  (A contiguous
// range per XCD was measured slower for these kernels.)
#define B_ROW_LOOP(NROWS)                                                                              \
    const int lane = threadIdx.x & (kWave - 1);                                                        \
    const bool on = lane < S;                                                                          \
    const int lc = min(lane, S - 1); /* gather lane: in bounds for every lane, so loads need no mask */ \
    for (int row = __builtin_amdgcn_readfirstlane(blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave)); \
         row < (NROWS); row += gridDim.x * kRowsPerBlock)

// ---- pipelined row walk for the solver kernels ----------------------------
// A row's entries sit one per lane (lane k = entry k, one vector load for
// indices and one for values); gather addresses are read from those lanes
// with readlane (wave-uniform), so a chunk of kGr gathers issues back to back.
// The next row's pointers and entries are fetched before the current row's
// gathers, so a wavefront has both in flight.  Rows longer than 64 entries
// are walked in 64-entry segments.
constexpr int kGr = 16;

struct RowEnt {
    int q0, q1;  // row extent (wave-uniform)
    int c;       // column of entry q0 + lane (lanes past the end: 0)
    double v;    // value of entry q0 + lane (lanes past the end: 0)
};

__device__ inline double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// acc = sum_k v_k xf(c_k) over one segment held in e (len <= 64 entries),
// in entry order.
template <typename T, typename XF>
__device__ inline T seg_sum(const RowEnt& e, int len, XF&& xf) {
    T acc = 0;
    for (int k0 = 0; k0 < len; k0 += kGr) {
        T xv[kGr];
        double w[kGr];
#pragma unroll
        for (int k = 0; k < kGr; ++k) {
            const int kk = min(k0 + k, len - 1);
            const int c = __builtin_amdgcn_readlane(e.c, kk);
            w[k] = readlane_f64(e.v, kk);
            xv[k] = xf(c);
        }
#pragma unroll
        for (int k = 0; k < kGr; ++k) acc = k0 + k < len ? acc + (T)w[k] * xv[k] : acc;
    }
    return acc;
}

template <typename T, typename V, typename XF>
__device__ inline T row_sum(const CsrArgs& A, const V* __restrict__ vals, const RowEnt& e, int lane, XF&& xf) {
    const int len = e.q1 - e.q0;
    if (len <= kWave) return seg_sum<T>(e, len, xf);
    T acc = 0;
    for (int q = e.q0; q < e.q1; q += kWave) {  // long rows: 64-entry segments
        RowEnt s;
        s.q0 = q;
        s.q1 = min(e.q1, q + kWave);
        const bool in = q + lane < s.q1;
        s.c = in ? A.ci[q + lane] : 0;
        s.v = in ? (double)vals[q + lane] : 0.0;
        acc += seg_sum<T>(s, s.q1 - s.q0, xf);
    }
    return acc;
}

template <typename V>
__device__ inline void row_entries(const CsrArgs& A, const V* __restrict__ vals, int q0, int q1, int lane,
                                   RowEnt& e) {
    e.q0 = q0;
    e.q1 = q1;
    const int q = q0 + lane;
    const bool in = q < q1;
    e.c = in ? A.ci[q] : 0;
    e.v = in ? (double)vals[q] : 0.0;
}

// Two rows' sums with all of their gathers in flight together (the walk is
// bound by gather latency, so a wavefront keeps twice as many loads
// outstanding).  mid() runs after the gathers are issued and before they are
// consumed (the next rows' entry loads go there).  Rows of more than kGr
// entries fall back to one row at a time.  An empty row (len 0) gathers
// entry 0 of its RowEnt (column 0, a valid row) and adds nothing.
template <typename T, typename V, typename XF, typename MID>
__device__ inline void row_sum2(const CsrArgs& A, const V* __restrict__ vals, const RowEnt& ea, const RowEnt& eb,
                                int lane, XF&& xf, MID&& mid, T& sa, T& sb) {
    const int la = ea.q1 - ea.q0, lb = eb.q1 - eb.q0;
    if (la > kGr || lb > kGr) {
        mid();
        sa = row_sum<T>(A, vals, ea, lane, xf);
        sb = row_sum<T>(A, vals, eb, lane, xf);
        return;
    }
    T xa[kGr], xb[kGr];
#pragma unroll
    for (int k = 0; k < kGr; ++k) {
        xa[k] = xf(__builtin_amdgcn_readlane(ea.c, max(0, min(k, la - 1))));
        xb[k] = xf(__builtin_amdgcn_readlane(eb.c, max(0, min(k, lb - 1))));
    }
    mid();
    T acc_a = 0, acc_b = 0;
#pragma unroll
    for (int k = 0; k < kGr; ++k) {
        const double wa = readlane_f64(ea.v, max(0, min(k, la - 1)));
        const double wb = readlane_f64(eb.v, max(0, min(k, lb - 1)));
        acc_a = k < la ? acc_a + (T)wa * xa[k] : acc_a;
        acc_b = k < lb ? acc_b + (T)wb * xb[k] : acc_b;
    }
    sa = acc_a;
    sb = acc_b;
}

__device__ inline void row_empty(RowEnt& e) {
    e.q0 = e.q1 = 0;
    e.c = 0;
    e.v = 0.0;
}

// Row-pair loop: a wavefront takes rows (row, row + stride), then advances by
// 2 stride.  Per pair: ld(row) loads what the row's epilogue needs, the
// gathers of both rows are issued, then the entries of the next pair (from
// row pointers loaded during the previous pair) and the row pointers of the
// pair after (one vector load: lanes 0-3 hold rp[ra], rp[ra+1], rp[rb],
// rp[rb+1], read with readlane one pair later, so no wait on them stalls the
// gathers), then the sums are formed and epi(row, acc, ld(row)) finishes
// each row.
#define B_ROW_PIPE2(T_, A, VALS, NROWS, LD, XF, EPI)                                                   \
    {                                                                                                   \
        const int stride_ = gridDim.x * kRowsPerBlock;                                                  \
        const int n_ = (NROWS);                                                                         \
        int row = __builtin_amdgcn_readfirstlane(blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave)); \
        auto rp_pair_ = [&](int ra) {                                                                   \
            const int rr = lane < 2 ? ra : ra + stride_;                                                \
            return (lane < 4 && rr < n_) ? (A).rp[rr + (lane & 1)] : 0;                                 \
        };                                                                                              \
        RowEnt ca_, cb_, na_, nb_;                                                                      \
        int rpv_ = rp_pair_(row);                                                                       \
        if (row < n_) row_entries((A), (VALS), __builtin_amdgcn_readlane(rpv_, 0),                      \
                                  __builtin_amdgcn_readlane(rpv_, 1), lane, na_);                       \
        else row_empty(na_);                                                                            \
        if (row + stride_ < n_) row_entries((A), (VALS), __builtin_amdgcn_readlane(rpv_, 2),            \
                                            __builtin_amdgcn_readlane(rpv_, 3), lane, nb_);             \
        else row_empty(nb_);                                                                            \
        rpv_ = rp_pair_(row + 2 * stride_);                                                             \
        /* the first entries and pointers arrive here, so at the loop head   \
           no wait on them is left for the back edge to inherit */            \
        asm volatile("" ::"v"(na_.c), "v"(na_.v), "v"(nb_.c), "v"(nb_.v), "v"(rpv_));                  \
        for (; row < n_; row += 2 * stride_) {                                                          \
            ca_ = na_;                                                                                  \
            cb_ = nb_;                                                                                  \
            const bool hb_ = row + stride_ < n_;                                                        \
            const auto la_ = LD(row);                                                                   \
            const auto lb_ = LD(hb_ ? row + stride_ : row);                                             \
            auto mid_ = [&]() {                                                                         \
                if (row + 2 * stride_ < n_) row_entries((A), (VALS), __builtin_amdgcn_readlane(rpv_, 0), \
                                                        __builtin_amdgcn_readlane(rpv_, 1), lane, na_); \
                else row_empty(na_);                                                                    \
                if (row + 3 * stride_ < n_) row_entries((A), (VALS), __builtin_amdgcn_readlane(rpv_, 2), \
                                                        __builtin_amdgcn_readlane(rpv_, 3), lane, nb_); \
                else row_empty(nb_);                                                                    \
                rpv_ = rp_pair_(row + 4 * stride_);                                                     \
            };                                                                                          \
            T_ sa_, sb_;                                                                                \
            row_sum2<T_>((A), (VALS), ca_, cb_, lane, XF, mid_, sa_, sb_);                              \
            EPI(row, sa_, la_);                                                                         \
            if (hb_) EPI(row + stride_, sb_, lb_);                                                      \
        }                                                                                               \
    }

// Row loop with look-ahead: while row i is gathered, the entries of row i+1
// (whose pointers arrived during row i-1) and the pointers of row i+2 are in
// flight.  body(row, e) handles one row.
#define B_ROW_PIPE(A, VALS, NROWS, BODY)                                                                \
    {                                                                                                   \
        const int stride_ = gridDim.x * kRowsPerBlock;                                                  \
        int row = __builtin_amdgcn_readfirstlane(blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave)); \
        RowEnt cur_, nxt_;                                                                              \
        int p0_ = 0, p1_ = 0;                                                                           \
        if (row < (NROWS)) row_entries((A), (VALS), (A).rp[row], (A).rp[row + 1], lane, cur_);          \
        if (row + stride_ < (NROWS)) {                                                                  \
            p0_ = (A).rp[row + stride_];                                                                \
            p1_ = (A).rp[row + stride_ + 1];                                                            \
        }                                                                                               \
        for (; row < (NROWS); row += stride_) {                                                         \
            if (row + stride_ < (NROWS)) row_entries((A), (VALS), p0_, p1_, lane, nxt_);                \
            if (row + 2 * stride_ < (NROWS)) {                                                          \
                p0_ = (A).rp[row + 2 * stride_];                                                        \
                p1_ = (A).rp[row + 2 * stride_ + 1];                                                    \
            }                                                                                           \
            BODY(row, cur_);                                                                            \
            cur_ = nxt_;                                                                                \
        }                                                                                               \
    }

// Row entries are walked in chunks of kCh: the chunk's column indices first,
// then kCh independent gathers, then the multiply-adds in the original entry
// order (so sums are unchanged).  Slots past the row end repeat the row's
// first column (an in-bounds address) and are skipped in the sums.
constexpr int kCh = 8;

template <int CH>
__device__ inline void chunk_cols(const int* __restrict__ ci, int q, int q1, int (&c)[CH]) {
#pragma unroll
    for (int k = 0; k < CH; ++k) c[k] = ci[q + k < q1 ? q + k : q];
}

// The chunk's matrix values, loaded together with the columns (past the row
// end: the row's first value, masked out by the caller's select).
template <int CH, typename V>
__device__ inline void chunk_vals(const V* __restrict__ v, int q, int q1, double (&w)[CH]) {
#pragma unroll
    for (int k = 0; k < CH; ++k) w[k] = (double)v[q + k < q1 ? q + k : q];
}

// acc = sum_q v[q] x[ci[q]][lc] over the row (lanes >= S read lane S-1's
// column and their results are discarded)
template <typename T, typename V>
__device__ inline T row_dot(const CsrArgs& A, const V* __restrict__ v, int row, int S, int lc,
                            const T* __restrict__ x) {
    const int q0 = A.rp[row], q1 = A.rp[row + 1];
    T acc = 0;
    for (int q = q0; q < q1; q += kCh) {
        int c[kCh];
        double w[kCh];
        chunk_cols<kCh>(A.ci, q, q1, c);
        chunk_vals<kCh>(v, q, q1, w);
        T xv[kCh];
#pragma unroll
        for (int k = 0; k < kCh; ++k) xv[k] = x[(size_t)c[k] * S + lc];
#pragma unroll
        for (int k = 0; k < kCh; ++k) acc = q + k < q1 ? acc + (T)w[k] * xv[k] : acc;
    }
    return acc;
}

// Writes the block's per-sample partial sum v (lane = sample) to part[s*G + block].
__device__ inline void b_partial(double* part, int S, double v) {
    __shared__ double sp[kRowsPerBlock][kWave];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    sp[wv][lane] = v;
    __syncthreads();
    if (wv == 0 && lane < S) {
        double t = 0.0;
        for (int k = 0; k < kRowsPerBlock; ++k) t += sp[k][lane];
        part[(size_t)lane * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
}

// Sum of part[s*G .. s*G+G) for sample s = blockIdx.x, returned to thread 0.
__device__ inline double b_total(const double* part, int G) {
    __shared__ double sr[kBlock / kWave];
    const int s = blockIdx.x;
    double t = 0.0;
    for (int g = threadIdx.x; g < G; g += kBlock) t += part[(size_t)s * G + g];
    t = wave_sum(t);
    if ((threadIdx.x & (kWave - 1)) == 0) sr[threadIdx.x / kWave] = t;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / kWave; ++k) r += sr[k];
    return r;
}

// ---------------------------------------------------------------- outer --
__global__ __launch_bounds__(kBlock) void k_b_fix(int n, int S, const unsigned char* __restrict__ fixed,
                                                  const double* __restrict__ b, double* __restrict__ x) {
    const size_t tot = (size_t)n * S;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        if (fixed[e / S]) x[e] = b[e];
}

// r = b - A x (FP64), rT = (T) r, partials of (b,b) and (r,r).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_resid(CsrArgs A, int S, const double* __restrict__ b,
                                                    const double* __restrict__ x, double* __restrict__ r64,
                                                    T* __restrict__ rT, double* part_bb, double* part_rr) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    double vb = 0.0, vr = 0.0;
    auto body = [&](int row, const RowEnt& e) {
        const double bi = b[(size_t)row * S + lc];
        const double acc = row_sum<double>(A, A.v, e, lane, [&](int c) { return x[(size_t)c * S + lc]; });
        if (on) {
            const size_t ee = (size_t)row * S + lane;
            const double ri = bi - acc;
            r64[ee] = ri;
            rT[ee] = (T)ri;
            vb += bi * bi;
            vr += ri * ri;
        }
    };
    B_ROW_PIPE(A, A.v, A.n, body);
    b_partial(part_bb, S, vb);
    b_partial(part_rr, S, vr);
}

// outer decision per sample; inner tolerance and initial inner state.
__global__ void k_b_outer_red(int G, int S, const double* part_bb, const double* part_rr, double tol, double itol,
                              OState* os, BState* bs) {
    const int s = blockIdx.x;
    const double bb = b_total(part_bb, G);
    __syncthreads();
    const double rr = b_total(part_rr, G);
    if (threadIdx.x == 0 && s < S) {
        os->bb[s] = bb;
        os->rr[s] = rr;
        const bool done = !(rr > tol * tol * bb) || !std::isfinite(rr);
        os->done[s] = done ? 1 : 0;
        // inner target: the final tolerance, or a relative itol for the
        // refinement steps, whichever is looser
        const double a2 = tol * tol * bb;
        const double i2 = itol * itol * rr;
        bs->itol2[s] = std::max(a2, i2);
        bs->rr[s] = rr;
        bs->active[s] = done ? 0 : 1;
        bs->iters[s] = 0;
        bs->beta[s] = 0.0;
        bs->alpha[s] = 0.0;
        bs->gam[s] = 0.0;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_add(size_t tot, const T* __restrict__ d, double* __restrict__ x) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        x[e] += (double)d[e];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_zero(size_t tot, T* __restrict__ x) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) x[e] = (T)0;
}

// ------------------------------------------------------------- inner CG --
// p = z + beta p_old (p_old not read for k = 0); q = A p; partial (p, q).
// p = z + beta p_old (p_old not read for k = 0)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_dir(size_t NS, int S, const T* __restrict__ z,
                                                     const T* __restrict__ pold, T* __restrict__ pnew, int k,
                                                     const BState* bs) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < NS; e += (size_t)gridDim.x * kBlock) {
        const int sm = (int)(e % S);
        pnew[e] = k > 0 ? z[e] + (T)bs->beta[sm] * pold[e] : z[e];
    }
}

// q = A p; partial (p, q)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_spmv(CsrArgs A, int S, const T* __restrict__ p, T* __restrict__ q,
                                                      double* part) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    double v = 0.0;
    auto ld = [&](int row) { return p[(size_t)row * S + lc]; };
    auto xf = [&](int c) { return p[(size_t)c * S + lc]; };
    auto epi = [&](int row, T acc, T pi) {
        if (on) {
            q[(size_t)row * S + lane] = acc;
            v += (double)pi * (double)acc;
        }
    };
    B_ROW_PIPE2(T, A, A.v, A.n, ld, xf, epi);
    b_partial(part, S, v);
}

__global__ void k_b_red_pq(int G, int S, const double* part, BState* bs) {
    const int s = blockIdx.x;
    const double pq = b_total(part, G);
    if (threadIdx.x == 0 && s < S) {
        const bool ok = bs->active[s] && pq > 0.0 && std::isfinite(pq);
        if (bs->active[s] && !ok) bs->active[s] = 0;  // breakdown: freeze the sample
        bs->alpha[s] = ok ? bs->gam[s] / pq : 0.0;
    }
}

// d += alpha p; r -= alpha q; partial (r, r)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_update(int n, int S, const T* __restrict__ p,
                                                        const T* __restrict__ q, T* __restrict__ d,
                                                        T* __restrict__ r, const BState* bs, double* part) {
    const int ln = threadIdx.x & (kWave - 1);
    const T alpha = ln < S ? (T)bs->alpha[ln] : (T)0;
    double v = 0.0;
    B_ROW_LOOP(n) {
        if (on) {
            const size_t e = (size_t)row * S + lane;
            d[e] += alpha * p[e];
            const T ri = r[e] - alpha * q[e];
            r[e] = ri;
            v += (double)ri * (double)ri;
        }
    }
    b_partial(part, S, v);
}

// after z = M r: gamma, convergence, beta
__global__ void k_b_red_rz(int G, int S, const double* part_rz, const double* part_rr, int k, BState* bs) {
    const int s = blockIdx.x;
    const double rz = b_total(part_rz, G);
    __syncthreads();
    const double rr = k >= 0 ? b_total(part_rr, G) : 0.0;
    if (threadIdx.x == 0 && s < S) {
        if (k >= 0) bs->rr[s] = rr;  // k < 0: initial call, rr comes from the outer residual
        const double r2 = bs->rr[s];
        bool act = bs->active[s] && r2 > bs->itol2[s] && std::isfinite(rz) && std::isfinite(r2);
        if (k >= 0 && bs->active[s]) bs->iters[s] += 1;
        bs->beta[s] = (act && k >= 0 && bs->gam[s] != 0.0) ? rz / bs->gam[s] : 0.0;
        bs->gam[s] = rz;
        bs->active[s] = act ? 1 : 0;
    }
}

// --------------------------------------------------------- AMG V-cycle --
// aw[k] = v[k] w dinv[ci[k]]: the operator with the Jacobi step folded into
// its columns, A w D^-1
__global__ __launch_bounds__(kBlock) void k_b_scale_cols(size_t nnz, const int* __restrict__ ci,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ dinv, double w,
                                                         double* __restrict__ aw) {
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (size_t)gridDim.x * kBlock)
        aw[k] = v[k] * w * dinv[ci[k]];
}

// x = w dinv b; r = b - A x, with AW = A w D^-1 (k_b_scale_cols) in A.v
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_pre(CsrArgs AW, int S, const double* __restrict__ dinv,
                                                  const T* __restrict__ b, double w, T* __restrict__ x,
                                                  T* __restrict__ r) {
    const CsrArgs& A = AW;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    auto ld = [&](int row) { return b[(size_t)row * S + lc]; };
    auto xf = [&](int c) { return b[(size_t)c * S + lc]; };
    auto epi = [&](int row, T acc, T bi) {
        if (on) {
            const size_t ee = (size_t)row * S + lane;
            x[ee] = (T)(w * dinv[row]) * bi;
            r[ee] = bi - acc;
        }
    };
    B_ROW_PIPE2(T, A, A.v, A.n, ld, xf, epi);
}

// y = M x (ACC: y += M x)
template <typename T, bool ACC>
__global__ __launch_bounds__(kBlock) void k_b_spmm(CsrArgs M, int S, const T* __restrict__ x, T* __restrict__ y) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    auto ld = [&](int row) { return ACC ? y[(size_t)row * S + lc] : (T)0; };
    auto xf = [&](int c) { return x[(size_t)c * S + lc]; };
    auto epi = [&](int row, T acc, T yi) {
        if (on) y[(size_t)row * S + lane] = ACC ? yi + acc : acc;
    };
    B_ROW_PIPE2(T, M, M.v, M.n, ld, xf, epi);
}

// xo = x + w dinv (b - A x); optional partial (b, xo)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_post(CsrArgs A, int S, const double* __restrict__ dinv,
                                                   const T* __restrict__ b, double w, const T* __restrict__ x,
                                                   T* __restrict__ xo, double* part) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    double v = 0.0;
    struct BX {
        T b, x;
    };
    auto ld = [&](int row) {
        const size_t er = (size_t)row * S + lc;
        return BX{b[er], x[er]};
    };
    auto xf = [&](int c) { return x[(size_t)c * S + lc]; };
    auto epi = [&](int row, T acc, const BX& bx) {
        if (on) {
            const T zi = bx.x + (T)(w * dinv[row]) * (bx.b - acc);
            xo[(size_t)row * S + lane] = zi;
            v += (double)bx.b * (double)zi;
        }
    };
    B_ROW_PIPE2(T, A, A.v, A.n, ld, xf, epi);
    if (part) b_partial(part, S, v);
}

// x = C b for the dense coarse inverse (row-major nc x nc)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dense(int nc, int S, const double* __restrict__ C,
                                                    const T* __restrict__ b, T* __restrict__ x) {
    B_ROW_LOOP(nc) {
        double acc = 0.0;
        for (int k0 = 0; k0 < nc; k0 += kGr) {
            double cv[kGr], bv[kGr];
#pragma unroll
            for (int k = 0; k < kGr; ++k) {
                const int kk = min(k0 + k, nc - 1);
                cv[k] = C[(size_t)row * nc + kk];
                bv[k] = (double)b[(size_t)kk * S + lc];
            }
#pragma unroll
            for (int k = 0; k < kGr; ++k) acc = k0 + k < nc ? acc + cv[k] * bv[k] : acc;
        }
        if (on) x[(size_t)row * S + lane] = (T)acc;
    }
}

// ------------------------------------------------------ fields and LLG --
__global__ __launch_bounds__(kBlock) void k_b_div(CsrArgs tx, const double* __restrict__ ty,
                                                  const double* __restrict__ tz, int S, int n,
                                                  const double* __restrict__ js, const double* __restrict__ m,
                                                  double scale, double* __restrict__ out) {
    const size_t NS = (size_t)n * S;
    B_ROW_LOOP(n) {
        double sx = 0, sy = 0, sz = 0;
        const int q0 = tx.rp[row], q1 = tx.rp[row + 1];
        for (int q = q0; q < q1; q += 4) {
            int c[4];
            chunk_cols<4>(tx.ci, q, q1, c);
            double mv[4][3];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const size_t e = (size_t)c[k] * S + lc;
                mv[k][0] = m[e];
                mv[k][1] = m[NS + e];
                mv[k][2] = m[2 * NS + e];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (q + k < q1) {
                    const double j = js[c[k]];
                    sx += tx.v[q + k] * (mv[k][0] * j);
                    sy += ty[q + k] * (mv[k][1] * j);
                    sz += tz[q + k] * (mv[k][2] * j);
                }
        }
        if (on) out[(size_t)row * S + lane] = (sx + sy + sz) * scale;
    }
}

struct BEvalArgs {
    CsrArgs A;                         // pattern + exchange values (v may be null)
    const double* gx; const double* gy; const double* gz;
    const double* cx; const double* cy; const double* cz;
    int S, n;
    const double* m; const double* u1; const double* u2;
    const double* invjs; const double* ku; const double* kaxis;
    const double* dmi_fac; const double* nvs;
    const double* hext;   // [3][S] per-sample uniform applied field
    double alpha;
    int variant;
    double* dmdt; double* heff;
    double* torque;       // [S], zeroed before
};

// Fused Heff + LLG for every (node, sample): TheLLG.cpp:238-254 and
// EffFieldGPU.cu:316-362, with the demag field -grad(u1 + u2).
template <bool DMI>
__global__ __launch_bounds__(kBlock) void k_b_eval(BEvalArgs a) {
    const int S = a.S;
    const size_t NS = (size_t)a.n * S;
    const CsrArgs A = a.A;
    double tmax = 0.0;
    B_ROW_LOOP(a.n) {
        double ex = 0, ey = 0, ez = 0, dx = 0, dy = 0, dz = 0, cxs = 0, cys = 0, czs = 0;
        const int q0 = A.rp[row], q1 = A.rp[row + 1];
        for (int q = q0; q < q1; q += 4) {
            int c[4];
            chunk_cols<4>(A.ci, q, q1, c);
            double mv[4][3], uv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const size_t e = (size_t)c[k] * S + lc;
                mv[k][0] = a.m[e];
                mv[k][1] = a.m[NS + e];
                mv[k][2] = a.m[2 * NS + e];
                uv[k] = a.u1[e] + a.u2[e];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (q + k >= q1) continue;
                const double mx = mv[k][0], my = mv[k][1], mz = mv[k][2];
                if (A.v) {
                    const double xc = A.v[q + k];
                    ex += xc * mx;
                    ey += xc * my;
                    ez += xc * mz;
                }
                dx += a.gx[q + k] * uv[k];
                dy += a.gy[q + k] * uv[k];
                dz += a.gz[q + k] * uv[k];
                if (DMI) {
                    const double X = a.cx[q + k], Y = a.cy[q + k], Z = a.cz[q + k];
                    cxs += Y * mz - Z * my;
                    cys += Z * mx - X * mz;
                    czs += X * my - Y * mx;
                }
            }
        }
        if (!on) continue;
        const size_t e = (size_t)row * S + lane;
        const double mx = a.m[e], my = a.m[NS + e], mz = a.m[2 * NS + e];
        double hx = ex, hy = ey, hz = ez;
        if (a.ku) {
            const double kx = a.kaxis[row], ky = a.kaxis[a.n + row], kz = a.kaxis[2 * a.n + row];
            const double f = 2.0 * a.ku[row] * (kx * mx + ky * my + kz * mz);
            hx += f * kx;
            hy += f * ky;
            hz += f * kz;
        }
        if (DMI) {
            const double sx = a.nvs[row], sy = a.nvs[a.n + row], sz = a.nvs[2 * a.n + row];
            const double f = a.dmi_fac[row];
            hx += ((my * sz - mz * sy) - 2.0 * cxs) * f;
            hy += ((mz * sx - mx * sz) - 2.0 * cys) * f;
            hz += ((mx * sy - my * sx) - 2.0 * czs) * f;
        }
        const double ij = a.invjs ? a.invjs[row] : 1.0;
        hx = hx * ij + a.hext[lane] - dx;
        hy = hy * ij + a.hext[S + lane] - dy;
        hz = hz * ij + a.hext[2 * S + lane] - dz;
        if (a.heff) {
            a.heff[e] = hx;
            a.heff[NS + e] = hy;
            a.heff[2 * NS + e] = hz;
        }
        const double tx = my * hz - mz * hy, ty = mz * hx - mx * hz, tz = mx * hy - my * hx;
        const double bx = my * tz - mz * ty, by = mz * tx - mx * tz, bz = mx * ty - my * tx;
        if (a.variant == TM_LLG_NOPREC) {
            a.dmdt[e] = -a.alpha * bx;
            a.dmdt[NS + e] = -a.alpha * by;
            a.dmdt[2 * NS + e] = -a.alpha * bz;
        } else {
            a.dmdt[e] = -tx - a.alpha * bx;
            a.dmdt[NS + e] = -ty - a.alpha * by;
            a.dmdt[2 * NS + e] = -tz - a.alpha * bz;
        }
        tmax = fmax(tmax, sqrt(tx * tx + ty * ty + tz * tz));
    }
    if (!a.torque) return;
    __shared__ double st[kRowsPerBlock][kWave];
    const int ln = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    st[wv][ln] = tmax;
    __syncthreads();
    if (wv == 0 && ln < S) {
        double t = st[0][ln];
        for (int k = 1; k < kRowsPerBlock; ++k) t = fmax(t, st[k][ln]);
        atomic_max_nonneg(a.torque + ln, t);
    }
}

// U[j][s] = u[bn[j]][s]  (S x b column-major, ld S)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_gather(int b, int S, const int* __restrict__ bn,
                                                     const double* __restrict__ u, T* __restrict__ U) {
    const size_t tot = (size_t)b * S;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        U[e] = (T)u[bn ? (size_t)bn[e / S] * S + e % S : e];
}

// out[bn[j]][s] = G[j + s b]  (G = B U^T, b x S column-major); bn null: out[j][s]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_scatter(int b, int S, const int* __restrict__ bn,
                                                      const T* __restrict__ G, double* __restrict__ out) {
    const size_t tot = (size_t)b * S;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        out[(size_t)(bn ? bn[e % b] : e % b) * S + e / b] = (double)G[e];
}

// ---- FP64 MFMA GEMM for the boundary product ---------------------------
// P[ks] (b x S, row-major: P[ks][row][j]) = B[:, K_ks] (b x |K_ks|,
// column-major, ld) * U^T[K_ks, :] with U S x b column-major (ld S), S = 64
// columns.  Block: 4 wavefronts x 16 MT rows, all 64 columns, one K range ks
// of the split; U^T chunks of KGK rows are staged in LDS.
// v_mfma_f64_16x16x4f64 operand layout (measured): A lane l = A[l%16][l/16],
// B lane l = B[l/16][l%16], D lane l register r = C[4r + l/16][l%16].
// PIPE = false: per chunk, stage U^T, then load the chunk's B values and run
// its MFMAs.  PIPE = true: software-pipelined; while the MFMAs of chunk c
// run, the B values of chunk c+1 are in flight into registers and its U^T
// rows into registers, stored to the other of two LDS buffers afterwards
// (one barrier per chunk).
typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int KGK, int MT, bool PIPE>
__global__ __launch_bounds__(kBlock) void k_b_gemm64(int b, int ld, const double* __restrict__ B,
                                                     const double* __restrict__ U, int kper,
                                                     double* __restrict__ P) {
    constexpr int NB = PIPE ? 2 : 1;
    constexpr int UPT = KGK * 64 / kBlock;  // U^T values per thread per chunk
    static_assert(KGK % 4 == 0 && (KGK * 64) % kBlock == 0, "chunk shape");
    __shared__ double us[NB][KGK * 64];  // U^T chunk: us[buf][k * 64 + j]
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int m0 = (blockIdx.x * (kBlock / kWave) + wv) * 16 * MT;  // this wavefront's rows
    const int ks = blockIdx.y;
    const int kb = ks * kper, ke = min(b, kb + kper);
    const int nc = ke > kb ? (ke - kb + KGK - 1) / KGK : 0;  // chunks, uniform over the block
    const int li = lane & 15, lk = lane >> 4;
    dbl4 acc[MT][4];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[t][u] = dbl4{0.0, 0.0, 0.0, 0.0};
    int rr[MT];
    bool ok[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        rr[t] = m0 + 16 * t + li;  // the A row of this lane in tile t
        ok[t] = rr[t] < b;
    }
    double a[MT][KGK / 4], an[MT][KGK / 4], ur[UPT];
    auto load_b = [&](int k0, double (&dst)[MT][KGK / 4]) {
#pragma unroll
        for (int q = 0; q < KGK / 4; ++q) {
            const int kc = k0 + 4 * q + lk;  // this lane's K column
            const bool kin = kc < ke;
#pragma unroll
            for (int t = 0; t < MT; ++t) dst[t][q] = (ok[t] && kin) ? B[(size_t)kc * ld + rr[t]] : 0.0;
        }
    };
    auto load_u = [&](int k0) {
#pragma unroll
        for (int i = 0; i < UPT; ++i) {
            const int t = threadIdx.x + i * kBlock;
            const int kk = k0 + t / 64;
            ur[i] = kk < ke ? U[(size_t)kk * 64 + (t & 63)] : 0.0;
        }
    };
    auto store_u = [&](int buf) {
#pragma unroll
        for (int i = 0; i < UPT; ++i) us[buf][threadIdx.x + i * kBlock] = ur[i];
    };
    auto mfmas = [&](const double (&av)[MT][KGK / 4], int buf) {
#pragma unroll
        for (int q = 0; q < KGK / 4; ++q) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double bu = us[buf][(4 * q + lk) * 64 + u * 16 + li];
#pragma unroll
                for (int t = 0; t < MT; ++t)
                    acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][q], bu, acc[t][u], 0, 0, 0);
            }
        }
    };
    if constexpr (PIPE) {
        if (nc > 0) {
            load_u(kb);
            load_b(kb, an);
            store_u(0);
        }
        __syncthreads();
        for (int c = 0; c < nc; ++c) {
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int q = 0; q < KGK / 4; ++q) a[t][q] = an[t][q];
            const bool more = c + 1 < nc;
            if (more) {  // next chunk's loads in flight during this chunk's MFMAs
                load_u(kb + (c + 1) * KGK);
                load_b(kb + (c + 1) * KGK, an);
            }
            mfmas(a, c & 1);
            if (more) store_u((c + 1) & 1);  // that buffer was last read in chunk c-1
            __syncthreads();
        }
    } else {
        for (int c = 0; c < nc; ++c) {
            const int k0 = kb + c * KGK;
            __syncthreads();  // the previous chunk has been consumed
            load_u(k0);
            store_u(0);
            __syncthreads();
            load_b(k0, a);  // all of the chunk's B loads in flight before its MFMAs
            mfmas(a, 0);
        }
    }
    double* out = P + (size_t)ks * b * 64;
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + t * 16 + 4 * r + lk;
                if (row < b) out[(size_t)row * 64 + u * 16 + li] = acc[t][u][r];
            }
}

struct GemmShape {
    int kgk, mt;
    bool pipe;
};
constexpr GemmShape kGemmShapes[] = {{32, 2, false}, {64, 2, false}, {32, 2, true}, {16, 2, true},
                                     {16, 4, true},  {32, 4, true},  {8, 2, true},  {8, 4, true},
                                     {16, 1, true},  {32, 1, true}};
constexpr int kNumGemmShapes = sizeof(kGemmShapes) / sizeof(kGemmShapes[0]);

inline int gemm_rows_per_block(int shape) { return (kBlock / kWave) * 16 * kGemmShapes[shape].mt; }

template <class F>
void gemm_dispatch(int shape, F&& f) {
    switch (shape) {
        case 0: f(k_b_gemm64<32, 2, false>); break;
        case 1: f(k_b_gemm64<64, 2, false>); break;
        case 2: f(k_b_gemm64<32, 2, true>); break;
        case 3: f(k_b_gemm64<16, 2, true>); break;
        case 4: f(k_b_gemm64<16, 4, true>); break;
        case 5: f(k_b_gemm64<32, 4, true>); break;
        case 6: f(k_b_gemm64<8, 2, true>); break;
        case 7: f(k_b_gemm64<8, 4, true>); break;
        case 8: f(k_b_gemm64<16, 1, true>); break;
        default: f(k_b_gemm64<32, 1, true>); break;
    }
}

void launch_gemm64(int shape, dim3 g, hipStream_t s, int b, int ld, const double* B, const double* U, int kper,
                   double* P) {
    gemm_dispatch(shape, [&](auto k) { hipLaunchKernelGGL(k, g, dim3(kBlock), 0, s, b, ld, B, U, kper, P); });
}

// resident blocks per CU of a tile shape (0 if the runtime cannot say)
int gemm_blocks_per_cu(int shape) {
    int nb = 0;
    gemm_dispatch(shape, [&](auto k) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kBlock, 0) != hipSuccess) nb = 0;
    });
    return nb;
}

// out[bn[j]][s] = sum_ks P[ks][j][s] (FP64 split-K partials of k_b_gemm64,
// row-major); without bn: out[j][s]
__global__ __launch_bounds__(kBlock) void k_b_ksum_scatter(int b, int S, int nks, const int* __restrict__ bn,
                                                           const double* __restrict__ P, double* __restrict__ out) {
    const size_t tot = (size_t)b * S;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) {
        double acc = 0.0;
        for (int c = 0; c < nks; ++c) acc += P[(size_t)c * tot + e];
        out[bn ? (size_t)bn[e / S] * S + e % S : e] = acc;
    }
}

// out[bn[j]][s] = sum_c P[c][j + s b] in FP64 (the split-K FP32 partials);
// bn null: out[j][s]
__global__ __launch_bounds__(kBlock) void k_b_chunk_sum(int b, int S, int nch, const int* __restrict__ bn,
                                                        const float* __restrict__ P, double* __restrict__ out) {
    const size_t tot = (size_t)b * S;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) {
        double acc = 0.0;
        for (int c = 0; c < nch; ++c) acc += (double)P[(size_t)c * tot + e];
        out[(size_t)(bn ? bn[e % b] : e % b) * S + e / b] = acc;
    }
}

__global__ __launch_bounds__(kBlock) void k_b_to_float(size_t tot, const double* __restrict__ a,
                                                       float* __restrict__ o) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        o[e] = (float)a[e];
}

// m <- normalize(m + dt dm) per (node, sample)
__global__ __launch_bounds__(kBlock) void k_b_euler(size_t NS, double* __restrict__ m, const double* __restrict__ dm,
                                                    double dt) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < NS; e += (size_t)gridDim.x * kBlock) {
        const double x = m[e] + dt * dm[e], y = m[NS + e] + dt * dm[NS + e], z = m[2 * NS + e] + dt * dm[2 * NS + e];
        const double l = sqrt(x * x + y * y + z * z);
        const double f = l > 0.0 ? 1.0 / l : 1.0;
        m[e] = x * f;
        m[NS + e] = y * f;
        m[2 * NS + e] = z * f;
    }
}

__global__ void k_b_zero_small(int cnt, double* p) {
    if ((int)threadIdx.x < cnt) p[threadIdx.x] = 0.0;
}

// RK4 helpers over [3][n][S]: out = base + c k;  m = normalise(base + dt/6 (k1 + 2k2 + 2k3 + k4))
__global__ __launch_bounds__(kBlock) void k_b_axpy(size_t tot, const double* __restrict__ base,
                                                   const double* __restrict__ k, double c, double* __restrict__ out) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        out[e] = base[e] + c * k[e];
}

__global__ __launch_bounds__(kBlock) void k_b_rk_final(size_t NS, const double* __restrict__ base,
                                                       const double* __restrict__ k1, const double* __restrict__ k2,
                                                       const double* __restrict__ k3, const double* __restrict__ k4,
                                                       double dt, double* __restrict__ m) {
    const double f = dt / 6.0;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < NS; e += (size_t)gridDim.x * kBlock) {
        double v[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t q = (size_t)d * NS + e;
            v[d] = base[q] + f * (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]);
        }
        const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const double g = l > 0.0 ? 1.0 / l : 1.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) m[(size_t)d * NS + e] = v[d] * g;
    }
}

inline int elem_grid(size_t tot) { return (int)std::min<size_t>(4096, (tot + kBlock - 1) / kBlock); }
// a multiple of the XCD count (see B_ROW_LOOP)
inline int row_grid(int rows) {
    return xcd_grid(std::min(2048, (rows + kRowsPerBlock - 1) / kRowsPerBlock));
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched solver for one demag system, sharing the context solver's operator
// and AMG hierarchy.
template <typename T>
struct BatchSolver {
    const CgSolver* cg = nullptr;
    int n = 0, S = 0, G = 0;
    DCsr A;                       // level-0 operator (FP64 values)
    DevBuf<double> dinv;
    DevBuf<unsigned char> fixed;
    DevBuf<double> r64;
    DevBuf<T> rT, d, z, p0, p1, q;
    struct Lv {
        DevBuf<T> x, x2, r, b;
    };
    std::vector<Lv> lv;
    std::vector<DevBuf<double>> aw;  // per level: A w D^-1 values (pre-smoother)
    DevBuf<T> xc;
    DevBuf<double> part;          // 4 slots x S x G
    DevBuf<OState> ost;
    DevBuf<BState> bst;
    OState* h_ost = nullptr;
    BState* h_bst = nullptr;
    int last_inner = 8;

    ~BatchSolver() {
        if (h_ost) (void)hipHostFree(h_ost);
        if (h_bst) (void)hipHostFree(h_bst);
    }

    int setup(const CgSolver& solver, int S_, hipStream_t s) {
        cg = &solver;
        n = solver.n;
        S = S_;
        G = row_grid(n);
        TM_CHECK(A.upload(n, solver.pat.h_rowptr.data(), solver.pat.h_csrcol.data(), solver.h_val.data(), s));
        TM_CHECK(dinv.upload(solver.h_dinv.data(), n, s));
        TM_CHECK(fixed.upload(solver.h_fixed.data(), n, s));
        const size_t NS = (size_t)n * S;
        TM_CHECK(r64.alloc(NS));
        TM_CHECK(rT.alloc(NS));
        TM_CHECK(d.alloc(NS));
        TM_CHECK(z.alloc(NS));
        TM_CHECK(p0.alloc(NS));
        TM_CHECK(p1.alloc(NS));
        TM_CHECK(q.alloc(NS));
        lv.clear();
        lv.resize(solver.amg.levels.size());
        for (size_t l = 0; l < solver.amg.levels.size(); ++l) {
            const auto& L = solver.amg.levels[l];
            TM_CHECK(lv[l].x.alloc((size_t)L.n * S));
            TM_CHECK(lv[l].x2.alloc((size_t)L.n * S));
            TM_CHECK(lv[l].r.alloc((size_t)L.n * S));
            TM_CHECK(lv[l].b.alloc((size_t)L.R.n * S));
        }
        aw.clear();
        aw.resize(solver.amg.levels.size());
        for (size_t l = 0; l < solver.amg.levels.size(); ++l) {
            const auto& L = solver.amg.levels[l];
            const size_t nnz = l == 0 ? (size_t)solver.pat.h_rowptr[n] : (size_t)L.nnz_a;
            const int* ci = l == 0 ? A.ci.p : L.A.ci.p;
            const double* v = l == 0 ? A.v.p : L.A.v.p;
            const double* di = l == 0 ? dinv.p : L.dinv.p;
            TM_CHECK(aw[l].alloc(std::max<size_t>(nnz, 1)));
            if (nnz > 0) {
                hipLaunchKernelGGL(k_b_scale_cols, dim3(elem_grid(nnz)), dim3(kBlock), 0, s, nnz, ci, v, di,
                                   L.omega, aw[l].p);
                TM_HIP(hipGetLastError());
            }
        }
        TM_CHECK(xc.alloc((size_t)std::max(1, solver.amg.nc) * S));
        TM_CHECK(part.alloc((size_t)4 * S * G));
        TM_CHECK(ost.alloc(1));
        TM_CHECK(bst.alloc(1));
        if (!h_ost) TM_HIP(hipHostMalloc((void**)&h_ost, sizeof(OState), hipHostMallocDefault));
        if (!h_bst) TM_HIP(hipHostMalloc((void**)&h_bst, sizeof(BState), hipHostMallocDefault));
        TM_HIP(hipStreamSynchronize(s));
        return TM_OK;
    }

    // z = M r (AMG V-cycle, batched), partial (r, z) into prz
    int precond(const T* r, T* zout, double* prz, hipStream_t s) {
        const AmgPrecond& M = cg->amg;
        const int L = (int)M.levels.size();
        if (L == 0) {
            hipLaunchKernelGGL((k_b_dense<T>), dim3(row_grid(M.nc)), dim3(kBlock), 0, s, M.nc, S, M.Cinv.p, r, zout);
            // (r, z) partial via a post-sweep with w = 0 on the identity: reuse k_b_cg_update-free path
            return TM_ERR_UNSUPPORTED;
        }
        const T* b = r;
        for (int l = 0; l < L; ++l) {
            const auto& V = M.levels[l];
            if (l == 0)
                hipLaunchKernelGGL((k_b_pre<T>), dim3(G), dim3(kBlock), 0, s, CsrArgs{A.n, A.rp.p, A.ci.p, aw[0].p},
                                   S, dinv.p, b, V.omega, lv[0].x.p, lv[0].r.p);
            else
                hipLaunchKernelGGL((k_b_pre<T>), dim3(row_grid(V.n)), dim3(kBlock), 0, s,
                                   CsrArgs{V.A.n, V.A.rp.p, V.A.ci.p, aw[l].p}, S, V.dinv.p, b, V.omega, lv[l].x.p,
                                   lv[l].r.p);
            hipLaunchKernelGGL((k_b_spmm<T, false>), dim3(row_grid(V.R.n)), dim3(kBlock), 0, s, args(V.R), S,
                               (const T*)lv[l].r.p, lv[l].b.p);
            b = lv[l].b.p;
        }
        hipLaunchKernelGGL((k_b_dense<T>), dim3(row_grid(M.nc)), dim3(kBlock), 0, s, M.nc, S, M.Cinv.p, b, xc.p);
        const T* xcoarse = xc.p;
        for (int l = L - 1; l >= 0; --l) {
            const auto& V = M.levels[l];
            const T* bl = l == 0 ? r : lv[l - 1].b.p;
            hipLaunchKernelGGL((k_b_spmm<T, true>), dim3(row_grid(V.P.n)), dim3(kBlock), 0, s, args(V.P), S, xcoarse,
                               lv[l].x.p);
            if (l == 0) {
                hipLaunchKernelGGL((k_b_post<T>), dim3(G), dim3(kBlock), 0, s, args(A), S, dinv.p, bl, V.omega,
                                   (const T*)lv[0].x.p, zout, prz);
                xcoarse = zout;
            } else {
                hipLaunchKernelGGL((k_b_post<T>), dim3(row_grid(V.n)), dim3(kBlock), 0, s, args(V.A), S, V.dinv.p, bl,
                                   V.omega, (const T*)lv[l].x.p, lv[l].x2.p, (double*)nullptr);
                xcoarse = lv[l].x2.p;
            }
        }
        TM_HIP(hipGetLastError());
        return TM_OK;
    }

    // Solves A x = b for all samples ([n][S] FP64); x is the warm start.
    int solve(const double* b, double* x, double tol, int maxit, hipStream_t s, int* outer_out, int* inner_out,
              double* maxrel) {
        if (cg->precond != TM_PRECOND_AMG || cg->amg.levels.empty()) return TM_ERR_UNSUPPORTED;
        const size_t NS = (size_t)n * S;
        double* pbb = part.p;
        double* prr = part.p + (size_t)S * G;
        double* prz = part.p + (size_t)2 * S * G;
        double* ppq = part.p + (size_t)3 * S * G;
        const double itol = sizeof(T) == 4 ? 1e-5 : 0.0;  // FP32 inner target per refinement step
        int outer = 0, inner_total = 0;
        hipLaunchKernelGGL(k_b_fix, dim3(elem_grid(NS)), dim3(kBlock), 0, s, n, S, fixed.p, b, x);
        for (;;) {
            hipLaunchKernelGGL((k_b_resid<T>), dim3(G), dim3(kBlock), 0, s, args(A), S, b, (const double*)x, r64.p,
                               rT.p, pbb, prr);
            hipLaunchKernelGGL(k_b_outer_red, dim3(S), dim3(kBlock), 0, s, G, S, pbb, prr, tol, itol, ost.p, bst.p);
            TM_HIP(hipMemcpyAsync(h_ost, ost.p, sizeof(OState), hipMemcpyDeviceToHost, s));
            TM_HIP(hipStreamSynchronize(s));
            bool all = true;
            double worst = 0.0;
            for (int k = 0; k < S; ++k) {
                all = all && h_ost->done[k];
                if (h_ost->bb[k] > 0) worst = std::max(worst, std::sqrt(h_ost->rr[k] / h_ost->bb[k]));
            }
            if (maxrel) *maxrel = worst;
            if (all || outer >= 8 || inner_total >= maxit) break;
            ++outer;
            // inner solve A d = rT from d = 0
            hipLaunchKernelGGL((k_b_zero<T>), dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, d.p);
            TM_CHECK(precond(rT.p, z.p, prz, s));
            hipLaunchKernelGGL(k_b_red_rz, dim3(S), dim3(kBlock), 0, s, G, S, prz, prr, -1, bst.p);
            int enq = 0;
            int chunk = std::max(4, last_inner);
            T* pc = p0.p;
            T* po = p1.p;
            for (;;) {
                for (int c = 0; c < chunk && enq < maxit; ++c, ++enq) {
                    hipLaunchKernelGGL((k_b_cg_dir<T>), dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, S,
                                       (const T*)z.p, (const T*)po, pc, enq, (const BState*)bst.p);
                    hipLaunchKernelGGL((k_b_cg_spmv<T>), dim3(G), dim3(kBlock), 0, s, args(A), S, (const T*)pc, q.p,
                                       ppq);
                    hipLaunchKernelGGL(k_b_red_pq, dim3(S), dim3(kBlock), 0, s, G, S, ppq, bst.p);
                    hipLaunchKernelGGL((k_b_cg_update<T>), dim3(G), dim3(kBlock), 0, s, n, S, (const T*)pc,
                                       (const T*)q.p, d.p, rT.p, (const BState*)bst.p, prr);
                    TM_CHECK(precond(rT.p, z.p, prz, s));
                    hipLaunchKernelGGL(k_b_red_rz, dim3(S), dim3(kBlock), 0, s, G, S, prz, prr, enq, bst.p);
                    std::swap(pc, po);
                }
                TM_HIP(hipGetLastError());
                TM_HIP(hipMemcpyAsync(h_bst, bst.p, sizeof(BState), hipMemcpyDeviceToHost, s));
                TM_HIP(hipStreamSynchronize(s));
                bool any = false;
                for (int k = 0; k < S; ++k) any = any || h_bst->active[k];
                if (!any || enq >= maxit) break;
                chunk = 2;
            }
            last_inner = std::max(4, enq - 1);
            inner_total += enq;
            hipLaunchKernelGGL((k_b_add<T>), dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, (const T*)d.p, x);
        }
        if (outer_out) *outer_out = outer;
        if (inner_out) *inner_out = inner_total;
        TM_HIP(hipGetLastError());
        bool all = true;
        for (int k = 0; k < S; ++k) all = all && h_ost->done[k];
        return all ? TM_OK : TM_ERR_MAXITER;
    }
};

}  // namespace tmh

using namespace tmh;

// tm_ctx internals needed here (defined in tm_field.hip)
namespace tmh {
}  // namespace tmh

struct tm_batch {
    int n = 0, S = 0, prec = TM_PREC_FP64;
    tm_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    CtxView cv{};
    DCsr ops;                          // union pattern with exchange values
    DevBuf<double> gx, gy, gz, tx, ty, tz, cx, cy, cz;
    bool exch = false, dmi = false;
    DevBuf<double> m, dmdt, heff, u1, u2, rhsN, rhsD, U, Gm, hext, torque;
    DevBuf<float> Bf, Uf, Gf;          // FP32 boundary product (TM_PREC_MIXED_BEM32)
    DevBuf<float> Pf;                  // its split-K partial products (nchunk x b x S)
    int kchunk = 512, nchunk = 0;
    DevBuf<double> rk_k[4], rk_base;   // RK4 stages and the step's start state
    int gemm_sol = 0;                  // tuned rocBLAS solution index (0: library default)
    bool gemm_own = false;             // the MFMA kernel k_b_gemm64 won the tuning
    int gemm_shape = 0, gemm_ks = 1;   // its tile shape and K split
    int gemm_own_candidates = 0;       // (shape, K split) pairs timed
    DevBuf<double> Pd;                 // its split-K partials (ks x b x S)
    double gemm_own_err = 0.0;         // its largest relative difference to rocBLAS at tuning
    int gemm_candidates = 0;
    float gemm_ms_default = 0.0f, gemm_ms_best = 0.0f;
    BatchSolver<double> sN64, sD64;
    BatchSolver<float> sN32, sD32;
    bool timers = false;
    hipEvent_t ev[6] = {};
    tm_phase_times tsum = {};
};

namespace {

// CSR values of an operator on the union pattern (pattern entries it lacks are 0).
int csr_on_pattern(const SellPattern& P, const std::vector<int>& rp, const std::vector<int>& col,
                   const std::vector<double>& val, std::vector<double>& out) {
    out.assign(P.h_csrcol.size(), 0.0);
    for (int i = 0; i < P.n; ++i) {
        int k = P.h_rowptr[i];
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            while (k < P.h_rowptr[i + 1] && P.h_csrcol[k] < col[q]) ++k;
            if (k >= P.h_rowptr[i + 1] || P.h_csrcol[k] != col[q]) return TM_ERR_BAD_ARG;
            out[k] += val[q];
        }
    }
    return TM_OK;
}

int upload_op(tm_batch* bt, int op, DevBuf<double>& dst) {
    std::vector<double> v;
    TM_CHECK(csr_on_pattern(*bt->cv.pat, *bt->cv.hop_rp[op], *bt->cv.hop_col[op], *bt->cv.hop_val[op], v));
    TM_CHECK(dst.upload(v.data(), v.size(), bt->stream));
    TM_HIP(hipStreamSynchronize(bt->stream));
    return TM_OK;
}

}  // namespace

namespace {

// The boundary GEMM G (b x S) = B (b x b) * U^T (S x b) is a tall, thin
// product; the library's default tile wastes half of a 128-wide tile on
// S = 64.  At batch creation every solution rocBLAS offers for this exact
// shape and layout is timed (HIP events, one warm-up and two timed runs) and
// the fastest is kept.
// FP32 products over K chunks of kchunk columns of B (strided-batched GEMM,
// FP32 accumulation inside a chunk only), partials left in Pf.
int gemm_split_f32(tm_batch* bt) {
    const CtxView& v = bt->cv;
    const int b = v.bnx, S = bt->S, kc = bt->kchunk;
    const int nfull = b / kc, tail = b - nfull * kc;
    const float one = 1.0f, zero = 0.0f;
    if (nfull > 0 &&
        rocblas_sgemm_strided_batched(bt->blas, rocblas_operation_none, rocblas_operation_transpose, b, S, kc, &one,
                                      bt->Bf.p, v.bld, (rocblas_stride)kc * v.bld, bt->Uf.p, S, (rocblas_stride)kc * S,
                                      &zero, bt->Pf.p, b, (rocblas_stride)b * S, nfull) != rocblas_status_success)
        return TM_ERR_HIP;
    if (tail > 0 &&
        rocblas_sgemm(bt->blas, rocblas_operation_none, rocblas_operation_transpose, b, S, tail, &one,
                      bt->Bf.p + (size_t)nfull * kc * v.bld, v.bld, bt->Uf.p + (size_t)nfull * kc * S, S, &zero,
                      bt->Pf.p + (size_t)nfull * b * S, b) != rocblas_status_success)
        return TM_ERR_HIP;
    return TM_OK;
}

int gemm_run(tm_batch* bt, int sol) {
    const CtxView& v = bt->cv;
    const int b = v.bnx, S = bt->S;
    const bool f32 = bt->prec == TM_PREC_MIXED_BEM32;
    if (f32) return gemm_split_f32(bt);
    const rocblas_datatype t = f32 ? rocblas_datatype_f32_r : rocblas_datatype_f64_r;
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    const void* al = f32 ? (const void*)&onef : (const void*)&one;
    const void* be = f32 ? (const void*)&zerof : (const void*)&zero;
    const void* A = f32 ? (const void*)bt->Bf.p : (const void*)v.B;
    const void* U = f32 ? (const void*)bt->Uf.p : (const void*)bt->U.p;
    void* G = f32 ? (void*)bt->Gf.p : (void*)bt->Gm.p;
    const rocblas_status r =
        rocblas_gemm_ex(bt->blas, rocblas_operation_none, rocblas_operation_transpose, b, S, b, al, A, t, v.bld, U, t,
                        S, be, G, t, b, G, t, b, t, sol > 0 ? rocblas_gemm_algo_solution_index : rocblas_gemm_algo_standard,
                        sol, 0);
    return r == rocblas_status_success ? TM_OK : TM_ERR_HIP;
}

__global__ __launch_bounds__(kBlock) void k_b_fill_pattern(size_t n, double* __restrict__ x) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (size_t)gridDim.x * kBlock)
        x[e] = sin(0.37 * (double)(e % 100003)) + 0.25;
}

// K splits worth timing for a tile shape: the one that gives ~1024 blocks,
// and those that fill j = 1..6 rounds of the resident block slots (all
// blocks carry the same work, so a last round that is only partly filled
// leaves CUs idle).  At most 16 partial products.
std::vector<int> gemm_ks_candidates(int b, int shape, int ncu) {
    const int rows = gemm_rows_per_block(shape);
    const int grid_m = (b + rows - 1) / rows;
    std::vector<int> ks = {std::max(1, std::min(16, (1024 + grid_m - 1) / grid_m))};
    const int slots = ncu * gemm_blocks_per_cu(shape);
    for (int j = 1; j <= 6 && slots > 0; ++j) {
        const int k = std::max(1, std::min(16, j * slots / grid_m));
        if (std::find(ks.begin(), ks.end(), k) == ks.end()) ks.push_back(k);
    }
    return ks;
}

int gemm_own_run(tm_batch* bt, double* G) {
    const CtxView& v = bt->cv;
    const int b = v.bnx;
    const int rows = gemm_rows_per_block(bt->gemm_shape);
    const int kper = (b + bt->gemm_ks - 1) / bt->gemm_ks;
    launch_gemm64(bt->gemm_shape, dim3((b + rows - 1) / rows, bt->gemm_ks), bt->stream, b, v.bld, v.B,
                  (const double*)bt->U.p, kper, bt->Pd.p);
    TM_HIP(hipGetLastError());
    if (G) {  // sum the partials into G (b x S, row-major) without scatter
        hipLaunchKernelGGL(k_b_ksum_scatter, dim3(elem_grid((size_t)b * bt->S)), dim3(kBlock), 0, bt->stream, b,
                           bt->S, bt->gemm_ks, (const int*)nullptr, (const double*)bt->Pd.p, G);
        TM_HIP(hipGetLastError());
    }
    return TM_OK;
}

int tune_gemm(tm_batch* bt) {
    const CtxView& v = bt->cv;
    const int b = v.bnx, S = bt->S;
    const bool f32 = bt->prec == TM_PREC_MIXED_BEM32;
    const rocblas_datatype t = f32 ? rocblas_datatype_f32_r : rocblas_datatype_f64_r;
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    const void* al = f32 ? (const void*)&onef : (const void*)&one;
    const void* be = f32 ? (const void*)&zerof : (const void*)&zero;
    const void* A = f32 ? (const void*)bt->Bf.p : (const void*)v.B;
    void* U = f32 ? (void*)bt->Uf.p : (void*)bt->U.p;
    void* G = f32 ? (void*)bt->Gf.p : (void*)bt->Gm.p;
    TM_HIP(hipMemsetAsync(U, 0, (size_t)b * S * (f32 ? sizeof(float) : sizeof(double)), bt->stream));
    if (f32) return TM_OK;  // split-K FP32 product: library defaults
    rocblas_int nsol = 0;
    if (rocblas_gemm_ex_get_solutions(bt->blas, rocblas_operation_none, rocblas_operation_transpose, b, S, b, al, A,
                                      t, v.bld, U, t, S, be, G, t, b, G, t, b, t, rocblas_gemm_algo_solution_index, 0,
                                      nullptr, &nsol) != rocblas_status_success ||
        nsol <= 0)
        return TM_OK;  // keep the default
    std::vector<rocblas_int> sols((size_t)nsol);
    if (rocblas_gemm_ex_get_solutions(bt->blas, rocblas_operation_none, rocblas_operation_transpose, b, S, b, al, A,
                                      t, v.bld, U, t, S, be, G, t, b, G, t, b, t, rocblas_gemm_algo_solution_index, 0,
                                      sols.data(), &nsol) != rocblas_status_success)
        return TM_OK;
    hipEvent_t e0, e1;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    auto time_sol = [&](int sol, float* ms) -> bool {
        if (gemm_run(bt, sol) != TM_OK) return false;
        if (hipEventRecord(e0, bt->stream) != hipSuccess) return false;
        for (int k = 0; k < 2; ++k)
            if (gemm_run(bt, sol) != TM_OK) return false;
        if (hipEventRecord(e1, bt->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return false;
        float t2 = 0.0f;
        (void)hipEventElapsedTime(&t2, e0, e1);
        *ms = 0.5f * t2;
        return true;
    };
    float best = 0.0f;
    if (!time_sol(0, &best)) best = 1e30f;
    bt->gemm_ms_default = best;
    bt->gemm_sol = 0;
    const int cap = std::min<int>((int)sols.size(), 64);
    for (int k = 0; k < cap; ++k) {
        float ms = 0.0f;
        if (time_sol(sols[k], &ms) && ms < best) {
            best = ms;
            bt->gemm_sol = sols[k];
        }
    }
    (void)hipGetLastError();  // a rejected candidate leaves no error behind
    bt->gemm_candidates = cap;
    // the MFMA kernel, each tile shape and K split: checked against the
    // library on a non-zero U, then timed
    if (!f32 && S == 64) {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 0;
        std::vector<std::pair<int, int>> cand;  // (shape, ks)
        int ksmax = 1;
        for (int sh = 0; sh < kNumGemmShapes; ++sh)
            for (int k : gemm_ks_candidates(b, sh, ncu)) {
                cand.push_back({sh, k});
                ksmax = std::max(ksmax, k);
            }
        if (bt->Pd.alloc((size_t)ksmax * b * S) == TM_OK) {
            hipLaunchKernelGGL(k_b_fill_pattern, dim3(elem_grid((size_t)b * S)), dim3(kBlock), 0, bt->stream,
                               (size_t)b * S, bt->U.p);
            std::vector<double> ref((size_t)b * S), own((size_t)b * S);
            DevBuf<double> G2;
            bool ok = G2.alloc((size_t)b * S) == TM_OK && gemm_run(bt, bt->gemm_sol) == TM_OK &&
                      hipMemcpyAsync(ref.data(), bt->Gm.p, ref.size() * sizeof(double), hipMemcpyDeviceToHost,
                                     bt->stream) == hipSuccess &&
                      hipStreamSynchronize(bt->stream) == hipSuccess;
            double rmax = 0.0;
            for (size_t e = 0; ok && e < ref.size(); ++e) rmax = std::max(rmax, std::fabs(ref[e]));
            const int keep_shape = bt->gemm_shape, keep_ks = bt->gemm_ks;
            int best_shape = -1, best_ks = 1;
            for (size_t ci = 0; ok && ci < cand.size(); ++ci) {
                bt->gemm_shape = cand[ci].first;
                bt->gemm_ks = cand[ci].second;
                if (gemm_own_run(bt, G2.p) != TM_OK ||
                    hipMemcpyAsync(own.data(), G2.p, own.size() * sizeof(double), hipMemcpyDeviceToHost,
                                   bt->stream) != hipSuccess ||
                    hipStreamSynchronize(bt->stream) != hipSuccess)
                    break;
                double dmax = 0.0;
                for (size_t e = 0; e < ref.size(); ++e)  // own: b x S row-major; ref: column-major
                    dmax = std::max(dmax, std::fabs(own[e] - ref[(e % S) * b + e / S]));
                const double err = rmax > 0 ? dmax / rmax : dmax;
                bt->gemm_own_err = std::max(bt->gemm_own_err, err);
                if (err > 1e-12) continue;  // never choose a kernel that disagrees with the library
                float ms = 0.0f;
                bool run_ok = hipEventRecord(e0, bt->stream) == hipSuccess;
                for (int k = 0; k < 2 && run_ok; ++k) run_ok = gemm_own_run(bt, nullptr) == TM_OK;
                if (run_ok && hipEventRecord(e1, bt->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess) {
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    ms *= 0.5f;
                    if (ms < best) {
                        best = ms;
                        best_shape = bt->gemm_shape;
                        best_ks = bt->gemm_ks;
                    }
                }
            }
            bt->gemm_own_candidates = (int)cand.size();
            if (best_shape >= 0) {
                bt->gemm_own = true;
                bt->gemm_shape = best_shape;
                bt->gemm_ks = best_ks;
            } else {
                bt->gemm_shape = keep_shape;
                bt->gemm_ks = keep_ks;
            }
            TM_HIP(hipMemsetAsync(bt->U.p, 0, (size_t)b * S * sizeof(double), bt->stream));
        }
    }
    bt->gemm_ms_best = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return TM_OK;
}

// out[bn[j]][s] = sum_k B[j][k] u[bn[k]][s] for every sample (bn null: the
// boundary values themselves, out[j][s] and u[k][s]), in the batch's
// precision (DemagField.cpp:145-147 for one vector).  Column-major G (b x S)
// = B (b x b, streamed in its storage order) * U^T, U = gathered boundary
// values (S x b).  FP64, except for TM_PREC_MIXED_BEM32: an FP32 copy of B,
// FP32 products accumulated in FP32 over 512-column chunks of B only, the
// chunk results summed in FP64 (FP32 accumulation over all 41k columns costs
// ~1e-5 relative field error).
int boundary_product(tm_batch* bt, const int* bn, const double* u, double* out) {
    const CtxView& v = bt->cv;
    const int b = v.bnx, S = bt->S;
    hipStream_t s = bt->stream;
    const dim3 g(elem_grid((size_t)b * S));
    if (bt->prec == TM_PREC_MIXED_BEM32) {
        hipLaunchKernelGGL(k_b_gather<float>, g, dim3(kBlock), 0, s, b, S, bn, u, bt->Uf.p);
        TM_CHECK(gemm_split_f32(bt));
        hipLaunchKernelGGL(k_b_chunk_sum, g, dim3(kBlock), 0, s, b, S, bt->nchunk, bn, (const float*)bt->Pf.p, out);
    } else {
        hipLaunchKernelGGL(k_b_gather<double>, g, dim3(kBlock), 0, s, b, S, bn, u, bt->U.p);
        if (bt->gemm_own) {
            TM_CHECK(gemm_own_run(bt, nullptr));
            hipLaunchKernelGGL(k_b_ksum_scatter, g, dim3(kBlock), 0, s, b, S, bt->gemm_ks, bn,
                               (const double*)bt->Pd.p, out);
        } else {
            TM_CHECK(gemm_run(bt, bt->gemm_sol));
            hipLaunchKernelGGL(k_b_scatter<double>, g, dim3(kBlock), 0, s, b, S, bn, (const double*)bt->Gm.p, out);
        }
    }
    TM_HIP(hipGetLastError());
    return TM_OK;
}

}  // namespace

extern "C" {

int tm_batch_create(tm_ctx* c, int nsamples, int precision, tm_batch** out) {
    if (!c || !out || nsamples < 1 || nsamples > kMaxS) return TM_ERR_BAD_ARG;
    if (precision != TM_PREC_FP64 && precision != TM_PREC_MIXED && precision != TM_PREC_MIXED_BEM32)
        return TM_ERR_BAD_ARG;
    tm_batch* bt = new tm_batch();
    int st = ctx_view(c, &bt->cv);
    if (st != TM_OK) {
        delete bt;
        return st;
    }
    const CtxView& v = bt->cv;
    if (!v.cgN || !v.cgD || !v.B || !v.js || !v.hop_set[TM_OP_GRAD_X] || !v.hop_set[TM_OP_TGRAD_X]) {
        delete bt;
        return TM_ERR_NOT_READY;
    }
    // the batched kernels cover exchange, uniaxial, bulk DMI, Zeeman and
    // demag; interfacial DMI contexts are evaluated per sample
    if (v.cgN->precond != TM_PRECOND_AMG || v.cgD->precond != TM_PRECOND_AMG || v.has_idmi) {
        delete bt;
        return TM_ERR_UNSUPPORTED;
    }
    bt->ctx = c;
    bt->n = v.n;
    bt->S = nsamples;
    bt->prec = precision;
    bt->stream = v.stream;
    auto fail = [&](int s) {
        tm_batch_destroy(bt);
        return s;
    };
    if (rocblas_create_handle(&bt->blas) != rocblas_status_success) return fail(TM_ERR_HIP);
    rocblas_set_stream(bt->blas, bt->stream);
    const SellPattern& P = *v.pat;
    std::vector<double> xc(P.h_csrcol.size(), 0.0);
    bt->exch = v.hop_set[TM_OP_EXCHANGE];
    if (bt->exch) {
        if ((st = csr_on_pattern(P, *v.hop_rp[TM_OP_EXCHANGE], *v.hop_col[TM_OP_EXCHANGE], *v.hop_val[TM_OP_EXCHANGE],
                                 xc)) != TM_OK)
            return fail(st);
    }
    if ((st = bt->ops.upload(P.n, P.h_rowptr.data(), P.h_csrcol.data(), xc.data(), bt->stream)) != TM_OK)
        return fail(st);
    int ops_ok = TM_OK;
    ops_ok |= upload_op(bt, TM_OP_GRAD_X, bt->gx);
    ops_ok |= upload_op(bt, TM_OP_GRAD_Y, bt->gy);
    ops_ok |= upload_op(bt, TM_OP_GRAD_Z, bt->gz);
    ops_ok |= upload_op(bt, TM_OP_TGRAD_X, bt->tx);
    ops_ok |= upload_op(bt, TM_OP_TGRAD_Y, bt->ty);
    ops_ok |= upload_op(bt, TM_OP_TGRAD_Z, bt->tz);
    bt->dmi = v.has_dmi && v.hop_set[TM_OP_CURL_X] && v.hop_set[TM_OP_CURL_Y] && v.hop_set[TM_OP_CURL_Z];
    if (bt->dmi) {
        ops_ok |= upload_op(bt, TM_OP_CURL_X, bt->cx);
        ops_ok |= upload_op(bt, TM_OP_CURL_Y, bt->cy);
        ops_ok |= upload_op(bt, TM_OP_CURL_Z, bt->cz);
    }
    if (ops_ok != TM_OK) return fail(TM_ERR_BAD_ARG);
    const size_t NS = (size_t)bt->n * nsamples;
    int a = TM_OK;
    a |= bt->m.alloc(3 * NS);
    a |= bt->dmdt.alloc(3 * NS);
    a |= bt->heff.alloc(3 * NS);
    a |= bt->u1.alloc(NS);
    a |= bt->u2.alloc(NS);
    a |= bt->rhsN.alloc(NS);
    a |= bt->rhsD.alloc(NS);
    if (precision == TM_PREC_MIXED_BEM32) {
        bt->nchunk = (v.bnx + bt->kchunk - 1) / bt->kchunk;
        a |= bt->Uf.alloc((size_t)v.bnx * nsamples);
        a |= bt->Gf.alloc((size_t)v.bnx * nsamples);
        a |= bt->Bf.alloc((size_t)v.bld * v.bnx);
        a |= bt->Pf.alloc((size_t)bt->nchunk * v.bnx * nsamples);
    } else {
        a |= bt->U.alloc((size_t)v.bnx * nsamples);
        a |= bt->Gm.alloc((size_t)v.bnx * nsamples);
    }
    a |= bt->hext.alloc((size_t)3 * nsamples);
    a |= bt->torque.alloc(nsamples);
    if (a != TM_OK) return fail(TM_ERR_HIP);
    if (precision == TM_PREC_MIXED_BEM32) {
        const size_t tot = (size_t)v.bld * v.bnx;
        hipLaunchKernelGGL(k_b_to_float, dim3(elem_grid(tot)), dim3(kBlock), 0, bt->stream, tot, v.B, bt->Bf.p);
        if (hipGetLastError() != hipSuccess) return fail(TM_ERR_HIP);
    }
    (void)hipMemsetAsync(bt->u1.p, 0, NS * sizeof(double), bt->stream);
    (void)hipMemsetAsync(bt->u2.p, 0, NS * sizeof(double), bt->stream);
    (void)hipMemsetAsync(bt->rhsD.p, 0, NS * sizeof(double), bt->stream);
    (void)hipMemsetAsync(bt->m.p, 0, 3 * NS * sizeof(double), bt->stream);
    std::vector<double> h((size_t)3 * nsamples);
    for (int s = 0; s < nsamples; ++s)
        for (int d = 0; d < 3; ++d) h[(size_t)d * nsamples + s] = v.hext[d];
    if ((st = bt->hext.upload(h.data(), h.size(), bt->stream)) != TM_OK) return fail(st);
    if (precision == TM_PREC_FP64) {
        if ((st = bt->sN64.setup(*v.cgN, nsamples, bt->stream)) != TM_OK) return fail(st);
        if ((st = bt->sD64.setup(*v.cgD, nsamples, bt->stream)) != TM_OK) return fail(st);
    } else {
        if ((st = bt->sN32.setup(*v.cgN, nsamples, bt->stream)) != TM_OK) return fail(st);
        if ((st = bt->sD32.setup(*v.cgD, nsamples, bt->stream)) != TM_OK) return fail(st);
    }
    if (hipStreamSynchronize(bt->stream) != hipSuccess) return fail(TM_ERR_HIP);
    if ((st = tune_gemm(bt)) != TM_OK) return fail(st);
    *out = bt;
    return TM_OK;
}

int tm_batch_destroy(tm_batch* bt) {
    if (!bt) return TM_OK;
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    if (bt->blas) rocblas_destroy_handle(bt->blas);
    for (auto& e : bt->ev)
        if (e) (void)hipEventDestroy(e);
    delete bt;
    return TM_OK;
}

int tm_batch_set_zeeman(tm_batch* bt, const double* h3s) {
    if (!bt || !h3s) return TM_ERR_BAD_ARG;
    TM_CHECK(bt->hext.upload(h3s, (size_t)3 * bt->S, bt->stream));
    TM_HIP(hipStreamSynchronize(bt->stream));
    return TM_OK;
}

int tm_batch_set_mag(tm_batch* bt, const double* m, int where) {
    if (!bt) return TM_ERR_BAD_ARG;
    return copy_in(bt->m.p, m, where, (size_t)3 * bt->n * bt->S, bt->stream);
}

int tm_batch_get_mag(tm_batch* bt, double* m, int where) {
    if (!bt) return TM_ERR_BAD_ARG;
    return copy_out(m, where, bt->m.p, (size_t)3 * bt->n * bt->S, bt->stream);
}

int tm_batch_get_potentials(tm_batch* bt, double* u1, double* u2, int where) {
    if (!bt) return TM_ERR_BAD_ARG;
    const size_t NS = (size_t)bt->n * bt->S;
    if (u1) TM_CHECK(copy_out(u1, where, bt->u1.p, NS, bt->stream));
    if (u2) TM_CHECK(copy_out(u2, where, bt->u2.p, NS, bt->stream));
    return TM_OK;
}

int tm_batch_enable_timers(tm_batch* bt, int on) {
    if (!bt) return TM_ERR_BAD_ARG;
    if (on && !bt->ev[0])
        for (auto& e : bt->ev) TM_HIP(hipEventCreate(&e));
    bt->timers = on != 0;
    return TM_OK;
}

int tm_batch_get_timers(tm_batch* bt, tm_phase_times* out, int reset) {
    if (!bt || !out) return TM_ERR_BAD_ARG;
    *out = bt->tsum;
    if (reset) bt->tsum = tm_phase_times{};
    return TM_OK;
}

int tm_batch_eval_rhs(tm_batch* bt, const double* m, int mwhere, double alpha, int variant, double* dmdt, int dwhere,
                      double* heff, int hwhere, tm_batch_stats* stats) {
    if (!bt || !dmdt || (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC)) return TM_ERR_BAD_ARG;
    if ((dwhere != TM_HOST && dwhere != TM_DEVICE) || (heff && hwhere != TM_HOST && hwhere != TM_DEVICE))
        return TM_ERR_BAD_ARG;
    const int n = bt->n, S = bt->S;
    const size_t NS = (size_t)n * S;
    hipStream_t s = bt->stream;
    const CtxView& v = bt->cv;
    if (m) TM_CHECK(copy_in(bt->m.p, m, mwhere, 3 * NS, s));
    if (bt->timers) TM_HIP(hipEventRecord(bt->ev[0], s));
    // Poisson right-hand side (DemagField.cpp:124-136), all samples
    hipLaunchKernelGGL(k_b_div, dim3(row_grid(n)), dim3(kBlock), 0, s,
                       CsrArgs{n, bt->ops.rp.p, bt->ops.ci.p, bt->tx.p}, (const double*)bt->ty.p,
                       (const double*)bt->tz.p, S, n, v.js, (const double*)bt->m.p, 1.0 / kMu0, bt->rhsN.p);
    TM_HIP(hipGetLastError());
    if (bt->timers) TM_HIP(hipEventRecord(bt->ev[1], s));
    int oN = 0, iN = 0, oD = 0, iD = 0;
    double rN = 0, rD = 0;
    int stN = bt->prec == TM_PREC_FP64 ? bt->sN64.solve(bt->rhsN.p, bt->u1.p, v.tol, v.maxit, s, &oN, &iN, &rN)
                                       : bt->sN32.solve(bt->rhsN.p, bt->u1.p, v.tol, v.maxit, s, &oN, &iN, &rN);
    if (stN != TM_OK && stN != TM_ERR_MAXITER) return stN;
    if (bt->timers) TM_HIP(hipEventRecord(bt->ev[2], s));
    // boundary integral (DemagField.cpp:184-201) as a GEMM: G = B U^T, U the
    // gathered boundary potentials of every sample
    TM_CHECK(boundary_product(bt, v.bnodes, bt->u1.p, bt->rhsD.p));
    if (bt->timers) TM_HIP(hipEventRecord(bt->ev[3], s));
    int stD = bt->prec == TM_PREC_FP64 ? bt->sD64.solve(bt->rhsD.p, bt->u2.p, v.tol, v.maxit, s, &oD, &iD, &rD)
                                       : bt->sD32.solve(bt->rhsD.p, bt->u2.p, v.tol, v.maxit, s, &oD, &iD, &rD);
    if (stD != TM_OK && stD != TM_ERR_MAXITER) return stD;
    if (bt->timers) TM_HIP(hipEventRecord(bt->ev[4], s));
    // fused fields + LLG
    BEvalArgs a = {};
    a.A = CsrArgs{n, bt->ops.rp.p, bt->ops.ci.p, bt->exch ? bt->ops.v.p : nullptr};
    a.gx = bt->gx.p;
    a.gy = bt->gy.p;
    a.gz = bt->gz.p;
    if (bt->dmi) {
        a.cx = bt->cx.p;
        a.cy = bt->cy.p;
        a.cz = bt->cz.p;
        a.dmi_fac = v.dmi_fac;
        a.nvs = v.nvs;
    }
    a.S = S;
    a.n = n;
    a.m = bt->m.p;
    a.u1 = bt->u1.p;
    a.u2 = bt->u2.p;
    a.invjs = v.invjs;
    a.ku = v.has_ku ? v.ku : nullptr;
    a.kaxis = v.has_ku ? v.kaxis : nullptr;
    a.hext = bt->hext.p;
    a.alpha = alpha;
    a.variant = variant;
    a.dmdt = dwhere == TM_DEVICE ? dmdt : bt->dmdt.p;
    a.heff = heff ? (hwhere == TM_DEVICE ? heff : bt->heff.p) : nullptr;
    a.torque = stats ? bt->torque.p : nullptr;
    if (stats) hipLaunchKernelGGL(k_b_zero_small, dim3(1), dim3(kMaxS), 0, s, S, bt->torque.p);
    if (bt->dmi) hipLaunchKernelGGL(k_b_eval<true>, dim3(row_grid(n)), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(k_b_eval<false>, dim3(row_grid(n)), dim3(kBlock), 0, s, a);
    TM_HIP(hipGetLastError());
    if (bt->timers) {
        TM_HIP(hipEventRecord(bt->ev[5], s));
        TM_HIP(hipEventSynchronize(bt->ev[5]));
        float t[5];
        for (int k = 0; k < 5; ++k) TM_HIP(hipEventElapsedTime(&t[k], bt->ev[k], bt->ev[k + 1]));
        bt->tsum.div_ms += t[0];
        bt->tsum.neumann_ms += t[1];
        bt->tsum.bem_ms += t[2];
        bt->tsum.dirichlet_ms += t[3];
        bt->tsum.fields_ms += t[4];
        bt->tsum.iters_neumann += iN;
        bt->tsum.iters_dirichlet += iD;
        bt->tsum.count += 1;
    }
    if (dwhere == TM_HOST) TM_CHECK(copy_out(dmdt, TM_HOST, a.dmdt, 3 * NS, s));
    if (heff && hwhere == TM_HOST) TM_CHECK(copy_out(heff, TM_HOST, a.heff, 3 * NS, s));
    if (stats) {
        stats->outer_neumann = oN;
        stats->outer_dirichlet = oD;
        stats->inner_neumann = iN;
        stats->inner_dirichlet = iD;
        stats->max_relres_neumann = rN;
        stats->max_relres_dirichlet = rD;
        stats->nsamples = S;
        TM_HIP(hipMemcpyAsync(stats->max_torque, bt->torque.p, sizeof(double) * S, hipMemcpyDeviceToHost, s));
        TM_HIP(hipStreamSynchronize(s));
    }
    return stN != TM_OK ? stN : stD;
}

int tm_batch_run_rk4(tm_batch* bt, double dtau, int nsteps, double alpha, int variant, tm_batch_stats* last) {
    if (!bt || nsteps < 0 || !(dtau > 0.0)) return TM_ERR_BAD_ARG;
    const size_t NS = (size_t)bt->n * bt->S, N3 = 3 * NS;
    for (auto& k : bt->rk_k) TM_CHECK(k.alloc(N3));
    TM_CHECK(bt->rk_base.alloc(N3));
    hipStream_t s = bt->stream;
    const dim3 g(elem_grid(N3));
    const double cst[4] = {0.0, 0.5 * dtau, 0.5 * dtau, dtau};
    int status = TM_OK;
    for (int step = 0; step < nsteps; ++step) {
        TM_HIP(hipMemcpyAsync(bt->rk_base.p, bt->m.p, N3 * sizeof(double), hipMemcpyDeviceToDevice, s));
        for (int st = 0; st < 4; ++st) {
            // the batch magnetisation holds the stage state
            if (st > 0)
                hipLaunchKernelGGL(k_b_axpy, g, dim3(kBlock), 0, s, N3, (const double*)bt->rk_base.p,
                                   (const double*)bt->rk_k[st - 1].p, cst[st], bt->m.p);
            const bool fin = step == nsteps - 1 && st == 3;
            const int r = tm_batch_eval_rhs(bt, nullptr, TM_DEVICE, alpha, variant, bt->rk_k[st].p, TM_DEVICE,
                                            nullptr, TM_DEVICE, fin ? last : nullptr);
            if (r != TM_OK && r != TM_ERR_MAXITER) return r;
            if (r == TM_ERR_MAXITER) status = r;
        }
        hipLaunchKernelGGL(k_b_rk_final, dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS,
                           (const double*)bt->rk_base.p, (const double*)bt->rk_k[0].p, (const double*)bt->rk_k[1].p,
                           (const double*)bt->rk_k[2].p, (const double*)bt->rk_k[3].p, dtau, bt->m.p);
        TM_HIP(hipGetLastError());
    }
    TM_HIP(hipStreamSynchronize(s));
    return status;
}

int tm_batch_gemm_info(tm_batch* bt, int* solution, int* candidates, double* ms_default, double* ms_tuned,
                       int* ksplit) {
    if (!bt) return TM_ERR_BAD_ARG;
    if (solution) *solution = bt->gemm_own ? -1 - bt->gemm_shape : bt->gemm_sol;
    if (candidates) *candidates = bt->gemm_candidates + bt->gemm_own_candidates;
    if (ksplit) *ksplit = bt->gemm_own ? bt->gemm_ks : 1;
    if (ms_default) *ms_default = bt->gemm_ms_default;
    if (ms_tuned) *ms_tuned = bt->gemm_ms_best;
    return TM_OK;
}

int tm_batch_bem_apply(tm_batch* bt, const double* x, int xwhere, double* y, int ywhere) {
    if (!bt || !x || !y || (xwhere != TM_HOST && xwhere != TM_DEVICE) || (ywhere != TM_HOST && ywhere != TM_DEVICE))
        return TM_ERR_BAD_ARG;
    const size_t bs = (size_t)bt->cv.bnx * bt->S;
    const double* xd = x;
    if (xwhere == TM_HOST) {  // staged in rhsN (n x S >= b x S)
        TM_CHECK(copy_in(bt->rhsN.p, x, TM_HOST, bs, bt->stream));
        xd = bt->rhsN.p;
    }
    double* yd = ywhere == TM_HOST ? bt->rhsD.p : y;
    TM_CHECK(boundary_product(bt, nullptr, xd, yd));
    if (ywhere == TM_HOST) TM_CHECK(copy_out(y, TM_HOST, yd, bs, bt->stream));
    return TM_OK;
}

int tm_batch_step_euler(tm_batch* bt, const double* dm, int where, double dtau) {
    if (!bt) return TM_ERR_BAD_ARG;
    const size_t NS = (size_t)bt->n * bt->S;
    const double* dd = dm;
    if (where == TM_HOST) {
        TM_CHECK(copy_in(bt->heff.p, dm, TM_HOST, 3 * NS, bt->stream));
        dd = bt->heff.p;
    } else if (where != TM_DEVICE) {
        return TM_ERR_BAD_ARG;
    }
    hipLaunchKernelGGL(k_b_euler, dim3(elem_grid(NS)), dim3(kBlock), 0, bt->stream, NS, bt->m.p, dd, dtau);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

}  // extern "C"
