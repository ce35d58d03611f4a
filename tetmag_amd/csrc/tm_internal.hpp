// tm_internal.hpp -- shared internals of the HIP library (not part of the ABI).
#pragma once

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

// "Last block finishes the reduction" handshake.  Every block writes its
// partials, fences, and bumps a counter; the block that sees count == grid-1
// reduces all partials (in block order, so the result is deterministic) and
// resets the counter for the next launch.
__device__ inline bool last_block_arrive(unsigned* counter) {
    __shared__ bool am_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        am_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (am_last) __threadfence();
    return am_last;
}

// Sums NV partial arrays (stride `stride`) of `count` entries inside one block.
template <int NV>
__device__ inline void block_reduce_partials(const double* partials, int stride, int count, double (&out)[NV],
                                             double* smem) {
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int i = threadIdx.x; i < count; i += kBlock) {
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += __builtin_nontemporal_load(partials + (size_t)q * stride + i);
    }
    block_sum<NV>(acc, smem);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = acc[q];
}

inline int grid_for(int64_t n, int block = kBlock) { return (int)((n + block - 1) / block); }

// Copies between a caller pointer (host or device) and a device buffer.
int copy_in(double* dst_dev, const double* src, int where, size_t count, hipStream_t s);
int copy_out(double* dst, int where, const double* src_dev, size_t count, hipStream_t s);

// ---------------------------------------------------------------------------
// Preconditioned CG on "symmetric matrix with identity rows" systems.
// ---------------------------------------------------------------------------
struct CgState {
    double gamma;    // (r, z) of the current residual
    double alpha;    // step length of the current iteration
    double beta;     // p update coefficient for the next iteration
    double rr;       // (r, r)
    double bb;       // (b, b)
    double tol2bb;   // tol^2 * (b, b)
    int iters;
    int done;        // 1 = converged, 2 = breakdown/non-finite
    unsigned counter;
    int pad;
};

struct CgSolver {
    int n = 0;
    SellPattern pat;
    DevBuf<double> val;      // matrix values in SELL order
    DevBuf<double> dinv;     // inverse diagonal (1 on fixed rows)
    DevBuf<unsigned char> fixed;  // 1 on identity rows
    DevBuf<double> r, z, q, p0, p1, xw;  // work vectors (xw: solver-owned solution)
    DevBuf<double> partials;  // 3 * nblocks
    DevBuf<CgState> st;
    CgState* h_st = nullptr;  // pinned mirror for the convergence check
    int nfixed = 0;
    int last_iters = 0;
    bool has_solution = false;
    ~CgSolver();
    int setup(int n, const int* rowptr, const int* colidx, const double* vals, int precond, hipStream_t s);
    // Solves A x = b for device vectors b, x (x is the warm start).  tol is
    // relative to ||b||.  Returns TM_OK, TM_ERR_MAXITER or TM_ERR_NONFINITE.
    int solve(const double* b, double* x, double tol, int maxit, hipStream_t s, int* iters, double* relres);
};

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
    tmh::DevBuf<double> b, x;
    hipStream_t stream = nullptr;
};
