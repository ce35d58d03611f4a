// tm_batch.hip -- many magnetisation samples on one mesh in one pass.
//
// The reference evaluates one magnetisation at a time (TheLLG, DemagField)
// and its hysteresis driver runs field values one after another
// (Hysteresis.cpp:43-169).  Here up to 64 samples share every operator read:
// vectors are stored node-major, sample-fastest ([node][sample]); each sparse
// product runs one wavefront per row with the samples across its lanes, so
// the row's offsets and weights (row-slot layout, below) are wavefront-uniform
// scalar loads read once for all samples and every gather is one contiguous
// access (64 FP32 samples: 16 lanes x 16 bytes per gathered row, four rows per
// load instruction).  The boundary-element product becomes a GEMM (B times the
// b x S boundary potentials) on the FP64 matrix cores.
//
// Both demag solves use the same scheme per sample:
//   outer (FP64):  x_fixed = b_fixed;  r = b - A x  (FP64 operator and vectors)
//   inner (T):     A d = r by AMG-preconditioned CG in precision T, from d = 0
//   x += d, repeat until ||b - A x|| <= tol ||b|| for every sample.
// T = double solves to the final tolerance in one outer step (the FP64
// batched form); T = float runs the inner iterations in FP32 to a relative
// 1e-5 and recovers FP64 accuracy by iterative refinement (the mixed form).
// Converged samples are frozen (their step lengths are zero); once all are,
// the iterations still queued return at once.
#define ROCBLAS_BETA_FEATURES_API  // rocblas_gemm_ex_get_solutions (the GEMM autotuner)
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <vector>

#include "tm_internal.hpp"

namespace tmh {

namespace {

constexpr int kMaxS = 64;  // any_active() reads 64 flags
constexpr int kRowsPerBlock = kBlock / kWave;  // one row per wavefront


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
    [[maybe_unused]] const int lc = min(lane, S - 1); /* gather lane: in bounds for every lane */  \
    for (int row = __builtin_amdgcn_readfirstlane(blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave)); \
         row < (NROWS); row += gridDim.x * kRowsPerBlock)

// As B_ROW_LOOP, over XCD-contiguous eighths of the rows (blocks with equal
// blockIdx % 8 share one XCD); needs lane, on, lc declared by the caller.
#define B_ROW_LOOP_XCD(NROWS)                                                                          \
    const int xr0_ = gridDim.x % kXcd ? 0 : (int)((int64_t)(NROWS) * (blockIdx.x % kXcd) / kXcd);      \
    const int xr1_ = gridDim.x % kXcd ? (NROWS) : (int)((int64_t)(NROWS) * (blockIdx.x % kXcd + 1) / kXcd); \
    const int xlb_ = gridDim.x % kXcd ? blockIdx.x : blockIdx.x / kXcd;                                \
    const int xst_ = (gridDim.x % kXcd ? gridDim.x : gridDim.x / kXcd) * kRowsPerBlock;                \
    for (int row = __builtin_amdgcn_readfirstlane(xr0_ + xlb_ * kRowsPerBlock + (int)(threadIdx.x / kWave)); \
         row < xr1_; row += xst_)

// ---- slot layout for the solver kernels ----------------------------------
// Every operator the batched solves apply is stored in row slots of kSlot
// entries: row i owns slots sp[i] .. sp[i+1]-1 (a multiple of kSlot, padded
// with offset 0 and weight 0), each slot holding the byte offset of the
// gathered row (column * S * sizeof(element)) and the weight.  A wavefront
// reads a chunk's offsets and weights with scalar loads (16 consecutive
// dwords each: wavefront-uniform data goes to SGPRs, no lane broadcasts),
// gathers with buffer loads whose row offset is the SGPR operand (no address
// arithmetic per entry) and accumulates with the weight as the SGPR operand
// of the FMA.  Per entry that is one load and one vector instruction; the
// earlier form (indices and values in lanes, read with readlane, 64-bit
// address adds) was bound by vector-instruction issue.
constexpr int kSlot = 16;

struct SlotArgs {
    int n;            // rows
    int xbytes;       // bytes of the gathered vector (ncols * S * element)
    const int* sp;    // n + 1 slot starts
    const int* off;   // per slot: byte offset of the gathered row
    const void* w;    // per slot: weight (float or double)
    const int* gate = nullptr;  // per-sample active flags: the kernel does nothing when none is set
    int gate_n = 0;
};

typedef const int __attribute__((address_space(4)))* cint_p;

// true when any of the flags is set (or there are no flags).  The batched
// CG enqueues iterations in chunks ahead of its convergence check; once every
// sample has converged, the iterations still queued return at once.  All
// kMaxS flags are read (four 16-dword scalar loads; flags past the batch size
// stay 0 from the solver's set-up).
__device__ inline bool any_active(const int* act, int) {
    if (!act) return true;
    const cint_p a = (cint_p)act;
    int any = 0;
#pragma unroll
    for (int s = 0; s < 64; ++s) any |= a[s];
    return any != 0;
}
template <typename W>
using cw_p = const W __attribute__((address_space(4)))*;

template <typename T>
__device__ inline T buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ inline float buf_ld<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ inline double buf_ld<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// raw buffer over the gathered vector (word 3: 32-bit data format, gfx9)
__device__ inline __amdgpu_buffer_rsrc_t gather_rsrc(const void* x, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(x), (short)0, bytes, 0x00020000);
}

// acc += sum over the kSlot slots at q of w * x[row]
template <typename T, typename W>
__device__ inline T slot_chunk(cint_p off, cw_p<W> w, int q, __amdgpu_buffer_rsrc_t xr, int voff, T acc) {
    int o[kSlot];
    T xv[kSlot];
#pragma unroll
    for (int k = 0; k < kSlot; ++k) o[k] = off[q + k];
#pragma unroll
    for (int k = 0; k < kSlot; ++k) xv[k] = buf_ld<T>(xr, voff, o[k]);
#pragma unroll
    for (int k = 0; k < kSlot; ++k) acc += (T)w[q + k] * xv[k];
    return acc;
}

// rows a wavefront walks at a time: two in FP32 (2 x 32 SGPRs of offsets and
// weights), one in FP64 (the weights take twice the SGPRs)
template <typename T>
constexpr int kRowsPerStep = sizeof(T) == 4 ? 2 : 1;

// Row walk: wavefront w of block b takes rows r0 + (b/8)*4 + w, then
// + 4*G/8, ... within its XCD's eighth [r0, r1); with RPI = 2 it takes two
// rows per step and issues both rows' first-chunk gathers before consuming
// any (twice the loads in flight).
// ld(row) loads what the row's epilogue needs (issued with the gathers);
// epi(row, acc, ld(row)) finishes the row.  Lanes >= S gather lane S-1's
// element (lc) and their results are discarded by epi.
template <typename T, typename W, int RPI, typename LDF, typename EPIF>
__device__ inline void slot_walk(const SlotArgs& M, const void* x, int lc, LDF&& ld, EPIF&& epi) {
    const int n = M.n;
    const __amdgpu_buffer_rsrc_t xr = gather_rsrc(x, M.xbytes);
    const cint_p sp = (cint_p)M.sp;
    const cint_p of = (cint_p)M.off;
    const cw_p<W> w = (cw_p<W>)M.w;
    const int vo = lc * (int)sizeof(T);
    // each XCD (blocks b with equal b % 8) walks its own contiguous eighth of
    // the rows, so neighbouring rows share that XCD's L2
    int r0 = 0, r1 = n, lb = blockIdx.x, gx = gridDim.x;
    if (gridDim.x % kXcd == 0) {
        const int x = blockIdx.x % kXcd;
        gx = gridDim.x / kXcd;
        lb = blockIdx.x / kXcd;
        r0 = (int)((int64_t)n * x / kXcd);
        r1 = (int)((int64_t)n * (x + 1) / kXcd);
    }
    const int stride = gx * kRowsPerBlock;
    for (int row = __builtin_amdgcn_readfirstlane(r0 + lb * kRowsPerBlock + (int)(threadIdx.x / kWave)); row < r1;
         row += RPI * stride) {
        if constexpr (RPI == 2) {
            const int rb = row + stride;
            const bool hb = rb < r1;
            const int a0 = sp[row], a1 = sp[row + 1];
            const int b0 = hb ? sp[rb] : a0, b1 = hb ? sp[rb + 1] : a0 + kSlot;
            const auto la = ld(row);
            const auto lb = ld(hb ? rb : row);
            int oa[kSlot], ob[kSlot];
            T xa[kSlot], xb[kSlot];
#pragma unroll
            for (int k = 0; k < kSlot; ++k) {
                oa[k] = of[a0 + k];
                ob[k] = of[b0 + k];
            }
#pragma unroll
            for (int k = 0; k < kSlot; ++k) {
                xa[k] = buf_ld<T>(xr, vo, oa[k]);
                xb[k] = buf_ld<T>(xr, vo, ob[k]);
            }
            T sa = 0, sb = 0;
#pragma unroll
            for (int k = 0; k < kSlot; ++k) {
                sa += (T)w[a0 + k] * xa[k];
                sb += (T)w[b0 + k] * xb[k];
            }
            for (int q = a0 + kSlot; q < a1; q += kSlot) sa = slot_chunk<T, W>(of, w, q, xr, vo, sa);
            for (int q = b0 + kSlot; q < b1; q += kSlot) sb = slot_chunk<T, W>(of, w, q, xr, vo, sb);
            epi(row, sa, la);
            if (hb) epi(rb, sb, lb);
        } else {
            const int a0 = sp[row], a1 = sp[row + 1];
            const auto la = ld(row);
            T sa = 0;
            for (int q = a0; q < a1; q += kSlot) sa = slot_chunk<T, W>(of, w, q, xr, vo, sa);
            epi(row, sa, la);
        }
    }
}


// ---- wide gathers for 64 FP32 samples -------------------------------------
// With S = 64 in FP32 a node's row of samples is 256 B.  Lanes gather 16 B
// each (4 samples), so 16 lanes cover a row and one load instruction gathers
// four different rows: lane group g = lane / 16 takes slots 4g .. 4g+3 of
// each 16-slot chunk (their offsets and weights arrive as one 16-byte load
// each), lane j = lane % 16 holds samples 4j .. 4j+3.  The groups' partial
// sums are added across lanes at the end of the row.  Against one dword
// gather per slot this issues a quarter of the vector-memory instructions
// (the per-instruction address work of the texture path bounds the 64-lane
// dword form).
typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef int i2 __attribute__((ext_vector_type(2)));

__device__ inline f4 buf_ld4(__amdgpu_buffer_rsrc_t r, int voff) {
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}

// the gather of a slot; where its weight is zero (padding slots) the offset
// is moved past the buffer's end, so the load returns zeros without a
// memory access (raw-buffer range check; gathered vectors are < 1 GiB, see
// SlotOp::build)
constexpr int kOutOfRange = 0x7ffffff0;
__device__ inline f4 ld4_if(__amdgpu_buffer_rsrc_t r, int voff, float w) {
    return buf_ld4(r, w != 0.f ? voff : kOutOfRange);
}

__device__ inline f4 shfl_xor4(f4 v, int m) {
    return f4{__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m), __shfl_xor(v.w, m)};
}

// Row walk as slot_walk (XCD-contiguous eighths).  BLK = false: one row per
// lane group, four consecutive rows per wavefront step; lane j of a group
// holds samples 4j .. 4j+3 of its row and walks all of the row's slots
// itself (per chunk of CH slots: the offsets and weights as CH/4 16-byte
// loads, then all CH gathers before any is used), so a wavefront has four
// rows' gathers in flight and no cross-group sum.  BLK = true (operators of
// long rows: the coarse levels' Galerkin products, the restrictions): one
// row per block, its 16-slot chunks dealt round-robin to the block's 4
// wavefronts (lane group g: slots q+4g .. q+4g+3 of chunk q), the groups'
// and wavefronts' sums added across lanes and in LDS.  ld(row, j) loads the
// epilogue's operands for samples 4j .. 4j+3; epi(row, j, acc, ld) runs on
// every lane of a group whose row exists (BLK = false), on lanes 0-15 of
// wavefront 0 (BLK = true).
template <bool BLK, int CH = kSlot, typename LDF, typename EPIF>
__device__ inline void wide_walk(const SlotArgs& M, const void* x, LDF&& ld, EPIF&& epi) {
    static_assert(CH == 16 || CH == 8, "chunks of 16 or 8 slots");
    const int n = M.n;
    const int lane = threadIdx.x & (kWave - 1), g = lane >> 4, j = lane & 15;
    const int wv = threadIdx.x / kWave;
    const __amdgpu_buffer_rsrc_t xr = gather_rsrc(x, M.xbytes);
    const i4* __restrict__ of = (const i4*)M.off;
    const f4* __restrict__ w = (const f4*)M.w;
    const int vo = j * 16;
    int r0 = 0, r1 = n, lb = blockIdx.x, gx = gridDim.x;
    if (gridDim.x % kXcd == 0) {
        const int xi = blockIdx.x % kXcd;
        gx = gridDim.x / kXcd;
        lb = blockIdx.x / kXcd;
        r0 = (int)((int64_t)n * xi / kXcd);
        r1 = (int)((int64_t)n * (xi + 1) / kXcd);
    }
    if constexpr (!BLK) {
        const int* __restrict__ sp = M.sp;
        const int stride = gx * kRowsPerBlock * 4;
        for (int base = __builtin_amdgcn_readfirstlane(r0 + (lb * kRowsPerBlock + wv) * 4); base < r1;
             base += stride) {
            const int row = base + g;
            const bool has = row < r1;
            const int rr = has ? row : base;
            const int a0 = sp[rr], a1 = has ? sp[rr + 1] : a0;
            const auto lv = ld(rr, j);
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int q = a0; q < a1; q += CH) {
                i4 o[CH / 4];
                f4 ww[CH / 4];
#pragma unroll
                for (int k = 0; k < CH / 4; ++k) {
                    o[k] = of[(q >> 2) + k];
                    ww[k] = w[(q >> 2) + k];
                }
                f4 xv[CH];
#pragma unroll
                for (int k = 0; k < CH / 4; ++k) {
                    xv[4 * k] = ld4_if(xr, vo + o[k].x, ww[k].x);
                    xv[4 * k + 1] = ld4_if(xr, vo + o[k].y, ww[k].y);
                    xv[4 * k + 2] = ld4_if(xr, vo + o[k].z, ww[k].z);
                    xv[4 * k + 3] = ld4_if(xr, vo + o[k].w, ww[k].w);
                }
#pragma unroll
                for (int k = 0; k < CH / 4; ++k) {
                    acc += ww[k].x * xv[4 * k];
                    acc += ww[k].y * xv[4 * k + 1];
                    acc += ww[k].z * xv[4 * k + 2];
                    acc += ww[k].w * xv[4 * k + 3];
                }
            }
            if (has) epi(row, j, acc, lv);
        }
    } else {
        static_assert(!BLK || CH == 16, "long rows in chunks of 16");
        auto chunk = [&](int q, f4 acc) {  // lane group g: slots q+4g .. q+4g+3
            const i4 o = of[(q >> 2) + g];
            const f4 ww = w[(q >> 2) + g];
            const f4 x0 = ld4_if(xr, vo + o.x, ww.x), x1 = ld4_if(xr, vo + o.y, ww.y);
            const f4 x2 = ld4_if(xr, vo + o.z, ww.z), x3 = ld4_if(xr, vo + o.w, ww.w);
            acc += ww.x * x0;
            acc += ww.y * x1;
            acc += ww.z * x2;
            acc += ww.w * x3;
            return acc;
        };
        const cint_p sp = (cint_p)M.sp;
        __shared__ f4 red[kRowsPerBlock][16];
        for (int row = r0 + lb; row < r1; row += gx) {  // block-uniform
            const int a0 = sp[row], a1 = sp[row + 1];
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int q = a0 + CH * wv; q < a1; q += CH * kRowsPerBlock) acc = chunk(q, acc);
            acc += shfl_xor4(acc, 16);
            acc += shfl_xor4(acc, 32);
            if (g == 0) red[wv][j] = acc;
            __syncthreads();
            if (wv == 0 && g == 0) {
                const auto lv = ld(row, j);
                f4 t = red[0][j];
#pragma unroll
                for (int k = 1; k < kRowsPerBlock; ++k) t += red[k][j];
                epi(row, j, t, lv);
            }
            __syncthreads();
        }
    }
}

// part[s*G + block] for samples 4j + c, v[c] held by the lanes j, j + 16,
// j + 32, j + 48 (summed across the four lane groups first)
__device__ inline void b_partial4(double* part, const double (&vin)[4]) {
    __shared__ double sp[kRowsPerBlock][kWave];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        v[c] = vin[c] + __shfl_xor(vin[c], 16);
        v[c] += __shfl_xor(v[c], 32);
    }
    if (lane < 16)
#pragma unroll
        for (int c = 0; c < 4; ++c) sp[wv][4 * lane + c] = v[c];
    __syncthreads();
    if (wv == 0) {
        double t = 0.0;
        for (int k = 0; k < kRowsPerBlock; ++k) t += sp[k][lane];
        part[(size_t)lane * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
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

// r = b - A x in FP64, stored as rT = (T) r (the inner solve's right-hand
// side), partials of (b,b) and (r,r).  A: FP64 weights, offsets of FP64 rows.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_resid(SlotArgs A, int S, const double* __restrict__ b,
                                                    const double* __restrict__ x, T* __restrict__ rT,
                                                    double* part_bb, double* part_rr) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    double vb = 0.0, vr = 0.0;
    auto ld = [&](int row) { return b[(size_t)row * S + lc]; };
    auto epi = [&](int row, double acc, double bi) {
        if (on) {
            const size_t ee = (size_t)row * S + lane;
            const double ri = bi - acc;
            rT[ee] = (T)ri;
            vb += bi * bi;
            vr += ri * ri;
        }
    };
    slot_walk<double, double, 1>(A, x, lc, ld, epi);
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

__global__ __launch_bounds__(kBlock) void k_b_sum2(size_t tot, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* __restrict__ c) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock)
        c[e] = a[e] + b[e];
}

// x += sum_k alpha_k p_k of an inner solve of K iterations, from the kept
// preconditioned residuals z_i (ring of [n][S] vectors): p_k = z_k +
// beta_k p_{k-1} gives sum_k alpha_k p_k = sum_i c_i z_i with c_{K-1} =
// alpha_{K-1}, c_i = alpha_i + beta_{i+1} c_{i+1} (alpha_i in ah[i * kMaxS +
// s], beta_i in bh[i * kMaxS + s]).  Iterations a sample did not take have
// alpha 0 (and beta 0 after them), so c_i = 0 selects their z_i (whose
// contents may be stale) out of the sum.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_combine(size_t NS, int S, int K, const T* __restrict__ ring,
                                                      const double* __restrict__ ah, const double* __restrict__ bh,
                                                      double* __restrict__ x) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < NS; e += (size_t)gridDim.x * kBlock) {
        const int sm = (int)(e % S);
        double acc = 0.0, c = 0.0;
        for (int i = K - 1; i >= 0; --i) {
            c = ah[i * kMaxS + sm] + bh[(i + 1) * kMaxS + sm] * c;
            const T v = ring[(size_t)i * NS + e];
            acc += c != 0.0 ? c * (double)v : 0.0;
        }
        x[e] += acc;
    }
}

// c_i of k_b_combine for every sample (thread s): ch[i * kMaxS + s], i < K
__global__ void k_b_coef(int S, int K, const double* __restrict__ ah, const double* __restrict__ bh,
                         double* __restrict__ ch) {
    const int sm = threadIdx.x;
    if (sm >= S) return;
    double c = 0.0;
    for (int i = K - 1; i >= 0; --i) {
        c = ah[i * kMaxS + sm] + bh[(i + 1) * kMaxS + sm] * c;
        ch[i * kMaxS + sm] = c;
    }
}

// k_b_combine for S = 64 with the c_i from k_b_coef (staged in LDS, K <=
// kCombineK): four samples per thread (vector loads of each z_i)
constexpr int kCombineK = 48;
// four consecutive elements as FP64 (one 16-byte load for float, two for double)
template <typename T>
__device__ inline void load4d(const T* p, double (&v)[4]) {
    if constexpr (sizeof(T) == 4) {
        const f4 a = *(const f4*)p;
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
    } else {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 a = ((const d2*)p)[0], b = ((const d2*)p)[1];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = b.x;
        v[3] = b.y;
    }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_combine4(size_t NS, int K, const T* __restrict__ ring,
                                                       const double* __restrict__ ch, double* __restrict__ x) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef double d4 __attribute__((ext_vector_type(4)));
    __shared__ d4 cs[kCombineK * kMaxS / 4];
    for (int t = threadIdx.x; t < K * kMaxS / 4; t += kBlock) cs[t] = ((const d4*)ch)[t];
    __syncthreads();
    const size_t N4 = NS / 4;
    for (size_t e4 = (size_t)blockIdx.x * kBlock + threadIdx.x; e4 < N4; e4 += (size_t)gridDim.x * kBlock) {
        const int q = (int)(e4 % (kMaxS / 4));  // samples 4q .. 4q+3
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < K; ++i) {
            double v[4];
            load4d(ring + (size_t)i * NS + 4 * e4, v);
            const d4 c = cs[i * (kMaxS / 4) + q];
            acc[0] += c.x != 0.0 ? c.x * v[0] : 0.0;
            acc[1] += c.y != 0.0 ? c.y * v[1] : 0.0;
            acc[2] += c.z != 0.0 ? c.z * v[2] : 0.0;
            acc[3] += c.w != 0.0 ? c.w * v[3] : 0.0;
        }
        d2* xp = (d2*)(x + 4 * e4);
        const d2 x0 = xp[0], x1 = xp[1];
        xp[0] = d2{x0.x + acc[0], x0.y + acc[1]};
        xp[1] = d2{x1.x + acc[2], x1.y + acc[3]};
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_zero(size_t tot, T* __restrict__ x) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) x[e] = (T)0;
}

// ------------------------------------------------------------- inner CG --
// Iteration k: q_k = A p_k = A z_k + beta_k q_{k-1} in one pass (the gathers
// read z_k; q_{k-1} is the row's own entry), with the partial (z_k, q_k):
// (p_k, A p_k) = (z_k, A p_k) as p_{k-1} is A-conjugate to p_k.  The
// directions are never formed: the preconditioned residuals z_k are kept (a
// ring of vectors) with alpha_k and beta_k, and the inner solution
// sum_k alpha_k p_k = sum_i c_i z_i is added once at the end (k_b_combine),
// so no iteration reads or writes a direction or the iterate.
// k = 0: q = A z (q_{-1} not read).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_spmv(SlotArgs A, int S, const T* __restrict__ z, T* __restrict__ q,
                                                      int k, const BState* bs, double* part) {
    if (!any_active(A.gate, A.gate_n)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    const T beta = k > 0 ? (T)bs->beta[lc] : (T)0;
    double v = 0.0;
    struct ZQ {
        T z, q;
    };
    auto ld = [&](int row) {
        const size_t e = (size_t)row * S + lc;
        return ZQ{z[e], k > 0 ? q[e] : (T)0};
    };
    auto epi = [&](int row, T acc, const ZQ& o) {
        if (on) {
            const T qi = acc + beta * o.q;
            q[(size_t)row * S + lane] = qi;
            v += (double)o.z * (double)qi;
        }
    };
    slot_walk<T, T, kRowsPerStep<T>>(A, z, lc, ld, epi);
    b_partial(part, S, v);
}

// k_b_cg_spmv for 64 FP32 samples, wide gathers
__global__ __launch_bounds__(kBlock) void k_b_cg_spmv_w(SlotArgs A, const float* __restrict__ dg,
                                                        const float* __restrict__ z, float* __restrict__ q, int k,
                                                        const BState* bs, double* part) {
    if (!any_active(A.gate, A.gate_n)) return;
    const int j = threadIdx.x & 15;
    const f4 beta = k > 0 ? f4{(float)bs->beta[4 * j], (float)bs->beta[4 * j + 1], (float)bs->beta[4 * j + 2],
                               (float)bs->beta[4 * j + 3]}
                          : f4{0.f, 0.f, 0.f, 0.f};
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    struct ZQ {
        f4 z, q;
    };
    auto ld = [&](int row, int jj) {
        const size_t e = (size_t)row * 64 + 4 * jj;
        return ZQ{*(const f4*)(z + e), k > 0 ? *(const f4*)(q + e) : f4{0.f, 0.f, 0.f, 0.f}};
    };
    auto epi = [&](int row, int jj, f4 acc, const ZQ& o) {
        acc += dg[row] * o.z;  // diagonal kept out of the slots
        const f4 qi = acc + beta * o.q;
        *(f4*)(q + (size_t)row * 64 + 4 * jj) = qi;
        v[0] += (double)o.z.x * (double)qi.x;
        v[1] += (double)o.z.y * (double)qi.y;
        v[2] += (double)o.z.z * (double)qi.z;
        v[3] += (double)o.z.w * (double)qi.w;
    };
    wide_walk<false>(A, z, ld, epi);
    b_partial4(part, v);
}

__global__ void k_b_red_pq(int G, int S, const double* part, BState* bs, double* ah) {
    const int s = blockIdx.x;
    const double pq = b_total(part, G);
    if (threadIdx.x == 0 && s < S) {
        const bool ok = bs->active[s] && pq > 0.0 && std::isfinite(pq);
        if (bs->active[s] && !ok) bs->active[s] = 0;  // breakdown: freeze the sample
        const double a = ok ? bs->gam[s] / pq : 0.0;
        bs->alpha[s] = a;
        ah[s] = a;
    }
}

// r -= alpha q; partial (r, r)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_cg_update(int n, int S, const T* __restrict__ q, T* __restrict__ r,
                                                        const BState* bs, double* part) {
    if (!any_active(bs->active, S)) return;
    const int ln = threadIdx.x & (kWave - 1);
    const T alpha = ln < S ? (T)bs->alpha[ln] : (T)0;
    double v = 0.0;
    B_ROW_LOOP(n) {
        if (on) {
            const size_t e = (size_t)row * S + lane;
            const T ri = r[e] - alpha * q[e];
            r[e] = ri;
            v += (double)ri * (double)ri;
        }
    }
    b_partial(part, S, v);
}

// after r -= alpha q: the iteration count and convergence on ||r||^2,
// decided before the preconditioner runs, so that a batch whose samples
// have all converged skips the V-cycle (its z would never be used: the
// correction sums alpha_k p_k over the iterations taken)
__global__ void k_b_red_rr(int G, int S, const double* part_rr, BState* bs) {
    const int s = blockIdx.x;
    const double rr = b_total(part_rr, G);
    if (threadIdx.x == 0 && s < S && bs->active[s]) {
        bs->iters[s] += 1;
        bs->rr[s] = rr;
        if (!(rr > bs->itol2[s]) || !std::isfinite(rr)) bs->active[s] = 0;
    }
}

// after z = M r: gamma and beta (bh[s]: the next iteration's slot of the
// beta history).  k < 0, the first call of an inner solve, also decides
// convergence on the outer residual's ||r||^2; later iterations decide it
// in k_b_red_rr.
__global__ void k_b_red_rz(int G, int S, const double* part_rz, int k, BState* bs, double* bh) {
    const int s = blockIdx.x;
    const double rz = b_total(part_rz, G);
    if (threadIdx.x == 0 && s < S) {
        const double r2 = bs->rr[s];
        bool act = bs->active[s] && std::isfinite(rz);
        if (k < 0) act = act && r2 > bs->itol2[s] && std::isfinite(r2);
        bs->beta[s] = (act && k >= 0 && bs->gam[s] != 0.0) ? rz / bs->gam[s] : 0.0;
        bh[s] = bs->beta[s];
        bs->gam[s] = rz;
        bs->active[s] = act ? 1 : 0;
    }
}

// --------------------------------------------------------- AMG V-cycle --
// y = M x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_spmm(SlotArgs M, int S, const T* __restrict__ x, T* __restrict__ y) {
    if (!any_active(M.gate, M.gate_n)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    auto ld = [&](int) { return (T)0; };
    auto epi = [&](int row, T acc, T) {
        if (on) y[(size_t)row * S + lane] = acc;
    };
    slot_walk<T, T, kRowsPerStep<T>>(M, x, lc, ld, epi);
}

// x = w dinv b + P xc: the pre-smoothed iterate plus the coarse correction
// (the pre-smoother from zero is folded into the restriction operator)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_prolong(SlotArgs P, int S, const double* __restrict__ dinv,
                                                      const T* __restrict__ b, double w, const T* __restrict__ xc,
                                                      T* __restrict__ x) {
    if (!any_active(P.gate, P.gate_n)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    auto ld = [&](int row) { return b[(size_t)row * S + lc]; };
    auto epi = [&](int row, T acc, T bi) {
        if (on) x[(size_t)row * S + lane] = (T)(w * dinv[row]) * bi + acc;
    };
    slot_walk<T, T, kRowsPerStep<T>>(P, xc, lc, ld, epi);
}

// xo = x + w dinv (b - A x); optional partial (b, xo)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_post(SlotArgs A, int S, const double* __restrict__ dinv,
                                                   const T* __restrict__ b, double w, const T* __restrict__ x,
                                                   T* __restrict__ xo, double* part) {
    if (!any_active(A.gate, A.gate_n)) return;
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
    auto epi = [&](int row, T acc, const BX& bx) {
        if (on) {
            const T zi = bx.x + (T)(w * dinv[row]) * (bx.b - acc);
            xo[(size_t)row * S + lane] = zi;
            v += (double)bx.b * (double)zi;
        }
    };
    slot_walk<T, T, kRowsPerStep<T>>(A, x, lc, ld, epi);
    if (part) b_partial(part, S, v);
}

// Wide-gather forms (FP32, S = 64) of the transfer products and the
// post-smoothers (the pre-smoother from zero is folded into the restriction
// operator R (I - A w D^-1)).
// x = w dinv b + P xc: the pre-smoothed iterate (not stored by the restriction pass)
// plus the coarse correction, in one pass (P in chunks of CH slots)
template <bool BLK, int CH>
__global__ __launch_bounds__(kBlock) void k_b_prolong_w(SlotArgs P, const double* __restrict__ dinv,
                                                        const float* __restrict__ b, double w,
                                                        const float* __restrict__ xc, float* __restrict__ x) {
    if (!any_active(P.gate, P.gate_n)) return;
    auto ld = [&](int row, int j) { return *(const f4*)(b + (size_t)row * 64 + 4 * j); };
    auto epi = [&](int row, int j, f4 acc, f4 bi) {
        *(f4*)(x + (size_t)row * 64 + 4 * j) = (float)(w * dinv[row]) * bi + acc;
    };
    wide_walk<BLK, CH>(P, xc, ld, epi);
}

// y = M x (the restriction operators)
template <bool BLK>
__global__ __launch_bounds__(kBlock) void k_b_spmm_w(SlotArgs M, const float* __restrict__ x, float* __restrict__ y) {
    if (!any_active(M.gate, M.gate_n)) return;
    auto ld = [&](int, int) { return 0; };
    auto epi = [&](int row, int j, f4 acc, int) { *(f4*)(y + (size_t)row * 64 + 4 * j) = acc; };
    wide_walk<BLK>(M, x, ld, epi);
}

template <bool BLK>
__global__ __launch_bounds__(kBlock) void k_b_post_w(SlotArgs A, const float* __restrict__ dg,
                                                     const double* __restrict__ dinv, const float* __restrict__ b,
                                                     double w, const float* __restrict__ x, float* __restrict__ xo,
                                                     double* part) {
    if (!any_active(A.gate, A.gate_n)) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    struct BX {
        f4 b, x;
    };
    auto ld = [&](int row, int j) {
        const size_t e = (size_t)row * 64 + 4 * j;
        return BX{*(const f4*)(b + e), *(const f4*)(x + e)};
    };
    auto epi = [&](int row, int j, f4 acc, const BX& bx) {
        if (dg) acc += dg[row] * bx.x;  // diagonal kept out of the slots
        const f4 zi = bx.x + (float)(w * dinv[row]) * (bx.b - acc);
        *(f4*)(xo + (size_t)row * 64 + 4 * j) = zi;
        v[0] += (double)bx.b.x * (double)zi.x;
        v[1] += (double)bx.b.y * (double)zi.y;
        v[2] += (double)bx.b.z * (double)zi.z;
        v[3] += (double)bx.b.w * (double)zi.w;
    };
    wide_walk<BLK>(A, x, ld, epi);
    if (part) b_partial4(part, v);
}

// The level-0 post-smoother with the diagonal in the slots: z = M x + w dinv b
// with M = I - w D^-1 A (only b is read per row); partial (b, z)
__global__ __launch_bounds__(kBlock) void k_b_post1_w(SlotArgs M, const double* __restrict__ dinv,
                                                      const float* __restrict__ b, double w,
                                                      const float* __restrict__ x, float* __restrict__ z,
                                                      double* part) {
    if (!any_active(M.gate, M.gate_n)) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    auto ld = [&](int row, int j) { return *(const f4*)(b + (size_t)row * 64 + 4 * j); };
    auto epi = [&](int row, int j, f4 acc, f4 bi) {
        const f4 zi = acc + (float)(w * dinv[row]) * bi;
        *(f4*)(z + (size_t)row * 64 + 4 * j) = zi;
        v[0] += (double)bi.x * (double)zi.x;
        v[1] += (double)bi.y * (double)zi.y;
        v[2] += (double)bi.z * (double)zi.z;
        v[3] += (double)bi.w * (double)zi.w;
    };
    wide_walk<false>(M, x, ld, epi);
    b_partial4(part, v);
}

// x = C b for the dense coarse inverse (row-major nc x nc): one row per block,
// its columns dealt to the 4 wavefronts in 16-column chunks (C's entries are
// wavefront-uniform scalar loads), the wavefronts' sums added in LDS.
constexpr int kDenseCh = 16;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_dense(int nc, int S, const double* __restrict__ C,
                                                    const T* __restrict__ b, T* __restrict__ x,
                                                    const int* gate) {
    if (!any_active(gate, S)) return;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int lc = min(lane, S - 1);
    const cw_p<double> Cc = (cw_p<double>)C;
    __shared__ double red[kRowsPerBlock][kWave];
    for (int row = blockIdx.x; row < nc; row += gridDim.x) {  // block-uniform
        double acc = 0.0;
        for (int k0 = kDenseCh * wv; k0 < nc; k0 += kDenseCh * kRowsPerBlock) {
            double cv[kDenseCh], bv[kDenseCh];
#pragma unroll
            for (int k = 0; k < kDenseCh; ++k) {
                const int kk = min(k0 + k, nc - 1);
                cv[k] = Cc[(size_t)row * nc + kk];
                bv[k] = (double)b[(size_t)kk * S + lc];
            }
#pragma unroll
            for (int k = 0; k < kDenseCh; ++k) acc = k0 + k < nc ? acc + cv[k] * bv[k] : acc;
        }
        red[wv][lane] = acc;
        __syncthreads();
        if (wv == 0 && lane < S) {
            double t = red[0][lane];
            for (int k = 1; k < kRowsPerBlock; ++k) t += red[k][lane];
            x[(size_t)row * S + lane] = (T)t;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ fields and LLG --
// The per-node operators (exchange, grad, tGrad, curl) share the union
// pattern, stored in row slots (SlotArgs) with FP64-row offsets; the slot
// weights of each operator and the slot columns are separate arrays.  A row
// is walked in halves of 8 slots: offsets, columns and weights by scalar
// loads, gathers by buffer loads with the row offset (plus the component's
// byte offset) as the SGPR operand.  Sums run in entry order (padding slots
// have weight 0).
constexpr int kHalf = 8;

__device__ inline double buf_ld8(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// b = sum_d tGrad_d (Js m_d) / mu0 for every sample (DemagField.cpp:124-136);
// the slot weights hold tGrad_d(i, j) Js_j / mu0 (set at batch creation)
__global__ __launch_bounds__(kBlock) void k_b_div(SlotArgs P, const double* __restrict__ wty,
                                                  const double* __restrict__ wtz, int S,
                                                  const double* __restrict__ m, double* __restrict__ out) {
    const int n = P.n;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    const int c8 = n * S * 8;  // bytes of one component
    const __amdgpu_buffer_rsrc_t mr = gather_rsrc(m, 3 * c8);
    const cint_p sp = (cint_p)P.sp, of = (cint_p)P.off;
    const cw_p<double> wx = (cw_p<double>)P.w, wy = (cw_p<double>)wty, wz = (cw_p<double>)wtz;
    const int vo = lc * 8;
    B_ROW_LOOP_XCD(n) {
        double sx = 0, sy = 0, sz = 0;
        const int a0 = sp[row], a1 = sp[row + 1];
        for (int q = a0; q < a1; q += kHalf) {
            int o[kHalf];
            double mx[kHalf], my[kHalf], mz[kHalf];
#pragma unroll
            for (int k = 0; k < kHalf; ++k) o[k] = of[q + k];
#pragma unroll
            for (int k = 0; k < kHalf; ++k) {
                mx[k] = buf_ld8(mr, vo, o[k]);
                my[k] = buf_ld8(mr, vo, o[k] + c8);
                mz[k] = buf_ld8(mr, vo, o[k] + 2 * c8);
            }
#pragma unroll
            for (int k = 0; k < kHalf; ++k) {
                sx += wx[q + k] * mx[k];
                sy += wy[q + k] * my[k];
                sz += wz[q + k] * mz[k];
            }
        }
        if (on) out[(size_t)row * S + lane] = sx + sy + sz;
    }
}

// ---- FP64 rows of 64 samples with 16-byte lanes (k_b_eval_w) --------------
// Lane l holds samples 2j, 2j+1 (j = l % 32) and takes the even (l < 32) or
// the odd (l >= 32) slots of the row; each half-wavefront picks its slot's
// offset and weight out of the pair read as wavefront-uniform scalars.
// pair_sum adds the two halves' sums and leaves sample l of the row in lane
// l.  Measured at 282,897 nodes: the field kernel 1.01 -> 0.94 ms; the same
// form of the divergence and of the FP64 residual was slower (0.62 -> 0.65,
// 0.21 -> 0.22 ms): their time follows the L2 misses of the gathered rows,
// not the lanes the texture path processes.
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ inline d2v buf_ld16d(__amdgpu_buffer_rsrc_t r, int voff) {
    return __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}

__device__ inline double pair_sum(d2v s) {
    s.x += __shfl_xor(s.x, 32);
    s.y += __shfl_xor(s.y, 32);
    const int l = threadIdx.x & (kWave - 1);
    const double vx = __shfl(s.x, l >> 1), vy = __shfl(s.y, l >> 1);
    return (l & 1) ? vy : vx;
}

struct BEvalArgs {
    SlotArgs P;                        // union pattern in slots, w = exchange weights (may be null)
    const double* gx; const double* gy; const double* gz;
    const double* cx; const double* cy; const double* cz;
    int S, n;
    const double* m; const double* u;   // u = u1 + u2
    const double* invjs; const double* ku; const double* kaxis;
    const double* dmi_fac; const double* nvs;
    const double* hext;   // [3][S] per-sample uniform applied field
    const double* hfield; // 3n per-node applied field shared by every sample (or null)
    double alpha;
    int variant;
    double* dmdt; double* heff;
    double* torque;       // [S], zeroed before
    double* mnext;        // if set: the explicit Euler step m + dt dm/dt, renormalised (a buffer apart from m)
    double dt;
};

// dm/dt of one (node, sample) into a.dmdt, and the Euler step into a.mnext
// when asked for (the arithmetic of k_b_euler)
__device__ inline void b_store_llg(const BEvalArgs& a, size_t e, size_t NS, double mx, double my, double mz,
                                   double dx, double dy, double dz) {
    a.dmdt[e] = dx;
    a.dmdt[NS + e] = dy;
    a.dmdt[2 * NS + e] = dz;
    if (a.mnext) {
        const double x = mx + a.dt * dx, y = my + a.dt * dy, z = mz + a.dt * dz;
        const double l = sqrt(x * x + y * y + z * z);
        const double f = l > 0.0 ? 1.0 / l : 1.0;
        a.mnext[e] = x * f;
        a.mnext[NS + e] = y * f;
        a.mnext[2 * NS + e] = z * f;
    }
}

// Fused Heff + LLG for every (node, sample): TheLLG.cpp:238-254 and
// EffFieldGPU.cu:316-362, with the demag field -grad(u1 + u2).
template <bool DMI>
__global__ __launch_bounds__(kBlock) void k_b_eval(BEvalArgs a) {
    const int S = a.S;
    const size_t NS = (size_t)a.n * S;
    const int n = a.n;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int lc = min(lane, S - 1);
    const int c8 = n * S * 8;
    const __amdgpu_buffer_rsrc_t mr = gather_rsrc(a.m, 3 * c8);
    const __amdgpu_buffer_rsrc_t ur = gather_rsrc(a.u, c8);
    const cint_p sp = (cint_p)a.P.sp, of = (cint_p)a.P.off;
    const cw_p<double> wx = (cw_p<double>)a.P.w;
    const cw_p<double> gx = (cw_p<double>)a.gx, gy = (cw_p<double>)a.gy, gz = (cw_p<double>)a.gz;
    const cw_p<double> cx = (cw_p<double>)a.cx, cy = (cw_p<double>)a.cy, cz = (cw_p<double>)a.cz;
    const bool exch = a.P.w != nullptr;
    const int vo = lc * 8;
    double tmax = 0.0;
    B_ROW_LOOP_XCD(n) {
        double ex = 0, ey = 0, ez = 0, dx = 0, dy = 0, dz = 0, cxs = 0, cys = 0, czs = 0;
        const int a0 = sp[row], a1 = sp[row + 1];
        for (int q = a0; q < a1; q += kHalf) {
            int o[kHalf];
            double mv[kHalf][3], uv[kHalf];
#pragma unroll
            for (int k = 0; k < kHalf; ++k) o[k] = of[q + k];
#pragma unroll
            for (int k = 0; k < kHalf; ++k) {
                mv[k][0] = buf_ld8(mr, vo, o[k]);
                mv[k][1] = buf_ld8(mr, vo, o[k] + c8);
                mv[k][2] = buf_ld8(mr, vo, o[k] + 2 * c8);
                uv[k] = buf_ld8(ur, vo, o[k]);
            }
#pragma unroll
            for (int k = 0; k < kHalf; ++k) {
                const double mx = mv[k][0], my = mv[k][1], mz = mv[k][2];
                if (exch) {
                    const double xc = wx[q + k];
                    ex += xc * mx;
                    ey += xc * my;
                    ez += xc * mz;
                }
                dx += gx[q + k] * uv[k];
                dy += gy[q + k] * uv[k];
                dz += gz[q + k] * uv[k];
                if (DMI) {
                    const double X = cx[q + k], Y = cy[q + k], Z = cz[q + k];
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
        if (a.hfield) {
            hx += a.hfield[row];
            hy += a.hfield[n + row];
            hz += a.hfield[2 * n + row];
        }
        if (a.heff) {
            a.heff[e] = hx;
            a.heff[NS + e] = hy;
            a.heff[2 * NS + e] = hz;
        }
        const double tx = my * hz - mz * hy, ty = mz * hx - mx * hz, tz = mx * hy - my * hx;
        const double bx = my * tz - mz * ty, by = mz * tx - mx * tz, bz = mx * ty - my * tx;
        if (a.variant == TM_LLG_NOPREC)
            b_store_llg(a, e, NS, mx, my, mz, -a.alpha * bx, -a.alpha * by, -a.alpha * bz);
        else
            b_store_llg(a, e, NS, mx, my, mz, -tx - a.alpha * bx, -ty - a.alpha * by, -tz - a.alpha * bz);
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

// k_b_eval for S = 64 (16-byte lanes, slots taken in pairs, two pairs per
// step)
template <bool DMI>
__global__ __launch_bounds__(kBlock) void k_b_eval_w(BEvalArgs a) {
    constexpr int S = kMaxS;
    constexpr int KP = 2;  // slot pairs per step
    const int n = a.n;
    const size_t NS = (size_t)n * S;
    const int lane = threadIdx.x & (kWave - 1);
    const bool odd = lane >= 32;
    const int vo = (lane & 31) * 16;
    const int c8 = n * S * 8;
    const __amdgpu_buffer_rsrc_t mr = gather_rsrc(a.m, 3 * c8);
    const __amdgpu_buffer_rsrc_t ur = gather_rsrc(a.u, c8);
    const cint_p sp = (cint_p)a.P.sp, of = (cint_p)a.P.off;
    const cw_p<double> wx = (cw_p<double>)a.P.w;
    const cw_p<double> gx = (cw_p<double>)a.gx, gy = (cw_p<double>)a.gy, gz = (cw_p<double>)a.gz;
    const cw_p<double> cx = (cw_p<double>)a.cx, cy = (cw_p<double>)a.cy, cz = (cw_p<double>)a.cz;
    const bool exch = a.P.w != nullptr;
    double tmax = 0.0;
    B_ROW_LOOP_XCD(n) {
        const d2v z2 = {0.0, 0.0};
        d2v ex = z2, ey = z2, ez = z2, dx = z2, dy = z2, dz = z2, cxs = z2, cys = z2, czs = z2;
        const int a0 = sp[row], a1 = sp[row + 1];
        for (int q = a0; q < a1; q += 2 * KP) {
            int o[KP];
#pragma unroll
            for (int k = 0; k < KP; ++k) o[k] = vo + (odd ? of[q + 2 * k + 1] : of[q + 2 * k]);
            d2v mv[KP][3], uv[KP];
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                mv[k][0] = buf_ld16d(mr, o[k]);
                mv[k][1] = buf_ld16d(mr, o[k] + c8);
                mv[k][2] = buf_ld16d(mr, o[k] + 2 * c8);
                uv[k] = buf_ld16d(ur, o[k]);
            }
            // the half's weight of pair k: both read as scalars, one picked
            auto pick = [&](cw_p<double> w, int k) { return odd ? w[q + 2 * k + 1] : w[q + 2 * k]; };
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const d2v mx = mv[k][0], my = mv[k][1], mz = mv[k][2];
                if (exch) {
                    const double xc = pick(wx, k);
                    ex += xc * mx;
                    ey += xc * my;
                    ez += xc * mz;
                }
                dx += pick(gx, k) * uv[k];
                dy += pick(gy, k) * uv[k];
                dz += pick(gz, k) * uv[k];
                if (DMI) {
                    const double X = pick(cx, k), Y = pick(cy, k), Z = pick(cz, k);
                    cxs += Y * mz - Z * my;
                    cys += Z * mx - X * mz;
                    czs += X * my - Y * mx;
                }
            }
        }
        const double Ex = pair_sum(ex), Ey = pair_sum(ey), Ez = pair_sum(ez);
        const double Dx = pair_sum(dx), Dy = pair_sum(dy), Dz = pair_sum(dz);
        double Cx = 0.0, Cy = 0.0, Cz = 0.0;
        if (DMI) {
            Cx = pair_sum(cxs);
            Cy = pair_sum(cys);
            Cz = pair_sum(czs);
        }
        const size_t e = (size_t)row * S + lane;
        const double mx = a.m[e], my = a.m[NS + e], mz = a.m[2 * NS + e];
        double hx = Ex, hy = Ey, hz = Ez;
        if (a.ku) {
            const double kx = a.kaxis[row], ky = a.kaxis[n + row], kz = a.kaxis[2 * n + row];
            const double f = 2.0 * a.ku[row] * (kx * mx + ky * my + kz * mz);
            hx += f * kx;
            hy += f * ky;
            hz += f * kz;
        }
        if (DMI) {
            const double sx = a.nvs[row], sy = a.nvs[n + row], sz = a.nvs[2 * n + row];
            const double f = a.dmi_fac[row];
            hx += ((my * sz - mz * sy) - 2.0 * Cx) * f;
            hy += ((mz * sx - mx * sz) - 2.0 * Cy) * f;
            hz += ((mx * sy - my * sx) - 2.0 * Cz) * f;
        }
        const double ij = a.invjs ? a.invjs[row] : 1.0;
        hx = hx * ij + a.hext[lane] - Dx;
        hy = hy * ij + a.hext[S + lane] - Dy;
        hz = hz * ij + a.hext[2 * S + lane] - Dz;
        if (a.hfield) {
            hx += a.hfield[row];
            hy += a.hfield[n + row];
            hz += a.hfield[2 * n + row];
        }
        if (a.heff) {
            a.heff[e] = hx;
            a.heff[NS + e] = hy;
            a.heff[2 * NS + e] = hz;
        }
        const double tx = my * hz - mz * hy, ty = mz * hx - mx * hz, tz = mx * hy - my * hx;
        const double bx = my * tz - mz * ty, by = mz * tx - mx * tz, bz = mx * ty - my * tx;
        if (a.variant == TM_LLG_NOPREC)
            b_store_llg(a, e, NS, mx, my, mz, -a.alpha * bx, -a.alpha * by, -a.alpha * bz);
        else
            b_store_llg(a, e, NS, mx, my, mz, -tx - a.alpha * bx, -ty - a.alpha * by, -tz - a.alpha * bz);
        tmax = fmax(tmax, sqrt(tx * tx + ty * ty + tz * tz));
    }
    if (!a.torque) return;
    __shared__ double st[kRowsPerBlock][kWave];
    const int wv = threadIdx.x / kWave;
    st[wv][lane] = tmax;
    __syncthreads();
    if (wv == 0) {
        double t = st[0][lane];
        for (int k = 1; k < kRowsPerBlock; ++k) t = fmax(t, st[k][lane]);
        atomic_max_nonneg(a.torque + lane, t);
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

// ---- boundary product from int8 digit slices (TM_PREC_MIXED_BEMI8) -------
// Every row j of B and every sample column s of U is scaled by a power of
// two (sigma_j >= max_k |B_jk|, tau_s >= max_k |U_ks|) and written as
// base-2^7 digits in [-64, 64], five for B and six for U:
//   B_jk = sigma_j sum_p d_p 2^-(6+7p),   U_ks = tau_s sum_q e_q 2^-(6+7q).
// Digit products accumulate exactly in int32 on the int8 matrix cores
// (v_mfma_i32_16x16x64_i8); the products with p + q <= 5 are kept and
// summed per p + q, so
//   (B U^T)_js = sigma_j tau_s 2^-12 sum_{d <= 5} 2^-7d acc_d(j, s),
//   acc_d = sum_{p+q=d} sum_k d_p(j, k) e_q(k, s).
// The error is then B's own rounding to five digits (~2^-35 of the row
// maximum per entry, a fixed perturbation of B, ~1e-8 relative on the
// product); U's rounding and the dropped products are ~2^-42, so the
// product varies as smoothly as U along a trajectory.  The splitting of a
// floating-point product into exact integer products follows Ozaki et al.
// (Numer. Algorithms 59 (2012) 95) and its integer-matrix-engine forms.
// |acc_d| <= 5 * 64 * 64 * K < 2^31 for K <= 104832 columns per K range.
typedef int i32x4 __attribute__((ext_vector_type(4)));
constexpr int kI8Digits = 5;            // digits of B
constexpr int kI8DigitsU = 6;           // digits of U
constexpr int kI8MaxD = 5;              // products with p + q <= kI8MaxD
constexpr int kI8Op = 64;               // i32x4 per MFMA operand (64 lanes x 16 bytes)
constexpr int kI8MaxK = 104832;         // columns per K range without int32 overflow
// Slice layouts (16-byte operands, lane l = row/column l % 16, k block l / 16):
//   Bq: [row tile T (16 rows)][k chunk c (64)][digit p][lane l] i32x4 =
//       d_p(B)[16T + l % 16][64c + 16(l / 16) + 0..15]
//   Uq: [k chunk c][digit q (6)][column tile u (16 samples)][lane l] i32x4 =
//       e_q(U)[64c + 16(l / 16) + 0..15][16u + l % 16]

// ND digits of r, |r| <= 1, each exact (power-of-two scalings and rint)
template <int ND>
__device__ inline void i8_digits(double r, int (&d)[ND]) {
#pragma unroll
    for (int p = 0; p < ND; ++p) {
        const double w = (double)(1LL << (6 + 7 * p));
        const double t = rint(r * w);
        d[p] = (int)t;
        r -= t / w;
    }
}

// the power of two 2^e >= m (m = mantissa 2^e, mantissa in [0.5, 1)); 1 for m = 0
__device__ inline double i8_scale(double m) {
    if (m == 0.0) return 1.0;
    if (!(m > 0.0) || !std::isfinite(m)) return m;  // NaN and inf pass through (non-finite products)
    int e;
    (void)frexp(m, &e);
    return ldexp(1.0, e);
}

// per-row maxima of |B| (B column-major, ld), K split over blockIdx.y
__global__ __launch_bounds__(kBlock) void k_b_i8_rowmax(int b, int ld, const double* __restrict__ B, int kper,
                                                        double* __restrict__ rmax) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= b) return;
    const int k0 = blockIdx.y * kper, k1 = min(b, k0 + kper);
    double m = 0.0;
    for (int k = k0; k < k1; ++k) m = fmax(m, fabs(B[(size_t)k * ld + j]));
    atomic_max_nonneg(rmax + j, m);
}

// Bq from B (once per batch): thread = (tile T, chunk c, lane l), 16 values
__global__ __launch_bounds__(kBlock) void k_b_i8_bslice(int b, int ld, const double* __restrict__ B, int nT, int nC,
                                                        const double* __restrict__ rmax, i32x4* __restrict__ Bq) {
    const size_t tot = (size_t)nT * nC * 64;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) {
        const int l = (int)(e & 63);
        const size_t tc = e >> 6;
        const int c = (int)(tc % nC), T = (int)(tc / nC);
        const int row = 16 * T + (l & 15), kb = 64 * c + 16 * (l >> 4);
        const double inv = row < b ? 1.0 / i8_scale(rmax[row]) : 0.0;
        i32x4 w[kI8Digits];
#pragma unroll
        for (int p = 0; p < kI8Digits; ++p) w[p] = i32x4{0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = kb + j;
            const double v = (row < b && k < b) ? B[(size_t)k * ld + row] * inv : 0.0;
            int d[kI8Digits];
            i8_digits(v, d);
#pragma unroll
            for (int p = 0; p < kI8Digits; ++p) w[p][j >> 2] |= (d[p] & 0xff) << (8 * (j & 3));
        }
#pragma unroll
        for (int p = 0; p < kI8Digits; ++p) Bq[(tc * kI8Digits + p) * kI8Op + l] = w[p];
    }
}

__global__ void k_b_i8_zero(int cnt, double* __restrict__ p) {
    if ((int)threadIdx.x < cnt) p[threadIdx.x] = 0.0;
}

// per-sample maxima of |u| over the boundary nodes (lane = sample)
__global__ __launch_bounds__(kBlock) void k_b_i8_umax(int b, int S, const int* __restrict__ bn,
                                                      const double* __restrict__ u, double* __restrict__ umax) {
    const int lane = threadIdx.x & 63;
    double m = 0.0;
    if (lane < S)
        for (int k = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6); k < b; k += gridDim.x * (kBlock / 64)) {
            const double v = u[(size_t)(bn ? bn[k] : k) * S + lane];
            m = fmax(m, v == v ? fabs(v) : INFINITY);  // a NaN counts as inf: the product turns non-finite
        }
    __shared__ double sm[kBlock / 64][64];
    sm[threadIdx.x >> 6][lane] = m;
    __syncthreads();
    if (threadIdx.x < 64 && (int)threadIdx.x < S) {
        double t = sm[0][lane];
        for (int w = 1; w < kBlock / 64; ++w) t = fmax(t, sm[w][lane]);
        atomic_max_nonneg(umax + lane, t);
    }
}

// Uq from the boundary values: thread = (chunk c, column tile u, lane l)
__global__ __launch_bounds__(kBlock) void k_b_i8_uslice(int b, int S, int nC, const int* __restrict__ bn,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ umax, i32x4* __restrict__ Uq) {
    const size_t tot = (size_t)nC * 4 * 64;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) {
        const int l = (int)(e & 63), ut = (int)((e >> 6) & 3), c = (int)(e >> 8);
        const int s = 16 * ut + (l & 15), kb = 64 * c + 16 * (l >> 4);
        const double inv = s < S ? 1.0 / i8_scale(umax[s]) : 0.0;
        i32x4 w[kI8DigitsU];
#pragma unroll
        for (int q = 0; q < kI8DigitsU; ++q) w[q] = i32x4{0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = kb + j;
            const double v = (s < S && k < b) ? u[(size_t)(bn ? bn[k] : k) * S + s] * inv : 0.0;
            int d[kI8DigitsU];
            i8_digits(v, d);
#pragma unroll
            for (int q = 0; q < kI8DigitsU; ++q) w[q][j >> 2] |= (d[q] & 0xff) << (8 * (j & 3));
        }
#pragma unroll
        for (int q = 0; q < kI8DigitsU; ++q) Uq[(((size_t)c * kI8DigitsU + q) * 4 + ut) * kI8Op + l] = w[q];
    }
}

// P[ks][row][col] (row-major, 64 columns) = sum_d 2^-7d acc_d over the K
// chunks [ks cper, (ks+1) cper) of the slices.  Block: 4 wavefronts x MT
// row tiles of 16 rows; the chunk's U digits (24 KB) are staged in LDS and
// shared by the four wavefronts, double-buffered: while the MFMAs of chunk
// c run, the B digits of chunk c+1 are in flight into registers and its U
// digits into registers, stored to the other buffer afterwards.
// v_mfma_i32_16x16x64_i8: A lane l = row l % 16, B lane l = column l % 16,
// both the 16 k of block l / 16; D lane l register r = C[4 (l / 16) + r][l % 16].
template <int MT>
__global__ __launch_bounds__(kBlock) void k_b_gemm_i8(int b, int nT, int nC, int cper,
                                                      const i32x4* __restrict__ Bq, const i32x4* __restrict__ Uq,
                                                      double* __restrict__ P) {
    constexpr int UPT = kI8DigitsU * 4 * kI8Op / kBlock;  // U operands per thread per chunk (6)
    __shared__ i32x4 us[2][kI8DigitsU * 4 * kI8Op];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int T0 = (blockIdx.x * (kBlock / 64) + wv) * MT;
    const int cb = blockIdx.y * cper, ce = min(nC, cb + cper);
    const int ncc = ce > cb ? ce - cb : 0;  // uniform over the block
    i32x4 acc[kI8MaxD + 1][MT][4];
#pragma unroll
    for (int d = 0; d <= kI8MaxD; ++d)
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[d][t][u] = i32x4{0, 0, 0, 0};
    bool tin[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) tin[t] = T0 + t < nT;
    i32x4 a[MT][kI8Digits], an[MT][kI8Digits], ur[UPT];
    auto load_a = [&](int c, i32x4 (&dst)[MT][kI8Digits]) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int p = 0; p < kI8Digits; ++p)
                dst[t][p] = tin[t] ? Bq[(((size_t)(T0 + t) * nC + c) * kI8Digits + p) * kI8Op + lane]
                                   : i32x4{0, 0, 0, 0};
    };
    auto load_u = [&](int c) {
#pragma unroll
        for (int i = 0; i < UPT; ++i) ur[i] = Uq[(size_t)c * (kI8DigitsU * 4 * kI8Op) + threadIdx.x + i * kBlock];
    };
    auto store_u = [&](int buf) {
#pragma unroll
        for (int i = 0; i < UPT; ++i) us[buf][threadIdx.x + i * kBlock] = ur[i];
    };
    auto mfmas = [&](const i32x4 (&av)[MT][kI8Digits], int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < kI8DigitsU; ++q) {
                const i32x4 bq = us[buf][(q * 4 + u) * kI8Op + lane];
#pragma unroll
                for (int t = 0; t < MT; ++t)
#pragma unroll
                    for (int p = 0; p < kI8Digits && p + q <= kI8MaxD; ++p)
                        acc[p + q][t][u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[t][p], bq, acc[p + q][t][u], 0, 0, 0);
            }
    };
    if (ncc > 0) {
        load_u(cb);
        load_a(cb, an);
        store_u(0);
    }
    __syncthreads();
    for (int c = 0; c < ncc; ++c) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int p = 0; p < kI8Digits; ++p) a[t][p] = an[t][p];
        const bool more = c + 1 < ncc;
        if (more) {
            load_u(cb + c + 1);
            load_a(cb + c + 1, an);
        }
        mfmas(a, c & 1);
        if (more) store_u((c + 1) & 1);  // that buffer was last read in chunk c-1
        __syncthreads();
    }
    double* out = P + (size_t)blockIdx.y * b * 64;
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = (T0 + t) * 16 + 4 * (lane >> 4) + r;
                double v = 0.0;
#pragma unroll
                for (int d = kI8MaxD; d >= 0; --d) v = v * (1.0 / 128.0) + (double)acc[d][t][u][r];
                if (tin[t] && row < b) out[(size_t)row * 64 + u * 16 + (lane & 15)] = v;
            }
}

// out[bn[j]][s] = sigma_j tau_s 2^-12 sum_ks P[ks][j][s] (s < S); without bn: out[j][s]
__global__ __launch_bounds__(kBlock) void k_b_i8_finish(int b, int S, int nks, const int* __restrict__ bn,
                                                        const double* __restrict__ P, const double* __restrict__ rmax,
                                                        const double* __restrict__ umax, double* __restrict__ out) {
    const size_t tot = (size_t)b * 64;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) {
        const int j = (int)(e >> 6), s = (int)(e & 63);
        if (s >= S) continue;
        double acc = 0.0;
        for (int c = 0; c < nks; ++c) acc += P[(size_t)c * tot + e];
        out[(size_t)(bn ? bn[j] : j) * S + s] = acc * i8_scale(rmax[j]) * i8_scale(umax[s]) * (1.0 / 4096.0);
    }
}

void launch_gemm_i8(int mt, dim3 g, hipStream_t s, int b, int nT, int nC, int cper, const i32x4* Bq,
                    const i32x4* Uq, double* P) {
    if (mt == 2)
        hipLaunchKernelGGL(k_b_gemm_i8<2>, g, dim3(kBlock), 0, s, b, nT, nC, cper, Bq, Uq, P);
    else
        hipLaunchKernelGGL(k_b_gemm_i8<1>, g, dim3(kBlock), 0, s, b, nT, nC, cper, Bq, Uq, P);
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

// *fail += number of samples whose outer residual missed the tolerance
__global__ void k_b_count_undone(int S, const OState* os, unsigned* fail) {
    const int s = threadIdx.x;
    if (s < S && !os->done[s]) atomicAdd(fail, 1u);
}

// every sample active with finite CG scalars (kernel timing only)
__global__ void k_b_activate(int S, BState* bs) {
    const int s = threadIdx.x;
    if (s < S) {
        bs->active[s] = 1;
        bs->gam[s] = 1.0;
        bs->alpha[s] = 0.0;
        bs->beta[s] = 0.0;
    }
}

__global__ void k_b_reset_fail(unsigned* fail) {
    if (threadIdx.x == 0 && blockIdx.x == 0) fail[0] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_b_copy(size_t tot, const double* __restrict__ a, double* __restrict__ b) {
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < tot; e += (size_t)gridDim.x * kBlock) b[e] = a[e];
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
// Host copy of a device CSR matrix (slot layouts are built on the host).
struct HostCsr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

int fetch_csr(const AmgPrecond::CsrDev& M, HostCsr& h, hipStream_t s) {
    h.n = M.n;
    h.rp.assign((size_t)M.n + 1, 0);
    TM_HIP(hipMemcpyAsync(h.rp.data(), M.rp.p, h.rp.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    TM_HIP(hipStreamSynchronize(s));
    const size_t nnz = (size_t)h.rp[M.n];
    h.ci.assign(nnz, 0);
    h.v.assign(nnz, 0.0);
    if (nnz) {
        TM_HIP(hipMemcpyAsync(h.ci.data(), M.ci.p, nnz * sizeof(int), hipMemcpyDeviceToHost, s));
        TM_HIP(hipMemcpyAsync(h.v.data(), M.v.p, nnz * sizeof(double), hipMemcpyDeviceToHost, s));
        TM_HIP(hipStreamSynchronize(s));
    }
    return TM_OK;
}

// A CSR pattern in row slots (see SlotArgs): every row gets a multiple of
// kSlot slots, at least one chunk, padded with offset 0 / weight 0; byte
// offsets for gathered elements of 4 and of 8 bytes.
struct SlotOp {
    int n = 0, ncols = 0, S = 0;
    int chunk = kSlot;            // rows padded to a multiple of this (16; 8 for the wide kernels' short rows)
    bool nodiag = false;          // diagonal entries left out (applied by the kernel's epilogue)
    std::vector<int> h_rp, h_ci, h_sp;
    DevBuf<int> sp, off4, off8;

    bool kept(int i, int q) const { return !(nodiag && h_ci[q] == i); }

    int build(int rows, int cols, const int* rp, const int* ci, int S_, hipStream_t s, bool drop_diag = false,
              int chunk_ = kSlot) {
        n = rows;
        ncols = cols;
        S = S_;
        nodiag = drop_diag;
        chunk = chunk_;
        if ((int64_t)cols * S * 8 > (int64_t)INT32_MAX) return TM_ERR_UNSUPPORTED;  // 32-bit byte offsets
        h_rp.assign(rp, rp + rows + 1);
        h_ci.assign(ci, ci + rp[rows]);
        h_sp.assign((size_t)rows + 1, 0);
        for (int i = 0; i < rows; ++i) {
            int len = 0;
            for (int q = rp[i]; q < rp[i + 1]; ++q) len += kept(i, q);
            h_sp[i + 1] = h_sp[i] + std::max(1, (len + chunk - 1) / chunk) * chunk;
        }
        const size_t ns = (size_t)h_sp[rows];
        std::vector<int> o4(ns, 0), o8(ns, 0);
        for (int i = 0; i < rows; ++i) {
            size_t k = (size_t)h_sp[i];
            for (int q = rp[i]; q < rp[i + 1]; ++q)
                if (kept(i, q)) {
                    o4[k] = ci[q] * S * 4;
                    o8[k] = ci[q] * S * 8;
                    ++k;
                }
        }
        TM_CHECK(sp.upload(h_sp.data(), h_sp.size(), s));
        TM_CHECK(off4.upload(o4.data(), ns, s));
        TM_CHECK(off8.upload(o8.data(), ns, s));
        TM_HIP(hipStreamSynchronize(s));
        return TM_OK;
    }

    // per-slot weights from CSR values v (optionally scaled by colscale[col] * w0)
    template <typename W>
    int values(const std::vector<double>& v, DevBuf<W>& out, hipStream_t s, const int* ci = nullptr,
               const double* colscale = nullptr, double w0 = 1.0) const {
        std::vector<W> h((size_t)h_sp[n], (W)0);
        for (int i = 0; i < n; ++i) {
            size_t k = (size_t)h_sp[i];
            for (int q = h_rp[i]; q < h_rp[i + 1]; ++q)
                if (kept(i, q)) h[k++] = (W)(colscale ? v[q] * w0 * colscale[ci[q]] : v[q]);
        }
        TM_CHECK(out.upload(h.data(), h.size(), s));
        TM_HIP(hipStreamSynchronize(s));
        return TM_OK;
    }

    // the diagonal (summed if repeated) of CSR values v, optionally scaled as in values()
    template <typename W>
    int diagonal(const std::vector<double>& v, DevBuf<W>& out, hipStream_t s, const double* colscale = nullptr,
                 double w0 = 1.0) const {
        std::vector<W> h((size_t)n, (W)0);
        for (int i = 0; i < n; ++i) {
            double d = 0.0;
            for (int q = h_rp[i]; q < h_rp[i + 1]; ++q)
                if (h_ci[q] == i) d += colscale ? v[q] * w0 * colscale[i] : v[q];
            h[i] = (W)d;
        }
        TM_CHECK(out.upload(h.data(), h.size(), s));
        TM_HIP(hipStreamSynchronize(s));
        return TM_OK;
    }

    // Chunk size for the wide kernels: 8 where it cuts the padded slots by a
    // quarter or more (the prolongators' rows, a few coarse columns each).
    // A padding slot costs a gather instruction's lanes like a real entry
    // (its load is dropped by the range check, not skipped).
    static int wide_chunk(int rows, const int* rp) {
        int64_t s16 = 0, s8 = 0;
        for (int i = 0; i < rows; ++i) {
            const int len = rp[i + 1] - rp[i];
            s16 += std::max(1, (len + 15) / 16) * 16;
            s8 += std::max(1, (len + 7) / 8) * 8;
        }
        return 4 * s8 <= 3 * s16 ? 8 : kSlot;
    }
    // long rows: more than 6 chunks of 16 on average (one block per row pays off there)
    bool long_rows() const { return n > 0 && (int64_t)h_sp[n] > (int64_t)kSlot * n * 6; }
    // every row one chunk (rows have at least one)
    bool one_chunk() const { return n > 0 && (int64_t)h_sp[n] == (int64_t)chunk * n; }
    SlotArgs args(int ebytes, const void* w) const {
        return SlotArgs{n, ncols * S * ebytes, sp.p, ebytes == 4 ? off4.p : off8.p, w};
    }
};

// C = R (I - A w D^-1): the level-0 pre-smoother from zero followed by the
// restriction, as one operator (rows of R, columns of A).  R: nr x n CSR;
// A: n x n CSR (rp, ci, v); dinv: A's inverse diagonal.
void restrict_presmoothed(const HostCsr& R, int n, const int* arp, const int* aci, const std::vector<double>& av,
                          const double* dinv, double w, HostCsr& C) {
    const int nr = (int)R.rp.size() - 1;
    C.n = nr;
    C.rp.assign((size_t)nr + 1, 0);
    C.ci.clear();
    C.v.clear();
    std::vector<double> acc((size_t)n, 0.0);
    std::vector<int> mark((size_t)n, -1), touched;
    for (int c = 0; c < nr; ++c) {
        touched.clear();
        for (int q = R.rp[c]; q < R.rp[c + 1]; ++q) {
            const int i = R.ci[q];
            const double r = R.v[q];
            for (int t = arp[i]; t < arp[i + 1]; ++t) {
                const int j = aci[t];
                if (mark[j] != c) {
                    mark[j] = c;
                    acc[j] = 0.0;
                    touched.push_back(j);
                }
                acc[j] += r * ((j == i ? 1.0 : 0.0) - av[t] * w * dinv[j]);
            }
        }
        std::sort(touched.begin(), touched.end());
        for (int j : touched) {
            C.ci.push_back(j);
            C.v.push_back(acc[j]);
        }
        C.rp[c + 1] = (int)C.ci.size();
    }
}

// ---- initial guesses from earlier solves (successive right-hand sides) ----
// Per sample as the single-sample solver: x0 = x_0 + sum_j c_j dX_j over the
// kept pairs, dB_j / dX_j backward differences at the newest pair and c the
// least-squares fit of b - b_0 (the method is described with kSolveWindow in
// tm_internal.hpp).  Window state and the small Gram systems live on the
// device (the kernels are capturable).
struct WinState {
    int head;                            // slot the next pair goes to
    int cnt;                             // pairs held (<= kSolveWindow)
    double a[kSolveWindow][kMaxS];       // x0 = sum_i a_i x_slot(i) per sample (newest first)
};

// per-block partials part[block][q][s] of the kWinDots dots, lane = sample.  The new b goes
// into the head slot as it is read: when the window is full that slot holds
// the oldest pair, whose value this thread has already loaded.
// The k-th row read is rows[k * st] (or k * st without a list), k < nr: with
// a list, every row outside it is zero in b and in the kept b (a Dirichlet
// right-hand side off the boundary); with st > 1 the dots are those of an
// evenly spread subset of the rows (a sampled least-squares fit; the kept b
// are only ever read, and written, on that subset).
__global__ __launch_bounds__(kBlock) void k_b_win_dots(int n, int S, const WinState* ws, double* __restrict__ Bw,
                                                       const double* __restrict__ b, double* part, int nr,
                                                       const int* __restrict__ rows, int st) {
    const int cnt = ws->cnt, head = ws->head;
    const size_t NS = (size_t)n * S;
    double acc[kWinDots] = {};
    const double* src[kSolveWindow];
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) src[i] = Bw + (size_t)win_slot(head, i) * NS;
    double* dst = Bw + (size_t)head * NS;
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int rstride = gridDim.x * kRowsPerBlock;
    // two rows per iteration: both rows' loads in flight together
    for (int k = __builtin_amdgcn_readfirstlane(blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave)); k < nr;
         k += 2 * rstride) {
        const bool two = k + rstride < nr;  // wavefront-uniform
        const int k2 = two ? k + rstride : k;
        const int row = rows ? rows[k * st] : k * st, row2 = rows ? rows[k2 * st] : k2 * st;
        if (on) {
            const size_t e1 = (size_t)row * S + lane, e2 = (size_t)row2 * S + lane;
            double v1[kSolveWindow], v2[kSolveWindow];
#pragma unroll
            for (int i = 0; i < kSolveWindow; ++i) {
                v1[i] = i < cnt ? src[i][e1] : 0.0;
                v2[i] = i < cnt ? src[i][e2] : 0.0;
            }
            const double b1 = b[e1], b2 = b[e2];
            const double r1 = b1 - v1[0], r2 = b2 - v2[0];
            win_differences(v1, cnt);
            win_accumulate(acc, v1, r1);
            if (two) {
                win_differences(v2, cnt);
                win_accumulate(acc, v2, r2);
            }
            dst[e1] = b1;  // after every load of the element (the head slot may be the oldest pair's)
            if (two) dst[e2] = b2;
        }
    }
    // the block's per-sample sums of all kWinDots dots, eight at a time
    __shared__ double sp[8][kRowsPerBlock][kWave];
    const int wv = threadIdx.x / kWave;
#pragma unroll
    for (int q0 = 0; q0 < kWinDots; q0 += 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (q0 + j < kWinDots) sp[j][wv][lane] = acc[q0 + j];
        __syncthreads();
        for (int i = threadIdx.x; i < 8 * kWave; i += kBlock) {
            const int j = i / kWave, sm = i % kWave;
            if (q0 + j < kWinDots && sm < S) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < kRowsPerBlock; ++k) t += sp[j][k][sm];
                part[((size_t)blockIdx.x * kWinDots + q0 + j) * S + sm] = t;  // block-major: contiguous per block
            }
        }
        __syncthreads();
    }
}

// second stage of the dots: block (q, c) sums dot q over chunk c of the
// block partials, lane = sample (rows of part are contiguous in s);
// red[c][q][s]
constexpr int kWinChunks = 8;
__global__ __launch_bounds__(kBlock) void k_b_win_red(int G, int S, const double* __restrict__ part,
                                                      double* __restrict__ red) {
    constexpr int kW = kBlock / kWave;
    const int q = blockIdx.x, c = blockIdx.y;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int sm = min(lane, S - 1);
    double v = 0.0;
    for (int g = c * kW + wv; g < G; g += kWinChunks * kW) v += part[((size_t)g * kWinDots + q) * S + sm];
    __shared__ double sr[kW][kWave];
    sr[wv][lane] = v;
    __syncthreads();
    if (wv == 0 && lane < S) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < kW; ++k) t += sr[k][lane];
        red[((size_t)c * kWinDots + q) * S + lane] = t;
    }
}

// the coefficients of sample s = blockIdx.x from the chunk sums (one wavefront)
__global__ __launch_bounds__(kWave) void k_b_win_solve(int S, WinState* ws, const double* red) {
    const int s = blockIdx.x;
    __shared__ double sd[kWinDots];
    const int q = threadIdx.x;
    if (q < kWinDots && s < S) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < kWinChunks; ++c) v += red[((size_t)c * kWinDots + q) * S + s];
        sd[q] = v;
    }
    __syncthreads();
    const int cnt = ws->cnt;
    if (threadIdx.x != 0 || s >= S || cnt == 0) return;
    double t[kWinDots];
#pragma unroll
    for (int k = 0; k < kWinDots; ++k) t[k] = sd[k];
    double a[kSolveWindow];
    win_coefficients(cnt, t, a);
    for (int i = 0; i < kSolveWindow; ++i) ws->a[i][s] = a[i];
}

// x0 = sum_i a_i x_slot(i) (nothing when the window is empty), lane = sample
// and the fixed rows' values x = b (also with an empty window)
__global__ __launch_bounds__(kBlock) void k_b_win_combine(int n, int S, const WinState* ws,
                                                          const double* __restrict__ Xw,
                                                          const unsigned char* __restrict__ fixed,
                                                          const double* __restrict__ b, double* __restrict__ x) {
    const int cnt = ws->cnt;
    const size_t NS = (size_t)n * S;
    double a[kSolveWindow];
    const double* src[kSolveWindow];
    const int ln = min((int)(threadIdx.x & (kWave - 1)), S - 1);
#pragma unroll
    for (int i = 0; i < kSolveWindow; ++i) {
        a[i] = i < cnt ? ws->a[i][ln] : 0.0;
        src[i] = Xw + (size_t)win_slot(ws->head, i) * NS;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const bool on = lane < S;
    const int rstride = gridDim.x * kRowsPerBlock;
    for (int row = __builtin_amdgcn_readfirstlane(blockIdx.x * kRowsPerBlock + (int)(threadIdx.x / kWave)); row < n;
         row += 2 * rstride) {
        const bool two = row + rstride < n;  // wavefront-uniform
        if (on) {
            const int row2 = two ? row + rstride : row;
            const size_t e1 = (size_t)row * S + lane, e2 = (size_t)row2 * S + lane;
            double acc1 = 0.0, acc2 = 0.0;
#pragma unroll
            for (int i = 0; i < kSolveWindow; ++i)
                if (i < cnt) {
                    acc1 += a[i] * src[i][e1];
                    acc2 += a[i] * src[i][e2];
                }
            if (fixed[row]) x[e1] = b[e1];
            else if (cnt > 0) x[e1] = acc1;
            if (two) {
                if (fixed[row2]) x[e2] = b[e2];
                else if (cnt > 0) x[e2] = acc2;
            }
        }
    }
}

// the solution into the head slot (its b is there already)
__global__ __launch_bounds__(kBlock) void k_b_win_store(size_t NS, const WinState* ws, const double* __restrict__ x,
                                                        double* __restrict__ Xw) {
    const size_t off = (size_t)ws->head * NS;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < NS; e += (size_t)gridDim.x * kBlock)
        Xw[off + e] = x[e];
}

// keep the pair and advance the head; a failed solve (done = 0 for some
// sample) empties the window instead
__global__ void k_b_win_advance(int S, WinState* ws, const OState* os) {
    const int s = threadIdx.x;
    __shared__ int bad;
    if (s == 0) bad = 0;
    __syncthreads();
    if (s < S && !os->done[s]) atomicOr(&bad, 1);
    __syncthreads();
    if (s != 0) return;
    if (bad) {
        ws->cnt = 0;
        ws->head = 0;
        return;
    }
    ws->head = (ws->head + 1) % kSolveWindow;
    ws->cnt = min(ws->cnt + 1, kSolveWindow);
}

template <typename T>
struct BatchSolver {
    const CgSolver* cg = nullptr;
    int n = 0, S = 0, G = 0;
    SlotOp sl0;                   // level-0 operator in slots
    DevBuf<T> wA0;                // its weights
    SlotOp sl0w;                  // the same without the diagonal, for the wide kernels
    DevBuf<T> wA0w, dA0;          // its weights and the diagonal
    DevBuf<T> wIDA0;              // on sl0: I - w D^-1 A (level-0 post-smoother, wide form)
    bool post1 = false;           // wIDA0 is set (every level-0 row fits one chunk with its diagonal)
    DevBuf<double> wA64;          // A in FP64 (outer residual)
    struct LvOps {
        SlotOp A, R, P;           // level l >= 1: A; every level: restriction, prolongation
        DevBuf<T> wA, wR, wP;
        SlotOp RA;                // R (I - A w D^-1): pre-smoothing from zero + restriction
        DevBuf<T> wRA;
    };
    std::vector<LvOps> lo;
    DevBuf<double> dinv;
    DevBuf<unsigned char> fixed;
    DevBuf<T> rT, q;
    static constexpr int kRing = 40;  // iterations per inner solve (a longer one is cut there; the outer loop goes on)
    DevBuf<T> ring;                   // (kRing + 1) x [n][S]: the preconditioned residuals z_k
    DevBuf<double> ahist, bhist;      // (kRing + 1) x kMaxS: alpha_k, beta_k
    DevBuf<double> chist;             // kRing x kMaxS: the combination coefficients c_i
    struct Lv {
        DevBuf<T> x, x2, b;
    };
    std::vector<Lv> lv;
    DevBuf<T> xc;
    DevBuf<double> part;          // 4 slots x S x G
    DevBuf<OState> ost;
    DevBuf<BState> bst;
    DevBuf<double> winX, winB;        // kSolveWindow x [n][S]: solved pairs for the initial guess
    DevBuf<WinState> win;
    bool use_window = true;
    // rows outside which every right-hand side is zero (the boundary nodes
    // of the Dirichlet problem), or null: the window's dots read only these
    const int* rhs_rows = nullptr;
    int rhs_nrows = 0;
    int win_stride = 1;  // the window's fit reads every win_stride-th of those rows
    OState* h_ost = nullptr;
    BState* h_bst = nullptr;
    static constexpr int kMaxOuter = 8;
    int last_inner[kMaxOuter] = {8, 8, 8, 8, 8, 8, 8, 8};  // per refinement step

    ~BatchSolver() {
        if (h_ost) (void)hipHostFree(h_ost);
        if (h_bst) (void)hipHostFree(h_bst);
    }

    int setup(const CgSolver& solver, int S_, hipStream_t s) {
        cg = &solver;
        n = solver.n;
        S = S_;
        G = row_grid(n);
        const AmgPrecond& M = solver.amg;
        const int L = (int)M.levels.size();
        TM_CHECK(sl0.build(n, n, solver.pat.h_rowptr.data(), solver.pat.h_csrcol.data(), S, s));
        TM_CHECK(sl0.values(solver.h_val, wA0, s));
        TM_CHECK(sl0.values(solver.h_val, wA64, s));
        if (wide() && L > 0) {
            const int* cc = solver.pat.h_csrcol.data();
            TM_CHECK(sl0w.build(n, n, solver.pat.h_rowptr.data(), cc, S, s, true));
            TM_CHECK(sl0w.values(solver.h_val, wA0w, s));
            TM_CHECK(sl0w.diagonal(solver.h_val, dA0, s));
            // the level-0 post-smoother's operator I - w D^-1 A with the
            // diagonal in the slots (rows of up to 16 entries: one chunk)
            if (sl0.one_chunk()) {
                const int* rp = solver.pat.h_rowptr.data();
                const double w = M.levels[0].omega;
                const std::vector<double>& dv = solver.h_dinv;
                std::vector<double> v2(solver.h_val.size());
                for (int i = 0; i < n; ++i)
                    for (int q = rp[i]; q < rp[i + 1]; ++q)
                        v2[q] = (cc[q] == i ? 1.0 : 0.0) - solver.h_val[q] * w * dv[i];
                TM_CHECK(sl0.values(v2, wIDA0, s));
                post1 = true;
            }
        }
        lo.clear();
        lo.resize((size_t)L);
        for (int l = 0; l < L; ++l) {
            const auto& V = M.levels[l];
            const int nnext = l + 1 < L ? M.levels[l + 1].n : M.nc;
            HostCsr h, ha;
            std::vector<double> dv;
            if (l > 0) {
                TM_CHECK(fetch_csr(V.A, ha, s));
                dv.resize((size_t)V.n);
                TM_HIP(hipMemcpyAsync(dv.data(), V.dinv.p, dv.size() * sizeof(double), hipMemcpyDeviceToHost, s));
                TM_HIP(hipStreamSynchronize(s));
                TM_CHECK(lo[l].A.build(V.n, V.n, ha.rp.data(), ha.ci.data(), S, s));
                TM_CHECK(lo[l].A.values(ha.v, lo[l].wA, s));
            }
            TM_CHECK(fetch_csr(V.R, h, s));
            TM_CHECK(lo[l].R.build(V.R.n, V.n, h.rp.data(), h.ci.data(), S, s));
            TM_CHECK(lo[l].R.values(h.v, lo[l].wR, s));
            {  // pre-smoothing from zero and restriction in one pass: R (I - A w D^-1)
                HostCsr c;
                if (l == 0)
                    restrict_presmoothed(h, n, solver.pat.h_rowptr.data(), solver.pat.h_csrcol.data(), solver.h_val,
                                         solver.h_dinv.data(), V.omega, c);
                else
                    restrict_presmoothed(h, V.n, ha.rp.data(), ha.ci.data(), ha.v, dv.data(), V.omega, c);
                TM_CHECK(lo[l].RA.build(V.R.n, V.n, c.rp.data(), c.ci.data(), S, s));
                TM_CHECK(lo[l].RA.values(c.v, lo[l].wRA, s));
            }
            TM_CHECK(fetch_csr(V.P, h, s));
            TM_CHECK(lo[l].P.build(V.P.n, nnext, h.rp.data(), h.ci.data(), S, s, false,
                                   wide() ? SlotOp::wide_chunk(V.P.n, h.rp.data()) : kSlot));
            TM_CHECK(lo[l].P.values(h.v, lo[l].wP, s));
        }
        TM_CHECK(dinv.upload(solver.h_dinv.data(), n, s));
        TM_CHECK(fixed.upload(solver.h_fixed.data(), n, s));
        const size_t NS = (size_t)n * S;
        TM_CHECK(rT.alloc(NS));
        TM_CHECK(ring.alloc((size_t)(kRing + 1) * NS));
        TM_CHECK(ahist.alloc((size_t)(kRing + 1) * kMaxS));
        TM_CHECK(bhist.alloc((size_t)(kRing + 1) * kMaxS));
        TM_CHECK(chist.alloc((size_t)(kRing + 1) * kMaxS));
        TM_CHECK(q.alloc(NS));
        lv.clear();
        lv.resize(solver.amg.levels.size());
        for (size_t l = 0; l < solver.amg.levels.size(); ++l) {
            const auto& L = solver.amg.levels[l];
            TM_CHECK(lv[l].x.alloc((size_t)L.n * S));
            TM_CHECK(lv[l].x2.alloc((size_t)L.n * S));
            TM_CHECK(lv[l].b.alloc((size_t)L.R.n * S));
        }
        TM_CHECK(xc.alloc((size_t)std::max(1, solver.amg.nc) * S));
        TM_CHECK(part.alloc((size_t)std::max(4, kWinDots) * S * G + (size_t)kWinDots * kWinChunks * S));
        TM_CHECK(winX.alloc((size_t)kSolveWindow * NS));
        TM_CHECK(winB.alloc((size_t)kSolveWindow * NS));
        TM_CHECK(win.alloc(1));
        TM_HIP(hipMemsetAsync(win.p, 0, sizeof(WinState), s));
        TM_CHECK(ost.alloc(1));
        TM_CHECK(bst.alloc(1));
        TM_HIP(hipMemsetAsync(bst.p, 0, sizeof(BState), s));  // flags past the batch size stay 0
        if (!h_ost) TM_HIP(hipHostMalloc((void**)&h_ost, sizeof(OState), hipHostMallocDefault));
        if (!h_bst) TM_HIP(hipHostMalloc((void**)&h_bst, sizeof(BState), hipHostMallocDefault));
        TM_HIP(hipStreamSynchronize(s));
        return TM_OK;
    }

    // z = M r (AMG V-cycle, batched), partial (r, z) into prz
    int precond(const T* r, T* zout, double* prz, hipStream_t s) {
        const AmgPrecond& M = cg->amg;
        const int L = (int)M.levels.size();
        if (L == 0) return TM_ERR_UNSUPPORTED;  // the batched solve needs at least one AMG level
        const int e = (int)sizeof(T);
        const T* b = r;
        for (int l = 0; l < L; ++l) {
            const auto& V = M.levels[l];
            const LvOps& O = lo[l];
            // b_{l+1} = R (I - A w D^-1) b: pre-smoothing from zero and
            // restriction in one pass over the coarse rows (the pre-smoothed
            // iterate w D^-1 b is formed by the prolongation pass)
            spmm(O.RA.long_rows(), dim3(row_grid(V.R.n)), O.RA.args(e, O.wRA.p), b, lv[l].b.p, s);
            b = lv[l].b.p;
        }
        hipLaunchKernelGGL((k_b_dense<T>), dim3(std::max(1, std::min(M.nc, 4096))), dim3(kBlock), 0, s, M.nc, S,
                           M.Cinv.p, b, xc.p, gate());
        const T* xcoarse = xc.p;
        for (int l = L - 1; l >= 0; --l) {
            const auto& V = M.levels[l];
            const LvOps& O = lo[l];
            const T* bl = l == 0 ? r : lv[l - 1].b.p;
            if (wide())
                prolong(O.P.long_rows(), O.P.chunk, dim3(row_grid(V.P.n)), O.P.args(e, O.wP.p),
                        l == 0 ? dinv.p : V.dinv.p, bl, V.omega, xcoarse, lv[l].x.p, s);
            else
                hipLaunchKernelGGL((k_b_prolong<T>), dim3(row_grid(V.P.n)), dim3(kBlock), 0, s,
                                   gated(O.P.args(e, O.wP.p)), S, l == 0 ? dinv.p : V.dinv.p, bl, V.omega, xcoarse,
                                   lv[l].x.p);
            if (l == 0) {
                if (wide() && post1) {  // z = (I - w D^-1 A) x + w D^-1 b
                    if constexpr (kWideT)
                        hipLaunchKernelGGL(k_b_post1_w, dim3(G), dim3(kBlock), 0, s, gated(sl0.args(e, wIDA0.p)),
                                           dinv.p, (const float*)bl, V.omega, (const float*)lv[0].x.p, (float*)zout,
                                           prz);
                } else if (wide())
                    post(false, dim3(G), sl0w.args(e, wA0w.p), dinv.p, bl, V.omega, lv[0].x.p, zout, prz, s, dA0.p);
                else
                    post(sl0.long_rows(), dim3(G), sl0.args(e, wA0.p), dinv.p, bl, V.omega, lv[0].x.p, zout, prz, s);
                xcoarse = zout;
            } else {
                post(O.A.long_rows(), dim3(row_grid(V.n)), O.A.args(e, O.wA.p), V.dinv.p, bl, V.omega, lv[l].x.p,
                     lv[l].x2.p, nullptr, s);
                xcoarse = lv[l].x2.p;
            }
        }
        TM_HIP(hipGetLastError());
        return TM_OK;
    }

    // the per-sample active flags of the inner solve (device address)
    const int* gate() const {
        return bst.p ? (const int*)((const char*)bst.p + offsetof(BState, active)) : nullptr;
    }
    SlotArgs gated(SlotArgs a) const {
        a.gate = gate();
        a.gate_n = S;
        return a;
    }

    // V-cycle kernels; the wide-gather forms for 64 FP32 samples
    static constexpr bool kWideT = sizeof(T) == 4;
    bool wide() const { return kWideT && S == kMaxS; }
    // one block per row for long-row operators (grid: a multiple of the XCD count)
    static dim3 blk_grid(int rows) { return dim3(xcd_grid(std::min(4096, std::max(rows, 1)))); }
    // y = M x; lg: one block per row (long rows, wide form)
    void spmm(bool lg, dim3 g, SlotArgs a, const T* x, T* y, hipStream_t s) {
        a = gated(a);
        if constexpr (kWideT)
            if (wide()) {
                const float* xf = (const float*)x;
                float* yf = (float*)y;
                if (lg) hipLaunchKernelGGL(k_b_spmm_w<true>, blk_grid(a.n), dim3(kBlock), 0, s, a, xf, yf);
                else hipLaunchKernelGGL(k_b_spmm_w<false>, g, dim3(kBlock), 0, s, a, xf, yf);
                return;
            }
        hipLaunchKernelGGL((k_b_spmm<T>), g, dim3(kBlock), 0, s, a, S, x, y);
    }
    // x = w dinv b + P xc (wide form only); ch: P's chunk size
    void prolong(bool lg, int ch, dim3 g, SlotArgs a, const double* di, const T* bb, double w, const T* xc, T* x,
                 hipStream_t s) {
        a = gated(a);
        if constexpr (kWideT) {
            const float* bf = (const float*)bb;
            const float* xf = (const float*)xc;
            if (ch == 8)
                hipLaunchKernelGGL((k_b_prolong_w<false, 8>), g, dim3(kBlock), 0, s, a, di, bf, w, xf, (float*)x);
            else if (lg)
                hipLaunchKernelGGL((k_b_prolong_w<true, kSlot>), blk_grid(a.n), dim3(kBlock), 0, s, a, di, bf, w, xf,
                                   (float*)x);
            else
                hipLaunchKernelGGL((k_b_prolong_w<false, kSlot>), g, dim3(kBlock), 0, s, a, di, bf, w, xf, (float*)x);
        }
    }
    void post(bool lg, dim3 g, SlotArgs a, const double* di, const T* bb, double w, const T* x, T* xo, double* part,
              hipStream_t s, const T* dg = nullptr) {
        a = gated(a);
        if constexpr (kWideT)
            if (wide()) {
                const float* df = (const float*)dg;
                if (lg && !part)  // partial sums are laid out per block of the row grid g
                    hipLaunchKernelGGL(k_b_post_w<true>, blk_grid(a.n), dim3(kBlock), 0, s, a, df, di,
                                       (const float*)bb, w, (const float*)x, (float*)xo, part);
                else
                    hipLaunchKernelGGL(k_b_post_w<false>, g, dim3(kBlock), 0, s, a, df, di, (const float*)bb, w,
                                       (const float*)x, (float*)xo, part);
                return;
            }
        hipLaunchKernelGGL((k_b_post<T>), g, dim3(kBlock), 0, s, a, S, di, bb, w, x, xo, part);
    }

    // initial guess of the solve of A x = b from the window of solved pairs;
    // b goes into the head slot
    void win_project(const double* b, double* x, hipStream_t s) {
        const int nr = ((rhs_rows ? rhs_nrows : n) + win_stride - 1) / win_stride;
        const int gw = std::min(row_grid(nr), 1024);  // fewer, longer blocks: half the partials, 1.2x faster at 283k rows
        hipLaunchKernelGGL(k_b_win_dots, dim3(gw), dim3(kBlock), 0, s, n, S, (const WinState*)win.p, winB.p, b,
                           part.p, nr, rhs_rows, win_stride);
        double* red = part.p + (size_t)kWinDots * S * gw;  // after the block partials
        hipLaunchKernelGGL(k_b_win_red, dim3(kWinDots, kWinChunks), dim3(kBlock), 0, s, gw, S, (const double*)part.p,
                           red);
        hipLaunchKernelGGL(k_b_win_solve, dim3(S), dim3(kWave), 0, s, S, win.p, (const double*)red);
        hipLaunchKernelGGL(k_b_win_combine, dim3(G), dim3(kBlock), 0, s, n, S, (const WinState*)win.p,
                           (const double*)winX.p, (const unsigned char*)fixed.p, b, x);
    }
    // the solution beside its b (after win_project); the window is emptied
    // when a sample did not converge
    void win_record(const double* x, hipStream_t s) {
        const size_t NS = (size_t)n * S;
        hipLaunchKernelGGL(k_b_win_store, dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, (const WinState*)win.p, x,
                           winX.p);
        hipLaunchKernelGGL(k_b_win_advance, dim3(1), dim3(kMaxS), 0, s, S, win.p, (const OState*)ost.p);
    }

    // r -= alpha q, partial (r, r)
    void cg_update(double* prr, hipStream_t s) {
        hipLaunchKernelGGL((k_b_cg_update<T>), dim3(G), dim3(kBlock), 0, s, n, S, (const T*)q.p, rT.p,
                           (const BState*)bst.p, prr);
    }
    // per-sample convergence after cg_update, before the preconditioner
    void conv_check(const double* prr, hipStream_t s) {
        hipLaunchKernelGGL(k_b_red_rr, dim3(S), dim3(kBlock), 0, s, G, S, prr, bst.p);
    }

    // r = b - A x in FP64 (b, x: [n][S]), rT = (T) r, partials (b, b), (r, r)
    void resid(const double* b, const double* x, double* pbb, double* prr, hipStream_t s) {
        hipLaunchKernelGGL((k_b_resid<T>), dim3(G), dim3(kBlock), 0, s, sl0.args(8, wA64.p), S, b, x, rT.p, pbb,
                           prr);
    }

    // iteration k: q = A p_k = A z_k + beta q, partial (z_k, q)
    void cg_spmv(const T* zk, int k, double* pzq, hipStream_t s) {
        if constexpr (kWideT)
            if (wide()) {
                hipLaunchKernelGGL(k_b_cg_spmv_w, dim3(G), dim3(kBlock), 0, s, gated(sl0w.args(4, wA0w.p)),
                                   (const float*)dA0.p, (const float*)zk, (float*)q.p, k, (const BState*)bst.p, pzq);
                return;
            }
        hipLaunchKernelGGL((k_b_cg_spmv<T>), dim3(G), dim3(kBlock), 0, s, gated(sl0.args((int)sizeof(T), wA0.p)), S,
                           zk, q.p, k, (const BState*)bst.p, pzq);
    }

    // Solves A x = b for all samples ([n][S] FP64); x is the warm start.
    int solve(const double* b, double* x, double tol, int maxit, hipStream_t s, int* outer_out, int* inner_out,
              double* maxrel) {
        if (cg->precond != TM_PRECOND_AMG || cg->amg.levels.empty()) return TM_ERR_UNSUPPORTED;
        const size_t NS = (size_t)n * S;
        double* pbb = part.p;
        double* prr = part.p + (size_t)S * G;
        double* prz = part.p + (size_t)2 * S * G;
        double* pzq = part.p + (size_t)3 * S * G;
        const double itol = sizeof(T) == 4 ? 1e-5 : 0.0;  // FP32 inner target per refinement step
        int outer = 0, inner_total = 0;
        if (use_window) win_project(b, x, s);  // (sets the fixed rows too)
        else hipLaunchKernelGGL(k_b_fix, dim3(elem_grid(NS)), dim3(kBlock), 0, s, n, S, fixed.p, b, x);
        // Host checks: after a solve that took one refinement step (the steady
        // state along a trajectory), the next one takes its first refinement
        // step without waiting for the initial residual check and enqueues
        // exactly the inner iterations the last solve needed without waiting
        // either (and no V-cycle after the last of them, whose z would go
        // unused): one host synchronisation per solve.  A sample still short of its
        // inner target after such a step is taken up by the next refinement
        // step, checked as usual, and the solve after it checks as usual too.
        bool spec_ok = spec_next;  // this solve may skip checks
        bool short_seen = false;
        bool pending = false;      // a speculative inner solve whose counts are read at the next check
        int pend_step = 0;
        for (;;) {
            resid(b, x, pbb, prr, s);
            hipLaunchKernelGGL(k_b_outer_red, dim3(S), dim3(kBlock), 0, s, G, S, pbb, prr, tol, itol, ost.p, bst.p);
            if (!(outer == 0 && spec_ok && last_inner[0] > 0)) {
                TM_HIP(hipMemcpyAsync(h_ost, ost.p, sizeof(OState), hipMemcpyDeviceToHost, s));
                TM_HIP(hipStreamSynchronize(s));
                if (pending) {
                    int used = 0;  // iterations the slowest sample ran
                    bool still = false;
                    for (int k = 0; k < S; ++k) {
                        used = std::max(used, h_bst->iters[k]);
                        still = still || h_bst->active[k];
                    }
                    last_inner[pend_step] = used;
                    inner_total += used;
                    pending = false;
                    if (still) {  // the enqueued iterations fell short: check as usual from here on
                        spec_ok = false;
                        short_seen = true;
                    }
                }
                bool all = true;
                double worst = 0.0;
                for (int k = 0; k < S; ++k) {
                    all = all && h_ost->done[k];
                    if (h_ost->bb[k] > 0) worst = std::max(worst, std::sqrt(h_ost->rr[k] / h_ost->bb[k]));
                }
                if (maxrel) *maxrel = worst;
                if (all || outer >= 8 || inner_total >= maxit) break;
            }
            ++outer;
            // inner solve A d = rT from d = 0; x += d at the end
            TM_CHECK(precond(rT.p, ring.p, prz, s));
            hipLaunchKernelGGL(k_b_red_rz, dim3(S), dim3(kBlock), 0, s, G, S, prz, -1, bst.p, bhist.p);
            int enq = 0;
            const int kmax = std::min(kRing, maxit);
            // iterations enqueued before the first check: what this refinement
            // step needed last time, plus one when the host checks after them
            // (queued iterations past convergence return at once, but their
            // launches still cost); unchecked, exactly that many: a sample
            // that needs more is taken up by the next refinement step
            const int step = std::min(outer, kMaxOuter) - 1;
            const bool spec = spec_ok && last_inner[step] > 0;
            int chunk = spec ? last_inner[step] : std::max(2, last_inner[step] + 1);
            for (;;) {
                for (int c = 0; c < chunk && enq < kmax; ++c, ++enq) {
                    T* zk = ring.p + (size_t)enq * NS;
                    cg_spmv(zk, enq, pzq, s);
                    hipLaunchKernelGGL(k_b_red_pq, dim3(S), dim3(kBlock), 0, s, G, S, pzq, bst.p,
                                       ahist.p + (size_t)enq * kMaxS);
                    cg_update(prr, s);
                    conv_check(prr, s);
                    // no iteration follows the last one of an unchecked chunk
                    // (or of the ring): its z would go unused
                    if ((spec && c + 1 == chunk) || enq + 1 == kmax) continue;
                    TM_CHECK(precond(rT.p, zk + NS, prz, s));
                    hipLaunchKernelGGL(k_b_red_rz, dim3(S), dim3(kBlock), 0, s, G, S, prz, enq, bst.p,
                                       bhist.p + (size_t)(enq + 1) * kMaxS);
                }
                TM_HIP(hipGetLastError());
                // the per-sample counts (read after the next wait for the speculative step)
                TM_HIP(hipMemcpyAsync(h_bst, bst.p, sizeof(BState), hipMemcpyDeviceToHost, s));
                if (spec) break;
                TM_HIP(hipStreamSynchronize(s));
                bool any = false;
                for (int k = 0; k < S; ++k) any = any || h_bst->active[k];
                if (!any || enq >= kmax) break;
                chunk = 2;
            }
            if (spec) {
                pending = true;
                pend_step = step;
            } else {
                int used = 0;  // iterations the slowest sample ran
                for (int k = 0; k < S; ++k) used = std::max(used, h_bst->iters[k]);
                last_inner[step] = used;
                inner_total += used;
            }
            if (S == kMaxS && enq <= kCombineK) {
                hipLaunchKernelGGL(k_b_coef, dim3(1), dim3(kMaxS), 0, s, S, enq, (const double*)ahist.p,
                                   (const double*)bhist.p, chist.p);
                hipLaunchKernelGGL((k_b_combine4<T>), dim3(elem_grid(NS / 4)), dim3(kBlock), 0, s, NS, enq,
                                   (const T*)ring.p, (const double*)chist.p, x);
            } else
                hipLaunchKernelGGL((k_b_combine<T>), dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, S, enq,
                                   (const T*)ring.p, (const double*)ahist.p, (const double*)bhist.p, x);
        }
        if (use_window) win_record(x, s);
        if (outer_out) *outer_out = outer;
        if (inner_out) *inner_out = inner_total;
        last_outer = outer;
        spec_next = outer == 1 && !short_seen;
        TM_HIP(hipGetLastError());
        bool all = true;
        for (int k = 0; k < S; ++k) all = all && h_ost->done[k];
        return all ? TM_OK : TM_ERR_MAXITER;
    }

    int last_outer = 0;  // refinement steps of the last host-polled solve
    bool spec_next = false;  // the last solve took one refinement step: the next may skip host checks

    // Times the level-0 kernels of one inner iteration and of the outer
    // residual, each launched `reps` times alone (all samples active), and
    // appends (name, ms per launch, algorithmic bytes per launch): slot
    // offsets (4 B) and weights once, every vector read or written once.
    template <class Timer>
    int time_kernels(hipStream_t s, Timer&& timeit, std::vector<tm_batch_kernel_time>& out) {
        if (cg->amg.levels.empty()) return TM_ERR_UNSUPPORTED;
        const int e = (int)sizeof(T);
        const double nS = (double)n * S;
        double* pzq = part.p + (size_t)3 * S * G;
        double* prz = part.p + (size_t)2 * S * G;
        auto add = [&out](const char* nm, double ms, double bytes) {
            tm_batch_kernel_time k = {};
            std::snprintf(k.name, sizeof(k.name), "%s", nm);
            k.ms = ms;
            k.bytes = bytes;
            out.push_back(k);
        };
        auto slots = [](const SlotOp& o) { return (double)o.h_sp[o.n]; };
        hipLaunchKernelGGL(k_b_activate, dim3(1), dim3(kMaxS), 0, s, S, bst.p);
        const SlotOp& A0 = wide() ? sl0w : sl0;
        add(wide() ? "k_b_cg_spmv_w" : "k_b_cg_spmv", timeit([&] { cg_spmv(ring.p, 0, pzq, s); }),
            slots(A0) * (4 + e) + (wide() ? (double)n * e : 0.0) + 3.0 * nS * e);
        add(sizeof(T) == 4 ? "k_b_cg_update<float>" : "k_b_cg_update<double>",
            timeit([&] { cg_update(part.p + (size_t)S * G, s); }), 3.0 * nS * e);
        const LvOps& O = lo[0];
        const int n1 = cg->amg.levels[0].R.n;
        add(O.RA.long_rows() ? "k_b_spmm_w<long> (R(I-AwD^-1), level 0)" : "k_b_spmm (R(I-AwD^-1), level 0)",
            timeit([&] { spmm(O.RA.long_rows(), dim3(row_grid(n1)), O.RA.args(e, O.wRA.p), rT.p, lv[0].b.p, s); }),
            slots(O.RA) * (4 + e) + nS * e + (double)n1 * S * e);
        const auto& V0 = cg->amg.levels[0];
        add(wide() ? "k_b_prolong_w (level 0)" : "k_b_prolong (level 0)", timeit([&] {
                if (wide())
                    prolong(O.P.long_rows(), O.P.chunk, dim3(row_grid(V0.P.n)), O.P.args(e, O.wP.p), dinv.p, rT.p,
                            V0.omega, lv[0].b.p, lv[0].x.p, s);
                else
                    hipLaunchKernelGGL((k_b_prolong<T>), dim3(row_grid(V0.P.n)), dim3(kBlock), 0, s,
                                       gated(O.P.args(e, O.wP.p)), S, dinv.p, rT.p, V0.omega, lv[0].b.p, lv[0].x.p);
            }),
            slots(O.P) * (4 + e) + 8.0 * n + 2.0 * nS * e + (double)n1 * S * e);
        if constexpr (kWideT) {
            if (wide() && post1) {
                add("k_b_post1_w (level 0)", timeit([&] {
                        hipLaunchKernelGGL(k_b_post1_w, dim3(G), dim3(kBlock), 0, s, gated(sl0.args(e, wIDA0.p)),
                                           dinv.p, (const float*)rT.p, V0.omega, (const float*)lv[0].x.p,
                                           (float*)ring.p, prz);
                    }),
                    slots(sl0) * (4 + e) + 8.0 * n + 3.0 * nS * e);
            }
        }
        double* pbb = part.p;
        double* prr = part.p + (size_t)S * G;
        add(sizeof(T) == 4 ? "k_b_resid<float>" : "k_b_resid<double>", timeit([&] {
                resid(lv0_b64(), lv0_x64(), pbb, prr, s);
            }),
            slots(sl0) * (4 + 8) + 2.0 * nS * 8 + nS * e);
        return TM_OK;
    }
    // FP64 [n][S] operands of the outer residual for kernel timing (the
    // solver's own FP64 vectors are the caller's; ring storage is large enough)
    const double* lv0_b64() const { return (const double*)ring.p; }
    const double* lv0_x64() const { return (const double*)(ring.p + (size_t)2 * n * S * 8 / sizeof(T)); }

    // The same solve with no host synchronisation (capturable in a HIP graph):
    // exactly `nouter` refinement steps of ninner[o] inner iterations each.
    // Every solver kernel is gated on the per-sample flags, so refinement
    // steps and iterations past convergence cost only their launches, and
    // iterations a sample did not take carry alpha = beta = 0 into the
    // combination.  A final residual check adds the number of samples that
    // missed the tolerance to *fail on the device.
    int enqueue(const double* b, double* x, double tol, int nouter, const int* ninner, hipStream_t s,
                unsigned* fail) {
        if (cg->precond != TM_PRECOND_AMG || cg->amg.levels.empty()) return TM_ERR_UNSUPPORTED;
        if (nouter < 0 || nouter > kMaxOuter || !fail) return TM_ERR_BAD_ARG;
        const size_t NS = (size_t)n * S;
        double* pbb = part.p;
        double* prr = part.p + (size_t)S * G;
        double* prz = part.p + (size_t)2 * S * G;
        double* pzq = part.p + (size_t)3 * S * G;
        const double itol = sizeof(T) == 4 ? 1e-5 : 0.0;
        if (use_window) win_project(b, x, s);  // (sets the fixed rows too)
        else hipLaunchKernelGGL(k_b_fix, dim3(elem_grid(NS)), dim3(kBlock), 0, s, n, S, fixed.p, b, x);
        for (int o = 0;; ++o) {
            resid(b, x, pbb, prr, s);
            hipLaunchKernelGGL(k_b_outer_red, dim3(S), dim3(kBlock), 0, s, G, S, pbb, prr, tol, itol, ost.p, bst.p);
            if (o == nouter) break;
            TM_CHECK(precond(rT.p, ring.p, prz, s));
            hipLaunchKernelGGL(k_b_red_rz, dim3(S), dim3(kBlock), 0, s, G, S, prz, -1, bst.p, bhist.p);
            const int K = std::max(1, std::min(kRing, ninner[o]));
            for (int k = 0; k < K; ++k) {
                T* zk = ring.p + (size_t)k * NS;
                cg_spmv(zk, k, pzq, s);
                hipLaunchKernelGGL(k_b_red_pq, dim3(S), dim3(kBlock), 0, s, G, S, pzq, bst.p,
                                   ahist.p + (size_t)k * kMaxS);
                cg_update(prr, s);
                conv_check(prr, s);
                if (k + 1 == K) break;  // the last z would go unused
                TM_CHECK(precond(rT.p, zk + NS, prz, s));
                hipLaunchKernelGGL(k_b_red_rz, dim3(S), dim3(kBlock), 0, s, G, S, prz, k, bst.p,
                                   bhist.p + (size_t)(k + 1) * kMaxS);
            }
            if (S == kMaxS && K <= kCombineK) {
                hipLaunchKernelGGL(k_b_coef, dim3(1), dim3(kMaxS), 0, s, S, K, (const double*)ahist.p,
                                   (const double*)bhist.p, chist.p);
                hipLaunchKernelGGL((k_b_combine4<T>), dim3(elem_grid(NS / 4)), dim3(kBlock), 0, s, NS, K,
                                   (const T*)ring.p, (const double*)chist.p, x);
            } else
                hipLaunchKernelGGL((k_b_combine<T>), dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, S, K,
                                   (const T*)ring.p, (const double*)ahist.p, (const double*)bhist.p, x);
        }
        hipLaunchKernelGGL(k_b_count_undone, dim3(1), dim3(kMaxS), 0, s, S, (const OState*)ost.p, fail);
        if (use_window) win_record(x, s);
        TM_HIP(hipGetLastError());
        return TM_OK;
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
    SlotOp eo;                         // union pattern of the per-node operators, in row slots
    DevBuf<double> xw, gx, gy, gz, tx, ty, tz, cx, cy, cz;  // slot weights: exchange, grad, tGrad, curl
    bool exch = false, dmi = false;
    DevBuf<double> m, dmdt, heff, u1, u2, usum, rhsN, rhsD, U, Gm, hext, torque;
    DevBuf<double> m2;                 // the next magnetisation of tm_batch_eval_euler (swapped with m)
    DevBuf<float> Bf, Uf, Gf;          // FP32 boundary product (TM_PREC_MIXED_BEM32)
    DevBuf<float> Pf;                  // its split-K partial products (nchunk x b x S)
    int kchunk = 512, nchunk = 0;
    // int8 digit-slice boundary product (TM_PREC_MIXED_BEMI8): B's digits
    // (Bq, 5 bytes per entry), the per-call digits of U, the row and sample
    // maxima, the split-K partials, and the tuned tile rows / K split
    DevBuf<int> Bq, Uq;
    DevBuf<double> i8rmax, i8umax, i8P;
    int i8nT = 0, i8nC = 0, i8mt = 1, i8ks = 1, i8candidates = 0;
    float i8ms = 0.0f, i8ms_first = 0.0f;
    DevBuf<double> rk_k[4], rk_base;   // RK4 stages and the step's start state
    // graph-replayed RK4 steps: samples whose fixed-budget solve missed the
    // tolerance (device counter), and the budgets learnt from host-polled
    // solves: refinement steps and inner iterations per step of the
    // Neumann (0) and Dirichlet (1) solves
    DevBuf<unsigned> rk_fail;
    bool bud_set = false;
    int bud_outer[2] = {0, 0};
    int bud_inner[2][8] = {};
    int gemm_sol = 0;                  // tuned rocBLAS solution index (0: library default)
    bool gemm_own = false;             // the MFMA kernel k_b_gemm64 won the tuning
    int gemm_shape = 0, gemm_ks = 1;   // its tile shape and K split
    int gemm_own_candidates = 0;       // (shape, K split) pairs timed
    bool own_ok = false;               // an own-kernel candidate agreed with rocBLAS (used in captured graphs)
    float own_ms = 0.0f;
    int own_shape = 0, own_ks = 1;
    DevBuf<double> Pd;                 // its split-K partials (ks x b x S)
    double gemm_own_err = 0.0;         // its largest relative difference to rocBLAS at tuning
    int gemm_candidates = 0;
    float gemm_ms_default = 0.0f, gemm_ms_best = 0.0f;
    BatchSolver<double> sN64, sD64;
    BatchSolver<float> sN32, sD32;
    bool timers = false;
    hipEvent_t ev[6] = {};
    tm_phase_times tsum = {};
    // in-pipeline event pairs around the boundary GEMM (tm_batch_gemm_live)
    static constexpr int kLivePairs = 256;
    hipEvent_t lev[2 * kLivePairs] = {};
    int lev_n = 0;
    bool lev_on = false;
    // marks the start (0) or end (1) of a boundary GEMM launch while recording
    void live_mark(int which) {
        if (!lev_on || lev_n >= kLivePairs) return;
        (void)hipEventRecord(lev[2 * lev_n + which], stream);
        if (which == 1) ++lev_n;
    }
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

// Slot weights of one operator; with colscale, entry (i, j) is scaled by w0 colscale[j].
int upload_op(tm_batch* bt, int op, DevBuf<double>& dst, const double* colscale = nullptr, double w0 = 1.0) {
    std::vector<double> v;
    TM_CHECK(csr_on_pattern(*bt->cv.pat, *bt->cv.hop_rp[op], *bt->cv.hop_col[op], *bt->cv.hop_val[op], v));
    return bt->eo.values(v, dst, bt->stream, bt->cv.pat->h_csrcol.data(), colscale, w0);
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
                        S, be, G, t, b, G, t, b, t, sol != 0 ? rocblas_gemm_algo_solution_index : rocblas_gemm_algo_standard,
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
    // every candidate (rocBLAS and own) is timed on the same non-zero operand:
    // matrix cores multiplying zeros draw less power and clock higher, which
    // would favour whichever candidates were timed on a zero U
    hipLaunchKernelGGL(k_b_fill_pattern, dim3(elem_grid((size_t)b * S)), dim3(kBlock), 0, bt->stream, (size_t)b * S,
                       bt->U.p);
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
                    if (!bt->own_ok || ms < bt->own_ms) {  // the fastest verified own kernel, for graphs
                        bt->own_ok = true;
                        bt->own_ms = ms;
                        bt->own_shape = bt->gemm_shape;
                        bt->own_ks = bt->gemm_ks;
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

// ---- int8 digit-slice boundary product: set-up, tuning, one product --------
inline const i32x4* i8ops(const DevBuf<int>& q) { return reinterpret_cast<const i32x4*>(q.p); }

int i8_gemm(tm_batch* bt, int mt, int ks) {
    const int cper = (bt->i8nC + ks - 1) / ks;
    const dim3 g((bt->i8nT + 4 * mt - 1) / (4 * mt), ks);
    launch_gemm_i8(mt, g, bt->stream, bt->cv.bnx, bt->i8nT, bt->i8nC, cper, i8ops(bt->Bq), i8ops(bt->Uq), bt->i8P.p);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

// U's digits from the boundary values u[bn[k]][s] (u[k][s] without bn)
int i8_uslice(tm_batch* bt, const int* bn, const double* u) {
    const int b = bt->cv.bnx, S = bt->S;
    hipStream_t s = bt->stream;
    hipLaunchKernelGGL(k_b_i8_zero, dim3(1), dim3(64), 0, s, 64, bt->i8umax.p);
    hipLaunchKernelGGL(k_b_i8_umax, dim3(std::min(1024, (b + 3) / 4)), dim3(kBlock), 0, s, b, S, bn, u,
                       bt->i8umax.p);
    hipLaunchKernelGGL(k_b_i8_uslice, dim3(elem_grid((size_t)bt->i8nC * 256)), dim3(kBlock), 0, s, b, S, bt->i8nC,
                       bn, u, (const double*)bt->i8umax.p, reinterpret_cast<i32x4*>(bt->Uq.p));
    TM_HIP(hipGetLastError());
    return TM_OK;
}

int i8_product(tm_batch* bt, const int* bn, const double* u, double* out) {
    TM_CHECK(i8_uslice(bt, bn, u));
    bt->live_mark(0);
    TM_CHECK(i8_gemm(bt, bt->i8mt, bt->i8ks));
    bt->live_mark(1);
    hipLaunchKernelGGL(k_b_i8_finish, dim3(elem_grid((size_t)bt->cv.bnx * 64)), dim3(kBlock), 0, bt->stream,
                       bt->cv.bnx, bt->S, bt->i8ks, bn, (const double*)bt->i8P.p, (const double*)bt->i8rmax.p,
                       (const double*)bt->i8umax.p, out);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

constexpr int kI8Splits[] = {1, 2, 3, 4, 6, 8};
constexpr int kI8MaxSplit = 8;

// B's digits (once per batch), then the tile rows (1 or 2 tiles of 16 per
// wavefront) and K split by timing each candidate on the digits of a
// non-zero U
int i8_setup(tm_batch* bt) {
    const CtxView& v = bt->cv;
    const int b = v.bnx;
    hipStream_t s = bt->stream;
    bt->i8nT = (b + 15) / 16;
    bt->i8nC = (b + 63) / 64;
    const size_t nT = bt->i8nT, nC = bt->i8nC;
    TM_CHECK(bt->Bq.alloc(nT * nC * kI8Digits * kI8Op * 4));
    TM_CHECK(bt->Uq.alloc(nC * kI8DigitsU * 4 * kI8Op * 4));
    TM_CHECK(bt->i8rmax.alloc(b));
    TM_CHECK(bt->i8umax.alloc(64));
    TM_CHECK(bt->i8P.alloc((size_t)kI8MaxSplit * b * 64));
    TM_HIP(hipMemsetAsync(bt->i8rmax.p, 0, sizeof(double) * b, s));
    const int kper = 1024;
    hipLaunchKernelGGL(k_b_i8_rowmax, dim3((b + kBlock - 1) / kBlock, (b + kper - 1) / kper), dim3(kBlock), 0, s, b,
                       v.bld, v.B, kper, bt->i8rmax.p);
    hipLaunchKernelGGL(k_b_i8_bslice, dim3(elem_grid(nT * nC * 64)), dim3(kBlock), 0, s, b, v.bld, v.B, bt->i8nT,
                       bt->i8nC, (const double*)bt->i8rmax.p, reinterpret_cast<i32x4*>(bt->Bq.p));
    TM_HIP(hipGetLastError());
    // tuning operand: a non-zero pattern as the boundary values
    hipLaunchKernelGGL(k_b_fill_pattern, dim3(elem_grid((size_t)b * bt->S)), dim3(kBlock), 0, s, (size_t)b * bt->S,
                       bt->rhsN.p);
    TM_CHECK(i8_uslice(bt, nullptr, bt->rhsN.p));
    hipEvent_t e0, e1;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    // each candidate timed alone five times (the median counts), over two
    // rounds in which the clock has settled under the MFMA load; a single
    // launch's time varies by a few per cent
    struct Cand {
        int mt, ks;
        float ms;
        std::vector<float> t;
    };
    std::vector<Cand> cands;
    for (int mt = 1; mt <= 2; ++mt)
        for (int ks : kI8Splits) {
            const int cper = (bt->i8nC + ks - 1) / ks;
            if (ks > bt->i8nC || (size_t)cper * 64 > (size_t)kI8MaxK) continue;  // int32 bound per K range
            cands.push_back({mt, ks, 1e30f, {}});
        }
    const int cand = (int)cands.size();
    for (int round = 0; round < 2; ++round)
        for (Cand& c : cands)
            for (int r = 0; r < 5; ++r) {
                (void)hipEventRecord(e0, s);
                if (i8_gemm(bt, c.mt, c.ks) != TM_OK) break;
                (void)hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess) break;
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (round == 1) c.t.push_back(ms);
            }
    (void)hipGetLastError();
    for (Cand& c : cands)
        if (!c.t.empty()) {
            std::sort(c.t.begin(), c.t.end());
            c.ms = c.t[c.t.size() / 2];
        }
    float best = 1e30f;
    for (const Cand& c : cands)
        if (c.ms < best) {
            best = c.ms;
            bt->i8mt = c.mt;
            bt->i8ks = c.ks;
        }
    if (cand > 0) bt->i8ms_first = cands[0].ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    bt->i8candidates = cand;
    bt->i8ms = best;
    if (cand == 0) return TM_ERR_UNSUPPORTED;  // K too long for one int32 range at any split
    TM_HIP(hipMemsetAsync(bt->rhsN.p, 0, sizeof(double) * (size_t)bt->n * bt->S, s));
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
    if (bt->prec == TM_PREC_MIXED_BEMI8) {
        TM_CHECK(i8_product(bt, bn, u, out));
    } else if (bt->prec == TM_PREC_MIXED_BEM32) {
        hipLaunchKernelGGL(k_b_gather<float>, g, dim3(kBlock), 0, s, b, S, bn, u, bt->Uf.p);
        bt->live_mark(0);
        TM_CHECK(gemm_split_f32(bt));
        bt->live_mark(1);
        hipLaunchKernelGGL(k_b_chunk_sum, g, dim3(kBlock), 0, s, b, S, bt->nchunk, bn, (const float*)bt->Pf.p, out);
    } else {
        hipLaunchKernelGGL(k_b_gather<double>, g, dim3(kBlock), 0, s, b, S, bn, u, bt->U.p);
        if (bt->gemm_own) {
            bt->live_mark(0);
            TM_CHECK(gemm_own_run(bt, nullptr));
            bt->live_mark(1);
            hipLaunchKernelGGL(k_b_ksum_scatter, g, dim3(kBlock), 0, s, b, S, bt->gemm_ks, bn,
                               (const double*)bt->Pd.p, out);
        } else {
            bt->live_mark(0);
            TM_CHECK(gemm_run(bt, bt->gemm_sol));
            bt->live_mark(1);
            hipLaunchKernelGGL(k_b_scatter<double>, g, dim3(kBlock), 0, s, b, S, bn, (const double*)bt->Gm.p, out);
        }
    }
    TM_HIP(hipGetLastError());
    return TM_OK;
}

// b = div(Js m) of every sample into the Neumann right-hand side
void launch_div(tm_batch* bt) {
    const int n = bt->n, S = bt->S;
    hipLaunchKernelGGL(k_b_div, dim3(row_grid(n)), dim3(kBlock), 0, bt->stream, bt->eo.args(8, bt->tx.p),
                       (const double*)bt->ty.p, (const double*)bt->tz.p, S, (const double*)bt->m.p, bt->rhsN.p);
}

// The fused field + LLG kernel of every (node, sample) on the batch
// magnetisation and the potentials in usum (TheLLG.cpp:238-254,
// EffFieldGPU.cu:316-362); with torque, the per-sample maximum torque too.
int launch_fields(tm_batch* bt, double alpha, int variant, double* dmdt, double* heff, bool torque,
                  double* mnext = nullptr, double dt = 0.0) {
    const CtxView& v = bt->cv;
    const int n = bt->n, S = bt->S;
    hipStream_t s = bt->stream;
    BEvalArgs a = {};
    a.P = bt->eo.args(8, bt->exch ? bt->xw.p : nullptr);
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
    a.u = bt->usum.p;
    a.invjs = v.invjs;
    a.ku = v.has_ku ? v.ku : nullptr;
    a.kaxis = v.has_ku ? v.kaxis : nullptr;
    a.hext = bt->hext.p;
    a.hfield = ctx_hext_field(bt->ctx);  // read per call: tm_ctx_set_zeeman may change it after creation
    a.alpha = alpha;
    a.variant = variant;
    a.dmdt = dmdt;
    a.heff = heff;
    a.torque = torque ? bt->torque.p : nullptr;
    a.mnext = mnext;
    a.dt = dt;
    if (torque) hipLaunchKernelGGL(k_b_zero_small, dim3(1), dim3(kMaxS), 0, s, S, bt->torque.p);
    const dim3 g(row_grid(n));
    if (S == kMaxS) {
        if (bt->dmi) hipLaunchKernelGGL(k_b_eval_w<true>, g, dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL(k_b_eval_w<false>, g, dim3(kBlock), 0, s, a);
    } else if (bt->dmi) {
        hipLaunchKernelGGL(k_b_eval<true>, g, dim3(kBlock), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_b_eval<false>, g, dim3(kBlock), 0, s, a);
    }
    TM_HIP(hipGetLastError());
    return TM_OK;
}

// One evaluation of the batch magnetisation into dmdt with fixed-budget
// solves (bt->bud_*): no host synchronisation, so it can be captured in a
// HIP graph.  Samples whose solve missed the tolerance are counted into
// bt->rk_fail on the device.
int eval_enqueue(tm_batch* bt, double alpha, int variant, double* dmdt) {
    const int n = bt->n, S = bt->S;
    const size_t NS = (size_t)n * S;
    hipStream_t s = bt->stream;
    const CtxView& v = bt->cv;
    launch_div(bt);
    TM_HIP(hipGetLastError());
    if (bt->prec == TM_PREC_FP64)
        TM_CHECK(bt->sN64.enqueue(bt->rhsN.p, bt->u1.p, v.tol, bt->bud_outer[0], bt->bud_inner[0], s, bt->rk_fail.p));
    else
        TM_CHECK(bt->sN32.enqueue(bt->rhsN.p, bt->u1.p, v.tol, bt->bud_outer[0], bt->bud_inner[0], s, bt->rk_fail.p));
    TM_CHECK(boundary_product(bt, v.bnodes, bt->u1.p, bt->rhsD.p));
    if (bt->prec == TM_PREC_FP64)
        TM_CHECK(bt->sD64.enqueue(bt->rhsD.p, bt->u2.p, v.tol, bt->bud_outer[1], bt->bud_inner[1], s, bt->rk_fail.p));
    else
        TM_CHECK(bt->sD32.enqueue(bt->rhsD.p, bt->u2.p, v.tol, bt->bud_outer[1], bt->bud_inner[1], s, bt->rk_fail.p));
    hipLaunchKernelGGL(k_b_sum2, dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, (const double*)bt->u1.p,
                       (const double*)bt->u2.p, bt->usum.p);
    return launch_fields(bt, alpha, variant, dmdt, nullptr, false);
}

}  // namespace

extern "C" {

int tm_batch_create(tm_ctx* c, int nsamples, int precision, tm_batch** out) {
    if (!c || !out || nsamples < 1 || nsamples > kMaxS) return TM_ERR_BAD_ARG;
    if (precision != TM_PREC_FP64 && precision != TM_PREC_MIXED && precision != TM_PREC_MIXED_BEM32 &&
        precision != TM_PREC_MIXED_BEMI8)
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
    // the field kernels address the three components of m and the potentials
    // with 32-bit byte offsets
    if ((int64_t)3 * P.n * nsamples * 8 > (int64_t)INT32_MAX) return fail(TM_ERR_UNSUPPORTED);
    if ((st = bt->eo.build(P.n, P.n, P.h_rowptr.data(), P.h_csrcol.data(), nsamples, bt->stream)) != TM_OK)
        return fail(st);
    bt->exch = v.hop_set[TM_OP_EXCHANGE];
    int ops_ok = TM_OK;
    if (bt->exch) ops_ok |= upload_op(bt, TM_OP_EXCHANGE, bt->xw);
    ops_ok |= upload_op(bt, TM_OP_GRAD_X, bt->gx);
    ops_ok |= upload_op(bt, TM_OP_GRAD_Y, bt->gy);
    ops_ok |= upload_op(bt, TM_OP_GRAD_Z, bt->gz);
    {  // the divergence weights carry Js of their column and 1/mu0 (the Js of the context now)
        std::vector<double> jsh((size_t)P.n);
        if (hipMemcpyAsync(jsh.data(), v.js, sizeof(double) * P.n, hipMemcpyDeviceToHost, bt->stream) != hipSuccess ||
            hipStreamSynchronize(bt->stream) != hipSuccess)
            return fail(TM_ERR_HIP);
        ops_ok |= upload_op(bt, TM_OP_TGRAD_X, bt->tx, jsh.data(), 1.0 / kMu0);
        ops_ok |= upload_op(bt, TM_OP_TGRAD_Y, bt->ty, jsh.data(), 1.0 / kMu0);
        ops_ok |= upload_op(bt, TM_OP_TGRAD_Z, bt->tz, jsh.data(), 1.0 / kMu0);
    }
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
    a |= bt->usum.alloc(NS);
    a |= bt->rhsN.alloc(NS);
    a |= bt->rhsD.alloc(NS);
    if (precision == TM_PREC_MIXED_BEM32) {
        bt->nchunk = (v.bnx + bt->kchunk - 1) / bt->kchunk;
        a |= bt->Uf.alloc((size_t)v.bnx * nsamples);
        a |= bt->Gf.alloc((size_t)v.bnx * nsamples);
        a |= bt->Bf.alloc((size_t)v.bld * v.bnx);
        a |= bt->Pf.alloc((size_t)bt->nchunk * v.bnx * nsamples);
    } else if (precision != TM_PREC_MIXED_BEMI8) {
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
    // the Dirichlet right-hand side is zero off the boundary nodes (Laplace,
    // no source; the boundary product scatters into those rows only)
    bt->sD64.rhs_rows = bt->sD32.rhs_rows = v.bnodes;
    bt->sD64.rhs_nrows = bt->sD32.rhs_nrows = v.bnx;
    bt->sN64.win_stride = bt->sN32.win_stride = 8;
    bt->sD64.win_stride = bt->sD32.win_stride = 4;
    if (hipStreamSynchronize(bt->stream) != hipSuccess) return fail(TM_ERR_HIP);
    if (precision == TM_PREC_MIXED_BEMI8) {
        if ((st = i8_setup(bt)) != TM_OK) return fail(st);
    } else if ((st = tune_gemm(bt)) != TM_OK) {
        return fail(st);
    }
    if (hipStreamSynchronize(bt->stream) != hipSuccess) return fail(TM_ERR_HIP);
    *out = bt;
    return TM_OK;
}

int tm_batch_destroy(tm_batch* bt) {
    if (!bt) return TM_OK;
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    if (bt->blas) rocblas_destroy_handle(bt->blas);
    for (auto& e : bt->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : bt->lev)
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

int tm_batch_gemm_live(tm_batch* bt, int on, double* ms_mean, int* count) {
    if (!bt || (on != 0 && on != 1)) return TM_ERR_BAD_ARG;
    if (on) {
        if (!bt->lev[0])
            for (auto& e : bt->lev) TM_HIP(hipEventCreate(&e));
        bt->lev_n = 0;
        bt->lev_on = true;
        return TM_OK;
    }
    bt->lev_on = false;
    double sum = 0.0;
    if (bt->lev_n > 0) TM_HIP(hipEventSynchronize(bt->lev[2 * bt->lev_n - 1]));
    for (int k = 0; k < bt->lev_n; ++k) {
        float ms = 0.0f;
        TM_HIP(hipEventElapsedTime(&ms, bt->lev[2 * k], bt->lev[2 * k + 1]));
        sum += ms;
    }
    if (ms_mean) *ms_mean = bt->lev_n > 0 ? sum / bt->lev_n : 0.0;
    if (count) *count = bt->lev_n;
    return TM_OK;
}

int tm_batch_get_timers(tm_batch* bt, tm_phase_times* out, int reset) {
    if (!bt || !out) return TM_ERR_BAD_ARG;
    *out = bt->tsum;
    if (reset) bt->tsum = tm_phase_times{};
    return TM_OK;
}

}  // extern "C"

namespace {

// tm_batch_eval_rhs; with euler_dt > 0 the field kernel also writes the
// Euler step of m into m2, which then becomes the batch's m
int batch_eval(tm_batch* bt, const double* m, int mwhere, double alpha, int variant, double* dmdt, int dwhere,
               double* heff, int hwhere, tm_batch_stats* stats, double euler_dt) {
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
    launch_div(bt);
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
    // total potential, gathered once per entry by the field kernel
    hipLaunchKernelGGL(k_b_sum2, dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, (const double*)bt->u1.p,
                       (const double*)bt->u2.p, bt->usum.p);
    // fused fields + LLG
    double* dout = dwhere == TM_DEVICE ? dmdt : bt->dmdt.p;
    double* hout = heff ? (hwhere == TM_DEVICE ? heff : bt->heff.p) : nullptr;
    if (euler_dt > 0.0) TM_CHECK(bt->m2.alloc(3 * NS));
    TM_CHECK(launch_fields(bt, alpha, variant, dout, hout, stats != nullptr, euler_dt > 0.0 ? bt->m2.p : nullptr,
                           euler_dt));
    if (euler_dt > 0.0) std::swap(bt->m, bt->m2);
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
    if (dwhere == TM_HOST) TM_CHECK(copy_out(dmdt, TM_HOST, dout, 3 * NS, s));
    if (heff && hwhere == TM_HOST) TM_CHECK(copy_out(heff, TM_HOST, hout, 3 * NS, s));
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

}  // namespace

extern "C" {

int tm_batch_eval_rhs(tm_batch* bt, const double* m, int mwhere, double alpha, int variant, double* dmdt, int dwhere,
                      double* heff, int hwhere, tm_batch_stats* stats) {
    return batch_eval(bt, m, mwhere, alpha, variant, dmdt, dwhere, heff, hwhere, stats, 0.0);
}

int tm_batch_eval_euler(tm_batch* bt, double alpha, int variant, double dtau, double* dmdt, int dwhere,
                        tm_batch_stats* stats) {
    if (!(dtau > 0.0)) return TM_ERR_BAD_ARG;
    return batch_eval(bt, nullptr, TM_DEVICE, alpha, variant, dmdt, dwhere, nullptr, TM_DEVICE, stats, dtau);
}

int tm_batch_run_rk4(tm_batch* bt, double dtau, int nsteps, double alpha, int variant, int use_graph,
                     tm_batch_stats* last, tm_rk4_stats* rk) {
    if (!bt || nsteps < 0 || !(dtau > 0.0) || (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC))
        return TM_ERR_BAD_ARG;
    const size_t NS = (size_t)bt->n * bt->S, N3 = 3 * NS;
    for (auto& k : bt->rk_k) TM_CHECK(k.alloc(N3));
    TM_CHECK(bt->rk_base.alloc(N3));
    TM_CHECK(bt->rk_fail.alloc(1));
    hipStream_t s = bt->stream;
    const dim3 g(elem_grid(N3));
    const double cst[4] = {0.0, 0.5 * dtau, 0.5 * dtau, dtau};
    // graphs need kernel-only boundary products (the own MFMA GEMM); the
    // rocBLAS paths run eagerly
    const bool graph_ok =
        use_graph && (bt->prec == TM_PREC_MIXED_BEMI8 || (bt->prec != TM_PREC_MIXED_BEM32 && bt->own_ok));
    tm_rk4_stats R = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    TM_HIP(hipEventRecord(e0, s));
    auto learn = [bt, &R]() {  // budgets from the last host-polled solves
        const int on = bt->prec == TM_PREC_FP64 ? bt->sN64.last_outer : bt->sN32.last_outer;
        const int od = bt->prec == TM_PREC_FP64 ? bt->sD64.last_outer : bt->sD32.last_outer;
        const int* inN = bt->prec == TM_PREC_FP64 ? bt->sN64.last_inner : bt->sN32.last_inner;
        const int* inD = bt->prec == TM_PREC_FP64 ? bt->sD64.last_inner : bt->sD32.last_inner;
        const int o[2] = {on, od};
        const int* in[2] = {inN, inD};
        for (int w = 0; w < 2; ++w) {
            bt->bud_outer[w] = std::min(8, std::max(bt->bud_outer[w], o[w]));
            for (int k = 0; k < o[w] && k < 8; ++k) bt->bud_inner[w][k] = std::max(bt->bud_inner[w][k], in[w][k] + 2);
            int tot = 0;
            for (int k = 0; k < bt->bud_outer[w]; ++k) tot += bt->bud_inner[w][k];
            if (w == 0) R.budget_neumann = tot;
            else R.budget_dirichlet = tot;
        }
        bt->bud_set = true;
    };
    auto polled_step = [&](bool final_step) -> int {
        int st = TM_OK;
        hipLaunchKernelGGL(k_b_copy, g, dim3(kBlock), 0, s, N3, (const double*)bt->m.p, bt->rk_base.p);
        for (int k = 0; k < 4; ++k) {
            if (k > 0)
                hipLaunchKernelGGL(k_b_axpy, g, dim3(kBlock), 0, s, N3, (const double*)bt->rk_base.p,
                                   (const double*)bt->rk_k[k - 1].p, cst[k], bt->m.p);
            tm_batch_stats es = {};
            const bool fin = final_step && k == 3;
            const int r = tm_batch_eval_rhs(bt, nullptr, TM_DEVICE, alpha, variant, bt->rk_k[k].p, TM_DEVICE, nullptr,
                                            TM_DEVICE, &es);
            if (r != TM_OK && r != TM_ERR_MAXITER) return r;
            if (r == TM_ERR_MAXITER) st = r;
            R.max_iters_neumann = std::max(R.max_iters_neumann, es.inner_neumann);
            R.max_iters_dirichlet = std::max(R.max_iters_dirichlet, es.inner_dirichlet);
            if (fin && last) *last = es;
            learn();
        }
        hipLaunchKernelGGL(k_b_rk_final, dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS, (const double*)bt->rk_base.p,
                           (const double*)bt->rk_k[0].p, (const double*)bt->rk_k[1].p, (const double*)bt->rk_k[2].p,
                           (const double*)bt->rk_k[3].p, dtau, bt->m.p);
        TM_HIP(hipGetLastError());
        R.sync_steps += 1;
        return st;
    };
    int status = TM_OK;
    hipGraphExec_t exec = nullptr;
    if (last) *last = tm_batch_stats{};
    for (int step = 0; step < nsteps; ++step) {
        const bool fin = step == nsteps - 1;
        if (graph_ok && bt->bud_set) {
            if (!exec) {
                hipGraph_t gr = nullptr;
                // the captured boundary products use the fastest verified own kernel
                const bool k_own = bt->gemm_own;
                const int k_shape = bt->gemm_shape, k_ks = bt->gemm_ks;
                bt->gemm_own = true;
                bt->gemm_shape = bt->own_shape;
                bt->gemm_ks = bt->own_ks;
                TM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int r = TM_OK;
                hipLaunchKernelGGL(k_b_reset_fail, dim3(1), dim3(64), 0, s, bt->rk_fail.p);
                hipLaunchKernelGGL(k_b_copy, g, dim3(kBlock), 0, s, N3, (const double*)bt->m.p, bt->rk_base.p);
                for (int k = 0; k < 4 && r == TM_OK; ++k) {
                    if (k > 0)
                        hipLaunchKernelGGL(k_b_axpy, g, dim3(kBlock), 0, s, N3, (const double*)bt->rk_base.p,
                                           (const double*)bt->rk_k[k - 1].p, cst[k], bt->m.p);
                    r = eval_enqueue(bt, alpha, variant, bt->rk_k[k].p);
                }
                hipLaunchKernelGGL(k_b_rk_final, dim3(elem_grid(NS)), dim3(kBlock), 0, s, NS,
                                   (const double*)bt->rk_base.p, (const double*)bt->rk_k[0].p,
                                   (const double*)bt->rk_k[1].p, (const double*)bt->rk_k[2].p,
                                   (const double*)bt->rk_k[3].p, dtau, bt->m.p);
                const hipError_t ec = hipStreamEndCapture(s, &gr);
                bt->gemm_own = k_own;
                bt->gemm_shape = k_shape;
                bt->gemm_ks = k_ks;
                if (r != TM_OK || ec != hipSuccess) {
                    if (gr) (void)hipGraphDestroy(gr);
                    status = r != TM_OK ? r : TM_ERR_HIP;
                    break;
                }
                const hipError_t ei = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (ei != hipSuccess) {
                    exec = nullptr;
                    status = TM_ERR_HIP;
                    break;
                }
                R.captures += 1;
            }
            unsigned fail = 0;
            if (hipGraphLaunch(exec, s) != hipSuccess ||
                hipMemcpyAsync(&fail, bt->rk_fail.p, sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                status = TM_ERR_HIP;
                break;
            }
            if (fail == 0) {
                R.graph_steps += 1;
                if (fin && last) last->nsamples = bt->S;
                continue;
            }
            // a sample's solve missed its tolerance within the budgets: restore
            // the step's start state and redo it with host-polled solves
            hipLaunchKernelGGL(k_b_copy, g, dim3(kBlock), 0, s, N3, (const double*)bt->rk_base.p, bt->m.p);
            R.redone_steps += 1;
            for (int w = 0; w < 2; ++w)
                for (int k = 0; k < bt->bud_outer[w]; ++k) bt->bud_inner[w][k] += 2;
            (void)hipGraphExecDestroy(exec);
            exec = nullptr;
        }
        const int r = polled_step(fin);
        if (r != TM_OK && r != TM_ERR_MAXITER) {
            status = r;
            break;
        }
        if (r == TM_ERR_MAXITER) status = r;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    TM_HIP(hipEventRecord(e1, s));
    TM_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    R.ms = ms;
    if (rk) *rk = R;
    return status;
}

int tm_batch_gemm_info(tm_batch* bt, int* solution, int* own_kernel, int* candidates, double* ms_default,
                       double* ms_tuned, int* ksplit) {
    if (!bt) return TM_ERR_BAD_ARG;
    if (bt->prec == TM_PREC_MIXED_BEMI8) {  // k_b_gemm_i8: solution = row tiles per wavefront
        if (solution) *solution = bt->i8mt;
        if (own_kernel) *own_kernel = 1;
        if (candidates) *candidates = bt->i8candidates;
        if (ksplit) *ksplit = bt->i8ks;
        if (ms_default) *ms_default = bt->i8ms_first;
        if (ms_tuned) *ms_tuned = bt->i8ms;
        return TM_OK;
    }
    if (solution) *solution = bt->gemm_own ? bt->gemm_shape : bt->gemm_sol;
    if (own_kernel) *own_kernel = bt->gemm_own ? 1 : 0;
    if (candidates) *candidates = bt->gemm_candidates + bt->gemm_own_candidates;
    if (ksplit) *ksplit = bt->gemm_own ? bt->gemm_ks : 1;
    if (ms_default) *ms_default = bt->gemm_ms_default;
    if (ms_tuned) *ms_tuned = bt->gemm_ms_best;
    return TM_OK;
}

int tm_batch_time_kernels(tm_batch* bt, int reps, tm_batch_kernel_time* out, int cap, int* count) {
    if (!bt || reps < 1 || !out || cap < 1 || !count) return TM_ERR_BAD_ARG;
    hipStream_t s = bt->stream;
    const int n = bt->n, S = bt->S;
    const double nS = (double)n * S;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    int status = TM_OK;
    auto timeit = [&](auto&& launch) -> double {
        launch();  // warm-up
        if (hipEventRecord(e0, s) != hipSuccess) status = TM_ERR_HIP;
        for (int r = 0; r < reps; ++r) launch();
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) status = TM_ERR_HIP;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return (double)ms / reps;
    };
    std::vector<tm_batch_kernel_time> v;
    auto add = [&v](const char* nm, double ms, double bytes) {
        tm_batch_kernel_time k = {};
        std::snprintf(k.name, sizeof(k.name), "%s", nm);
        k.ms = ms;
        k.bytes = bytes;
        v.push_back(k);
    };
    const double eslots = (double)bt->eo.h_sp[n];
    add("k_b_div", timeit([&] {
            launch_div(bt);
        }),
        eslots * (4 + 3 * 8) + 3 * nS * 8 + nS * 8);
    const int nw = (bt->exch ? 1 : 0) + 3 + (bt->dmi ? 3 : 0);
    const double per_node = 8.0 * n * (1 + (bt->cv.has_ku ? 4 : 0) + (bt->dmi ? 4 : 0));
    add(S == kMaxS ? (bt->dmi ? "k_b_eval_w<DMI>" : "k_b_eval_w") : (bt->dmi ? "k_b_eval<DMI>" : "k_b_eval"),
        timeit([&] {
            (void)launch_fields(bt, 0.5, TM_LLG_CLASSIC, bt->dmdt.p, nullptr, false);
        }),
        eslots * (4 + 8.0 * nw) + 3 * nS * 8 + nS * 8 + 3 * nS * 8 + per_node);
    // the boundary GEMM as the batch runs it (the tuned kernel, without the
    // gather and scatter); bytes: B once plus U and the product
    if (bt->prec == TM_PREC_MIXED_BEMI8) {
        // B's digit slices once, U's digits and the FP64 partials
        const double bytes = (double)bt->i8nT * bt->i8nC * kI8Digits * 1024.0 +
                             (double)bt->i8nC * kI8DigitsU * 4 * 1024.0 + 8.0 * bt->i8ks * bt->cv.bnx * 64.0;
        add("boundary GEMM (k_b_gemm_i8, int8 digits, tuned)", timeit([&] { (void)i8_gemm(bt, bt->i8mt, bt->i8ks); }),
            bytes);
    } else if (bt->prec != TM_PREC_MIXED_BEM32) {
        const double bb = (double)bt->cv.bnx;
        add(bt->gemm_own ? "boundary GEMM (k_b_gemm64, tuned)" : "boundary GEMM (rocBLAS, tuned)", timeit([&] {
                if (bt->gemm_own) (void)gemm_own_run(bt, nullptr);
                else (void)gemm_run(bt, bt->gemm_sol);
            }),
            8.0 * bb * bb + 2.0 * 8.0 * bb * S);
    }
    int st = bt->prec == TM_PREC_FP64 ? bt->sN64.time_kernels(s, timeit, v) : bt->sN32.time_kernels(s, timeit, v);
    if (status == TM_OK) status = st;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *count = 0;
    for (size_t k = 0; k < v.size() && (int)k < cap; ++k) out[(*count)++] = v[k];
    return status;
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
    // staged in usum (rewritten by every evaluation before it is read), not
    // in rhsD: the Dirichlet right-hand side must stay zero off the boundary
    double* yd = ywhere == TM_HOST ? bt->usum.p : y;
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
