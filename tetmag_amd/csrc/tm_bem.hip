// tm_bem.hip -- dense boundary-element matrix: assembly and product on the GPU.
//
// The reference assembles the double-layer matrix on the host, one boundary
// node at a time (BEMprocessing.cpp:375-411, Lindholm.cpp:63-128), and
// multiplies it on the host too (DemagField.cpp:139-147, Eigen GEMV or
// H2Lib).  Here both run on the device: B (bnx x bnx, FP64, column-major,
// leading dimension padded to a multiple of 64) stays resident in HBM.
#include <cmath>
#include <cstdint>
#include <mutex>

#include "tm_internal.hpp"

namespace tmh {

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));

struct V3 {
    double x, y, z;
};
// Helpers evaluate left to right with FP contraction off, matching the
// operation order of the reference's Eigen expressions.
__device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline double dot(V3 a, V3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ inline V3 cross(V3 a, V3 b) {
#pragma clang fp contract(off)
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Double-layer weights of the three vertices of a flat triangle seen from an
// observation point off the triangle: D.A. Lindholm, IEEE Trans. Magn. 20,
// 2025 (1984), the linear-density double layer of Lindholm.cpp:63-128.
//
// Two forms of the subtended solid angle:
//  * ACCURATE = false ("reference"): 2 acos(numer/denom) with the areEqual
//    clamp, exactly the expression of Lindholm.cpp:41-60, evaluated in the
//    same operation order with FP contraction off.  For a distant triangle
//    the argument is 1 - O(Omega^2), so Omega carries a relative rounding
//    error of ~1e-16/Omega^2, and the weights (which cancel by ~(rho/s)^2)
//    inherit it; this form reproduces the reference's matrix, noise included.
//  * ACCURATE = true: Van Oosterom-Strackee 2 atan2(|r1.(r2 x r3)|, ...),
//    which keeps full relative precision (B 1 = -1 holds to ~1e-14).
__device__ inline bool are_equal_ulp100(double x, double y) {  // auxiliaries.cpp:42-47
    const double eps = 2.220446049250313080847e-16;
    bool almost = fabs(x - y) < (eps * fabs(x + y) * 100);
    if (fabs(x) < eps * 100 && fabs(y) < eps * 100) almost = true;
    return almost;
}

template <bool ACCURATE>
__device__ __noinline__ void lindholm(V3 x0, V3 p1, V3 p2, V3 p3, double area, V3 nv, double* L) {
#pragma clang fp contract(off)
    const V3 r[3] = {p1 - x0, p2 - x0, p3 - x0};
    V3 chi = cross(p1 - p2, p1 - p3);
    const double cl = sqrt(dot(chi, chi));
    chi = V3{chi.x / cl, chi.y / cl, chi.z / cl};
    double s[3], rho[3], eta0[3];
    V3 xi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const V3 e = r[(i + 1) % 3] - r[i];
        s[i] = sqrt(dot(e, e));
        xi[i] = V3{e.x / s[i], e.y / s[i], e.z / s[i]};
        eta0[i] = dot(cross(chi, xi[i]), r[i]);
        rho[i] = sqrt(dot(r[i], r[i]));
    }
    const double chi0 = dot(chi, r[0]);
    double gam[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) gam[i][j] = dot(xi[(i + 1) % 3], xi[j]);
    double P[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3;
        P[i] = log((rho[i] + rho[i1] + s[i]) / (rho[i] + rho[i1] - s[i]));
    }
    double Om;
    if (ACCURATE) {
        const double trip = fabs(dot(r[0], cross(r[1], r[2])));
        const double den = rho[0] * rho[1] * rho[2] + dot(r[0], r[1]) * rho[2] + dot(r[1], r[2]) * rho[0] +
                           dot(r[2], r[0]) * rho[1];
        double half = atan2(trip, den);
        if (half < 0) half += kPi;
        Om = 2.0 * half;
    } else {
        const double r123 = rho[0] * rho[1] * rho[2];
        const double r23 = dot(r[1], r[2]), r12 = dot(r[0], r[1]), r31 = dot(r[2], r[0]);
        const double numer = r123 + rho[0] * r23 + rho[1] * r31 + rho[2] * r12;
        const double denom = sqrt(2.0 * (rho[1] * rho[2] + r23) * (rho[2] * rho[0] + r31) * (rho[0] * rho[1] + r12));
        double arg = numer / denom;
        if (are_equal_ulp100(arg, 1.0)) arg = 1.0;
        Om = 2.0 * acos(arg);
    }
    if (chi0 < 0.0) Om = -Om;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3;
        const double gP = gam[i][0] * P[0] + gam[i][1] * P[1] + gam[i][2] * P[2];
        L[i] = s[i1] / (8. * kPi * area) * (eta0[i1] * Om - chi0 * gP);
    }
    if (dot(nv, chi) < 0) {
        L[0] = -L[0];
        L[1] = -L[1];
        L[2] = -L[2];
    }
}

// One thread per matrix row (observation node); all lanes walk the faces in
// the same order, so face data loads are uniform and each lane's updates of
// B(row, vertex) touch only its own row (no atomics).  Rows are written
// coalesced across the wavefront.
template <bool ACCURATE>
__global__ __launch_bounds__(kBlock) void k_bem_assemble(int b, int ld, int nbel, const double* __restrict__ bxyz,
                                                         const int* __restrict__ faces,
                                                         const double* __restrict__ farea,
                                                         const double* __restrict__ fnorm,
                                                         const double* __restrict__ omega, double* __restrict__ B) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= b) return;
    for (int c = 0; c < b; ++c) B[(size_t)c * ld + row] = 0.0;
    B[(size_t)row * ld + row] = omega[row] / (4.0 * kPi) - 1.0;  // BEMprocessing.cpp:387-389
    const V3 x0 = {bxyz[3 * row], bxyz[3 * row + 1], bxyz[3 * row + 2]};
    for (int f = 0; f < nbel; ++f) {
        const int v0 = faces[3 * f], v1 = faces[3 * f + 1], v2 = faces[3 * f + 2];
        if (row == v0 || row == v1 || row == v2) continue;  // BEMprocessing.cpp:398-399
        const V3 p1 = {bxyz[3 * v0], bxyz[3 * v0 + 1], bxyz[3 * v0 + 2]};
        const V3 p2 = {bxyz[3 * v1], bxyz[3 * v1 + 1], bxyz[3 * v1 + 2]};
        const V3 p3 = {bxyz[3 * v2], bxyz[3 * v2 + 1], bxyz[3 * v2 + 2]};
        const V3 nv = {fnorm[3 * f], fnorm[3 * f + 1], fnorm[3 * f + 2]};
        double L[3];
        lindholm<ACCURATE>(x0, p1, p2, p3, farea[f], nv, L);
        B[(size_t)v0 * ld + row] += L[0];
        B[(size_t)v1 * ld + row] += L[1];
        B[(size_t)v2 * ld + row] += L[2];
    }
}

constexpr int kGemvRowsPerThread = 8;  // 4 x double2, at a stride of 512 rows
constexpr int kGemvTile = kBlock * kGemvRowsPerThread;  // 2048 rows per block
constexpr int kGemvMaxCols = 2048;  // columns per block (LDS: 16 KB)

// Split-K partial products of y = B x for column-major B (ld even, 16-byte
// aligned columns).  Block (tile, split) multiplies a 2048-row tile by a
// column range and writes one partial per row.  x is gathered through `gidx`
// (x_c = src[gidx[c]]) when gidx != NULL, which fuses DemagField's boundary
// gather (DemagField.cpp:187-193) into the product.
__global__ __launch_bounds__(kBlock) void k_bem_gemv_part(int b, int ld, const double* __restrict__ B,
                                                          const double* __restrict__ src,
                                                          const int* __restrict__ gidx, int cols_per_split,
                                                          double* __restrict__ part) {
    __shared__ double xs[kGemvMaxCols];
    const int tile = blockIdx.x, split = blockIdx.y;
    const int c0 = split * cols_per_split;
    const int c1 = min(b, c0 + cols_per_split);
    for (int c = c0 + threadIdx.x; c < c1; c += kBlock) xs[c - c0] = gidx ? src[gidx[c]] : src[c];
    __syncthreads();
    const int r0 = tile * kGemvTile + 2 * threadIdx.x;  // rows r0 + 512 k, r0 + 1 + 512 k
    dbl2 acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = dbl2{0.0, 0.0};
    bool valid[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) valid[k] = (r0 + 512 * k) < ld;  // rows in [b, ld) are padding (zero)
    int c = c0;
    // four columns per step: 16 loads of 16 bytes in flight per thread
    for (; c + 3 < c1; c += 4) {
        const double x4[4] = {xs[c - c0], xs[c + 1 - c0], xs[c + 2 - c0], xs[c + 3 - c0]};
        dbl2 v[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double* cj = B + (size_t)(c + j) * ld + r0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                v[j][k] = valid[k] ? __builtin_nontemporal_load((const dbl2*)(cj + 512 * k)) : dbl2{0.0, 0.0};
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] += v[j][k] * x4[j];
    }
    for (; c + 1 < c1; c += 2) {
        const double xa = xs[c - c0], xb = xs[c + 1 - c0];
        const double* ca = B + (size_t)c * ld + r0;
        const double* cb = ca + ld;
        dbl2 va[4], vb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (valid[k]) {
                va[k] = __builtin_nontemporal_load((const dbl2*)(ca + 512 * k));
                vb[k] = __builtin_nontemporal_load((const dbl2*)(cb + 512 * k));
            } else {
                va[k] = vb[k] = dbl2{0.0, 0.0};
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc[k] += va[k] * xa;
            acc[k] += vb[k] * xb;
        }
    }
    if (c < c1) {
        const double xa = xs[c - c0];
        const double* ca = B + (size_t)c * ld + r0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (valid[k]) acc[k] += __builtin_nontemporal_load((const dbl2*)(ca + 512 * k)) * xa;
    }
    double* out = part + (size_t)split * ld;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int r = r0 + 512 * k;
        if (r < b) out[r] = acc[k].x;
        if (r + 1 < b) out[r + 1] = acc[k].y;
    }
}

// y_r = sum_s part[s][r]; optionally also scattered to dst[sidx[r]]
// (DemagField.cpp:195-199, boundary values of the Dirichlet problem).
__global__ __launch_bounds__(kBlock) void k_bem_gemv_finish(int b, int ld, int splits,
                                                            const double* __restrict__ part,
                                                            double* __restrict__ y, const int* __restrict__ sidx,
                                                            double* __restrict__ dst) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= b) return;
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * ld + r];
    if (y) y[r] = s;
    if (dst) dst[sidx[r]] = s;
}

// Transposed product y = B^T x for the adjoint of the boundary integral
// (x, y on the boundary).  Column-major B makes each output a column dot
// product: block = 16 columns (4 per wavefront), x staged in LDS in chunks of
// 2048 rows and shared by the 16 columns; lanes read 16-byte pairs of rows so
// each wavefront load covers 1 KB of one column.  With dst != NULL the result
// is also added to dst[sidx[j]] (the scatter back to the nodes).
constexpr int kGemvTCols = 16;
constexpr int kGemvTChunk = 2048;

__global__ __launch_bounds__(kBlock) void k_bem_gemvT(int b, int ld, const double* __restrict__ B,
                                                      const double* __restrict__ x, double* __restrict__ y,
                                                      const int* __restrict__ sidx, double* __restrict__ dst) {
    __shared__ double xs[kGemvTChunk];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int j0 = blockIdx.x * kGemvTCols + wave * 4;
    dbl2 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = dbl2{0.0, 0.0};
    for (int r0 = 0; r0 < b; r0 += kGemvTChunk) {
        const int rn = min(kGemvTChunk, b - r0);
        __syncthreads();  // the previous chunk has been consumed
        for (int i = threadIdx.x; i < kGemvTChunk; i += kBlock) xs[i] = i < rn ? x[r0 + i] : 0.0;
        __syncthreads();
        for (int i = 2 * lane; i < rn; i += 2 * kWave) {
            // rows r0+i, r0+i+1: the second is a zero padding row or (x = 0)
            // the next column's first entry only when rn is odd, which needs
            // b odd and therefore ld > b.
            const double x0 = xs[i], x1 = xs[i + 1];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int j = j0 + c;
                if (j < b) {
                    const dbl2 v = __builtin_nontemporal_load((const dbl2*)(B + (size_t)j * ld + r0 + i));
                    acc[c].x += v.x * x0;
                    acc[c].y += v.y * x1;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double sum = wave_sum(acc[c].x + acc[c].y);
        const int j = j0 + c;
        if (lane == 0 && j < b) {
            if (y) y[j] = sum;
            if (dst) dst[sidx[j]] += sum;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_bem_gather(int b, const double* __restrict__ src,
                                                       const int* __restrict__ gidx, double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < b) out[i] = src[gidx[i]];
}

}  // namespace

int bem_assemble(int b, int ld, int nbel, const double* bxyz, const int* faces, const double* farea,
                 const double* fnorm, const double* omega, double* B, int formula, hipStream_t s) {
    if (formula == TM_BEM_ACCURATE)
        hipLaunchKernelGGL(k_bem_assemble<true>, dim3(grid_for(b)), dim3(kBlock), 0, s, b, ld, nbel, bxyz, faces,
                           farea, fnorm, omega, B);
    else
        hipLaunchKernelGGL(k_bem_assemble<false>, dim3(grid_for(b)), dim3(kBlock), 0, s, b, ld, nbel, bxyz, faces,
                           farea, fnorm, omega, B);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

int bem_gemv_splits(int b) {
    const int tiles = (b + kGemvTile - 1) / kGemvTile;
    // aim for >= 1024 blocks so all 256 CUs have several to overlap
    int splits = (1024 + tiles - 1) / tiles;
    const int min_cols = 256;
    splits = std::min(splits, std::max(1, b / min_cols));
    while ((b + splits - 1) / splits > kGemvMaxCols) ++splits;
    return std::max(1, splits);
}

// y = B x with x optionally gathered through gidx, y optionally scattered to
// dst via sidx.  part needs splits * ld doubles.
int bem_gemv(int b, int ld, const double* B, const double* src, const int* gidx, double* part, double* y,
             const int* sidx, double* dst, hipStream_t s) {
    if (ld % 2 != 0) return TM_ERR_BAD_ARG;
    const int tiles = (b + kGemvTile - 1) / kGemvTile;
    const int splits = bem_gemv_splits(b);
    const int cps = (b + splits - 1) / splits;
    hipLaunchKernelGGL(k_bem_gemv_part, dim3(tiles, splits), dim3(kBlock), 0, s, b, ld, B, src, gidx, cps, part);
    hipLaunchKernelGGL(k_bem_gemv_finish, dim3(grid_for(b)), dim3(kBlock), 0, s, b, ld, splits, part, y, sidx, dst);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

// y = B^T x with x optionally gathered through gidx (into xbuf, b doubles),
// the result optionally added to dst[sidx[j]].
int bem_gemvT(int b, int ld, const double* B, const double* src, const int* gidx, double* xbuf, double* y,
              const int* sidx, double* dst, hipStream_t s) {
    if (ld % 2 != 0 || ld < b) return TM_ERR_BAD_ARG;
    const double* x = src;
    if (gidx) {
        hipLaunchKernelGGL(k_bem_gather, dim3(grid_for(b)), dim3(kBlock), 0, s, b, src, gidx, xbuf);
        x = xbuf;
    }
    hipLaunchKernelGGL(k_bem_gemvT, dim3((b + kGemvTCols - 1) / kGemvTCols), dim3(kBlock), 0, s, b, ld, B, x, y,
                       sidx, dst);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

}  // namespace tmh

// ---------------------------------------------------------------------------
// Boundary product that allocates nothing (tm_bem_apply_noalloc): the column
// splits' partial products live in a file-scope device array instead of a
// context buffer.  In host code `g_noalloc_scratch` is only a host-side
// stand-in for the symbol; its device address (one copy per GPU) is looked up
// once per device with hipGetSymbolAddress and only that address is used.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoallocScratchDoubles = (size_t)4 << 20;  // 32 MiB per GPU
__device__ __align__(256) double g_noalloc_scratch[kNoallocScratchDoubles];
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoallocMaxDevices] = {};

// Device address of the scratch array on the current device (looked up on the
// first call per device).
int noalloc_scratch(double** out) {
    int dev = 0;
    TM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoallocMaxDevices) return TM_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_noalloc_scratch));
        if (e != hipSuccess || !p) return TM_ERR_HIP;
        g_noalloc_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return TM_OK;
}
}  // namespace

namespace tmh {
int noalloc_scratch_init() {
    double* p = nullptr;
    return noalloc_scratch(&p);
}
}  // namespace tmh

extern "C" {

int tm_bem_apply_noalloc(int bnx, int ld, const double* B, const double* x, double* y, void* stream) {
    if (bnx <= 0 || ld < bnx || ld % 2 != 0 || !B || !x || !y) return TM_ERR_BAD_ARG;
    if (reinterpret_cast<uintptr_t>(B) % 16 != 0) return TM_ERR_BAD_ARG;  // 16-byte loads of column pairs
    if ((size_t)tmh::bem_gemv_splits(bnx) * (size_t)ld > kNoallocScratchDoubles) return TM_ERR_BAD_ARG;
    double* part = nullptr;
    TM_CHECK(noalloc_scratch(&part));
    return tmh::bem_gemv(bnx, ld, B, x, nullptr, part, y, nullptr, nullptr, static_cast<hipStream_t>(stream));
}

}  // extern "C"
