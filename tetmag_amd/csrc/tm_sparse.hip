// tm_sparse.hip -- SELL-64 sparse storage, SpMV, device selection.
//
// Replaces SpMatCUDA (SpMatCUDA.h:49-79, SpMatCUDA.cu:59-181): the reference
// stores CSC arrays and calls cuSPARSE's transposed CSR SpMV on them, copies x
// in and allocates a new device vector per product.  Here the matrix is kept
// as SELL-64 (one wavefront per 64-row slice, column-major within a slice) so
// every load of the SpMV is a 64-lane contiguous read, and products write
// into caller-owned buffers.
#include <algorithm>
#include <cstring>

#include "tm_internal.hpp"

namespace tmh {

void SellPattern::build(int rows, int cols, const int* rowptr, const int* colidx) {
    n = rows;
    ncols = cols;
    nslices = (rows + kWave - 1) / kWave;
    nnz = rowptr[rows];
    h_rowptr.assign(rowptr, rowptr + rows + 1);
    h_csrcol.assign(colidx, colidx + nnz);
    h_off.assign((size_t)nslices + 1, 0);
    h_w.assign((size_t)nslices, 0);
    for (int s = 0; s < nslices; ++s) {
        int wmax = 0;
        for (int l = 0; l < kWave; ++l) {
            int r = s * kWave + l;
            if (r < rows) wmax = std::max(wmax, rowptr[r + 1] - rowptr[r]);
        }
        h_w[s] = wmax;
        h_off[s + 1] = h_off[s] + (int64_t)kWave * wmax;
    }
    stored = h_off[nslices];
    h_col.assign((size_t)stored, 0);
    for (int s = 0; s < nslices; ++s) {
        for (int l = 0; l < kWave; ++l) {
            int r = s * kWave + l;
            int len = r < rows ? rowptr[r + 1] - rowptr[r] : 0;
            int pad = r < rows ? std::min(r, cols - 1) : 0;
            for (int k = 0; k < h_w[s]; ++k)
                h_col[(size_t)(h_off[s] + (int64_t)k * kWave + l)] = k < len ? colidx[rowptr[r] + k] : pad;
        }
    }
}

int SellPattern::upload(hipStream_t s) {
    TM_CHECK(off.upload(h_off.data(), h_off.size(), s));
    TM_CHECK(w.upload(h_w.data(), h_w.size(), s));
    TM_CHECK(col.upload(h_col.data(), h_col.size(), s));
    return TM_OK;
}

int SellPattern::scatter(const int* rowptr, const int* colidx, const double* vals, std::vector<double>& out) const {
    out.assign((size_t)stored, 0.0);
    for (int s = 0; s < nslices; ++s) {
        for (int l = 0; l < kWave; ++l) {
            int r = s * kWave + l;
            if (r >= n) continue;
            const int plo = h_rowptr[r], phi = h_rowptr[r + 1];
            int k = 0;
            for (int q = rowptr[r]; q < rowptr[r + 1]; ++q) {
                const int c = colidx[q];
                while (plo + k < phi && h_csrcol[plo + k] < c) ++k;
                if (plo + k >= phi || h_csrcol[plo + k] != c) return TM_ERR_BAD_ARG;
                out[(size_t)(h_off[s] + (int64_t)k * kWave + l)] += vals[q];
            }
        }
    }
    return TM_OK;
}

int copy_in(double* dst_dev, const double* src, int where, size_t count, hipStream_t s) {
    if (!count) return TM_OK;
    if (!src || !dst_dev) return TM_ERR_BAD_ARG;
    if (dst_dev == src) return TM_OK;
    if (where == TM_HOST) {
        TM_HIP(hipMemcpyAsync(dst_dev, src, count * sizeof(double), hipMemcpyHostToDevice, s));
        TM_HIP(hipStreamSynchronize(s));
    } else if (where == TM_DEVICE) {
        TM_HIP(hipMemcpyAsync(dst_dev, src, count * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        return TM_ERR_BAD_ARG;
    }
    return TM_OK;
}

int copy_out(double* dst, int where, const double* src_dev, size_t count, hipStream_t s) {
    if (!count) return TM_OK;
    if (!src_dev || !dst) return TM_ERR_BAD_ARG;
    if (dst == src_dev) return TM_OK;
    if (where == TM_HOST) {
        TM_HIP(hipMemcpyAsync(dst, src_dev, count * sizeof(double), hipMemcpyDeviceToHost, s));
        TM_HIP(hipStreamSynchronize(s));
    } else if (where == TM_DEVICE) {
        TM_HIP(hipMemcpyAsync(dst, src_dev, count * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        return TM_ERR_BAD_ARG;
    }
    return TM_OK;
}

// y = A x, one row per lane.
__global__ __launch_bounds__(kBlock) void k_sell_spmv(SellView A, const double* __restrict__ val,
                                                      const double* __restrict__ x, double* __restrict__ y) {
    const int row = xcd_block() * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    if (s >= (A.n + kWave - 1) / kWave) return;
    const int64_t base = A.off[s] + lane;
    const int w = A.w[s];
    double acc = 0.0;
    for (int k = 0; k < w; ++k) {
        const int64_t q = base + (int64_t)k * kWave;
        acc += val[q] * x[A.col[q]];
    }
    if (row < A.n) y[row] = acc;
}

int spmv(const SellPattern& P, const double* val, const double* x, double* y, hipStream_t s) {
    if (P.n == 0) return TM_OK;
    hipLaunchKernelGGL(k_sell_spmv, dim3(xcd_grid(grid_for((int64_t)P.nslices * kWave))), dim3(kBlock), 0, s, view(P),
                       val, x, y);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

}  // namespace tmh

using namespace tmh;

extern "C" {

const char* tm_version(void) { return "tetmag_hip 0.1 (gfx950)"; }

int tm_device_count(int* ndev) {
    if (!ndev) return TM_ERR_BAD_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) {
        *ndev = 0;
        return TM_ERR_NO_DEVICE;
    }
    *ndev = c;
    return TM_OK;
}

int tm_set_device(int dev, int* chosen) {
    int c = 0;
    int st = tm_device_count(&c);
    if (st != TM_OK) return st;
    if (dev < 0 || dev > c - 1) dev = 0;  // SetDevice.cu:54-57
    TM_HIP(hipSetDevice(dev));
    if (chosen) *chosen = dev;
    // set-up of tm_bem_apply_noalloc on this device (its scratch address)
    return tmh::noalloc_scratch_init();
}

static bool csr_ok(int rows, int cols, const int* rowptr, const int* colidx) {
    if (rows <= 0 || cols <= 0 || !rowptr || (!colidx && rowptr[rows] > 0)) return false;
    if (rowptr[0] != 0) return false;
    for (int r = 0; r < rows; ++r) {
        if (rowptr[r + 1] < rowptr[r]) return false;
        for (int q = rowptr[r]; q < rowptr[r + 1]; ++q) {
            if (colidx[q] < 0 || colidx[q] >= cols) return false;
            if (q > rowptr[r] && colidx[q] <= colidx[q - 1]) return false;
        }
    }
    return true;
}

int tm_spmat_create(int rows, int cols, const int* rowptr, const int* colidx, const double* vals, tm_spmat** out) {
    if (!out || !csr_ok(rows, cols, rowptr, colidx) || (!vals && rowptr[rows] > 0)) return TM_ERR_BAD_ARG;
    int nd = 0;
    TM_CHECK(tm_device_count(&nd));
    tm_spmat* A = new tm_spmat();
    A->rows = rows;
    A->cols = cols;
    if (hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking) != hipSuccess) {
        delete A;
        return TM_ERR_HIP;
    }
    A->pat.build(rows, cols, rowptr, colidx);
    std::vector<double> v;
    int st = A->pat.scatter(rowptr, colidx, vals, v);
    if (st == TM_OK) st = A->pat.upload(A->stream);
    if (st == TM_OK) st = A->val.upload(v.data(), v.size(), A->stream);
    if (st == TM_OK) st = A->x.alloc(cols);
    if (st == TM_OK) st = A->y.alloc(rows);
    if (st == TM_OK && hipStreamSynchronize(A->stream) != hipSuccess) st = TM_ERR_HIP;
    if (st != TM_OK) {
        tm_spmat_destroy(A);
        return st;
    }
    *out = A;
    return TM_OK;
}

int tm_spmat_destroy(tm_spmat* A) {
    if (!A) return TM_OK;
    if (A->stream) (void)hipStreamSynchronize(A->stream);
    if (A->stream) (void)hipStreamDestroy(A->stream);
    delete A;
    return TM_OK;
}

int tm_spmat_mvp(tm_spmat* A, const double* x, int xwhere, double* y, int ywhere) {
    if (!A || !x || !y) return TM_ERR_BAD_ARG;
    const double* xd = x;
    if (xwhere == TM_HOST) {
        TM_CHECK(copy_in(A->x.p, x, TM_HOST, A->cols, A->stream));
        xd = A->x.p;
    } else if (xwhere != TM_DEVICE) {
        return TM_ERR_BAD_ARG;
    }
    double* yd = (ywhere == TM_DEVICE) ? y : A->y.p;
    if (ywhere != TM_DEVICE && ywhere != TM_HOST) return TM_ERR_BAD_ARG;
    TM_CHECK(spmv(A->pat, A->val.p, xd, yd, A->stream));
    if (ywhere == TM_HOST) TM_CHECK(copy_out(y, TM_HOST, yd, A->rows, A->stream));
    else TM_HIP(hipStreamSynchronize(A->stream));
    return TM_OK;
}

}  // extern "C"
