// tm_field.hip -- effective fields, LLG right-hand side and the FEM-BEM demag
// pipeline on one device-resident context.
//
// Replaces EffFieldGPU (EffFieldGPU.cu), GpuPotential (GpuPotential.cu) and
// the host composition in TheLLG (TheLLG.cpp:238-262) / LLGWrapper
// (LLGWrapper.cu:70-82, 213-228) plus DemagField (DemagField.cpp:92-240).
// The reference moves every field over PCIe and sums Heff on the host; here m,
// all operators, the dense boundary matrix and every intermediate stay in
// HBM and one evaluation is:
//   k_div             b = (1/mu0) sum_d tGrad_d (Js m_d)
//   Neumann CG        u1   (tm_cg.hip)
//   BEM product       g = B u1|Gamma, scattered into the Dirichlet RHS (tm_bem.hip)
//   Dirichlet CG      u2
//   k_eval            Hexc + Hdem(-grad(u1+u2)) + Hdmi + Hku + Hext -> Heff,
//                     LLG right-hand side and max torque, one SELL pass.
// All sparse operators of the context share one SELL-64 pattern (the union
// of the patterns set; for tetmag's P1 operators they are identical), so the
// fused kernel reads the column indices once for all operators.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tm_ctx.hpp"

namespace tmh {
int spmv(const SellPattern& P, const double* val, const double* x, double* y, hipStream_t s);
int bem_assemble(int b, int ld, int nbel, const double* bxyz, const int* faces, const double* farea,
                 const double* fnorm, const double* omega, double* B, int formula, hipStream_t s);
int bem_gemv(int b, int ld, const double* B, const double* src, const int* gidx, double* part, double* y,
             const int* sidx, double* dst, hipStream_t s);
int bem_gemv_splits(int b);
int amg_info(const CgSolver& cg, int* nlevels, int* sizes, int cap);
}  // namespace tmh

using namespace tmh;


namespace tmh {

int ctx_ok(tm_ctx* c) { return c && c->n > 0 ? TM_OK : TM_ERR_BAD_ARG; }

int commit_ops(tm_ctx* c) {
    if (!c->ops_dirty) return TM_OK;
    const int n = c->n;
    // union pattern over every operator set
    int nset = 0;
    int ref = -1;
    for (int o = 0; o < TM_NUM_OPS; ++o)
        if (c->hop[o].set) {
            ++nset;
            if (ref < 0) ref = o;
        }
    if (nset == 0) return TM_ERR_NOT_READY;
    bool same = true;
    for (int o = 0; o < TM_NUM_OPS && same; ++o)
        if (c->hop[o].set && (c->hop[o].rp != c->hop[ref].rp || c->hop[o].col != c->hop[ref].col)) same = false;
    if (same) {
        c->pat.build(n, n, c->hop[ref].rp.data(), c->hop[ref].col.data());
    } else {
        std::vector<int> rp((size_t)n + 1, 0), col;
        std::vector<int> tmp;
        for (int i = 0; i < n; ++i) {
            tmp.clear();
            for (int o = 0; o < TM_NUM_OPS; ++o)
                if (c->hop[o].set)
                    for (int q = c->hop[o].rp[i]; q < c->hop[o].rp[i + 1]; ++q) tmp.push_back(c->hop[o].col[q]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            col.insert(col.end(), tmp.begin(), tmp.end());
            rp[i + 1] = (int)col.size();
        }
        c->pat.build(n, n, rp.data(), col.data());
    }
    TM_CHECK(c->pat.upload(c->stream));
    std::vector<double> v;
    for (int o = 0; o < TM_NUM_OPS; ++o) {
        c->has_op[o] = c->hop[o].set;
        if (!c->hop[o].set) {
            c->opv[o].release();
            continue;
        }
        TM_CHECK(c->pat.scatter(c->hop[o].rp.data(), c->hop[o].col.data(), c->hop[o].val.data(), v));
        TM_CHECK(c->opv[o].upload(v.data(), v.size(), c->stream));
        TM_HIP(hipStreamSynchronize(c->stream));  // v is reused for the next operator
    }
    TM_HIP(hipStreamSynchronize(c->stream));
    c->ops_dirty = false;
    ++c->ops_gen;  // transposed values built for the previous operators are stale
    return TM_OK;
}

}  // namespace tmh

namespace {

// --------------------------------------------------------------------------
// kernels
// --------------------------------------------------------------------------
#define SELL_ROW_PROLOGUE                                          \
    const int row = xcd_block() * kBlock + threadIdx.x;             \
    const int s = row / kWave, lane = row & (kWave - 1);           \
    if (s >= (A.n + kWave - 1) / kWave) return;                    \
    const int64_t base = A.off[s] + lane;                          \
    const int w = A.w[s];

// EffFieldGPU::ExchangeFieldGPU (EffFieldGPU.cu:71-86): H_c = XC m_c.
__global__ __launch_bounds__(kBlock) void k_exchange(SellView A, const double* __restrict__ v,
                                                     const double* __restrict__ m, double* __restrict__ out) {
    SELL_ROW_PROLOGUE
    const int n = A.n;
    double hx = 0, hy = 0, hz = 0;
    for (int k = 0; k < w; ++k) {
        const int64_t q = base + (int64_t)k * kWave;
        const int cc = A.col[q];
        const double e = v[q];
        hx += e * m[cc];
        hy += e * m[n + cc];
        hz += e * m[2 * n + cc];
    }
    if (row < n) {
        out[row] = hx;
        out[n + row] = hy;
        out[2 * n + row] = hz;
    }
}

// EffFieldGPU::UniaxialAnisotropyField (EffFieldGPU.cu:538-560): 2 Ku (k.m) k.
__global__ __launch_bounds__(kBlock) void k_uniaxial(int n, const double* __restrict__ ku,
                                                     const double* __restrict__ ax, const double* __restrict__ m,
                                                     double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double kx = ax[i], ky = ax[n + i], kz = ax[2 * n + i];
    const double f = 2.0 * ku[i] * (kx * m[i] + ky * m[n + i] + kz * m[2 * n + i]);
    out[i] = f * kx;
    out[n + i] = f * ky;
    out[2 * n + i] = f * kz;
}

// EffFieldGPU::calcCurlM + DMIField (EffFieldGPU.cu:150-162, 563-622):
// H = (m x (S nv) - 2 curl m) D/V with curl_x = Cy m_z - Cz m_y (cyclic).
__global__ __launch_bounds__(kBlock) void k_dmi(SellView A, const double* __restrict__ cx,
                                                const double* __restrict__ cy, const double* __restrict__ cz,
                                                const double* __restrict__ m, const double* __restrict__ fac,
                                                const double* __restrict__ nvs, double* __restrict__ out) {
    SELL_ROW_PROLOGUE
    const int n = A.n;
    double ax = 0, ay = 0, az = 0;
    for (int k = 0; k < w; ++k) {
        const int64_t q = base + (int64_t)k * kWave;
        const int cc = A.col[q];
        const double X = cx[q], Y = cy[q], Z = cz[q];
        const double mx = m[cc], my = m[n + cc], mz = m[2 * n + cc];
        ax += Y * mz - Z * my;
        ay += Z * mx - X * mz;
        az += X * my - Y * mx;
    }
    if (row < n) {
        const double mx = m[row], my = m[n + row], mz = m[2 * n + row];
        const double sx = nvs[row], sy = nvs[n + row], sz = nvs[2 * n + row];
        const double f = fac[row];
        out[row] = ((my * sz - mz * sy) - 2.0 * ax) * f;
        out[n + row] = ((mz * sx - mx * sz) - 2.0 * ay) * f;
        out[2 * n + row] = ((mx * sy - my * sx) - 2.0 * az) * f;
    }
}

// EffFieldGPU::UTermSTT_GPU (EffFieldGPU.cu:107-148): u_c = sum_d eta_d (Grad_d m_c).
__global__ __launch_bounds__(kBlock) void k_stt_uterm(SellView A, const double* __restrict__ gx,
                                                      const double* __restrict__ gy, const double* __restrict__ gz,
                                                      const double* __restrict__ eta, const double* __restrict__ m,
                                                      double* __restrict__ out) {
    SELL_ROW_PROLOGUE
    const int n = A.n;
    double d[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // d[c][dir]
    for (int k = 0; k < w; ++k) {
        const int64_t q = base + (int64_t)k * kWave;
        const int cc = A.col[q];
        const double X = gx[q], Y = gy[q], Z = gz[q];
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
            const double mc = m[c3 * n + cc];
            d[c3][0] += X * mc;
            d[c3][1] += Y * mc;
            d[c3][2] += Z * mc;
        }
    }
    if (row < n) {
        const double ex = eta[row], ey = eta[n + row], ez = eta[2 * n + row];
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) out[c3 * n + row] = ex * d[c3][0] + ey * d[c3][1] + ez * d[c3][2];
    }
}

// GpuPotential::calcDivM (GpuPotential.cu:86-106): sum_d tGrad_d (Js m_d), times `scale`.
__global__ __launch_bounds__(kBlock) void k_div(SellView A, const double* __restrict__ tx,
                                                const double* __restrict__ ty, const double* __restrict__ tz,
                                                const double* __restrict__ js, const double* __restrict__ m,
                                                double scale, double* __restrict__ out) {
    SELL_ROW_PROLOGUE
    const int n = A.n;
    double sx = 0, sy = 0, sz = 0;
    (void)base;
    (void)w;
    sell_row(A, s, lane, [&](int64_t q, int cc) {
        const double j = js[cc];
        sx += tx[q] * (m[cc] * j);
        sy += ty[q] * (m[n + cc] * j);
        sz += tz[q] * (m[2 * n + cc] * j);
    });
    if (row < n) out[row] = (sx + sy + sz) * scale;
}

// GpuPotential::calcGradientField (GpuPotential.cu:109-122): H = -grad(u1 [+ u2]).
__global__ __launch_bounds__(kBlock) void k_grad(SellView A, const double* __restrict__ gx,
                                                 const double* __restrict__ gy, const double* __restrict__ gz,
                                                 const double* __restrict__ u1, const double* __restrict__ u2,
                                                 double* __restrict__ out) {
    SELL_ROW_PROLOGUE
    const int n = A.n;
    double sx = 0, sy = 0, sz = 0;
    for (int k = 0; k < w; ++k) {
        const int64_t q = base + (int64_t)k * kWave;
        const int cc = A.col[q];
        const double u = u2 ? u1[cc] + u2[cc] : u1[cc];
        sx += gx[q] * u;
        sy += gy[q] * u;
        sz += gz[q] * u;
    }
    if (row < n) {
        out[row] = -sx;
        out[n + row] = -sy;
        out[2 * n + row] = -sz;
    }
}

__device__ inline void llg_point(double mx, double my, double mz, double hx, double hy, double hz, double alpha,
                                 int variant, double* o) {
    // m x H, then m x (m x H)   (EffFieldGPU.cu:316-362)
    const double ax = my * hz - mz * hy, ay = mz * hx - mx * hz, az = mx * hy - my * hx;
    const double bx = my * az - mz * ay, by = mz * ax - mx * az, bz = mx * ay - my * ax;
    if (variant == TM_LLG_NOPREC) {
        o[0] = -alpha * bx;
        o[1] = -alpha * by;
        o[2] = -alpha * bz;
    } else {
        o[0] = -ax - alpha * bx;
        o[1] = -ay - alpha * by;
        o[2] = -az - alpha * bz;
    }
}

__global__ __launch_bounds__(kBlock) void k_llg(int n, const double* __restrict__ m, const double* __restrict__ h,
                                                double alpha, int variant, double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double o[3];
    llg_point(m[i], m[n + i], m[2 * n + i], h[i], h[n + i], h[2 * n + i], alpha, variant, o);
    out[i] = o[0];
    out[n + i] = o[1];
    out[2 * n + i] = o[2];
}

// STT_functor (EffFieldGPU.cu:434-459): -(beta-alpha) m x u - (1+alpha beta) m x (m x u).
__global__ __launch_bounds__(kBlock) void k_stt_llg(int n, const double* __restrict__ m,
                                                    const double* __restrict__ u, double alpha, double beta,
                                                    double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double mx = m[i], my = m[n + i], mz = m[2 * n + i];
    const double ux = u[i], uy = u[n + i], uz = u[2 * n + i];
    const double ax = my * uz - mz * uy, ay = mz * ux - mx * uz, az = mx * uy - my * ux;
    const double bx = my * az - mz * ay, by = mz * ax - mx * az, bz = mx * ay - my * ax;
    out[i] = -(beta - alpha) * ax - (1.0 + alpha * beta) * bx;
    out[n + i] = -(beta - alpha) * ay - (1.0 + alpha * beta) * by;
    out[2 * n + i] = -(beta - alpha) * az - (1.0 + alpha * beta) * bz;
}

// max_i |m_i x H_i| (EffFieldGPU.cu:484-512): block max, then one integer
// atomic max per block into *out (zeroed beforehand).
__global__ __launch_bounds__(kBlock) void k_max_torque(int n, const double* __restrict__ m,
                                                       const double* __restrict__ h, double* out) {
    __shared__ double smem[kBlock / kWave];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double t = 0.0;
    if (i < n) {
        const double mx = m[i], my = m[n + i], mz = m[2 * n + i];
        const double hx = h[i], hy = h[n + i], hz = h[2 * n + i];
        const double ax = my * hz - mz * hy, ay = mz * hx - mx * hz, az = mx * hy - my * hx;
        t = sqrt(ax * ax + ay * ay + az * az);
    }
    t = wave_max(t);
    if ((threadIdx.x & (kWave - 1)) == 0) smem[threadIdx.x / kWave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = smem[0];
        for (int k = 1; k < kBlock / kWave; ++k) v = fmax(v, smem[k]);
        atomic_max_nonneg(out, v);
    }
}

__global__ void k_zero1(double* p) { *p = 0.0; }

// o += a
__global__ __launch_bounds__(kBlock) void k_add(int n, const double* __restrict__ a, double* __restrict__ o) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) o[i] += a[i];
}

// m <- normalize(m + dt dm)
__global__ __launch_bounds__(kBlock) void k_euler(int n, double* __restrict__ m, const double* __restrict__ dm,
                                                  double dt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double x = m[i] + dt * dm[i], y = m[n + i] + dt * dm[n + i], z = m[2 * n + i] + dt * dm[2 * n + i];
    const double l = sqrt(x * x + y * y + z * z);
    const double f = l > 0.0 ? 1.0 / l : 1.0;  // a zero vector stays zero (Eigen normalize())
    m[i] = x * f;
    m[n + i] = y * f;
    m[2 * n + i] = z * f;
}

// normalize functor (EffFieldGPU.cu:625-663)
__global__ __launch_bounds__(kBlock) void k_normalize(int n, double* __restrict__ m) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double x = m[i], y = m[n + i], z = m[2 * n + i];
    const double l = sqrt(x * x + y * y + z * z);
    if (l > 0.0) {  // Eigen's normalize() leaves a zero vector unchanged
        m[i] = x / l;
        m[n + i] = y / l;
        m[2 * n + i] = z / l;
    }
}

struct EvalArgs {
    SellView A;
    const double* xc;           // exchange values (or null)
    const double* gx; const double* gy; const double* gz;  // grad values (demag)
    const double* cx; const double* cy; const double* cz;  // curl values (DMI)
    const double* m;
    const double* u1; const double* u2;
    const double* invjs;        // may be null: invJs = 1
    const double* ku; const double* kaxis;
    const double* dmi_fac; const double* nvs;
    // interfacial DMI (null idmi: off): D per node, grad_x/y and tGrad_x/y
    // values, node volumes
    const double* idmi; const double* igx; const double* igy; const double* itx; const double* ity;
    const double* ivol;
    const double* hext;         // per-node (or null)
    double hx, hy, hz;          // uniform applied field
    double alpha;
    int variant;
    double* dmdt;               // may be null
    double* heff;               // may be null
    double* torque;             // may be null (zeroed before the launch)
};

// Interfacial DMI (interface normal z; no reference counterpart), the
// gradient of the lumped P1 energy sum_i D_i [m_z (Q_x m_x + Q_y m_y) -
// m_x Q_x m_z - m_y Q_y m_z]_i with Q_d = tGrad_d^T = diag(V) grad_d:
//   H_x = D grad_x m_z - tGrad_x (D m_z) / V,  H_y likewise with y,
//   H_z = -D (grad_x m_x + grad_y m_y) + (tGrad_x (D m_x) + tGrad_y (D m_y)) / V.
// ia: grad_x m_z, grad_y m_z, grad_x m_x + grad_y m_y, tGrad_x(D m_z),
//     tGrad_y(D m_z), tGrad_x(D m_x) + tGrad_y(D m_y)
__device__ inline void idmi_accumulate(const EvalArgs& a, int64_t q, int cc, double mx, double my, double mz,
                                       double (&ia)[8]) {
    const double gx = a.igx[q], gy = a.igy[q], tx = a.itx[q], ty = a.ity[q];
    const double Dc = a.idmi[cc];
    ia[0] += gx * mz;
    ia[1] += gy * mz;
    ia[2] += gx * mx + gy * my;
    ia[3] += tx * Dc * mz;
    ia[4] += ty * Dc * mz;
    ia[5] += (tx * mx + ty * my) * Dc;
}

__device__ inline void idmi_field(const EvalArgs& a, int row, const double (&ia)[8], double& hx, double& hy,
                                  double& hz) {
    const double D = a.idmi[row], iv = 1.0 / a.ivol[row];
    hx = D * ia[0] - iv * ia[3];
    hy = D * ia[1] - iv * ia[4];
    hz = -D * ia[2] + iv * ia[5];
}

// Fused Heff + LLG (TheLLG.cpp:238-254 and EffFieldGPU.cu:316-362):
// one pass over the shared pattern computes exchange, demag gradient and DMI
// curl for the row, then the pointwise terms, the LLG right-hand side and the
// torque maximum.
template <bool EXCH, bool DEMAG, bool DMI>
__global__ __launch_bounds__(kBlock) void k_eval(EvalArgs a) {
    __shared__ double smem[kBlock / kWave];
    const SellView A = a.A;
    const int n = A.n;
    const int row = xcd_block() * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    double t = 0.0;
    if (s < (n + kWave - 1) / kWave) {
        const double* __restrict__ m = a.m;
        double ex = 0, ey = 0, ez = 0, dx = 0, dy = 0, dz = 0, cx = 0, cy = 0, cz = 0;
        double ia[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // interfacial DMI row sums
        const bool idmi = a.idmi != nullptr;
        sell_row(A, s, lane, [&](int64_t q, int cc) {
            double mx = 0, my = 0, mz = 0;
            if (EXCH || DMI || idmi) {
                mx = m[cc];
                my = m[n + cc];
                mz = m[2 * n + cc];
            }
            if (idmi) idmi_accumulate(a, q, cc, mx, my, mz, ia);
            if (EXCH) {
                const double e = a.xc[q];
                ex += e * mx;
                ey += e * my;
                ez += e * mz;
            }
            if (DEMAG) {
                const double u = a.u1[cc] + a.u2[cc];
                dx += a.gx[q] * u;
                dy += a.gy[q] * u;
                dz += a.gz[q] * u;
            }
            if (DMI) {
                const double X = a.cx[q], Y = a.cy[q], Z = a.cz[q];
                cx += Y * mz - Z * my;
                cy += Z * mx - X * mz;
                cz += X * my - Y * mx;
            }
        });
        if (row < n) {
            const double mx = m[row], my = m[n + row], mz = m[2 * n + row];
            double hx = ex, hy = ey, hz = ez;  // J/m^3 terms
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
                hx += ((my * sz - mz * sy) - 2.0 * cx) * f;
                hy += ((mz * sx - mx * sz) - 2.0 * cy) * f;
                hz += ((mx * sy - my * sx) - 2.0 * cz) * f;
            }
            if (idmi) {
                double qx, qy, qz;
                idmi_field(a, row, ia, qx, qy, qz);
                hx += qx;
                hy += qy;
                hz += qz;
            }
            const double ij = a.invjs ? a.invjs[row] : 1.0;
            hx = hx * ij + a.hx;
            hy = hy * ij + a.hy;
            hz = hz * ij + a.hz;
            if (a.hext) {
                hx += a.hext[row];
                hy += a.hext[n + row];
                hz += a.hext[2 * n + row];
            }
            if (DEMAG) {
                hx -= dx;
                hy -= dy;
                hz -= dz;
            }
            if (a.heff) {
                a.heff[row] = hx;
                a.heff[n + row] = hy;
                a.heff[2 * n + row] = hz;
            }
            if (a.dmdt) {
                double o[3];
                llg_point(mx, my, mz, hx, hy, hz, a.alpha, a.variant, o);
                a.dmdt[row] = o[0];
                a.dmdt[n + row] = o[1];
                a.dmdt[2 * n + row] = o[2];
            }
            const double tx = my * hz - mz * hy, ty = mz * hx - mx * hz, tz = mx * hy - my * hx;
            t = sqrt(tx * tx + ty * ty + tz * tz);
        }
    }
    if (!a.torque) return;
    t = wave_max(t);
    if ((threadIdx.x & (kWave - 1)) == 0) smem[threadIdx.x / kWave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = smem[0];
        for (int k = 1; k < kBlock / kWave; ++k) v = fmax(v, smem[k]);
        atomic_max_nonneg(a.torque, v);
    }
}

template <bool E, bool D, bool M>
void launch_eval(const EvalArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_eval<E, D, M>), dim3(grid), dim3(kBlock), 0, s, a);
}

int eval_dispatch(const EvalArgs& a, bool e, bool d, bool m, int grid, hipStream_t s) {
    const int key = (e ? 4 : 0) | (d ? 2 : 0) | (m ? 1 : 0);
    switch (key) {
        case 0: launch_eval<false, false, false>(a, grid, s); break;
        case 1: launch_eval<false, false, true>(a, grid, s); break;
        case 2: launch_eval<false, true, false>(a, grid, s); break;
        case 3: launch_eval<false, true, true>(a, grid, s); break;
        case 4: launch_eval<true, false, false>(a, grid, s); break;
        case 5: launch_eval<true, false, true>(a, grid, s); break;
        case 6: launch_eval<true, true, false>(a, grid, s); break;
        default: launch_eval<true, true, true>(a, grid, s); break;
    }
    TM_HIP(hipGetLastError());
    return TM_OK;
}

// Energies and volume averages (TheSimulation.cpp:100-102, 384-396):
//   exchange  -1/2 sum V Hxc . m          (EffFieldCalc.cpp:65-69, Hxc in J/m^3)
//   uniaxial  sum V Ku - sum V Ku (k.m)^2 (EffFieldCalc.cpp:86-95)
//   zeeman    -sum Js V Hext . m          (TheLLG.cpp:345-348)
//   dmi       -1/2 sum V Hdmi . m         (EffFieldCalc.cpp:172-175)
//   demag     -1/2 sum Js V Hdem . m      (DemagField.cpp:97-99)
//   <m> = sum V m / sum V,  <H> = sum V Hext / sum V  (TheLLG.cpp:421-423)
// One SELL pass recomputes the row terms; each block writes kEnergyTerms
// partial sums, and k_energy_finish adds them in a fixed order (FP64).
constexpr int kEnergyTerms = 12;  // xc, ku, zee, dmi, dem, m(3), hext(3), vol

template <bool EXCH, bool DEMAG, bool DMI>
__global__ __launch_bounds__(kBlock) void k_energy(EvalArgs a, const double* __restrict__ vol,
                                                   const double* __restrict__ js, double* __restrict__ part) {
    __shared__ double smem[kEnergyTerms * (kBlock / kWave)];
    const SellView A = a.A;
    const int n = A.n;
    const int row = xcd_block() * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    double e[kEnergyTerms];
#pragma unroll
    for (int q = 0; q < kEnergyTerms; ++q) e[q] = 0.0;
    if (s < (n + kWave - 1) / kWave) {
        const double* __restrict__ m = a.m;
        double ex = 0, ey = 0, ez = 0, dx = 0, dy = 0, dz = 0, cx = 0, cy = 0, cz = 0;
        double ia[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // interfacial DMI row sums
        const bool idmi = a.idmi != nullptr;
        sell_row(A, s, lane, [&](int64_t q, int cc) {
            double mx = 0, my = 0, mz = 0;
            if (EXCH || DMI || idmi) {
                mx = m[cc];
                my = m[n + cc];
                mz = m[2 * n + cc];
            }
            if (idmi) idmi_accumulate(a, q, cc, mx, my, mz, ia);
            if (EXCH) {
                const double v = a.xc[q];
                ex += v * mx;
                ey += v * my;
                ez += v * mz;
            }
            if (DEMAG) {
                const double u = a.u1[cc] + a.u2[cc];
                dx += a.gx[q] * u;
                dy += a.gy[q] * u;
                dz += a.gz[q] * u;
            }
            if (DMI) {
                const double X = a.cx[q], Y = a.cy[q], Z = a.cz[q];
                cx += Y * mz - Z * my;
                cy += Z * mx - X * mz;
                cz += X * my - Y * mx;
            }
        });
        if (row < n) {
            const double V = vol[row];
            const double J = js ? js[row] : 0.0;
            const double mx = m[row], my = m[n + row], mz = m[2 * n + row];
            if (EXCH) e[0] = V * (ex * mx + ey * my + ez * mz);
            if (a.ku) {
                const double kx = a.kaxis[row], ky = a.kaxis[n + row], kz = a.kaxis[2 * n + row];
                const double d = kx * mx + ky * my + kz * mz;
                e[1] = V * a.ku[row] - a.ku[row] * V * d * d;
            }
            double hx = a.hx, hy = a.hy, hz = a.hz;
            if (a.hext) {
                hx += a.hext[row];
                hy += a.hext[n + row];
                hz += a.hext[2 * n + row];
            }
            e[2] = J * V * (hx * mx + hy * my + hz * mz);
            if (DMI) {
                const double sx = a.nvs[row], sy = a.nvs[n + row], sz = a.nvs[2 * n + row];
                const double f = a.dmi_fac[row];
                const double px = ((my * sz - mz * sy) - 2.0 * cx) * f;
                const double py = ((mz * sx - mx * sz) - 2.0 * cy) * f;
                const double pz = ((mx * sy - my * sx) - 2.0 * cz) * f;
                e[3] = V * (px * mx + py * my + pz * mz);
            }
            if (idmi) {
                double qx, qy, qz;
                idmi_field(a, row, ia, qx, qy, qz);
                e[3] += V * (qx * mx + qy * my + qz * mz);
            }
            if (DEMAG) e[4] = -J * V * (dx * mx + dy * my + dz * mz);  // Hdem = -grad u
            e[5] = V * mx;
            e[6] = V * my;
            e[7] = V * mz;
            e[8] = V * hx;
            e[9] = V * hy;
            e[10] = V * hz;
            e[11] = V;
        }
    }
    block_sum<kEnergyTerms>(e, smem);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < kEnergyTerms; ++q) part[(size_t)q * gridDim.x + blockIdx.x] = e[q];
}

__global__ __launch_bounds__(kBlock) void k_energy_finish(int G, const double* __restrict__ part,
                                                          double* __restrict__ out) {
    double t[kEnergyTerms];
    reduce_partials_all<kEnergyTerms>(part, G, G, t);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < kEnergyTerms; ++q) out[q] = t[q];
}

template <bool E, bool D, bool M>
void launch_energy(const EvalArgs& a, const double* vol, const double* js, double* part, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_energy<E, D, M>), dim3(grid), dim3(kBlock), 0, s, a, vol, js, part);
}

int energy_dispatch(const EvalArgs& a, bool e, bool d, bool m, const double* vol, const double* js, double* part,
                    int grid, hipStream_t s) {
    const int key = (e ? 4 : 0) | (d ? 2 : 0) | (m ? 1 : 0);
    switch (key) {
        case 0: launch_energy<false, false, false>(a, vol, js, part, grid, s); break;
        case 1: launch_energy<false, false, true>(a, vol, js, part, grid, s); break;
        case 2: launch_energy<false, true, false>(a, vol, js, part, grid, s); break;
        case 3: launch_energy<false, true, true>(a, vol, js, part, grid, s); break;
        case 4: launch_energy<true, false, false>(a, vol, js, part, grid, s); break;
        case 5: launch_energy<true, false, true>(a, vol, js, part, grid, s); break;
        case 6: launch_energy<true, true, false>(a, vol, js, part, grid, s); break;
        default: launch_energy<true, true, true>(a, vol, js, part, grid, s); break;
    }
    TM_HIP(hipGetLastError());
    return TM_OK;
}

}  // namespace

namespace tmh {
int sell_grid(const tm_ctx* c) { return xcd_grid(grid_for((int64_t)c->pat.nslices * kWave)); }

// Stage a caller pointer as a device pointer (copying host data into `tmp`).
int stage_in(tm_ctx* c, const double* src, int where, size_t count, DevBuf<double>& tmp, const double** out) {
    if (!src) return TM_ERR_BAD_ARG;
    if (where == TM_DEVICE) {
        *out = src;
        return TM_OK;
    }
    if (where != TM_HOST) return TM_ERR_BAD_ARG;
    TM_CHECK(tmp.alloc(std::max(tmp.n, count)));
    TM_CHECK(copy_in(tmp.p, src, TM_HOST, count, c->stream));
    *out = tmp.p;
    return TM_OK;
}

int finish_out(tm_ctx* c, double* dst, int where, const double* dev, size_t count) {
    if (where == TM_DEVICE) {
        if (dst != dev) TM_HIP(hipMemcpyAsync(dst, dev, count * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        return TM_OK;
    }
    if (where != TM_HOST) return TM_ERR_BAD_ARG;
    return copy_out(dst, TM_HOST, dev, count, c->stream);
}

// Output target: the caller's device buffer or the context scratch w3a.
double* out_target(tm_ctx* c, double* dst, int where) { return where == TM_DEVICE ? dst : c->w3a.p; }

bool idmi_ready(const tm_ctx* c) {
    return c->has_idmi && c->has_vol && c->has_op[TM_OP_GRAD_X] && c->has_op[TM_OP_GRAD_Y] &&
           c->has_op[TM_OP_TGRAD_X] && c->has_op[TM_OP_TGRAD_Y];
}

void fill_idmi(const tm_ctx* c, EvalArgs& a) {
    if (!idmi_ready(c)) return;
    a.idmi = c->idmi.p;
    a.igx = c->opv[TM_OP_GRAD_X].p;
    a.igy = c->opv[TM_OP_GRAD_Y].p;
    a.itx = c->opv[TM_OP_TGRAD_X].p;
    a.ity = c->opv[TM_OP_TGRAD_Y].p;
    a.ivol = c->vol.p;
}

int upload_vec(tm_ctx* c, DevBuf<double>& d, const double* h, size_t count) {
    if (!h) return TM_ERR_BAD_ARG;
    TM_CHECK(d.upload(h, count, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    return TM_OK;
}

}  // namespace tmh

namespace {
}  // namespace

extern "C" {

int tm_ctx_create(int n, tm_ctx** out) {
    if (n <= 0 || !out) return TM_ERR_BAD_ARG;
    int nd = 0;
    TM_CHECK(tm_device_count(&nd));
    tm_ctx* c = new tm_ctx();
    c->n = n;
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return TM_ERR_HIP;
    }
    c->stream = c->own_stream;
    int st = c->m.alloc((size_t)3 * n);
    if (st == TM_OK) st = c->w3a.alloc((size_t)3 * n);
    if (st == TM_OK) st = c->w3b.alloc((size_t)3 * n);
    if (st == TM_OK) st = c->wn.alloc(n);
    if (st == TM_OK) st = c->scal.alloc(4);
    if (st == TM_OK && hipMemsetAsync(c->m.p, 0, 3 * sizeof(double) * n, c->stream) != hipSuccess) st = TM_ERR_HIP;
    if (st == TM_OK && hipStreamSynchronize(c->stream) != hipSuccess) st = TM_ERR_HIP;
    if (st != TM_OK) {
        tm_ctx_destroy(c);
        return st;
    }
    *out = c;
    return TM_OK;
}

int tm_ctx_destroy(tm_ctx* c) {
    if (!c) return TM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return TM_OK;
}

int tm_ctx_set_stream(tm_ctx* c, void* s) {
    TM_CHECK(ctx_ok(c));
    TM_HIP(hipStreamSynchronize(c->stream));
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return TM_OK;
}

int tm_ctx_synchronize(tm_ctx* c) {
    TM_CHECK(ctx_ok(c));
    TM_HIP(hipStreamSynchronize(c->stream));
    return TM_OK;
}

int tm_ctx_set_op(tm_ctx* c, int op, const int* rowptr, const int* colidx, const double* vals) {
    TM_CHECK(ctx_ok(c));
    if (op < 0 || op >= TM_NUM_OPS || !rowptr || !colidx || !vals) return TM_ERR_BAD_ARG;
    const int n = c->n;
    if (rowptr[0] != 0) return TM_ERR_BAD_ARG;
    for (int i = 0; i < n; ++i) {
        if (rowptr[i + 1] < rowptr[i]) return TM_ERR_BAD_ARG;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q)
            if (colidx[q] < 0 || colidx[q] >= n || (q > rowptr[i] && colidx[q] <= colidx[q - 1]))
                return TM_ERR_BAD_ARG;
    }
    auto& h = c->hop[op];
    h.set = true;
    h.rp.assign(rowptr, rowptr + n + 1);
    h.col.assign(colidx, colidx + rowptr[n]);
    h.val.assign(vals, vals + rowptr[n]);
    c->ops_dirty = true;
    return TM_OK;
}

int tm_ctx_set_js(tm_ctx* c, const double* Js) {
    TM_CHECK(ctx_ok(c));
    if (!Js) return TM_ERR_BAD_ARG;
    std::vector<double> inv((size_t)c->n);
    for (int i = 0; i < c->n; ++i) {
        const double v = 1.0 / Js[i];
        inv[i] = std::isfinite(v) ? v : 0.0;  // TheLLG.cpp:110
    }
    TM_CHECK(upload_vec(c, c->js, Js, c->n));
    TM_CHECK(upload_vec(c, c->invjs, inv.data(), c->n));
    c->has_js = true;
    return TM_OK;
}

int tm_ctx_set_uniaxial(tm_ctx* c, const double* axis3n, const double* Ku) {
    TM_CHECK(ctx_ok(c));
    if (!axis3n || !Ku) return TM_ERR_BAD_ARG;
    TM_CHECK(upload_vec(c, c->kaxis, axis3n, (size_t)3 * c->n));
    TM_CHECK(upload_vec(c, c->ku, Ku, c->n));
    c->has_ku = true;
    return TM_OK;
}

int tm_ctx_set_dmi(tm_ctx* c, const double* D, const double* invNodeVol, const double* nv3n, const double* nodeArea) {
    TM_CHECK(ctx_ok(c));
    if (!D || !invNodeVol || !nv3n || !nodeArea) return TM_ERR_BAD_ARG;
    const int n = c->n;
    std::vector<double> fac((size_t)n), nvs((size_t)3 * n);
    for (int i = 0; i < n; ++i) {
        fac[i] = invNodeVol[i] * D[i];  // EffFieldGPU.cu:305-306
        for (int d = 0; d < 3; ++d) nvs[(size_t)d * n + i] = nv3n[(size_t)d * n + i] * nodeArea[i];
    }
    TM_CHECK(upload_vec(c, c->dmi_fac, fac.data(), n));
    TM_CHECK(upload_vec(c, c->nvs, nvs.data(), (size_t)3 * n));
    c->has_dmi = true;
    return TM_OK;
}

int tm_ctx_set_stt_eta(tm_ctx* c, const double* ex, const double* ey, const double* ez) {
    TM_CHECK(ctx_ok(c));
    if (!ex || !ey || !ez) return TM_ERR_BAD_ARG;
    const int n = c->n;
    std::vector<double> e((size_t)3 * n);
    std::memcpy(e.data(), ex, sizeof(double) * n);
    std::memcpy(e.data() + n, ey, sizeof(double) * n);
    std::memcpy(e.data() + 2 * (size_t)n, ez, sizeof(double) * n);
    TM_CHECK(upload_vec(c, c->eta, e.data(), e.size()));
    c->has_eta = true;
    return TM_OK;
}

int tm_ctx_set_zeeman(tm_ctx* c, const double* H3n, double hx, double hy, double hz) {
    TM_CHECK(ctx_ok(c));
    c->hext_u[0] = hx;
    c->hext_u[1] = hy;
    c->hext_u[2] = hz;
    if (H3n) {
        TM_CHECK(upload_vec(c, c->hext, H3n, (size_t)3 * c->n));
        c->hext_field = true;
    } else {
        c->hext_field = false;
    }
    return TM_OK;
}

int tm_ctx_set_mag(tm_ctx* c, const double* m3n, int where) {
    TM_CHECK(ctx_ok(c));
    return copy_in(c->m.p, m3n, where, (size_t)3 * c->n, c->stream);
}

int tm_ctx_get_mag(tm_ctx* c, double* m3n, int where) {
    TM_CHECK(ctx_ok(c));
    return copy_out(m3n, where, c->m.p, (size_t)3 * c->n, c->stream);
}

int tm_ctx_normalize_mag(tm_ctx* c, double* m3n, int where) {
    TM_CHECK(ctx_ok(c));
    const int n = c->n;
    TM_CHECK(copy_in(c->m.p, m3n, where, (size_t)3 * n, c->stream));
    hipLaunchKernelGGL(k_normalize, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, n, c->m.p);
    TM_HIP(hipGetLastError());
    return copy_out(m3n, where, c->m.p, (size_t)3 * n, c->stream);
}

int tm_exchange_field(tm_ctx* c, double* out, int where) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    TM_CHECK(commit_ops(c));
    if (!c->has_op[TM_OP_EXCHANGE]) return TM_ERR_NOT_READY;
    double* o = out_target(c, out, where);
    hipLaunchKernelGGL(k_exchange, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat),
                       c->opv[TM_OP_EXCHANGE].p, c->m.p, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, out, where, o, (size_t)3 * c->n);
}

int tm_uniaxial_field(tm_ctx* c, double* out, int where) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    if (!c->has_ku) return TM_ERR_NOT_READY;
    double* o = out_target(c, out, where);
    hipLaunchKernelGGL(k_uniaxial, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->ku.p, c->kaxis.p,
                       c->m.p, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, out, where, o, (size_t)3 * c->n);
}

int tm_cubic_field(tm_ctx* c, double* out, int where) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    // The reference GPU class returns a zero matrix (EffFieldGPU.cu:514-535).
    double* o = out_target(c, out, where);
    TM_HIP(hipMemsetAsync(o, 0, sizeof(double) * 3 * c->n, c->stream));
    return finish_out(c, out, where, o, (size_t)3 * c->n);
}

int tm_dmi_field(tm_ctx* c, double* out, int where) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    TM_CHECK(commit_ops(c));
    if (!c->has_dmi || !c->has_op[TM_OP_CURL_X] || !c->has_op[TM_OP_CURL_Y] || !c->has_op[TM_OP_CURL_Z])
        return TM_ERR_NOT_READY;
    double* o = out_target(c, out, where);
    hipLaunchKernelGGL(k_dmi, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat), c->opv[TM_OP_CURL_X].p,
                       c->opv[TM_OP_CURL_Y].p, c->opv[TM_OP_CURL_Z].p, c->m.p, c->dmi_fac.p, c->nvs.p, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, out, where, o, (size_t)3 * c->n);
}

int tm_stt_uterm(tm_ctx* c, double* out, int where) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    TM_CHECK(commit_ops(c));
    if (!c->has_eta || !c->has_op[TM_OP_STT_GRAD_X] || !c->has_op[TM_OP_STT_GRAD_Y] || !c->has_op[TM_OP_STT_GRAD_Z])
        return TM_ERR_NOT_READY;
    double* o = out_target(c, out, where);
    hipLaunchKernelGGL(k_stt_uterm, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat),
                       c->opv[TM_OP_STT_GRAD_X].p, c->opv[TM_OP_STT_GRAD_Y].p, c->opv[TM_OP_STT_GRAD_Z].p,
                       c->eta.p, c->m.p, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, out, where, o, (size_t)3 * c->n);
}

int tm_llg_rhs(tm_ctx* c, const double* H, int hwhere, double alpha, int variant, double* out, int owhere) {
    TM_CHECK(ctx_ok(c));
    if (!out || (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC)) return TM_ERR_BAD_ARG;
    const double* hd = nullptr;
    TM_CHECK(stage_in(c, H, hwhere, (size_t)3 * c->n, c->w3b, &hd));
    double* o = out_target(c, out, owhere);
    hipLaunchKernelGGL(k_llg, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->m.p, hd, alpha, variant, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, out, owhere, o, (size_t)3 * c->n);
}

int tm_stt_llg(tm_ctx* c, const double* U, int uwhere, double alpha, double beta, double* out, int owhere) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    const double* ud = nullptr;
    TM_CHECK(stage_in(c, U, uwhere, (size_t)3 * c->n, c->w3b, &ud));
    double* o = out_target(c, out, owhere);
    hipLaunchKernelGGL(k_stt_llg, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->m.p, ud, alpha, beta, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, out, owhere, o, (size_t)3 * c->n);
}

int tm_max_torque(tm_ctx* c, const double* H, int hwhere, double* out) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    const double* hd = nullptr;
    TM_CHECK(stage_in(c, H, hwhere, (size_t)3 * c->n, c->w3b, &hd));
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, c->stream, c->scal.p);
    hipLaunchKernelGGL(k_max_torque, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->m.p, hd, c->scal.p);
    TM_HIP(hipGetLastError());
    TM_HIP(hipMemcpyAsync(out, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    return TM_OK;
}

int tm_div_m(tm_ctx* c, const double* m3n, int mwhere, double* b, int bwhere) {
    TM_CHECK(ctx_ok(c));
    if (!b) return TM_ERR_BAD_ARG;
    TM_CHECK(commit_ops(c));
    if (!c->has_js || !c->has_op[TM_OP_TGRAD_X] || !c->has_op[TM_OP_TGRAD_Y] || !c->has_op[TM_OP_TGRAD_Z])
        return TM_ERR_NOT_READY;
    const double* md = nullptr;
    TM_CHECK(stage_in(c, m3n, mwhere, (size_t)3 * c->n, c->w3b, &md));
    double* o = bwhere == TM_DEVICE ? b : c->wn.p;
    hipLaunchKernelGGL(k_div, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat), c->opv[TM_OP_TGRAD_X].p,
                       c->opv[TM_OP_TGRAD_Y].p, c->opv[TM_OP_TGRAD_Z].p, c->js.p, md, 1.0, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, b, bwhere, o, (size_t)c->n);
}

int tm_grad_field(tm_ctx* c, const double* u, int uwhere, double* H, int hwhere) {
    TM_CHECK(ctx_ok(c));
    if (!H) return TM_ERR_BAD_ARG;
    TM_CHECK(commit_ops(c));
    if (!c->has_op[TM_OP_GRAD_X] || !c->has_op[TM_OP_GRAD_Y] || !c->has_op[TM_OP_GRAD_Z]) return TM_ERR_NOT_READY;
    const double* ud = nullptr;
    TM_CHECK(stage_in(c, u, uwhere, (size_t)c->n, c->w3b, &ud));
    double* o = out_target(c, H, hwhere);
    hipLaunchKernelGGL(k_grad, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat), c->opv[TM_OP_GRAD_X].p,
                       c->opv[TM_OP_GRAD_Y].p, c->opv[TM_OP_GRAD_Z].p, ud, (const double*)nullptr, o);
    TM_HIP(hipGetLastError());
    return finish_out(c, H, hwhere, o, (size_t)3 * c->n);
}

int tm_ctx_set_boundary(tm_ctx* c, int bnx, const int* bnodes) {
    TM_CHECK(ctx_ok(c));
    if (bnx <= 0 || bnx > c->n || !bnodes) return TM_ERR_BAD_ARG;
    for (int i = 0; i < bnx; ++i)
        if (bnodes[i] < 0 || bnodes[i] >= c->n || (i > 0 && bnodes[i] <= bnodes[i - 1])) return TM_ERR_BAD_ARG;
    if (c->bnx != bnx) {
        c->has_B = false;
        c->B.release();
    }
    c->bnx = bnx;
    c->bld = (bnx + 63) / 64 * 64;
    TM_CHECK(c->bnodes.upload(bnodes, bnx, c->stream));
    TM_CHECK(c->bpart.alloc((size_t)bem_gemv_splits(bnx) * c->bld));
    TM_CHECK(c->bvec_a.alloc(c->bld));
    TM_CHECK(c->bvec_b.alloc(c->bld));
    TM_HIP(hipStreamSynchronize(c->stream));
    return TM_OK;
}

int tm_ctx_set_bem_dense(tm_ctx* c, const double* B, int ld, int where) {
    TM_CHECK(ctx_ok(c));
    if (c->bnx <= 0) return TM_ERR_NOT_READY;
    if (!B || ld < c->bnx || (where != TM_HOST && where != TM_DEVICE)) return TM_ERR_BAD_ARG;
    const int b = c->bnx;
    TM_CHECK(c->B.alloc((size_t)c->bld * b));
    TM_HIP(hipMemsetAsync(c->B.p, 0, sizeof(double) * (size_t)c->bld * b, c->stream));
    TM_HIP(hipMemcpy2DAsync(c->B.p, sizeof(double) * c->bld, B, sizeof(double) * ld, sizeof(double) * b, b,
                            where == TM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    c->has_B = true;
    return TM_OK;
}

int tm_ctx_assemble_bem(tm_ctx* c, int nbel, const double* bxyz, const int* faces, const double* face_area,
                        const double* face_normal, const double* solid_angle, int formula) {
    TM_CHECK(ctx_ok(c));
    if (c->bnx <= 0) return TM_ERR_NOT_READY;
    if (nbel <= 0 || !bxyz || !faces || !face_area || !face_normal || !solid_angle) return TM_ERR_BAD_ARG;
    if (formula != TM_BEM_REFERENCE && formula != TM_BEM_ACCURATE) return TM_ERR_BAD_ARG;
    const int b = c->bnx;
    for (int f = 0; f < 3 * nbel; ++f)
        if (faces[f] < 0 || faces[f] >= b) return TM_ERR_BAD_ARG;
    DevBuf<double> dx, da, dn, dom;
    DevBuf<int> df;
    TM_CHECK(dx.upload(bxyz, (size_t)3 * b, c->stream));
    TM_CHECK(df.upload(faces, (size_t)3 * nbel, c->stream));
    TM_CHECK(da.upload(face_area, nbel, c->stream));
    TM_CHECK(dn.upload(face_normal, (size_t)3 * nbel, c->stream));
    TM_CHECK(dom.upload(solid_angle, b, c->stream));
    TM_CHECK(c->B.alloc((size_t)c->bld * b));
    TM_HIP(hipMemsetAsync(c->B.p, 0, sizeof(double) * (size_t)c->bld * b, c->stream));
    TM_CHECK(bem_assemble(b, c->bld, nbel, dx.p, df.p, da.p, dn.p, dom.p, c->B.p, formula, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    c->has_B = true;
    return TM_OK;
}

int tm_ctx_get_bem_dense(tm_ctx* c, double* B, int where) {
    TM_CHECK(ctx_ok(c));
    if (!c->has_B) return TM_ERR_NOT_READY;
    if (!B || (where != TM_HOST && where != TM_DEVICE)) return TM_ERR_BAD_ARG;
    const int b = c->bnx;
    TM_HIP(hipMemcpy2DAsync(B, sizeof(double) * b, c->B.p, sizeof(double) * c->bld, sizeof(double) * b, b,
                            where == TM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    return TM_OK;
}

int tm_bem_apply(tm_ctx* c, const double* ub, int uwhere, double* gb, int gwhere) {
    TM_CHECK(ctx_ok(c));
    if (!c->has_B) return TM_ERR_NOT_READY;
    if (!gb) return TM_ERR_BAD_ARG;
    const double* ud = nullptr;
    TM_CHECK(stage_in(c, ub, uwhere, (size_t)c->bnx, c->bvec_a, &ud));
    double* o = gwhere == TM_DEVICE ? gb : c->bvec_b.p;
    TM_CHECK(bem_gemv(c->bnx, c->bld, c->B.p, ud, nullptr, c->bpart.p, o, nullptr, nullptr, c->stream));
    return finish_out(c, gb, gwhere, o, (size_t)c->bnx);
}

int tm_ctx_set_demag_solvers(tm_ctx* c, const int* rpN, const int* colN, const double* valN, const int* rpD,
                             const int* colD, const double* valD, int precond) {
    TM_CHECK(ctx_ok(c));
    if (!rpN || !colN || !valN || !rpD || !colD || !valD) return TM_ERR_BAD_ARG;
    const int n = c->n;
    TM_CHECK(c->cgN.setup(n, rpN, colN, valN, precond, c->stream));
    TM_CHECK(c->cgD.setup(n, rpD, colD, valD, precond, c->stream));
    TM_CHECK(c->rhsN.alloc(n));
    TM_CHECK(c->rhsD.alloc(n));
    TM_CHECK(c->u1.alloc(n));
    TM_CHECK(c->u2.alloc(n));
    TM_HIP(hipMemsetAsync(c->rhsD.p, 0, sizeof(double) * n, c->stream));
    TM_HIP(hipMemsetAsync(c->u1.p, 0, sizeof(double) * n, c->stream));
    TM_HIP(hipMemsetAsync(c->u2.p, 0, sizeof(double) * n, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    c->has_solvers = true;
    return TM_OK;
}

int tm_ctx_solver_amg_info(tm_ctx* c, int which, int* nlevels, int* sizes, int cap) {
    TM_CHECK(ctx_ok(c));
    if (!c->has_solvers) return TM_ERR_NOT_READY;
    if (which != 0 && which != 1) return TM_ERR_BAD_ARG;
    return amg_info(which == 0 ? c->cgN : c->cgD, nlevels, sizes, cap);
}

int tm_ctx_set_solver_params(tm_ctx* c, double tol, int maxit) {
    TM_CHECK(ctx_ok(c));
    if (!(tol > 0.0) || maxit < 1) return TM_ERR_BAD_ARG;
    c->tol = tol;
    c->maxit = maxit;
    return TM_OK;
}

}  // extern "C"

namespace tmh {

// DemagField::calcPotential (DemagField.cpp:163-181) on device data; leaves
// u1, u2 in the context.
int demag_potential_dev(tm_ctx* c, const double* md, tm_eval_stats* stats) {
    if (!c->has_solvers || !c->has_B || !c->has_js) return TM_ERR_NOT_READY;
    TM_CHECK(commit_ops(c));
    if (!c->has_op[TM_OP_TGRAD_X] || !c->has_op[TM_OP_TGRAD_Y] || !c->has_op[TM_OP_TGRAD_Z]) return TM_ERR_NOT_READY;
    if (c->timers) TM_HIP(hipEventRecord(c->ev[0], c->stream));
    // rhsPoisson (DemagField.cpp:124-136): divM / mu0
    hipLaunchKernelGGL(k_div, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat), c->opv[TM_OP_TGRAD_X].p,
                       c->opv[TM_OP_TGRAD_Y].p, c->opv[TM_OP_TGRAD_Z].p, c->js.p, md, 1.0 / kMu0, c->rhsN.p);
    TM_HIP(hipGetLastError());
    int itN = 0, itD = 0;
    double rN = 0, rD = 0;
    if (c->timers) TM_HIP(hipEventRecord(c->ev[1], c->stream));
    // A non-finite solve leaves NaN/Inf in the warm-start vectors; they are
    // cleared so that the next call with a finite m starts from zero.
    auto clear_potentials = [c]() {
        (void)hipMemsetAsync(c->u1.p, 0, sizeof(double) * c->n, c->stream);
        (void)hipMemsetAsync(c->u2.p, 0, sizeof(double) * c->n, c->stream);
        c->have_u = false;
    };
    int st = c->cgN.solve(c->rhsN.p, c->u1.p, c->tol, c->maxit, c->stream, &itN, &rN, true);
    if (st == TM_ERR_NONFINITE) clear_potentials();
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    int stN = st;
    if (c->timers) TM_HIP(hipEventRecord(c->ev[2], c->stream));
    // boundaryIntegral (DemagField.cpp:184-201): gather, B product, scatter
    TM_CHECK(bem_gemv(c->bnx, c->bld, c->B.p, c->u1.p, c->bnodes.p, c->bpart.p, nullptr, c->bnodes.p, c->rhsD.p,
                      c->stream));
    if (c->timers) TM_HIP(hipEventRecord(c->ev[3], c->stream));
    st = c->cgD.solve(c->rhsD.p, c->u2.p, c->tol, c->maxit, c->stream, &itD, &rD, true);
    if (c->timers) TM_HIP(hipEventRecord(c->ev[4], c->stream));
    c->tsum.iters_neumann += itN;
    c->tsum.iters_dirichlet += itD;
    if (st == TM_ERR_NONFINITE) clear_potentials();
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    c->have_u = true;
    if (stats) {
        stats->iters_neumann = itN;
        stats->iters_dirichlet = itD;
        stats->relres_neumann = rN;
        stats->relres_dirichlet = rD;
    }
    return stN != TM_OK ? stN : st;
}

}  // namespace

namespace tmh {
// One right-hand side evaluation from a device magnetisation `md` into dmdt,
// on the context's stream.  budgetN > 0 selects fixed-budget solves (no host
// synchronisation; failures counted on the device), otherwise the solves poll
// the host and `stats` receives the iteration counts.  Demag is skipped
// (frozen potentials kept) when freeze_demag is on and potentials exist.
int eval_device(tm_ctx* c, const double* md, double* dmdt, double alpha, int variant, int budgetN, int budgetD,
                unsigned* fails, int* imax2, tm_eval_stats* stats) {
    TM_CHECK(commit_ops(c));
    const bool demag = c->has_solvers && c->has_B;
    const bool solve = demag && !(c->freeze_demag && c->have_u);
    if (demag && (!c->has_op[TM_OP_GRAD_X] || !c->has_op[TM_OP_GRAD_Y] || !c->has_op[TM_OP_GRAD_Z]))
        return TM_ERR_NOT_READY;
    int st = TM_OK;
    if (solve) {
        if (budgetN > 0) {
            if (!c->has_js || !c->has_op[TM_OP_TGRAD_X] || !c->has_op[TM_OP_TGRAD_Y] || !c->has_op[TM_OP_TGRAD_Z])
                return TM_ERR_NOT_READY;
            hipLaunchKernelGGL(k_div, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat),
                               c->opv[TM_OP_TGRAD_X].p, c->opv[TM_OP_TGRAD_Y].p, c->opv[TM_OP_TGRAD_Z].p, c->js.p, md,
                               1.0 / kMu0, c->rhsN.p);
            TM_HIP(hipGetLastError());
            TM_CHECK(c->cgN.enqueue(c->rhsN.p, c->u1.p, c->tol, budgetN, c->stream, fails, imax2, true));
            TM_CHECK(bem_gemv(c->bnx, c->bld, c->B.p, c->u1.p, c->bnodes.p, c->bpart.p, nullptr, c->bnodes.p,
                              c->rhsD.p, c->stream));
            TM_CHECK(c->cgD.enqueue(c->rhsD.p, c->u2.p, c->tol, budgetD, c->stream, fails, imax2 + 1, true));
            c->have_u = true;
        } else {
            st = demag_potential_dev(c, md, stats);
            if (st != TM_OK && st != TM_ERR_MAXITER) return st;
        }
    }
    const bool exch = c->has_op[TM_OP_EXCHANGE];
    const bool dmi = c->has_dmi && c->has_op[TM_OP_CURL_X] && c->has_op[TM_OP_CURL_Y] && c->has_op[TM_OP_CURL_Z];
    EvalArgs a = {};
    a.A = view(c->pat);
    a.xc = exch ? c->opv[TM_OP_EXCHANGE].p : nullptr;
    if (demag) {
        a.gx = c->opv[TM_OP_GRAD_X].p;
        a.gy = c->opv[TM_OP_GRAD_Y].p;
        a.gz = c->opv[TM_OP_GRAD_Z].p;
        a.u1 = c->u1.p;
        a.u2 = c->u2.p;
    }
    if (dmi) {
        a.cx = c->opv[TM_OP_CURL_X].p;
        a.cy = c->opv[TM_OP_CURL_Y].p;
        a.cz = c->opv[TM_OP_CURL_Z].p;
        a.dmi_fac = c->dmi_fac.p;
        a.nvs = c->nvs.p;
    }
    fill_idmi(c, a);
    a.m = md;
    a.invjs = c->has_js ? c->invjs.p : nullptr;
    a.ku = c->has_ku ? c->ku.p : nullptr;
    a.kaxis = c->has_ku ? c->kaxis.p : nullptr;
    a.hext = c->hext_field ? c->hext.p : nullptr;
    a.hx = c->hext_u[0];
    a.hy = c->hext_u[1];
    a.hz = c->hext_u[2];
    a.alpha = alpha;
    a.variant = variant;
    a.dmdt = dmdt;
    TM_CHECK(eval_dispatch(a, exch, demag, dmi, sell_grid(c), c->stream));
    return st;
}
}  // namespace tmh

extern "C" {

int tm_demag_potential(tm_ctx* c, const double* m3n, int mwhere, double* u, int uwhere, tm_eval_stats* stats) {
    TM_CHECK(ctx_ok(c));
    if (!u) return TM_ERR_BAD_ARG;
    const double* md = c->m.p;
    if (m3n) TM_CHECK(stage_in(c, m3n, mwhere, (size_t)3 * c->n, c->w3b, &md));
    int st = demag_potential_dev(c, md, stats);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    // u = u1 + u2
    double* o = uwhere == TM_DEVICE ? u : c->wn.p;
    TM_HIP(hipMemcpyAsync(o, c->u1.p, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_add, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, (const double*)c->u2.p, o);
    TM_HIP(hipGetLastError());
    int st2 = finish_out(c, u, uwhere, o, (size_t)c->n);
    return st != TM_OK ? st : st2;
}

int tm_demag_field(tm_ctx* c, const double* m3n, int mwhere, double* H, int hwhere, tm_eval_stats* stats) {
    TM_CHECK(ctx_ok(c));
    if (!H) return TM_ERR_BAD_ARG;
    const double* md = c->m.p;
    if (m3n) TM_CHECK(stage_in(c, m3n, mwhere, (size_t)3 * c->n, c->w3b, &md));
    int st = demag_potential_dev(c, md, stats);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    if (!c->has_op[TM_OP_GRAD_X] || !c->has_op[TM_OP_GRAD_Y] || !c->has_op[TM_OP_GRAD_Z]) return TM_ERR_NOT_READY;
    double* o = out_target(c, H, hwhere);
    hipLaunchKernelGGL(k_grad, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat), c->opv[TM_OP_GRAD_X].p,
                       c->opv[TM_OP_GRAD_Y].p, c->opv[TM_OP_GRAD_Z].p, (const double*)c->u1.p,
                       (const double*)c->u2.p, o);
    TM_HIP(hipGetLastError());
    int st2 = finish_out(c, H, hwhere, o, (size_t)3 * c->n);
    return st != TM_OK ? st : st2;
}

int tm_demag_last_potentials(tm_ctx* c, double* u1, double* u2, int where) {
    TM_CHECK(ctx_ok(c));
    if (!c->has_solvers) return TM_ERR_NOT_READY;
    if (u1) TM_CHECK(copy_out(u1, where, c->u1.p, c->n, c->stream));
    if (u2) TM_CHECK(copy_out(u2, where, c->u2.p, c->n, c->stream));
    return TM_OK;
}

int tm_eval_rhs(tm_ctx* c, const double* m3n, int mwhere, double alpha, int variant, double* dmdt, int dwhere,
                double* Heff, int hwhere, tm_eval_stats* stats) {
    TM_CHECK(ctx_ok(c));
    if (!dmdt || (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC)) return TM_ERR_BAD_ARG;
    if ((dwhere != TM_HOST && dwhere != TM_DEVICE) || (Heff && hwhere != TM_HOST && hwhere != TM_DEVICE))
        return TM_ERR_BAD_ARG;
    const int n = c->n;
    if (m3n) TM_CHECK(copy_in(c->m.p, m3n, mwhere, (size_t)3 * n, c->stream));  // setMagDev
    TM_CHECK(commit_ops(c));
    tm_eval_stats local = {};
    const bool demag = c->has_solvers && c->has_B;
    const bool solve = demag && !(c->freeze_demag && c->have_u);
    int st = TM_OK;
    if (demag) {
        if (solve) st = demag_potential_dev(c, c->m.p, &local);
        if (st != TM_OK && st != TM_ERR_MAXITER) return st;
        if (!c->has_op[TM_OP_GRAD_X] || !c->has_op[TM_OP_GRAD_Y] || !c->has_op[TM_OP_GRAD_Z]) return TM_ERR_NOT_READY;
    }
    const bool exch = c->has_op[TM_OP_EXCHANGE];
    const bool dmi = c->has_dmi && c->has_op[TM_OP_CURL_X] && c->has_op[TM_OP_CURL_Y] && c->has_op[TM_OP_CURL_Z];
    EvalArgs a = {};
    a.A = view(c->pat);
    a.xc = exch ? c->opv[TM_OP_EXCHANGE].p : nullptr;
    if (demag) {
        a.gx = c->opv[TM_OP_GRAD_X].p;
        a.gy = c->opv[TM_OP_GRAD_Y].p;
        a.gz = c->opv[TM_OP_GRAD_Z].p;
        a.u1 = c->u1.p;
        a.u2 = c->u2.p;
    }
    if (dmi) {
        a.cx = c->opv[TM_OP_CURL_X].p;
        a.cy = c->opv[TM_OP_CURL_Y].p;
        a.cz = c->opv[TM_OP_CURL_Z].p;
        a.dmi_fac = c->dmi_fac.p;
        a.nvs = c->nvs.p;
    }
    fill_idmi(c, a);
    a.m = c->m.p;
    a.invjs = c->has_js ? c->invjs.p : nullptr;
    a.ku = c->has_ku ? c->ku.p : nullptr;
    a.kaxis = c->has_ku ? c->kaxis.p : nullptr;
    a.hext = c->hext_field ? c->hext.p : nullptr;
    a.hx = c->hext_u[0];
    a.hy = c->hext_u[1];
    a.hz = c->hext_u[2];
    a.alpha = alpha;
    a.variant = variant;
    a.dmdt = dwhere == TM_DEVICE ? dmdt : c->w3a.p;
    a.heff = Heff ? (hwhere == TM_DEVICE ? Heff : c->w3b.p) : nullptr;
    a.torque = stats ? c->scal.p : nullptr;
    if (stats) hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, c->stream, c->scal.p);
    TM_CHECK(eval_dispatch(a, exch, demag, dmi, sell_grid(c), c->stream));
    if (c->timers && solve) {
        TM_HIP(hipEventRecord(c->ev[5], c->stream));
        TM_HIP(hipEventSynchronize(c->ev[5]));
        float t[5];
        for (int k = 0; k < 5; ++k) TM_HIP(hipEventElapsedTime(&t[k], c->ev[k], c->ev[k + 1]));
        c->tsum.div_ms += t[0];
        c->tsum.neumann_ms += t[1];
        c->tsum.bem_ms += t[2];
        c->tsum.dirichlet_ms += t[3];
        c->tsum.fields_ms += t[4];
        c->tsum.count += 1;
    }
    if (dwhere == TM_HOST) TM_CHECK(copy_out(dmdt, TM_HOST, a.dmdt, (size_t)3 * n, c->stream));
    if (Heff && hwhere == TM_HOST) TM_CHECK(copy_out(Heff, TM_HOST, a.heff, (size_t)3 * n, c->stream));
    if (stats) {
        double mt = 0.0;
        TM_HIP(hipMemcpyAsync(&mt, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        TM_HIP(hipStreamSynchronize(c->stream));
        *stats = local;
        stats->max_torque = mt;
    }
    return st;
}

int tm_ctx_set_node_volume(tm_ctx* c, const double* V) {
    TM_CHECK(ctx_ok(c));
    TM_CHECK(upload_vec(c, c->vol, V, c->n));
    c->has_vol = true;
    return TM_OK;
}

int tm_ctx_set_dmi_interfacial(tm_ctx* c, const double* D) {
    TM_CHECK(ctx_ok(c));
    if (!D) {
        c->has_idmi = false;
        return TM_OK;
    }
    TM_CHECK(upload_vec(c, c->idmi, D, c->n));
    c->has_idmi = true;
    return TM_OK;
}

int tm_idmi_field(tm_ctx* c, double* out, int where) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    TM_CHECK(commit_ops(c));
    if (!idmi_ready(c)) return TM_ERR_NOT_READY;
    EvalArgs a = {};
    a.A = view(c->pat);
    fill_idmi(c, a);
    a.m = c->m.p;
    double* o = out_target(c, out, where);
    a.heff = o;
    TM_CHECK(eval_dispatch(a, false, false, false, sell_grid(c), c->stream));
    return finish_out(c, out, where, o, (size_t)3 * c->n);
}

int tm_ctx_set_freeze_demag(tm_ctx* c, int on) {
    TM_CHECK(ctx_ok(c));
    c->freeze_demag = on != 0;
    return TM_OK;
}

int tm_ctx_refresh_demag(tm_ctx* c, const double* m3n, int mwhere, tm_eval_stats* stats) {
    TM_CHECK(ctx_ok(c));
    if (m3n) TM_CHECK(copy_in(c->m.p, m3n, mwhere, (size_t)3 * c->n, c->stream));
    int st = demag_potential_dev(c, c->m.p, stats);
    if (st == TM_OK || st == TM_ERR_MAXITER) TM_HIP(hipStreamSynchronize(c->stream));
    return st;
}

int tm_energies_eval(tm_ctx* c, const double* m3n, int mwhere, int recompute_demag, tm_energies* out) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    if (!c->has_vol) return TM_ERR_NOT_READY;
    const int n = c->n;
    if (m3n) TM_CHECK(copy_in(c->m.p, m3n, mwhere, (size_t)3 * n, c->stream));
    TM_CHECK(commit_ops(c));
    const bool demag = c->has_solvers && c->has_B;
    int st = TM_OK;
    if (demag) {
        if (recompute_demag || !c->have_u) st = demag_potential_dev(c, c->m.p, nullptr);
        if (st != TM_OK && st != TM_ERR_MAXITER) return st;
        if (!c->has_op[TM_OP_GRAD_X] || !c->has_op[TM_OP_GRAD_Y] || !c->has_op[TM_OP_GRAD_Z]) return TM_ERR_NOT_READY;
    }
    const bool exch = c->has_op[TM_OP_EXCHANGE];
    const bool dmi = c->has_dmi && c->has_op[TM_OP_CURL_X] && c->has_op[TM_OP_CURL_Y] && c->has_op[TM_OP_CURL_Z];
    EvalArgs a = {};
    a.A = view(c->pat);
    a.xc = exch ? c->opv[TM_OP_EXCHANGE].p : nullptr;
    if (demag) {
        a.gx = c->opv[TM_OP_GRAD_X].p;
        a.gy = c->opv[TM_OP_GRAD_Y].p;
        a.gz = c->opv[TM_OP_GRAD_Z].p;
        a.u1 = c->u1.p;
        a.u2 = c->u2.p;
    }
    if (dmi) {
        a.cx = c->opv[TM_OP_CURL_X].p;
        a.cy = c->opv[TM_OP_CURL_Y].p;
        a.cz = c->opv[TM_OP_CURL_Z].p;
        a.dmi_fac = c->dmi_fac.p;
        a.nvs = c->nvs.p;
    }
    fill_idmi(c, a);
    a.m = c->m.p;
    a.ku = c->has_ku ? c->ku.p : nullptr;
    a.kaxis = c->has_ku ? c->kaxis.p : nullptr;
    a.hext = c->hext_field ? c->hext.p : nullptr;
    a.hx = c->hext_u[0];
    a.hy = c->hext_u[1];
    a.hz = c->hext_u[2];
    const int G = sell_grid(c);
    TM_CHECK(c->epart.alloc(std::max(c->epart.n, (size_t)kEnergyTerms * G)));
    TM_CHECK(c->eout.alloc(std::max(c->eout.n, (size_t)kEnergyTerms)));
    TM_CHECK(energy_dispatch(a, exch, demag, dmi, c->vol.p, c->has_js ? c->js.p : nullptr, c->epart.p, G, c->stream));
    hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(kBlock), 0, c->stream, G, (const double*)c->epart.p, c->eout.p);
    TM_HIP(hipGetLastError());
    double t[kEnergyTerms];
    TM_HIP(hipMemcpyAsync(t, c->eout.p, sizeof(t), hipMemcpyDeviceToHost, c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    const double V = t[11];
    *out = tm_energies{};
    out->volume = V;
    out->exchange = -0.5 * t[0] / V;
    out->uniaxial = t[1] / V;
    out->zeeman = -t[2] / V;
    out->dmi = -0.5 * t[3] / V;
    out->demag = -0.5 * t[4] / V;
    out->total = out->exchange + out->uniaxial + out->zeeman + out->dmi + out->demag;
    for (int d = 0; d < 3; ++d) {
        out->mean_m[d] = t[5 + d] / V;
        out->mean_h[d] = t[8 + d] / V;
    }
    return st;
}

int tm_ctx_step_euler(tm_ctx* c, const double* dm, int where, double dtau) {
    TM_CHECK(ctx_ok(c));
    const double* dd = nullptr;
    TM_CHECK(stage_in(c, dm, where, (size_t)3 * c->n, c->w3b, &dd));
    hipLaunchKernelGGL(k_euler, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->m.p, dd, dtau);
    TM_HIP(hipGetLastError());
    return TM_OK;
}

int tm_ctx_enable_timers(tm_ctx* c, int on) {
    TM_CHECK(ctx_ok(c));
    if (on && !c->ev[0])
        for (auto& e : c->ev) TM_HIP(hipEventCreate(&e));
    c->timers = on != 0;
    return TM_OK;
}

int tm_ctx_get_timers(tm_ctx* c, tm_phase_times* out, int reset) {
    TM_CHECK(ctx_ok(c));
    if (!out) return TM_ERR_BAD_ARG;
    *out = c->tsum;
    if (reset) c->tsum = tm_phase_times{};
    return TM_OK;
}

}  // extern "C"

namespace tmh {
// Read-only view of a context for the batched evaluator (tm_batch.hip).
int ctx_view(tm_ctx* c, CtxView* v) {
    TM_CHECK(ctx_ok(c));
    TM_CHECK(commit_ops(c));
    v->n = c->n;
    v->stream = c->stream;
    v->pat = &c->pat;
    for (int o = 0; o < TM_NUM_OPS; ++o) {
        v->hop_set[o] = c->hop[o].set;
        v->hop_rp[o] = &c->hop[o].rp;
        v->hop_col[o] = &c->hop[o].col;
        v->hop_val[o] = &c->hop[o].val;
    }
    v->js = c->has_js ? c->js.p : nullptr;
    v->invjs = c->has_js ? c->invjs.p : nullptr;
    v->has_ku = c->has_ku;
    v->ku = c->ku.p;
    v->kaxis = c->kaxis.p;
    v->has_dmi = c->has_dmi;
    v->has_idmi = c->has_idmi;
    v->dmi_fac = c->dmi_fac.p;
    v->nvs = c->nvs.p;
    v->bnx = c->bnx;
    v->bld = c->bld;
    v->bnodes = c->bnodes.p;
    v->B = c->has_B ? c->B.p : nullptr;
    v->cgN = c->has_solvers ? &c->cgN : nullptr;
    v->cgD = c->has_solvers ? &c->cgD : nullptr;
    v->tol = c->tol;
    v->maxit = c->maxit;
    for (int d = 0; d < 3; ++d) v->hext[d] = c->hext_u[d];
    return TM_OK;
}

const double* ctx_hext_field(tm_ctx* c) { return c && c->hext_field ? c->hext.p : nullptr; }
}  // namespace tmh
