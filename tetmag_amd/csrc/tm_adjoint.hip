// tm_adjoint.hip -- vector-Jacobian products of the effective-field + LLG
// right-hand side, for gradients of objectives of the magnetisation
// trajectory (no counterpart in the reference, which has no derivatives).
//
// For f(m; p) = LLG(m, Heff(m; p)) and a cotangent lambda, the VJP is
//   g_H  = (df/dH)^T lambda                         pointwise
//   g_m  = (df/dm |_H)^T lambda + (dHeff/dm)^T g_H
// with the transposes of every linear piece of Heff:
//   exchange    XC^T (invJs g_H)            transposed values on the pattern
//   uniaxial    2 Ku invJs (k . g_H) k      (symmetric)
//   bulk DMI    f (s x g') + curl^T(-2 f g') with g' = invJs g_H, curl^T from
//               the transposed -tGrad values
//   demag       Hdem = -G (I + K_D^-1 S B S^T) K_N^-1 T (Js m) / mu0, so
//               w0 = -G^T g_H; y2 = K_D^-T w0; t = w0 + S B^T S^T y2;
//               y1 = K_N^-T t; g_m += Js/mu0 T^T y1
// and the parameter gradients
//   uniform applied field   sum_i g_H,i
//   per-node applied field  g_H
//   per-node Ku             2 invJs (k . m)(k . g_H)
//   per-node D              invJs / V g_H . (m x s - 2 curl m)
//   per-node interfacial D  the transposes of its two D factors (row, column)
// The transposed solves run on the same identity-row matrices
// (CgSolver::solve_transposed); B^T is a column-dot-product kernel.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tm_ctx.hpp"

namespace tmh {
int bem_gemvT(int b, int ld, const double* B, const double* src, const int* gidx, double* xbuf, double* y,
              const int* sidx, double* dst, hipStream_t s);
}  // namespace tmh

using namespace tmh;

namespace {

// Transposed values of operator `op`, scattered into the context's union
// SELL pattern.  The pattern of the transpose must lie inside the union
// pattern (true for the symmetric P1 patterns); otherwise TM_ERR_UNSUPPORTED.
int ensure_transposed(tm_ctx* c, int op) {
    if (!c->has_op[op]) return TM_ERR_NOT_READY;
    if (c->opT_gen[op] == c->ops_gen && c->opvT[op].p) return TM_OK;
    const auto& h = c->hop[op];
    const int n = c->n;
    std::vector<int> rp((size_t)n + 1, 0), col(h.col.size());
    std::vector<double> val(h.val.size());
    for (size_t q = 0; q < h.col.size(); ++q) rp[h.col[q] + 1]++;
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int q = h.rp[i]; q < h.rp[i + 1]; ++q) {
            const int p = fill[h.col[q]]++;
            col[p] = i;
            val[p] = h.val[q];
        }
    std::vector<double> v;
    if (c->pat.scatter(rp.data(), col.data(), val.data(), v) != TM_OK) return TM_ERR_UNSUPPORTED;
    TM_CHECK(c->opvT[op].upload(v.data(), v.size(), c->stream));
    TM_HIP(hipStreamSynchronize(c->stream));
    c->opT_gen[op] = c->ops_gen;
    return TM_OK;
}

// Pointwise LLG VJP.  Classic f = -m x H - alpha m x (m x H):
//   g_H = m x l - alpha m x (m x l)
//   g_m = l x H + alpha l x (m x H) - alpha H x (l x m)
// No precession f = -alpha m x (m x H): the terms with alpha only.
__device__ inline void cross3(double ax, double ay, double az, double bx, double by, double bz, double& cx,
                              double& cy, double& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}

__global__ __launch_bounds__(kBlock) void k_vjp_point(int n, const double* __restrict__ m,
                                                      const double* __restrict__ H, const double* __restrict__ l,
                                                      double alpha, int variant, double* __restrict__ gH,
                                                      double* __restrict__ gm) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double mx = m[i], my = m[n + i], mz = m[2 * n + i];
    const double hx = H[i], hy = H[n + i], hz = H[2 * n + i];
    const double lx = l[i], ly = l[n + i], lz = l[2 * n + i];
    double ax, ay, az, bx, by, bz;
    cross3(mx, my, mz, lx, ly, lz, ax, ay, az);  // m x l
    cross3(mx, my, mz, ax, ay, az, bx, by, bz);  // m x (m x l)
    double gx = -alpha * bx, gy = -alpha * by, gz = -alpha * bz;
    double mhx, mhy, mhz, p1x, p1y, p1z, lmx, lmy, lmz, p2x, p2y, p2z;
    cross3(mx, my, mz, hx, hy, hz, mhx, mhy, mhz);      // m x H
    cross3(lx, ly, lz, mhx, mhy, mhz, p1x, p1y, p1z);   // l x (m x H)
    cross3(lx, ly, lz, mx, my, mz, lmx, lmy, lmz);      // l x m
    cross3(hx, hy, hz, lmx, lmy, lmz, p2x, p2y, p2z);   // H x (l x m)
    double ox = alpha * (p1x - p2x), oy = alpha * (p1y - p2y), oz = alpha * (p1z - p2z);
    if (variant == TM_LLG_CLASSIC) {
        gx += ax;
        gy += ay;
        gz += az;
        double lhx, lhy, lhz;
        cross3(lx, ly, lz, hx, hy, hz, lhx, lhy, lhz);  // l x H
        ox += lhx;
        oy += lhy;
        oz += lhz;
    }
    gH[i] = gx;
    gH[n + i] = gy;
    gH[2 * n + i] = gz;
    gm[i] = ox;
    gm[n + i] = oy;
    gm[2 * n + i] = oz;
}

struct VjpArgs {
    SellView A;
    const double* xcT;                          // exchange^T (or null)
    const double *cxT, *cyT, *czT;              // (-tGrad)^T for the DMI curl
    const double *cx, *cy, *cz;                 // forward curl values (D gradient)
    const double *gxT, *gyT, *gzT;              // grad^T (demag)
    const double* m;
    const double* gH;
    const double* invjs;                        // null: 1
    const double *ku, *kaxis;                   // null: no anisotropy
    const double *dmi_fac, *nvs;                // DMI data
    const double* vol;                          // node volumes (D gradient)
    double* gm;                                 // in: pointwise part; out: += linear transposes
    double* w0;                                 // -G^T g_H (demag) or null
    double* g_ku;                               // n or null
    double* g_d;                                // n or null
    // interfacial DMI (idmi null: off): D, transposed grad_x/y and tGrad_x/y,
    // forward grad_x/y (Di gradient), node volumes
    const double *idmi, *igxT, *igyT, *itxT, *ityT, *igx, *igy, *ivol;
    double* g_di;                               // n or null
    double* part;                               // 3 x gridDim partial sums of g_H
};

template <bool EXCH, bool DMI, bool DEMAG, bool GD>
__global__ __launch_bounds__(kBlock) void k_vjp_lin(VjpArgs a) {
    __shared__ double smem[3 * (kBlock / kWave)];
    const SellView A = a.A;
    const int n = A.n;
    const int row = xcd_block() * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    double hs[3] = {0.0, 0.0, 0.0};
    if (s < (n + kWave - 1) / kWave) {
        const double* __restrict__ gH = a.gH;
        const double* __restrict__ ij = a.invjs;
        double ex = 0, ey = 0, ez = 0, dx = 0, dy = 0, dz = 0, w = 0, cx = 0, cy = 0, cz = 0;
        double ib[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // interfacial DMI row sums
        const bool idmi = a.idmi != nullptr;
        sell_row(A, s, lane, [&](int64_t q, int cc) {
            const double gx = gH[cc], gy = gH[n + cc], gz = gH[2 * n + cc];
            const double sj = ij ? ij[cc] : 1.0;
            if (idmi) {
                // transposes of H_x = D G_x m_z - T_x (D m_z) / V (and y), and
                // H_z = -D (G_x m_x + G_y m_y) + (T_x D m_x + T_y D m_y) / V
                const double wx = sj * gx, wy = sj * gy, wz = sj * gz;
                const double Dc = a.idmi[cc], ivc = 1.0 / a.ivol[cc];
                const double GX = a.igxT[q], GY = a.igyT[q], TX = a.itxT[q], TY = a.ityT[q];
                ib[0] += GX * Dc * wz;
                ib[1] += GY * Dc * wz;
                ib[2] += (GX * wx + GY * wy) * Dc;
                ib[3] += TX * wz * ivc;
                ib[4] += TY * wz * ivc;
                ib[5] += TX * wx * ivc;
                ib[6] += TY * wy * ivc;
                if (a.g_di) {
                    const double Gx = a.igx[q], Gy = a.igy[q];
                    const double mx = a.m[cc], my = a.m[n + cc], mz = a.m[2 * n + cc];
                    ib[7] += Gx * mz;
                    ib[8] += Gy * mz;
                    ib[9] += Gx * mx + Gy * my;
                }
            }
            if (EXCH) {
                const double e = a.xcT[q];
                ex += e * sj * gx;
                ey += e * sj * gy;
                ez += e * sj * gz;
            }
            if (DMI) {
                const double f = -2.0 * a.dmi_fac[cc] * sj;
                const double wx = f * gx, wy = f * gy, wz = f * gz;
                const double X = a.cxT[q], Y = a.cyT[q], Z = a.czT[q];
                dx += Z * wy - Y * wz;
                dy += X * wz - Z * wx;
                dz += Y * wx - X * wy;
                if (GD) {
                    const double mx = a.m[cc], my = a.m[n + cc], mz = a.m[2 * n + cc];
                    const double Xf = a.cx[q], Yf = a.cy[q], Zf = a.cz[q];
                    cx += Yf * mz - Zf * my;
                    cy += Zf * mx - Xf * mz;
                    cz += Xf * my - Yf * mx;
                }
            }
            if (DEMAG) w -= a.gxT[q] * gx + a.gyT[q] * gy + a.gzT[q] * gz;
        });
        if (row < n) {
            const double gx = gH[row], gy = gH[n + row], gz = gH[2 * n + row];
            hs[0] = gx;
            hs[1] = gy;
            hs[2] = gz;
            const double sj = ij ? ij[row] : 1.0;
            const double px = sj * gx, py = sj * gy, pz = sj * gz;  // g'
            double ox = ex + dx, oy = ey + dy, oz = ez + dz;
            const double mx = a.m[row], my = a.m[n + row], mz = a.m[2 * n + row];
            if (a.ku) {
                const double kx = a.kaxis[row], ky = a.kaxis[n + row], kz = a.kaxis[2 * n + row];
                const double kg = kx * px + ky * py + kz * pz;
                const double f = 2.0 * a.ku[row] * kg;
                ox += f * kx;
                oy += f * ky;
                oz += f * kz;
                if (a.g_ku) a.g_ku[row] = 2.0 * (kx * mx + ky * my + kz * mz) * kg;
            } else if (a.g_ku) {
                a.g_ku[row] = 0.0;
            }
            if (DMI) {
                const double sx = a.nvs[row], sy = a.nvs[n + row], sz = a.nvs[2 * n + row];
                const double f = a.dmi_fac[row];
                double tx, ty, tz;
                cross3(sx, sy, sz, px, py, pz, tx, ty, tz);  // s x g'
                ox += f * tx;
                oy += f * ty;
                oz += f * tz;
                if (GD) {
                    double qx, qy, qz;
                    cross3(mx, my, mz, sx, sy, sz, qx, qy, qz);  // m x s
                    a.g_d[row] = (px * (qx - 2.0 * cx) + py * (qy - 2.0 * cy) + pz * (qz - 2.0 * cz)) / a.vol[row];
                }
            } else if (a.g_d) {
                a.g_d[row] = 0.0;
            }
            if (idmi) {
                const double D = a.idmi[row];
                ox += -ib[0] + D * ib[3];
                oy += -ib[1] + D * ib[4];
                oz += ib[2] - D * (ib[5] + ib[6]);
                if (a.g_di)
                    a.g_di[row] = px * ib[7] + py * ib[8] - pz * ib[9] - mz * (ib[5] + ib[6]) + mx * ib[3] +
                                  my * ib[4];
            } else if (a.g_di) {
                a.g_di[row] = 0.0;
            }
            a.gm[row] += ox;
            a.gm[n + row] += oy;
            a.gm[2 * n + row] += oz;
            if (DEMAG) a.w0[row] = w;
        }
    }
    block_sum<3>(hs, smem);
    if (threadIdx.x == 0 && a.part)
        for (int d = 0; d < 3; ++d) a.part[(size_t)d * gridDim.x + blockIdx.x] = hs[d];
}

template <bool E, bool M, bool D, bool GD>
void launch_lin(const VjpArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_vjp_lin<E, M, D, GD>), dim3(grid), dim3(kBlock), 0, s, a);
}

int lin_dispatch(const VjpArgs& a, bool e, bool m, bool d, bool gd, int grid, hipStream_t s) {
    const int key = (e ? 8 : 0) | (m ? 4 : 0) | (d ? 2 : 0) | (gd && m ? 1 : 0);
    switch (key) {
        case 0: launch_lin<false, false, false, false>(a, grid, s); break;
        case 2: launch_lin<false, false, true, false>(a, grid, s); break;
        case 4: launch_lin<false, true, false, false>(a, grid, s); break;
        case 5: launch_lin<false, true, false, true>(a, grid, s); break;
        case 6: launch_lin<false, true, true, false>(a, grid, s); break;
        case 7: launch_lin<false, true, true, true>(a, grid, s); break;
        case 8: launch_lin<true, false, false, false>(a, grid, s); break;
        case 10: launch_lin<true, false, true, false>(a, grid, s); break;
        case 12: launch_lin<true, true, false, false>(a, grid, s); break;
        case 13: launch_lin<true, true, false, true>(a, grid, s); break;
        case 14: launch_lin<true, true, true, false>(a, grid, s); break;
        default: launch_lin<true, true, true, true>(a, grid, s); break;
    }
    TM_HIP(hipGetLastError());
    return TM_OK;
}

// g_m,d += (Js / mu0) tGrad_d^T y1  (the transpose of the Poisson right-hand
// side b = (1/mu0) sum_d tGrad_d (Js m_d)).
__global__ __launch_bounds__(kBlock) void k_vjp_div_T(SellView A, const double* __restrict__ txT,
                                                      const double* __restrict__ tyT,
                                                      const double* __restrict__ tzT,
                                                      const double* __restrict__ js, double scale,
                                                      const double* __restrict__ y, double* __restrict__ g) {
    const int n = A.n;
    const int row = xcd_block() * kBlock + threadIdx.x;
    const int s = row / kWave, lane = row & (kWave - 1);
    if (s >= (n + kWave - 1) / kWave) return;
    double ax = 0, ay = 0, az = 0;
    sell_row(A, s, lane, [&](int64_t q, int cc) {
        const double v = y[cc];
        ax += txT[q] * v;
        ay += tyT[q] * v;
        az += tzT[q] * v;
    });
    if (row < n) {
        const double f = js[row] * scale;
        g[row] += f * ax;
        g[n + row] += f * ay;
        g[2 * n + row] += f * az;
    }
}

__global__ __launch_bounds__(kBlock) void k_vjp_red3(int G, const double* __restrict__ part,
                                                     double* __restrict__ out) {
    double t[3];
    reduce_partials_all<3>(part, G, G, t);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) out[d] = t[d];
}

__global__ __launch_bounds__(kBlock) void k_copy(int n, const double* __restrict__ a, double* __restrict__ o) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) o[i] = a[i];
}

// Transposed demag chain on device: w0 (n) -> g (3n) += Js/mu0 T^T K_N^-T
// (w0 + S B^T S^T K_D^-T w0).  Warm starts in a_y1/a_y2.
int demag_transpose(tm_ctx* c, const double* w0, double* g, tm_eval_stats* stats) {
    const int n = c->n;
    int itD = 0, itN = 0;
    double rD = 0, rN = 0;
    int st = c->cgD.solve_transposed(w0, c->a_y2.p, c->tol, c->maxit, c->stream, &itD, &rD);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    const int stD = st;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, n, w0, c->a_t.p);
    TM_HIP(hipGetLastError());
    TM_CHECK(bem_gemvT(c->bnx, c->bld, c->B.p, c->a_y2.p, c->bnodes.p, c->a_xb.p, nullptr, c->bnodes.p, c->a_t.p,
                       c->stream));
    st = c->cgN.solve_transposed(c->a_t.p, c->a_y1.p, c->tol, c->maxit, c->stream, &itN, &rN);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    hipLaunchKernelGGL(k_vjp_div_T, dim3(sell_grid(c)), dim3(kBlock), 0, c->stream, view(c->pat),
                       c->opvT[TM_OP_TGRAD_X].p, c->opvT[TM_OP_TGRAD_Y].p, c->opvT[TM_OP_TGRAD_Z].p, c->js.p,
                       1.0 / kMu0, (const double*)c->a_y1.p, g);
    TM_HIP(hipGetLastError());
    if (stats) {
        stats->iters_neumann = itN;
        stats->iters_dirichlet = itD;
        stats->relres_neumann = rN;
        stats->relres_dirichlet = rD;
    }
    return stD != TM_OK ? stD : st;
}

int ensure_adjoint_buffers(tm_ctx* c, bool demag) {
    const int n = c->n;
    TM_CHECK(c->a_dm.alloc((size_t)3 * n));
    TM_CHECK(c->a_h.alloc((size_t)3 * n));
    TM_CHECK(c->a_gh.alloc((size_t)3 * n));
    TM_CHECK(c->a_part.alloc((size_t)3 * std::max(1, sell_grid(c))));
    TM_CHECK(c->a_red.alloc(4));
    if (demag) {
        const bool fresh = !c->a_y1.p;
        TM_CHECK(c->a_w0.alloc(n));
        TM_CHECK(c->a_t.alloc(n));
        TM_CHECK(c->a_y1.alloc(n));
        TM_CHECK(c->a_y2.alloc(n));
        TM_CHECK(c->a_xb.alloc(std::max(1, c->bnx)));
        if (fresh) {  // zero warm starts
            TM_HIP(hipMemsetAsync(c->a_y1.p, 0, sizeof(double) * n, c->stream));
            TM_HIP(hipMemsetAsync(c->a_y2.p, 0, sizeof(double) * n, c->stream));
        }
    }
    return TM_OK;
}

int demag_ready(tm_ctx* c) {
    if (!c->has_solvers || !c->has_B || !c->has_js) return TM_ERR_NOT_READY;
    for (int op : {TM_OP_GRAD_X, TM_OP_GRAD_Y, TM_OP_GRAD_Z, TM_OP_TGRAD_X, TM_OP_TGRAD_Y, TM_OP_TGRAD_Z})
        TM_CHECK(ensure_transposed(c, op));
    return TM_OK;
}

}  // namespace

extern "C" {

int tm_eval_rhs_vjp(tm_ctx* c, const double* m3n, int mwhere, double alpha, int variant, const double* lam3n,
                    int lwhere, double* gm3n, int gwhere, double* g_hext3, double* g_hfield3n, double* g_ku,
                    double* g_d, double* g_di, tm_eval_stats* stats) {
    TM_CHECK(ctx_ok(c));
    if (!lam3n || !gm3n || (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC)) return TM_ERR_BAD_ARG;
    if ((gwhere != TM_HOST && gwhere != TM_DEVICE) || (lwhere != TM_HOST && lwhere != TM_DEVICE))
        return TM_ERR_BAD_ARG;
    const int n = c->n;
    TM_CHECK(commit_ops(c));
    const bool demag_on = c->has_solvers && c->has_B;
    const bool demag = demag_on && !c->freeze_demag;  // frozen potentials do not depend on m
    const bool exch = c->has_op[TM_OP_EXCHANGE];
    const bool dmi = c->has_dmi && c->has_op[TM_OP_CURL_X] && c->has_op[TM_OP_CURL_Y] && c->has_op[TM_OP_CURL_Z];
    if (g_d && (!dmi || !c->has_vol)) return TM_ERR_NOT_READY;
    const bool idmi = idmi_ready(c);
    if (g_di && !idmi) return TM_ERR_NOT_READY;
    TM_CHECK(ensure_adjoint_buffers(c, demag));
    if (idmi)
        for (int op : {TM_OP_GRAD_X, TM_OP_GRAD_Y, TM_OP_TGRAD_X, TM_OP_TGRAD_Y}) TM_CHECK(ensure_transposed(c, op));
    if (exch) TM_CHECK(ensure_transposed(c, TM_OP_EXCHANGE));
    if (dmi)
        for (int op : {TM_OP_CURL_X, TM_OP_CURL_Y, TM_OP_CURL_Z}) TM_CHECK(ensure_transposed(c, op));
    if (demag) TM_CHECK(demag_ready(c));
    // forward at m: Heff (the context's magnetisation becomes m)
    tm_eval_stats fst = {};
    int st = tm_eval_rhs(c, m3n, mwhere, alpha, variant, c->a_dm.p, TM_DEVICE, c->a_h.p, TM_DEVICE, &fst);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    const double* ld = nullptr;
    TM_CHECK(stage_in(c, lam3n, lwhere, (size_t)3 * n, c->w3b, &ld));
    double* g = gwhere == TM_DEVICE ? gm3n : c->w3a.p;
    hipLaunchKernelGGL(k_vjp_point, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, n, (const double*)c->m.p,
                       (const double*)c->a_h.p, ld, alpha, variant, c->a_gh.p, g);
    TM_HIP(hipGetLastError());
    VjpArgs a = {};
    a.A = view(c->pat);
    a.xcT = exch ? c->opvT[TM_OP_EXCHANGE].p : nullptr;
    if (dmi) {
        a.cxT = c->opvT[TM_OP_CURL_X].p;
        a.cyT = c->opvT[TM_OP_CURL_Y].p;
        a.czT = c->opvT[TM_OP_CURL_Z].p;
        a.cx = c->opv[TM_OP_CURL_X].p;
        a.cy = c->opv[TM_OP_CURL_Y].p;
        a.cz = c->opv[TM_OP_CURL_Z].p;
        a.dmi_fac = c->dmi_fac.p;
        a.nvs = c->nvs.p;
    }
    if (demag) {
        a.gxT = c->opvT[TM_OP_GRAD_X].p;
        a.gyT = c->opvT[TM_OP_GRAD_Y].p;
        a.gzT = c->opvT[TM_OP_GRAD_Z].p;
        a.w0 = c->a_w0.p;
    }
    a.m = c->m.p;
    a.gH = c->a_gh.p;
    a.invjs = c->has_js ? c->invjs.p : nullptr;
    a.ku = c->has_ku ? c->ku.p : nullptr;
    a.kaxis = c->has_ku ? c->kaxis.p : nullptr;
    a.vol = c->has_vol ? c->vol.p : nullptr;
    a.gm = g;
    double* gku_d = nullptr;
    double* gd_d = nullptr;
    // parameter gradients go to the caller's device buffers, or the context's
    // scratch (a_dm, free after the forward) when host output is asked for
    if (g_ku) gku_d = gwhere == TM_DEVICE ? g_ku : c->a_dm.p;
    if (g_d) gd_d = gwhere == TM_DEVICE ? g_d : c->a_dm.p + n;
    double* gdi_d = nullptr;
    if (g_di) gdi_d = gwhere == TM_DEVICE ? g_di : c->a_dm.p + 2 * (size_t)n;
    a.g_ku = gku_d;
    a.g_d = gd_d;
    a.g_di = gdi_d;
    if (idmi) {
        a.idmi = c->idmi.p;
        a.igxT = c->opvT[TM_OP_GRAD_X].p;
        a.igyT = c->opvT[TM_OP_GRAD_Y].p;
        a.itxT = c->opvT[TM_OP_TGRAD_X].p;
        a.ityT = c->opvT[TM_OP_TGRAD_Y].p;
        a.igx = c->opv[TM_OP_GRAD_X].p;
        a.igy = c->opv[TM_OP_GRAD_Y].p;
        a.ivol = c->vol.p;
    }
    const int G = sell_grid(c);
    a.part = c->a_part.p;
    TM_CHECK(lin_dispatch(a, exch, dmi, demag, g_d != nullptr, G, c->stream));
    hipLaunchKernelGGL(k_vjp_red3, dim3(1), dim3(kBlock), 0, c->stream, G, (const double*)c->a_part.p, c->a_red.p);
    TM_HIP(hipGetLastError());
    tm_eval_stats ast = {};
    int sta = TM_OK;
    if (demag) {
        sta = demag_transpose(c, c->a_w0.p, g, &ast);
        if (sta != TM_OK && sta != TM_ERR_MAXITER) return sta;
    }
    if (gwhere == TM_HOST) {
        TM_CHECK(copy_out(gm3n, TM_HOST, g, (size_t)3 * n, c->stream));
        if (g_ku) TM_CHECK(copy_out(g_ku, TM_HOST, gku_d, n, c->stream));
        if (g_d) TM_CHECK(copy_out(g_d, TM_HOST, gd_d, n, c->stream));
        if (g_di) TM_CHECK(copy_out(g_di, TM_HOST, gdi_d, n, c->stream));
    }
    if (g_hfield3n) TM_CHECK(finish_out(c, g_hfield3n, gwhere, c->a_gh.p, (size_t)3 * n));
    if (g_hext3) {
        TM_HIP(hipMemcpyAsync(g_hext3, c->a_red.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        TM_HIP(hipStreamSynchronize(c->stream));
    }
    if (stats) {
        TM_HIP(hipStreamSynchronize(c->stream));
        *stats = ast;
        stats->max_torque = fst.max_torque;
    }
    if (st != TM_OK) return st;
    return sta;
}

int tm_demag_field_vjp(tm_ctx* c, const double* v3n, int vwhere, double* g3n, int gwhere, tm_eval_stats* stats) {
    TM_CHECK(ctx_ok(c));
    if (!v3n || !g3n) return TM_ERR_BAD_ARG;
    const int n = c->n;
    TM_CHECK(commit_ops(c));
    TM_CHECK(demag_ready(c));
    TM_CHECK(ensure_adjoint_buffers(c, true));
    const double* vd = nullptr;
    TM_CHECK(stage_in(c, v3n, vwhere, (size_t)3 * n, c->w3b, &vd));
    double* g = gwhere == TM_DEVICE ? g3n : c->w3a.p;
    TM_HIP(hipMemsetAsync(g, 0, sizeof(double) * 3 * n, c->stream));
    VjpArgs a = {};
    a.A = view(c->pat);
    a.gxT = c->opvT[TM_OP_GRAD_X].p;
    a.gyT = c->opvT[TM_OP_GRAD_Y].p;
    a.gzT = c->opvT[TM_OP_GRAD_Z].p;
    a.w0 = c->a_w0.p;
    a.m = c->m.p;  // read by the pointwise part only, which has no terms here
    a.gH = vd;
    a.gm = g;
    a.part = nullptr;
    TM_CHECK(lin_dispatch(a, false, false, true, false, sell_grid(c), c->stream));
    const int st = demag_transpose(c, c->a_w0.p, g, stats);
    if (st != TM_OK && st != TM_ERR_MAXITER) return st;
    if (gwhere == TM_HOST) TM_CHECK(copy_out(g3n, TM_HOST, g, (size_t)3 * n, c->stream));
    else TM_HIP(hipStreamSynchronize(c->stream));
    return st;
}

int tm_bem_apply_transposed(tm_ctx* c, const double* yb, int ywhere, double* xb, int xwhere) {
    TM_CHECK(ctx_ok(c));
    if (!c->has_B || !yb || !xb) return !c->has_B ? TM_ERR_NOT_READY : TM_ERR_BAD_ARG;
    const int b = c->bnx;
    const double* yd = nullptr;
    TM_CHECK(stage_in(c, yb, ywhere, (size_t)b, c->w3b, &yd));
    double* o = xwhere == TM_DEVICE ? xb : c->w3a.p;
    TM_CHECK(bem_gemvT(b, c->bld, c->B.p, yd, nullptr, nullptr, o, nullptr, nullptr, c->stream));
    if (xwhere == TM_HOST) return copy_out(xb, TM_HOST, o, b, c->stream);
    TM_HIP(hipStreamSynchronize(c->stream));
    return TM_OK;
}

}  // extern "C"
