// tm_integrate.hip -- classical RK4 time stepping of the context's
// magnetisation on the device, each step captured once in a HIP graph and
// replayed (the reference integrates with CVODE on the host around its GPU
// right-hand side, LLGWrapper.cu:139-228 / TheLLG.cpp:578-711).
//
// One step: k1 = f(m), k2 = f(m + dt/2 k1), k3 = f(m + dt/2 k2),
// k4 = f(m + dt k3), m <- normalise(m + dt/6 (k1 + 2 k2 + 2 k3 + k4)), where
// f is the full right-hand side (demag recomputed at every stage unless
// frozen).  In the graph the CG solves run a fixed iteration budget and stop
// early on the device once converged; a device counter records any solve that
// missed the tolerance.  After each replay the host reads that counter: on a
// miss the step is restored from its backup, redone with host-polled solves,
// and the budgets grow before the graph is captured again.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tm_ctx.hpp"

using namespace tmh;

namespace {

__global__ __launch_bounds__(kBlock) void k_rk_begin(size_t n3, const double* __restrict__ m,
                                                     double* __restrict__ bak, unsigned* fail, int* imax) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n3) bak[i] = m[i];
    if (i == 0) {
        fail[0] = 0u;
        imax[0] = 0;
        imax[1] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_rk_stage(size_t n3, const double* __restrict__ m,
                                                     const double* __restrict__ k, double c,
                                                     double* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n3) out[i] = m[i] + c * k[i];
}

__global__ __launch_bounds__(kBlock) void k_rk_final(int n, double* __restrict__ m, const double* __restrict__ k1,
                                                     const double* __restrict__ k2, const double* __restrict__ k3,
                                                     const double* __restrict__ k4, double dt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double f = dt / 6.0;
    double v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t q = (size_t)d * n + i;
        v[d] = m[q] + f * (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]);
    }
    const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0)
#pragma unroll
        for (int d = 0; d < 3; ++d) m[(size_t)d * n + i] = v[d] / l;
}

// Enqueues one step.  budget[0] > 0: fixed-budget solves (capturable);
// otherwise host-polled solves, and iters[0..1] receive the largest counts.
int rk4_step(tm_ctx* c, double dt, double alpha, int variant, const int* budget, int* iters) {
    const int n = c->n;
    const size_t n3 = (size_t)3 * n;
    const dim3 g3((unsigned)((n3 + kBlock - 1) / kBlock));
    double* m = c->m.p;
    hipLaunchKernelGGL(k_rk_begin, g3, dim3(kBlock), 0, c->stream, n3, (const double*)m, c->rk_bak.p, c->rk_fail.p,
                       c->rk_imax.p);
    const double* src[4] = {m, c->rk_tmp.p, c->rk_tmp.p, c->rk_tmp.p};
    const double cst[4] = {0.0, 0.5 * dt, 0.5 * dt, dt};
    int st = TM_OK;
    for (int s = 0; s < 4; ++s) {
        if (s > 0)
            hipLaunchKernelGGL(k_rk_stage, g3, dim3(kBlock), 0, c->stream, n3, (const double*)m,
                               (const double*)c->rk_k[s - 1].p, cst[s], c->rk_tmp.p);
        tm_eval_stats es = {};
        const int r = eval_device(c, src[s], c->rk_k[s].p, alpha, variant, budget[0], budget[1], c->rk_fail.p,
                                  c->rk_imax.p, budget[0] > 0 ? nullptr : &es);
        if (r != TM_OK && r != TM_ERR_MAXITER) return r;
        if (r == TM_ERR_MAXITER) st = r;
        if (iters) {
            iters[0] = std::max(iters[0], es.iters_neumann);
            iters[1] = std::max(iters[1], es.iters_dirichlet);
        }
    }
    hipLaunchKernelGGL(k_rk_final, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, n, m,
                       (const double*)c->rk_k[0].p, (const double*)c->rk_k[1].p, (const double*)c->rk_k[2].p,
                       (const double*)c->rk_k[3].p, dt);
    TM_HIP(hipGetLastError());
    return st;
}

int grow(int iters) { return (int)std::ceil(1.25 * iters) + 4; }

}  // namespace

extern "C" {

int tm_ctx_run_rk4(tm_ctx* c, double dtau, int nsteps, double alpha, int variant, int use_graph,
                   tm_rk4_stats* out) {
    TM_CHECK(ctx_ok(c));
    if (nsteps < 0 || !(dtau > 0.0) || (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC))
        return TM_ERR_BAD_ARG;
    const size_t n3 = (size_t)3 * c->n;
    for (auto& k : c->rk_k) TM_CHECK(k.alloc(n3));
    TM_CHECK(c->rk_tmp.alloc(n3));
    TM_CHECK(c->rk_bak.alloc(n3));
    TM_CHECK(c->rk_fail.alloc(1));
    TM_CHECK(c->rk_imax.alloc(2));
    TM_CHECK(commit_ops(c));
    tm_rk4_stats S = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    TM_HIP(hipEventRecord(e0, c->stream));
    hipGraphExec_t exec = nullptr;
    int status = TM_OK;
    int* budget = c->rk_budget;
    for (int step = 0; step < nsteps && status == TM_OK; ++step) {
        if (!use_graph || budget[0] == 0) {
            int it[2] = {0, 0};
            const int zero[2] = {0, 0};
            int r = rk4_step(c, dtau, alpha, variant, zero, it);
            if (r != TM_OK) {
                status = r;
                break;
            }
            S.sync_steps += 1;
            S.max_iters_neumann = std::max(S.max_iters_neumann, it[0]);
            S.max_iters_dirichlet = std::max(S.max_iters_dirichlet, it[1]);
            budget[0] = std::max(budget[0], grow(it[0]));
            budget[1] = std::max(budget[1], grow(it[1]));
            continue;
        }
        if (!exec) {
            hipGraph_t g = nullptr;
            TM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const int r = rk4_step(c, dtau, alpha, variant, budget, nullptr);
            const hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (r != TM_OK || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                status = r != TM_OK ? r : TM_ERR_HIP;
                break;
            }
            const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                exec = nullptr;
                status = TM_ERR_HIP;
                break;
            }
            S.captures += 1;
        }
        TM_HIP(hipGraphLaunch(exec, c->stream));
        unsigned fail = 0;
        int imax[2] = {0, 0};
        TM_HIP(hipMemcpyAsync(&fail, c->rk_fail.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
        TM_HIP(hipMemcpyAsync(imax, c->rk_imax.p, sizeof(imax), hipMemcpyDeviceToHost, c->stream));
        TM_HIP(hipStreamSynchronize(c->stream));
        S.max_iters_neumann = std::max(S.max_iters_neumann, imax[0]);
        S.max_iters_dirichlet = std::max(S.max_iters_dirichlet, imax[1]);
        if (fail == 0) {
            S.graph_steps += 1;
            continue;
        }
        // a solve missed the tolerance within its budget: redo the step
        TM_HIP(hipMemcpyAsync(c->m.p, c->rk_bak.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        int it[2] = {0, 0};
        const int zero[2] = {0, 0};
        const int r = rk4_step(c, dtau, alpha, variant, zero, it);
        if (r != TM_OK) {
            status = r;
            break;
        }
        S.redone_steps += 1;
        S.max_iters_neumann = std::max(S.max_iters_neumann, it[0]);
        S.max_iters_dirichlet = std::max(S.max_iters_dirichlet, it[1]);
        budget[0] = std::max(budget[0] + 4, grow(it[0]));
        budget[1] = std::max(budget[1] + 4, grow(it[1]));
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    TM_HIP(hipEventRecord(e1, c->stream));
    TM_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    S.ms = ms;
    S.budget_neumann = budget[0];
    S.budget_dirichlet = budget[1];
    if (out) *out = S;
    return status;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adaptive Dormand-Prince 5(4) with first-same-as-last stages (J. R. Dormand,
// P. J. Prince, J. Comput. Appl. Math. 6 (1980) 19; the step-size control of
// Hairer, Norsett, Wanner, Solving ODEs I, II.4), replacing the reference's
// CVODE integration of the same right-hand side (LLGWrapper.cu:139-211,
// TheLLG.cpp:578-654, rtol = atol = 1e-6 there).  The error is measured in
// CVODE's weighted RMS norm, weights 1 / (rtol |y_i| + atol) (with the larger
// of the old and new |y_i|).  Step control runs on the device: the step size,
// time and acceptance live in a device state, so a whole step (seven stage
// evaluations, error norm, accept/reject, commit) is one capturable sequence
// of kernels and the host only reads the small state after each step.
// ---------------------------------------------------------------------------
namespace {

struct OdeState {
    double t, t_end, h, h_step, h_next, err, rtol, atol;
    double t_new;
    int accept, done, nonfinite, pad;
    long long n_accept, n_reject;
};

// Butcher tableau (rows a_s, s = 2..7), 5th-order weights = row 7, and the
// error weights e = b5 - b4.
constexpr double kA[6][6] = {
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
constexpr double kE[7] = {71.0 / 57600, 0.0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200, 22.0 / 525, -1.0 / 40};

struct KPtrs {
    const double* k[7];
};
struct Coefs {
    double a[7];
};

// begin: h_step = min(h, t_end - t); clear the failure counter of the step's
// fixed-budget solves and the iteration maxima
__global__ void k_ode_begin(OdeState* s, unsigned* fail, int* imax) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double rem = s->t_end - s->t;
        s->h_step = s->h < rem ? s->h : rem;
        fail[0] = 0u;
        imax[0] = 0;
        imax[1] = 0;
    }
}

// out = y + h_step sum_{j < ns} a_j k_j
__global__ __launch_bounds__(kBlock) void k_ode_stage(size_t n3, const double* __restrict__ y, KPtrs K, Coefs c,
                                                      int ns, const OdeState* s, double* __restrict__ out) {
    const double h = s->h_step;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (size_t)gridDim.x * kBlock) {
        double acc = 0.0;
        for (int j = 0; j < ns; ++j)
            if (c.a[j] != 0.0) acc += c.a[j] * K.k[j][i];
        out[i] = y[i] + h * acc;
    }
}

// partial sums of (e_i / (atol + rtol max(|y_i|, |ynew_i|)))^2, e = h sum_j E_j k_j
__global__ __launch_bounds__(kBlock) void k_ode_err(size_t n3, const double* __restrict__ y,
                                                    const double* __restrict__ ynew, KPtrs K, const OdeState* s,
                                                    double* __restrict__ part) {
    __shared__ double red[kBlock / kWave];
    const double h = s->h_step, rtol = s->rtol, atol = s->atol;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (size_t)gridDim.x * kBlock) {
        double e = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j)
            if (kE[j] != 0.0) e += kE[j] * K.k[j][i];
        const double sc = atol + rtol * fmax(fabs(y[i]), fabs(ynew[i]));
        const double r = h * e / sc;
        acc += r * r;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

// error norm, accept / reject, proposed next step (one block; partials summed
// in a fixed order)
__global__ __launch_bounds__(kBlock) void k_ode_control(int nparts, double n3, OdeState* s,
                                                        const double* __restrict__ part) {
    __shared__ double red[kBlock / kWave];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) acc += part[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
    const double err = sqrt(t / n3);
    const double h = s->h_step;
    s->err = err;
    if (!isfinite(err)) {
        s->accept = 0;
        s->nonfinite = 1;
        s->h_next = 0.1 * h;
        return;
    }
    s->nonfinite = 0;
    const bool ok = err <= 1.0;
    double fac = err > 0.0 ? 0.9 * pow(err, -0.2) : 5.0;
    fac = fmin(ok ? 5.0 : 1.0, fmax(0.2, fac));
    s->accept = ok ? 1 : 0;
    s->h_next = h * fac;
    s->t_new = s->t + h;
}

// commit (skipped when a fixed-budget solve of the step missed its
// tolerance: the host redoes the step): on acceptance m <- y7 and k1 <- k7
__global__ __launch_bounds__(kBlock) void k_ode_commit(size_t n3, double* __restrict__ m,
                                                       const double* __restrict__ y7, double* __restrict__ k1,
                                                       const double* __restrict__ k7, OdeState* s,
                                                       const unsigned* fail) {
    if (fail[0] != 0u) return;
    const bool acc = s->accept != 0;
    if (acc)
        for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (size_t)gridDim.x * kBlock) {
            m[i] = y7[i];
            k1[i] = k7[i];
        }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (acc) {
            s->n_accept += 1;
            const bool last = s->t_new >= s->t_end * (1.0 - 1e-14) - 1e-300 || s->h_step >= s->t_end - s->t;
            s->t = last ? s->t_end : s->t_new;
            s->done = last ? 1 : 0;
        } else {
            s->n_reject += 1;
        }
        s->h = s->h_next;
    }
}

int ode_grid(size_t n3) { return (int)std::min<size_t>((n3 + kBlock - 1) / kBlock, 2048); }

// One Dormand-Prince step from (m, k1 = f(m)); budget[0] > 0: fixed-budget
// solves (capturable), else host-polled solves with the largest iteration
// counts in iters[0..1].
int dp_step(tm_ctx* c, OdeState* s, double alpha, int variant, const int* budget, int* iters) {
    const size_t n3 = (size_t)3 * c->n;
    const int G = ode_grid(n3);
    hipLaunchKernelGGL(k_ode_begin, dim3(1), dim3(64), 0, c->stream, s, c->rk_fail.p, c->rk_imax.p);
    KPtrs K;
    for (int j = 0; j < 7; ++j) K.k[j] = c->ode_k[j].p;
    int st = TM_OK;
    for (int stg = 1; stg < 7; ++stg) {  // stages 2..7 (index stg), k[stg] = f(y + h sum a k)
        Coefs cf = {};
        for (int j = 0; j < stg; ++j) cf.a[j] = kA[stg - 1][j];
        hipLaunchKernelGGL(k_ode_stage, dim3(G), dim3(kBlock), 0, c->stream, n3, (const double*)c->m.p, K, cf, stg,
                           (const OdeState*)s, c->ode_y.p);
        TM_HIP(hipGetLastError());
        tm_eval_stats es = {};
        const int r = eval_device(c, c->ode_y.p, c->ode_k[stg].p, alpha, variant, budget[0], budget[1], c->rk_fail.p,
                                  c->rk_imax.p, budget[0] > 0 ? nullptr : &es);
        if (r != TM_OK && r != TM_ERR_MAXITER) return r;
        if (r == TM_ERR_MAXITER) st = r;
        if (iters) {
            iters[0] = std::max(iters[0], es.iters_neumann);
            iters[1] = std::max(iters[1], es.iters_dirichlet);
        }
    }
    // ode_y now holds the 5th-order solution y7 (a_7j = b_j)
    hipLaunchKernelGGL(k_ode_err, dim3(G), dim3(kBlock), 0, c->stream, n3, (const double*)c->m.p,
                       (const double*)c->ode_y.p, K, (const OdeState*)s, c->ode_part.p);
    hipLaunchKernelGGL(k_ode_control, dim3(1), dim3(kBlock), 0, c->stream, G, (double)n3, s,
                       (const double*)c->ode_part.p);
    hipLaunchKernelGGL(k_ode_commit, dim3(G), dim3(kBlock), 0, c->stream, n3, c->m.p, (const double*)c->ode_y.p,
                       c->ode_k[0].p, (const double*)c->ode_k[6].p, s, (const unsigned*)c->rk_fail.p);
    TM_HIP(hipGetLastError());
    return st;
}

// per-block max |x_i|
__global__ __launch_bounds__(kBlock) void k_ode_maxabs(size_t n3, const double* __restrict__ x,
                                                       double* __restrict__ part) {
    __shared__ double red[kBlock];
    double v = 0.0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (size_t)gridDim.x * kBlock)
        v = fmax(v, fabs(x[i]));
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_ode_normalize(int n, double* __restrict__ m) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double x = m[i], y = m[(size_t)n + i], z = m[2 * (size_t)n + i];
        const double l = sqrt(x * x + y * y + z * z);
        if (l > 0.0) {
            m[i] = x / l;
            m[(size_t)n + i] = y / l;
            m[2 * (size_t)n + i] = z / l;
        }
    }
}

}  // namespace

extern "C" {

int tm_ctx_integrate(tm_ctx* c, double tau_span, double rtol, double atol, double* h_io, int max_steps,
                     double alpha, int variant, int use_graph, int normalize, tm_ode_stats* out) {
    TM_CHECK(ctx_ok(c));
    if (!(tau_span >= 0.0) || !(rtol > 0.0) || !(atol > 0.0) || max_steps < 1 ||
        (variant != TM_LLG_CLASSIC && variant != TM_LLG_NOPREC))
        return TM_ERR_BAD_ARG;
    const size_t n3 = (size_t)3 * c->n;
    for (auto& k : c->ode_k) TM_CHECK(k.alloc(n3));
    TM_CHECK(c->ode_y.alloc(n3));
    TM_CHECK(c->ode_part.alloc((size_t)ode_grid(n3)));
    TM_CHECK(c->ode_state.alloc(sizeof(OdeState)));
    TM_CHECK(c->rk_fail.alloc(1));
    TM_CHECK(c->rk_imax.alloc(2));
    TM_CHECK(commit_ops(c));
    OdeState* sd = reinterpret_cast<OdeState*>(c->ode_state.p);
    tm_ode_stats S = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    TM_HIP(hipEventRecord(e0, c->stream));
    // k1 = f(m) at the start (parameters may have changed since the last call)
    tm_eval_stats es = {};
    int status = eval_device(c, c->m.p, c->ode_k[0].p, alpha, variant, 0, 0, nullptr, nullptr, &es);
    if (status == TM_ERR_MAXITER) status = TM_OK;
    OdeState h = {};
    h.t = 0.0;
    h.t_end = tau_span;
    h.rtol = rtol;
    h.atol = atol;
    double h0 = h_io ? *h_io : 0.0;
    if (status == TM_OK && !(h0 > 0.0)) {  // first step: |dm| <= 1e-3 where dm/dtau is largest
        const int G = ode_grid(n3);
        hipLaunchKernelGGL(k_ode_maxabs, dim3(G), dim3(kBlock), 0, c->stream, n3, (const double*)c->ode_k[0].p,
                           c->ode_part.p);
        std::vector<double> pm((size_t)G);
        TM_HIP(hipMemcpyAsync(pm.data(), c->ode_part.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
        TM_HIP(hipStreamSynchronize(c->stream));
        double f = 0.0;
        for (double v : pm) f = std::max(f, v);
        h0 = f > 0.0 ? 1e-3 / f : tau_span;
    }
    h.h = std::min(h0, tau_span);
    h.done = tau_span == 0.0 ? 1 : 0;
    if (status == TM_OK) TM_HIP(hipMemcpyAsync(sd, &h, sizeof(OdeState), hipMemcpyHostToDevice, c->stream));
    S.evaluations = 1;
    hipGraphExec_t exec = nullptr;
    int* budget = c->rk_budget;
    OdeState hs = h;
    for (int step = 0; status == TM_OK && !hs.done && step < max_steps; ++step) {
        bool polled = !use_graph || budget[0] == 0;
        unsigned fail = 0;
        int imax[2] = {0, 0};
        if (!polled) {
            if (!exec) {
                hipGraph_t g = nullptr;
                TM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                const int r = dp_step(c, sd, alpha, variant, budget, nullptr);
                const hipError_t ec = hipStreamEndCapture(c->stream, &g);
                if (r != TM_OK || ec != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    status = r != TM_OK ? r : TM_ERR_HIP;
                    break;
                }
                const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    exec = nullptr;
                    status = TM_ERR_HIP;
                    break;
                }
                S.captures += 1;
            }
            TM_HIP(hipGraphLaunch(exec, c->stream));
            TM_HIP(hipMemcpyAsync(&fail, c->rk_fail.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
            TM_HIP(hipMemcpyAsync(imax, c->rk_imax.p, sizeof(imax), hipMemcpyDeviceToHost, c->stream));
            TM_HIP(hipMemcpyAsync(&hs, sd, sizeof(OdeState), hipMemcpyDeviceToHost, c->stream));
            TM_HIP(hipStreamSynchronize(c->stream));
            S.evaluations += 6;
            S.max_iters_neumann = std::max(S.max_iters_neumann, imax[0]);
            S.max_iters_dirichlet = std::max(S.max_iters_dirichlet, imax[1]);
            if (fail == 0) {
                S.graph_steps += 1;
            } else {  // the commit was skipped: redo the step with host-polled solves
                S.redone_steps += 1;
                polled = true;
                (void)hipGraphExecDestroy(exec);
                exec = nullptr;
            }
        }
        if (polled) {
            int it[2] = {0, 0};
            const int zero[2] = {0, 0};
            const int r = dp_step(c, sd, alpha, variant, zero, it);
            if (r != TM_OK && r != TM_ERR_MAXITER) {
                status = r;
                break;
            }
            TM_HIP(hipMemcpyAsync(&hs, sd, sizeof(OdeState), hipMemcpyDeviceToHost, c->stream));
            TM_HIP(hipStreamSynchronize(c->stream));
            S.evaluations += 6;
            S.sync_steps += 1;
            S.max_iters_neumann = std::max(S.max_iters_neumann, it[0]);
            S.max_iters_dirichlet = std::max(S.max_iters_dirichlet, it[1]);
            budget[0] = std::max(budget[0] + (fail ? 4 : 0), grow(it[0]));
            budget[1] = std::max(budget[1] + (fail ? 4 : 0), grow(it[1]));
        }
        if (hs.nonfinite && hs.h < 1e-300) status = TM_ERR_NONFINITE;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (status == TM_OK && !hs.done) status = TM_ERR_MAXITER;  // max_steps reached before tau_span
    if (normalize && (status == TM_OK || status == TM_ERR_MAXITER))
        hipLaunchKernelGGL(k_ode_normalize, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->n, c->m.p);
    TM_HIP(hipEventRecord(e1, c->stream));
    TM_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    S.accepted = (int)hs.n_accept;
    S.rejected = (int)hs.n_reject;
    S.tau = hs.t;
    S.h_next = hs.h;
    S.last_err = hs.err;
    S.ms = ms;
    S.budget_neumann = budget[0];
    S.budget_dirichlet = budget[1];
    if (h_io) *h_io = hs.h;
    if (out) *out = S;
    return status;
}

}  // extern "C"
