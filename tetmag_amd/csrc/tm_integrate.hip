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
