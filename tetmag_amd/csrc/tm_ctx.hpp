// tm_ctx.hpp -- the device-resident field context shared by the field,
// demag and adjoint sources (tm_field.hip, tm_adjoint.hip), and the helpers
// they share.
#pragma once

#include <vector>

#include "tm_internal.hpp"

struct tm_ctx {
    using SellPattern = tmh::SellPattern;
    template <class T>
    using DevBuf = tmh::DevBuf<T>;
    using CgSolver = tmh::CgSolver;
    int n = 0;
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // sparse operators: host CSR copies until committed
    struct HostCsr {
        bool set = false;
        std::vector<int> rp, col;
        std::vector<double> val;
    } hop[TM_NUM_OPS];
    bool ops_dirty = false;
    SellPattern pat;
    DevBuf<double> opv[TM_NUM_OPS];
    bool has_op[TM_NUM_OPS] = {};
    // per-node data
    DevBuf<double> js, invjs;      // n
    DevBuf<double> ku, kaxis;      // n, 3n
    DevBuf<double> dmi_fac, nvs;   // n (D / V), 3n (node normal * node area)
    DevBuf<double> eta;            // 3n
    DevBuf<double> vol;            // n, node volumes (energies, mean m / mean H)
    DevBuf<double> idmi;           // n, interfacial DMI constant per node
    DevBuf<double> epart, eout;    // energy partial sums and totals
    DevBuf<double> hext;           // 3n, when per-node
    double hext_u[3] = {0, 0, 0};
    bool has_js = false, has_ku = false, has_dmi = false, has_eta = false, hext_field = false, has_vol = false,
         has_idmi = false;
    bool have_u = false;           // u1/u2 hold the potentials of an earlier demag solve
    bool freeze_demag = false;     // reuse u1/u2 in tm_eval_rhs (ProgramSpecs.cpp:81)
    // state and work
    DevBuf<double> m, w3a, w3b;    // 3n
    DevBuf<double> wn;             // n
    DevBuf<double> scal;           // device scalars (max torque)
    // demag
    int bnx = 0;
    DevBuf<int> bnodes;
    DevBuf<double> B;
    int bld = 0;
    bool has_B = false;
    DevBuf<double> bpart, bvec_a, bvec_b;
    CgSolver cgN, cgD;
    bool has_solvers = false;
    DevBuf<double> rhsN, rhsD, u1, u2;
    double tol = 1e-10;
    int maxit = 100000;
    // adjoint: transposed operator values on the union pattern, built on first
    // use (opT_gen[o] == ops_gen when current), and work vectors
    DevBuf<double> opvT[TM_NUM_OPS];
    unsigned opT_gen[TM_NUM_OPS] = {};
    unsigned ops_gen = 1;
    DevBuf<double> a_dm, a_h, a_gh;      // 3n
    DevBuf<double> a_w0, a_t, a_y1, a_y2;  // n
    DevBuf<double> a_xb, a_part, a_red;  // b, partial sums, totals
    // on-device RK4 (tm_integrate.hip): stages, backup of the step's start,
    // device failure counter and iteration maxima of the fixed-budget solves
    DevBuf<double> rk_k[4], rk_tmp, rk_bak;
    DevBuf<unsigned> rk_fail;
    DevBuf<int> rk_imax;
    int rk_budget[2] = {0, 0};
    // adaptive Dormand-Prince 5(4) (tm_integrate.hip): the seven stage
    // derivatives (k[0] = f(m), carried from the last stage of an accepted
    // step: first same as last), the stage argument, the error-norm partial
    // sums and the device step state
    DevBuf<double> ode_k[7], ode_y, ode_part;
    DevBuf<unsigned char> ode_state;
    // phase timers
    bool timers = false;
    hipEvent_t ev[6] = {};
    tm_phase_times tsum = {};
};

namespace tmh {
int ctx_ok(tm_ctx* c);
// Builds the union SELL pattern and uploads the operator values set since the
// last commit.
int commit_ops(tm_ctx* c);
int sell_grid(const tm_ctx* c);
// Stage a caller pointer as a device pointer (copying host data into `tmp`).
int stage_in(tm_ctx* c, const double* src, int where, size_t count, DevBuf<double>& tmp, const double** out);
int finish_out(tm_ctx* c, double* dst, int where, const double* dev, size_t count);
// Output target: the caller's device buffer or the context scratch w3a.
double* out_target(tm_ctx* c, double* dst, int where);
int upload_vec(tm_ctx* c, DevBuf<double>& d, const double* h, size_t count);
// Interfacial DMI configured with everything it needs (D, node volumes,
// grad_x/y, tGrad_x/y).
bool idmi_ready(const tm_ctx* c);
// DemagField::calcPotential on device data; leaves u1, u2 in the context.
int demag_potential_dev(tm_ctx* c, const double* mdev, tm_eval_stats* stats);
// Right-hand side from a device magnetisation; budgetN > 0: fixed-budget
// (graph-capturable) solves, else host-polled solves reporting into stats.
int eval_device(tm_ctx* c, const double* md, double* dmdt, double alpha, int variant, int budgetN, int budgetD,
                unsigned* fails, int* imax2, tm_eval_stats* stats);
}  // namespace tmh
