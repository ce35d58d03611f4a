// facade_test.cpp -- drives the C++ facades (include/tetmag_gpu.hpp) the way
// TheLLG (TheLLG.cpp:55-115, 150-178) and DemagField (DemagField.cpp:92-282)
// call the reference GPU classes, and checks every result against the FP64
// oracle (oracle/refcpu.c, linked as the checker).  Exit code 0 = all checks
// passed.  Needs a GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "tetmag_gpu.hpp"

extern "C" {
// oracle (test infrastructure)
int ref_fem_pattern(int n, int ntet, const int* fel, int* rowptr);
void ref_fem_assemble(int n, int ntet, const double* xyz, const int* fel, const int* rowptr, int* col, double* K,
                      double* gX, double* gY, double* gZ, double* tX, double* tY, double* tZ, double* nodevol);
void ref_exchange_field(int n, const int* rp, const int* col, const double* K, const double* A, const double* V,
                        const double* m, double* H);
void ref_uniaxial_field(int n, const double* Ku, const double* axis, const double* m, double* H);
void ref_div_m(int n, const int* rp, const int* col, const double* tX, const double* tY, const double* tZ,
               const double* Js, const double* m, double* b);
void ref_grad_field(int n, const int* rp, const int* col, const double* gX, const double* gY, const double* gZ,
                    const double* u, double* H);
void ref_llg_rhs(int n, double alpha, int variant, const double* m, const double* H, double* out);
double ref_max_torque(int n, const double* m, const double* H);
int ref_replace_rows_identity(int n, const int* rp, const int* col, const double* val, const char* fixed, int* rp_out,
                              int* col_out, double* val_out);
void* ref_solver_create(int n, const int* rp, const int* col, const double* val);
int ref_solver_solve(void* h, const double* b, double* x, double tol, int maxit, int restart, double* relres);
void ref_solver_destroy(void* h);
void ref_spmv(int n, const int* rp, const int* col, const double* val, const double* x, double* y);
void ref_stt_uterm(int n, const int* rp, const int* col, const double* gX, const double* gY, const double* gZ,
                   const double* ex, const double* ey, const double* ez, const double* m, double* U);
void ref_stt_llg(int n, double alpha, double beta, const double* m, const double* U, double* out);
}

// The aliases a tetmag build would make (with Eigen::MatrixXd, SpMat,
// Eigen::VectorXd in place of the standalone types).
using Dense = tmg::DenseMatrix;
using Sparse = tmg::CsrMatrix;
using VectorXd = tmg::DenseVector;
using dev_vec = tmg::DeviceVector;
using EffFieldGPU = tmg::EffFieldGPU_T<Dense, Sparse>;
using GpuPotential = tmg::GpuPotential_T<Dense, Sparse, VectorXd>;
using SolverFactory = tmg::SolverFactory_T<Sparse, VectorXd>;
using SpMatCUDA = tmg::SpMatCUDA_T<Sparse, VectorXd>;

// A host-side SolverFactory of its own (as tetmag keeps SolverFactory.h with
// its CPU solvers): GPUSolver derives from it, as GPUSolver.h:65 does.
class HostSolverFactory {
   protected:
    double cgTol = 1e-7;
    Sparse A;
    VectorXd b, x;
    std::string preconditionerType;

   public:
    virtual ~HostSolverFactory() {}
    virtual void solve() = 0;
    virtual bool compute() = 0;
    virtual void report() = 0;
    virtual bool wasSuccessful() = 0;
    void setMatrix(const Sparse& M) { A = M; }
    void setLoadVector(const VectorXd& v) { b = v; }
    void setCGTolerance(const double t) { cgTol = t; }
    void definePreconditioner(const std::string p) { preconditionerType = p; }
    VectorXd result() { return x; }
};
class GPUSolver : public tmg::GPUSolver_T<HostSolverFactory, Sparse, VectorXd> {};

// A stand-in for SUNDIALS' N_Vector and its device-array accessor.
struct FakeNVector {
    double* dev;
};
static double* fake_device_pointer(FakeNVector* v) { return v->dev; }

static int failures = 0;

static double rel(const double* a, const double* b, size_t n) {
    double d = 0, r = 0;
    for (size_t i = 0; i < n; ++i) {
        d += (a[i] - b[i]) * (a[i] - b[i]);
        r += b[i] * b[i];
    }
    return std::sqrt(d / (r > 0 ? r : 1.0));
}

static void expect(const char* what, double err, double tol) {
    const bool ok = err <= tol;
    std::printf("%-28s rel_err %.3e (tol %.0e) %s\n", what, err, tol, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

int main() {
    tmg::SetDevice dev(0);
    if (!dev.set()) {
        std::printf("no GPU\n");
        return 2;
    }
    const int nx = 9, ny = 7, nz = 4;
    int n = 0, nt = 0;
    tm_mesh_kuhn_cuboid_size(nx, ny, nz, &n, &nt);
    std::vector<double> xyz((size_t)3 * n);
    std::vector<int> fel((size_t)4 * nt);
    tmg::check(tm_mesh_kuhn_cuboid(nx, ny, nz, 3.0, 0.1, 77, xyz.data(), fel.data()), "mesh");
    // oracle operators
    std::vector<int> rp((size_t)n + 1);
    ref_fem_pattern(n, nt, fel.data(), rp.data());
    const int nnz = rp[n];
    std::vector<int> col(nnz);
    std::vector<double> K(nnz), gX(nnz), gY(nnz), gZ(nnz), tX(nnz), tY(nnz), tZ(nnz), V(n);
    ref_fem_assemble(n, nt, xyz.data(), fel.data(), rp.data(), col.data(), K.data(), gX.data(), gY.data(), gZ.data(),
                     tX.data(), tY.data(), tZ.data(), V.data());
    auto sparse = [&](const std::vector<double>& v, double s) {
        Sparse S;
        S.r = S.c = n;
        S.rp = rp;
        S.ci = col;
        S.val = v;
        for (auto& x : S.val) x *= s;
        return S;
    };
    std::vector<double> A(n, 1.3e-11 / 1e-18), Js(n, 1.00530964914873), Ku(n, 4e4), m((size_t)3 * n), axis((size_t)3 * n);
    for (int i = 0; i < n; ++i) {
        double a = 0.37 * i, b = 0.11 * i + 0.3;
        double v[3] = {std::cos(a) * std::sin(b), std::sin(a) * std::sin(b), std::cos(b)};
        for (int d = 0; d < 3; ++d) {
            m[(size_t)d * n + i] = v[d];
            axis[(size_t)d * n + i] = d == 2 ? 0.8 : (d == 0 ? 0.6 : 0.0);
        }
    }
    // TheLLG constructor: XC = -2 diag(A/V) K (TheLLG.cpp:88-93)
    Sparse XC = sparse(K, 1.0);
    for (int i = 0; i < n; ++i)
        for (int q = rp[i]; q < rp[i + 1]; ++q) XC.val[q] = -2.0 * (A[i] * (1.0 / V[i])) * K[q];
    EffFieldGPU gpucalc(n);
    gpucalc.setExchangeMatOnDev(XC);
    Dense axisM(n, 3), magM(n, 3);
    std::copy(axis.begin(), axis.end(), axisM.v.begin());
    std::copy(m.begin(), m.end(), magM.v.begin());
    gpucalc.setUniaxialAnisotropy(axisM, Ku);
    gpucalc.setMagDev(magM);
    std::vector<double> ref((size_t)3 * n);
    Dense Hx = gpucalc.ExchangeFieldGPU();
    ref_exchange_field(n, rp.data(), col.data(), K.data(), A.data(), V.data(), m.data(), ref.data());
    expect("ExchangeFieldGPU", rel(Hx.data(), ref.data(), ref.size()), 1e-12);
    Dense Hk = gpucalc.UniaxialAnisotropyField();
    ref_uniaxial_field(n, Ku.data(), axis.data(), m.data(), ref.data());
    expect("UniaxialAnisotropyField", rel(Hk.data(), ref.data(), ref.size()), 1e-14);
    // Heff on the host, then ClassicLLG_hst (TheLLG.cpp:150-178)
    Dense H(n, 3);
    for (size_t q = 0; q < H.v.size(); ++q) H.v[q] = (Hx.v[q] + Hk.v[q]) / Js[q % n];
    std::vector<double> dm = gpucalc.ClassicLLG_hst(H, 0.5);
    ref_llg_rhs(n, 0.5, 1, m.data(), H.data(), ref.data());
    expect("ClassicLLG_hst", rel(dm.data(), ref.data(), ref.size()), 1e-14);
    dm = gpucalc.LLG_noPrec_hst(H, 0.5);
    ref_llg_rhs(n, 0.5, 0, m.data(), H.data(), ref.data());
    expect("LLG_noPrec_hst", rel(dm.data(), ref.data(), ref.size()), 1e-14);
    const double mt = gpucalc.MaxTorque(H), mtr = ref_max_torque(n, m.data(), H.data());
    expect("MaxTorque", std::fabs(mt - mtr) / mtr, 1e-14);
    // LLGWrapper.cu:61, 81, 87 dereference the returned pointer into a dev_vec
    dev_vec dd = *gpucalc.ClassicLLG_dev(H, 0.5);
    std::vector<double> dh = dd.host();
    ref_llg_rhs(n, 0.5, 1, m.data(), H.data(), ref.data());
    expect("ClassicLLG_dev", rel(dh.data(), ref.data(), ref.size()), 1e-14);
    {
        std::shared_ptr<dev_vec> p1 = gpucalc.ClassicLLG_dev(H, 0.5);
        std::shared_ptr<dev_vec> p2 = gpucalc.LLG_noPrec_dev(H, 0.5);
        if (p1.get() != p2.get()) {  // one dxdt buffer per object, as EffFieldGPU.h:70
            std::printf("the *_dev results do not alias one buffer\n");
            ++failures;
        }
        std::vector<double> h2 = p2->host();
        ref_llg_rhs(n, 0.5, 0, m.data(), H.data(), ref.data());
        expect("LLG_noPrec_dev", rel(h2.data(), ref.data(), ref.size()), 1e-14);
    }

    // LLGWrapper::rhs_d: whole right-hand side on the device (no Js set, so
    // invJs = 1 and Heff = Hx + Hk; no demag on this context).
    {
        tmg::LLGWrapper w;
        w.init(n);
        w.bind(gpucalc.handle(), 0.5, false);
        tmg::DeviceVector u((size_t)3 * n), udot;
        if (hipMemcpy(u.data(), m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 1;
        w(u, udot);
        std::vector<double> got = udot.host(), Hs((size_t)3 * n);
        for (size_t q = 0; q < Hs.size(); ++q) Hs[q] = Hx.v[q] + Hk.v[q];
        ref_llg_rhs(n, 0.5, 1, m.data(), Hs.data(), ref.data());
        expect("LLGWrapper::rhs_d", rel(got.data(), ref.data(), ref.size()), 1e-13);
        // the CVODE callback of TheLLG::rhs_d (TheLLG.h:109)
        int (*cb)(double, FakeNVector*, FakeNVector*, void*) = tmg::rhs_d<FakeNVector*, fake_device_pointer>;
        tmg::UserData ud{&w, n};
        dev_vec udot2(u.size());
        FakeNVector nu{u.data()}, nud{udot2.data()};
        if (cb(0.0, &nu, &nud, &ud) != 0) ++failures;
        got = udot2.host();
        expect("rhs_d (CVODE callback)", rel(got.data(), ref.data(), ref.size()), 1e-13);
        if (cb(0.0, &nu, &nud, nullptr) != -1) ++failures;  // bad user data: error status, no throw
    }

    // GpuPotential (DemagField::initializeSolvers, DemagField.cpp:243-250)
    GpuPotential gp;
    gp.setGradientMatrices(sparse(tX, 1), sparse(tY, 1), sparse(tZ, 1), sparse(gX, 1), sparse(gY, 1), sparse(gZ, 1), Js);
    VectorXd b = gp.calcDivM(magM);  // DemagField.cpp:128
    std::vector<double> bref(n);
    ref_div_m(n, rp.data(), col.data(), tX.data(), tY.data(), tZ.data(), Js.data(), m.data(), bref.data());
    for (auto& v : bref) v *= 1.256637061435917295376e-6;  // the oracle divides by mu0 like rhsPoisson
    expect("GpuPotential::calcDivM", rel(b.data(), bref.data(), b.size()), 1e-13);
    Dense Hg = gp.calcGradientField(b);
    ref_grad_field(n, rp.data(), col.data(), gX.data(), gY.data(), gZ.data(), b.data(), ref.data());
    // both return H = -grad u (GpuPotential.cu:120, DemagField.cpp:108-110)
    expect("GpuPotential::calcGradField", rel(Hg.data(), ref.data(), ref.size()), 1e-13);

    // GPUSolver on the Neumann matrix (DemagField.cpp:204-220, 233-240)
    std::vector<char> fixed(n, 0);
    fixed[n / 2] = 1;
    Sparse KN;
    KN.r = KN.c = n;
    KN.rp.resize(n + 1);
    KN.ci.resize(nnz);
    KN.val.resize(nnz);
    const int nn = ref_replace_rows_identity(n, rp.data(), col.data(), K.data(), fixed.data(), KN.rp.data(),
                                             KN.ci.data(), KN.val.data());
    KN.ci.resize(nn);
    KN.val.resize(nn);
    for (auto& v : b) v /= 1.256637061435917295376e-6;
    std::vector<double> xo(n, 0.0);
    {
        void* h = ref_solver_create(n, KN.rp.data(), KN.ci.data(), KN.val.data());
        double rr = 0;
        ref_solver_solve(h, b.data(), xo.data(), 1e-14, 20000, 60, &rr);
        ref_solver_destroy(h);
    }
    for (const char* pre : {"amg", "jacobi"}) {
        // DemagField::initializeSolvers (DemagField.cpp:252-266) and solvePoisson (:233-240)
        std::shared_ptr<SolverFactory> s = SolverFactory::makeSolver("gpu");
        s->setMatrix(KN);
        s->setCGTolerance(1e-13);
        s->definePreconditioner(pre);
        if (!s->compute()) {
            std::printf("GPUSolver::compute failed\n");
            return 1;
        }
        s->setLoadVector(b);
        s->solve();
        VectorXd u1 = s->result();
        s->report();
        expect(pre[0] == 'a' ? "makeSolver(gpu) amg" : "makeSolver(gpu) jacobi", rel(u1.data(), xo.data(), xo.size()),
               1e-10);
        if (!s->wasSuccessful()) ++failures;
    }
    {
        GPUSolver s;  // derived from the host's own SolverFactory
        s.setMatrix(KN);
        s.setCGTolerance(1e-13);
        s.definePreconditioner("amg");
        if (!s.compute()) return 1;
        s.setLoadVector(b);
        s.solve();
        VectorXd u1 = s.result();
        expect("GPUSolver : host factory", rel(u1.data(), xo.data(), xo.size()), 1e-10);
        s.solve();  // warm start from the previous solution: already converged
        if (s.iterations() != 0 || !s.wasSuccessful()) ++failures;
    }
    // SpMatCUDA (SpMatCUDA.h:65-78)
    SpMatCUDA sm(sparse(K, 1.0));
    std::vector<double> yr(n);
    ref_spmv(n, rp.data(), col.data(), K.data(), b.data(), yr.data());
    sm.setX(b);
    sm.mvp();
    VectorXd y = sm.mvpResEig();
    expect("SpMatCUDA setX/mvp/mvpResEig", rel(y.data(), yr.data(), yr.size()), 1e-14);
    dev_vec bd(std::vector<double>(b.data(), b.data() + n));
    dev_vec yd = sm.mvp(bd);
    std::vector<double> yh = yd.host();
    expect("SpMatCUDA mvp(devVecD)", rel(yh.data(), yr.data(), yr.size()), 1e-14);
    sm.setX(std::vector<double>(b.data(), b.data() + n));
    sm.mvp();
    yh = sm.mvpResDev().host();
    expect("SpMatCUDA mvpResDev", rel(yh.data(), yr.data(), yr.size()), 1e-14);

    // Zhang-Li STT (TheLLG.cpp:402-416 -> UTermSTT_GPU, TheLLG.cpp:207 -> STT_term_LLG_hst)
    {
        std::vector<double> ex(n), ey(n, 0.0), ez(n);
        for (int i = 0; i < n; ++i) {
            ex[i] = 1e-3 * (1.0 + 0.1 * std::sin(0.3 * i));
            ez[i] = -4e-4;
        }
        gpucalc.setSTTDataOnDevice(sparse(gX, 1), sparse(gY, 1), sparse(gZ, 1), ex, ey, ez);
        Dense U = gpucalc.UTermSTT_GPU();
        std::vector<double> Ur((size_t)3 * n);
        ref_stt_uterm(n, rp.data(), col.data(), gX.data(), gY.data(), gZ.data(), ex.data(), ey.data(), ez.data(),
                      m.data(), Ur.data());
        expect("UTermSTT_GPU", rel(U.data(), Ur.data(), Ur.size()), 1e-13);
        std::vector<double> st = gpucalc.STT_term_LLG_hst(U, 0.5, 0.04);
        ref_stt_llg(n, 0.5, 0.04, m.data(), U.data(), ref.data());
        expect("STT_term_LLG_hst", rel(st.data(), ref.data(), ref.size()), 1e-14);
        dev_vec sd = *gpucalc.STT_term_LLG_dev(U, 0.5, 0.04);
        std::vector<double> sh = sd.host();
        expect("STT_term_LLG_dev", rel(sh.data(), ref.data(), ref.size()), 1e-14);
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
