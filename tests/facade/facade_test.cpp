// facade_test.cpp -- drives the C++ facades (include/tetmag_gpu.hpp) the way
// TheLLG (TheLLG.cpp:55-115, 150-178) and DemagField (DemagField.cpp:92-282)
// call the reference GPU classes, and checks every result against the FP64
// oracle (oracle/refcpu.c, linked as the checker).  Exit code 0 = all checks
// passed.  Needs a GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
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
}

using Dense = tmg::DenseMatrix;
using Sparse = tmg::CsrMatrix;
using EffFieldGPU = tmg::EffFieldGPU_T<Dense, Sparse>;
using GpuPotential = tmg::GpuPotential_T<Dense, Sparse>;
using GPUSolver = tmg::GPUSolver_T<Sparse>;
using SpMatCUDA = tmg::SpMatCUDA_T<Sparse>;

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
    tmg::DeviceVector dd;
    gpucalc.ClassicLLG_dev(H, 0.5, dd);
    std::vector<double> dh = dd.host();
    ref_llg_rhs(n, 0.5, 1, m.data(), H.data(), ref.data());
    expect("ClassicLLG_dev", rel(dh.data(), ref.data(), ref.size()), 1e-14);

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
    }

    // GpuPotential (DemagField::initializeSolvers, DemagField.cpp:243-250)
    GpuPotential gp;
    gp.setGradientMatrices(sparse(tX, 1), sparse(tY, 1), sparse(tZ, 1), sparse(gX, 1), sparse(gY, 1), sparse(gZ, 1), Js);
    std::vector<double> b = gp.calcDivM(magM), bref(n);
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
    for (const char* pre : {"amg", "jacobi"}) {
        GPUSolver s;
        s.setMatrix(KN);
        s.setCGTolerance(1e-13);
        s.definePreconditioner(pre);
        if (!s.compute()) {
            std::printf("GPUSolver::compute failed\n");
            return 1;
        }
        s.setLoadVector(b);
        s.solve();
        std::vector<double> x(n, 0.0);
        void* h = ref_solver_create(n, KN.rp.data(), KN.ci.data(), KN.val.data());
        double rr = 0;
        ref_solver_solve(h, b.data(), x.data(), 1e-14, 20000, 60, &rr);
        ref_solver_destroy(h);
        std::printf("GPUSolver(%s) iterations %d\n", pre, s.iterations());
        expect(pre[0] == 'a' ? "GPUSolver(amg)" : "GPUSolver(jacobi)", rel(s.result().data(), x.data(), x.size()),
               1e-10);
        if (!s.wasSuccessful()) ++failures;
    }
    // SpMatCUDA
    SpMatCUDA sm(sparse(K, 1.0));
    std::vector<double> y = sm.mvp(b), yr(n);
    ref_spmv(n, rp.data(), col.data(), K.data(), b.data(), yr.data());
    expect("SpMatCUDA::mvp", rel(y.data(), yr.data(), y.size()), 1e-14);
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
