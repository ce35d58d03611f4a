// tetmag_gpu.hpp -- header-only C++ facades with the reference GPU class
// interfaces, implemented over the C ABI of tetmag_hip.h.
//
// The classes keep the names, method names, argument meaning and return
// conventions of tetmag's gpu/ classes (EffFieldGPU.h:54-127,
// GpuPotential.h:50-75, GPUSolver.h:65-87 + SolverFactory.h:42-65,
// SpMatCUDA.h:49-79, LLGWrapper.h:57-63, SetDevice.h:36-44), so TheLLG,
// DemagField and TheSimulation compile against them unchanged apart from the
// type aliases below.  They are templates over the dense and sparse matrix
// types so that no Eigen is needed to build them:
//
//   * Dense  : column-major matrix with Dense(rows, cols), rows(), cols(),
//              data() (Eigen::MatrixXd fits);
//   * Sparse : compressed sparse matrix with rows(), cols(), nonZeros(),
//              outerIndexPtr(), innerIndexPtr(), valuePtr() and a static
//              IsRowMajor (Eigen::SparseMatrix<double, RowMajor> fits);
//   * Vector : host column vector with Vector(n), size(), data()
//              (Eigen::VectorXd fits; tmg::DenseVector standalone);
//   * the device vector is tmg::DeviceVector, with the value semantics of
//     the reference's thrust::device_vector<double> (dev_vec / devVecD).
//
// In tetmag:   using EffFieldGPU  = tmg::EffFieldGPU_T<Eigen::MatrixXd, SpMat>;
//              using GpuPotential = tmg::GpuPotential_T<Eigen::MatrixXd, SpMat, Eigen::VectorXd>;
//              using SpMatCUDA    = tmg::SpMatCUDA_T<SpMat, Eigen::VectorXd>;
//              class GPUSolver : public tmg::GPUSolver_T<SolverFactory, SpMat, Eigen::VectorXd> {};
//              using dev_vec = tmg::DeviceVector;
// Standalone:  using EffFieldGPU = tmg::EffFieldGPU_T<tmg::DenseMatrix, tmg::CsrMatrix>;  (etc.)
//
// Errors: the reference prints and continues or exits; the facades throw
// tmg::Error carrying the TM_ERR_* status, except where the reference returns
// a success flag (SetDevice::set, GPUSolver::compute, wasSuccessful).
#ifndef TETMAG_GPU_HPP_
#define TETMAG_GPU_HPP_

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tetmag_hip.h"

namespace tmg {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& what) : std::runtime_error(what + " failed with status " + std::to_string(st)), status(st) {}
};

inline void check(int st, const char* what) {
    if (st != TM_OK) throw Error(st, what);
}

// ---- standalone matrix types (used when Eigen is not available) ----------
struct DenseMatrix {  // column-major
    int r = 0, c = 0;
    std::vector<double> v;
    DenseMatrix() = default;
    DenseMatrix(int rows_, int cols_) : r(rows_), c(cols_), v((size_t)rows_ * cols_, 0.0) {}
    int rows() const { return r; }
    int cols() const { return c; }
    double* data() { return v.data(); }
    const double* data() const { return v.data(); }
    double& operator()(int i, int j) { return v[(size_t)j * r + i]; }
    double operator()(int i, int j) const { return v[(size_t)j * r + i]; }
};

struct CsrMatrix {  // compressed row storage
    static constexpr bool IsRowMajor = true;
    int r = 0, c = 0;
    std::vector<int> rp, ci;
    std::vector<double> val;
    int rows() const { return r; }
    int cols() const { return c; }
    int nonZeros() const { return rp.empty() ? 0 : rp.back(); }
    const int* outerIndexPtr() const { return rp.data(); }
    const int* innerIndexPtr() const { return ci.data(); }
    const double* valuePtr() const { return val.data(); }
};

// Column vector on the host (the Eigen::VectorXd of the reference when Eigen
// is not available): Vector(n), size(), data(), resize(n), operator[].
struct DenseVector {
    std::vector<double> v;
    DenseVector() = default;
    explicit DenseVector(int n) : v((size_t)n, 0.0) {}
    int size() const { return (int)v.size(); }
    void resize(int n) { v.assign((size_t)n, 0.0); }
    double* data() { return v.data(); }
    const double* data() const { return v.data(); }
    double& operator[](int i) { return v[(size_t)i]; }
    double operator[](int i) const { return v[(size_t)i]; }
    double* begin() { return v.data(); }
    double* end() { return v.data() + v.size(); }
    const double* begin() const { return v.data(); }
    const double* end() const { return v.data() + v.size(); }
};

// Owning device vector with the value semantics of the reference's
// thrust::device_vector<double> (deep copies on copy, moves are cheap), so
// that expressions such as `dev_vec r = *gpucalc->ClassicLLG_dev(H, a)`
// (LLGWrapper.cu:61, 81, 87) compile unchanged.
class DeviceVector {
    double* p_ = nullptr;
    size_t n_ = 0;

   public:
    DeviceVector() = default;
    explicit DeviceVector(size_t n) { resize(n); }
    explicit DeviceVector(const std::vector<double>& h) {
        resize(h.size());
        if (n_ && hipMemcpy(p_, h.data(), n_ * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw Error(TM_ERR_HIP, "hipMemcpy");
    }
    DeviceVector(const DeviceVector& o) { *this = o; }
    DeviceVector& operator=(const DeviceVector& o) {
        if (this == &o) return *this;
        resize(o.n_);
        if (n_ && hipMemcpy(p_, o.p_, n_ * sizeof(double), hipMemcpyDeviceToDevice) != hipSuccess)
            throw Error(TM_ERR_HIP, "hipMemcpy");
        return *this;
    }
    DeviceVector(DeviceVector&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceVector& operator=(DeviceVector&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceVector() {
        if (p_) (void)hipFree(p_);
    }
    void resize(size_t n) {
        if (n == n_) return;
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
        if (n && hipMalloc(&p_, n * sizeof(double)) != hipSuccess) throw Error(TM_ERR_HIP, "hipMalloc");
        n_ = n;
    }
    double* data() { return p_; }
    const double* data() const { return p_; }
    size_t size() const { return n_; }
    std::vector<double> host() const {
        std::vector<double> h(n_);
        if (n_ && hipMemcpy(h.data(), p_, n_ * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw Error(TM_ERR_HIP, "hipMemcpy");
        return h;
    }
};

namespace detail {
// CSR arrays of a sparse matrix (row-major as is, column-major transposed).
template <class Sparse>
struct CsrView {
    std::vector<int> rp, ci;
    std::vector<double> val;
    int rows = 0, cols = 0;
    explicit CsrView(const Sparse& A) {
        rows = (int)A.rows();
        cols = (int)A.cols();
        const int outer = Sparse::IsRowMajor ? rows : cols;
        const int* op = A.outerIndexPtr();
        const int* ip = A.innerIndexPtr();
        const double* vp = A.valuePtr();
        if (Sparse::IsRowMajor) {
            rp.assign(op, op + outer + 1);
            ci.assign(ip, ip + op[outer]);
            val.assign(vp, vp + op[outer]);
        } else {  // CSC -> CSR
            rp.assign((size_t)rows + 1, 0);
            for (int q = 0; q < op[outer]; ++q) rp[ip[q] + 1]++;
            for (int i = 0; i < rows; ++i) rp[i + 1] += rp[i];
            ci.resize(op[outer]);
            val.resize(op[outer]);
            std::vector<int> fill(rp.begin(), rp.end() - 1);
            for (int j = 0; j < outer; ++j)
                for (int q = op[j]; q < op[j + 1]; ++q) {
                    const int p = fill[ip[q]]++;
                    ci[p] = j;
                    val[p] = vp[q];
                }
        }
        // sorted columns within rows
        for (int i = 0; i < rows; ++i) {
            std::vector<std::pair<int, double>> t;
            for (int q = rp[i]; q < rp[i + 1]; ++q) t.push_back({ci[q], val[q]});
            if (std::is_sorted(t.begin(), t.end())) continue;
            std::sort(t.begin(), t.end());
            for (size_t k = 0; k < t.size(); ++k) {
                ci[rp[i] + k] = t[k].first;
                val[rp[i] + k] = t[k].second;
            }
        }
    }
};

// Column-major n x 3 dense matrix -> 3n SoA pointer (they are the same layout).
template <class Dense>
const double* soa(const Dense& M) {
    return M.data();
}
}  // namespace detail

// ---- SetDevice (SetDevice.h:36-44) ---------------------------------------
class SetDevice {
    int ndev = 0;
    int mydev;
    bool success = false;

   public:
    explicit SetDevice(int dev) : mydev(dev) {}
    // true on success; false when no device exists (the caller falls back to
    // the CPU path, tetmag.cpp:82-90).  Out-of-range ids select device 0.
    bool set() {
        int chosen = 0;
        const int st = tm_set_device(mydev, &chosen);
        success = st == TM_OK;
        if (success) mydev = chosen;
        tm_device_count(&ndev);
        return success;
    }
    int device() const { return mydev; }
};

// ---- SpMatCUDA (SpMatCUDA.h:49-79) ---------------------------------------
// Vector: the host column vector type (Eigen::VectorXd in tetmag).
template <class Sparse, class Vector = DenseVector>
class SpMatCUDA_T {
    tm_spmat* h_ = nullptr;
    int rows_ = 0, cols_ = 0;
    DeviceVector x_, y_;

   public:
    const double alpha = 1.0;  // SAXPY parameters, as the reference exposes them
    const double beta = 0.0;
    double* dY = nullptr;  // output (device), valid after mvp()
    double* dX = nullptr;  // input (device), valid after setX()
    SpMatCUDA_T() = default;
    explicit SpMatCUDA_T(const Sparse& A) {
        detail::CsrView<Sparse> c(A);
        rows_ = c.rows;
        cols_ = c.cols;
        check(tm_spmat_create(rows_, cols_, c.rp.data(), c.ci.data(), c.val.data(), &h_), "tm_spmat_create");
        x_.resize((size_t)cols_);
        y_.resize((size_t)rows_);
        dX = x_.data();
        dY = y_.data();
    }
    SpMatCUDA_T(const SpMatCUDA_T&) = delete;
    SpMatCUDA_T& operator=(const SpMatCUDA_T&) = delete;
    virtual ~SpMatCUDA_T() { tm_spmat_destroy(h_); }
    // setX (SpMatCUDA.cu:150-165): the input of the next mvp()
    void setX(const Vector& x) { upload(x.data(), (size_t)x.size()); }
    void setX(const std::vector<double>& x) { upload(x.data(), x.size()); }
    void setX(const DeviceVector& x) {
        if ((int)x.size() != cols_) throw Error(TM_ERR_BAD_ARG, "SpMatCUDA::setX");
        x_ = x;
        dX = x_.data();
    }
    // mvp() (SpMatCUDA.cu:131-141): y = A x on the device
    void mvp() { check(tm_spmat_mvp(h_, x_.data(), TM_DEVICE, y_.data(), TM_DEVICE), "tm_spmat_mvp"); }
    // devVecD mvp(const devVecD&) (SpMatCUDA.cu:143-148): returns A x by value
    DeviceVector mvp(const DeviceVector& x) {
        setX(x);
        mvp();
        return mvpResDev();
    }
    // mvpResEig / mvpResDev (SpMatCUDA.cu:160-169): the result of the last mvp()
    Vector mvpResEig() {
        Vector y(rows_);
        check(hipDeviceSynchronize() == hipSuccess ? TM_OK : TM_ERR_HIP, "hipDeviceSynchronize");
        if (rows_ && hipMemcpy(y.data(), y_.data(), (size_t)rows_ * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw Error(TM_ERR_HIP, "hipMemcpy");
        return y;
    }
    DeviceVector mvpResDev() {
        check(hipDeviceSynchronize() == hipSuccess ? TM_OK : TM_ERR_HIP, "hipDeviceSynchronize");
        return y_;
    }

   private:
    void upload(const double* x, size_t n) {
        if ((int)n != cols_) throw Error(TM_ERR_BAD_ARG, "SpMatCUDA::setX");
        if (n && hipMemcpy(x_.data(), x, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw Error(TM_ERR_HIP, "hipMemcpy");
    }
};

// ---- SolverFactory (SolverFactory.h:42-65) --------------------------------
// The base of the solver family with the reference's protected state; tetmag
// keeps its own SolverFactory (whose CPU solvers are not part of this
// library) and derives GPUSolver_T from it instead.  makeSolver("gpu")
// returns the GPU solver; the CPU solver types are the host's own.
template <class Sparse, class Vector = DenseVector>
class SolverFactory_T {
   protected:
    double cgTol = 1e-7;  // "CG tolerance" default, ProgramSpecs.cpp:70
    Sparse A;
    Vector b;
    Vector x;
    std::string type;
    Vector previous;
    std::string preconditionerType;

   public:
    static std::shared_ptr<SolverFactory_T> makeSolver(std::string type);
    template <class V>
    void setLoadVector(const V& v) {
        b = Vector((int)v.size());
        std::copy(v.data(), v.data() + v.size(), b.data());
    }
    virtual void solve() = 0;
    virtual bool compute() = 0;
    virtual void report() = 0;
    virtual bool wasSuccessful() = 0;
    virtual ~SolverFactory_T() {}
    Vector result() { return x; }
    void setMatrix(const Sparse& M) { A = M; }
    void definePreconditioner(const std::string p) { preconditionerType = p; }
    void setCGTolerance(const double t) { cgTol = t; }
};

// ---- GPUSolver : SolverFactory (GPUSolver.h:65-87) ------------------------
// Base: SolverFactory_T<Sparse, Vector> or the host's own SolverFactory (it
// must provide the protected members A, b, x, cgTol, preconditionerType of
// SolverFactory.h:44-53).  compute() builds the solver once from A; solve()
// solves A x = b warm-started from the previous solution (GPUSolver keeps
// x_d between calls, GPUSolver.cu:82-86) to cgTol.
template <class Base, class Sparse, class Vector = DenseVector>
class GPUSolver_T : public Base {
    tm_solver* h_ = nullptr;
    int n_ = 0, iters_ = 0, status_ = TM_OK;
    double error_ = NAN;

   public:
    int maxit = 100000;
    GPUSolver_T() = default;
    GPUSolver_T(const GPUSolver_T&) = delete;
    GPUSolver_T& operator=(const GPUSolver_T&) = delete;
    virtual ~GPUSolver_T() { tm_solver_destroy(h_); }
    // GPUSolver::compute / setup (GPUSolver.cu:57-79, 88-91): false when the
    // matrix cannot be handled (setup throws in the reference and is caught).
    bool compute() override {
        detail::CsrView<Sparse> c(this->A);
        const std::string& p = this->preconditionerType;
        const int pre = (p == "jacobi" || p == "diagonal") ? TM_PRECOND_JACOBI : TM_PRECOND_AMG;
        tm_solver_destroy(h_);
        h_ = nullptr;
        n_ = c.rows;
        status_ = tm_solver_create(c.rows, c.rp.data(), c.ci.data(), c.val.data(), pre, &h_);
        this->x = Vector(n_);
        std::fill(this->x.data(), this->x.data() + n_, 0.0);
        return status_ == TM_OK;
    }
    void solve() override {
        if (!h_) throw Error(TM_ERR_NOT_READY, "GPUSolver::solve before compute");
        if ((int)this->b.size() != n_) throw Error(TM_ERR_BAD_ARG, "GPUSolver::solve: load vector size");
        if ((int)this->x.size() != n_) this->x = Vector(n_);
        status_ = tm_solver_solve(h_, this->b.data(), TM_HOST, this->x.data(), TM_HOST, this->cgTol, maxit, 2, &iters_,
                                  &error_);
        if (status_ != TM_OK && status_ != TM_ERR_MAXITER && status_ != TM_ERR_NONFINITE)
            check(status_, "tm_solver_solve");
    }
    // GPUSolver::report (GPUSolver.cu:52-54) prints the solver state
    void report() override {
        int nl = 0, sizes[16] = {0};
        if (h_) tm_solver_amg_info(h_, &nl, sizes, 16);
        std::string lv;
        for (int l = 0; l < nl && l < 16; ++l) lv += (l ? " " : "") + std::to_string(sizes[l]);
        std::printf("GPUSolver: PCG (%s), unknowns %d, AMG levels [%s], iterations %d, error %.3e\n",
                    this->preconditionerType.empty() ? "amg" : this->preconditionerType.c_str(), n_, lv.c_str(),
                    iters_, error_);
    }
    // GPUSolver::wasSuccessful (GPUSolver.cu:93-95): the residual is finite
    bool wasSuccessful() override { return std::isfinite(error_) && status_ == TM_OK; }
    int iterations() const { return iters_; }
    double error() const { return error_; }
    int status() const { return status_; }
};

template <class Sparse, class Vector>
std::shared_ptr<SolverFactory_T<Sparse, Vector>> SolverFactory_T<Sparse, Vector>::makeSolver(std::string t) {
    // SolverFactory.cpp:124-134 ("gpu"); the other types are CPU solvers of the host
    if (t == "gpu")
        return std::shared_ptr<SolverFactory_T>(std::make_shared<GPUSolver_T<SolverFactory_T, Sparse, Vector>>());
    throw Error(TM_ERR_UNSUPPORTED, "SolverFactory::makeSolver(" + t + ")");
}

// ---- EffFieldGPU (EffFieldGPU.h:54-127) ----------------------------------
template <class Dense, class Sparse>
class EffFieldGPU_T {
    int nx_;
    tm_ctx* h_ = nullptr;
    // dxdt of the reference (EffFieldGPU.h:70): the *_dev results alias it
    // and are overwritten by the next call, as in EffFieldGPU.cu:336-339
    std::shared_ptr<DeviceVector> dxdt_;
    Dense field(int (*fn)(tm_ctx*, double*, int), const char* what) {
        Dense M(nx_, 3);
        check(fn(h_, M.data(), TM_HOST), what);
        return M;
    }
    void setOp(int op, const Sparse& A) {
        detail::CsrView<Sparse> c(A);
        check(tm_ctx_set_op(h_, op, c.rp.data(), c.ci.data(), c.val.data()), "tm_ctx_set_op");
    }
    std::vector<double> llg(const Dense& H, double alpha, int variant) {
        std::vector<double> out((size_t)3 * nx_);
        check(tm_llg_rhs(h_, detail::soa(H), TM_HOST, alpha, variant, out.data(), TM_HOST), "tm_llg_rhs");
        return out;
    }
    std::shared_ptr<DeviceVector> llg_dev(const Dense& H, double alpha, int variant) {
        check(tm_llg_rhs(h_, detail::soa(H), TM_HOST, alpha, variant, dxdt_->data(), TM_DEVICE), "tm_llg_rhs");
        check(tm_ctx_synchronize(h_), "tm_ctx_synchronize");
        return dxdt_;
    }

   public:
    explicit EffFieldGPU_T(const int nx) : nx_(nx) { init(nx); }
    EffFieldGPU_T(const EffFieldGPU_T&) = delete;
    EffFieldGPU_T& operator=(const EffFieldGPU_T&) = delete;
    ~EffFieldGPU_T() { tm_ctx_destroy(h_); }
    void init(int nx) {
        if (h_) return;
        nx_ = nx;
        check(tm_ctx_create(nx, &h_), "tm_ctx_create");
        dxdt_ = std::make_shared<DeviceVector>((size_t)3 * nx);
    }
    tm_ctx* handle() { return h_; }

    std::vector<double> ClassicLLG_hst(const Dense& Heff, const double alpha) {
        return llg(Heff, alpha, TM_LLG_CLASSIC);
    }
    std::vector<double> LLG_noPrec_hst(const Dense& Heff, const double alpha) {
        return llg(Heff, alpha, TM_LLG_NOPREC);
    }
    std::vector<double> STT_term_LLG_hst(const Dense& U, const double alpha, const double beta) {
        std::vector<double> out((size_t)3 * nx_);
        check(tm_stt_llg(h_, detail::soa(U), TM_HOST, alpha, beta, out.data(), TM_HOST), "tm_stt_llg");
        return out;
    }

    void setMagDev(const Dense& M) { check(tm_ctx_set_mag(h_, detail::soa(M), TM_HOST), "tm_ctx_set_mag"); }
    void setMagDev(const std::vector<double>& m) { check(tm_ctx_set_mag(h_, m.data(), TM_HOST), "tm_ctx_set_mag"); }
    void setMagDev(const DeviceVector& m) {
        check(tm_ctx_set_mag(h_, m.data(), TM_DEVICE), "tm_ctx_set_mag");
        check(tm_ctx_synchronize(h_), "tm_ctx_synchronize");
    }
    void setExchangeMatOnDev(const Sparse& XC) { setOp(TM_OP_EXCHANGE, XC); }
    void setGradientMatsOnDev(const Sparse& a, const Sparse& b, const Sparse& c) {
        setOp(TM_OP_CURL_X, a);
        setOp(TM_OP_CURL_Y, b);
        setOp(TM_OP_CURL_Z, c);
    }

    Dense ExchangeFieldGPU() { return field(tm_exchange_field, "tm_exchange_field"); }
    Dense UniaxialAnisotropyField() { return field(tm_uniaxial_field, "tm_uniaxial_field"); }
    Dense CubicAnisotropyField() { return field(tm_cubic_field, "tm_cubic_field"); }

    template <class Vec>
    void setSTTDataOnDevice(const Sparse& Gx, const Sparse& Gy, const Sparse& Gz, const Vec& ex, const Vec& ey,
                            const Vec& ez) {
        setOp(TM_OP_STT_GRAD_X, Gx);
        setOp(TM_OP_STT_GRAD_Y, Gy);
        setOp(TM_OP_STT_GRAD_Z, Gz);
        check(tm_ctx_set_stt_eta(h_, ex.data(), ey.data(), ez.data()), "tm_ctx_set_stt_eta");
    }
    Dense UTermSTT_GPU() { return field(tm_stt_uterm, "tm_stt_uterm"); }

    // ClassicLLG_dev / LLG_noPrec_dev / STT_term_LLG_dev (EffFieldGPU.cu:316-362,
    // 434-459): the object's dxdt buffer, overwritten by the next call
    std::shared_ptr<DeviceVector> ClassicLLG_dev(const Dense& Heff, const double alpha) {
        return llg_dev(Heff, alpha, TM_LLG_CLASSIC);
    }
    std::shared_ptr<DeviceVector> LLG_noPrec_dev(const Dense& Heff, const double alpha) {
        return llg_dev(Heff, alpha, TM_LLG_NOPREC);
    }
    std::shared_ptr<DeviceVector> STT_term_LLG_dev(const Dense& U, const double alpha, const double beta) {
        check(tm_stt_llg(h_, detail::soa(U), TM_HOST, alpha, beta, dxdt_->data(), TM_DEVICE), "tm_stt_llg");
        check(tm_ctx_synchronize(h_), "tm_ctx_synchronize");
        return dxdt_;
    }

    double MaxTorque(const Dense& Heff) {
        double t = 0.0;
        check(tm_max_torque(h_, detail::soa(Heff), TM_HOST, &t), "tm_max_torque");
        return t;
    }
    template <class Vec>
    void setUniaxialAnisotropy(const Dense& axis, const Vec& Ku) {
        check(tm_ctx_set_uniaxial(h_, detail::soa(axis), Ku.data()), "tm_ctx_set_uniaxial");
    }
    template <class Axes, class Vec>
    void setCubicAnisotropy(const Axes&, const Vec&, const Vec&) {}  // not implemented on the GPU in the reference
    template <class Vec>
    void setDMIdata(const Vec& D, const Vec& invNodeVol, const Dense& nv, const Vec& nodeArea) {
        check(tm_ctx_set_dmi(h_, D.data(), invNodeVol.data(), detail::soa(nv), nodeArea.data()), "tm_ctx_set_dmi");
    }
    Dense DMIField() { return field(tm_dmi_field, "tm_dmi_field"); }
    void NormalizeMag(Dense& M, int nx) {
        (void)nx;
        check(tm_ctx_normalize_mag(h_, M.data(), TM_HOST), "tm_ctx_normalize_mag");
    }
    void displayTimer() {}
};

// ---- GpuPotential (GpuPotential.h:50-75) ---------------------------------
template <class Dense, class Sparse, class Vector = DenseVector>
class GpuPotential_T {
    tm_ctx* h_ = nullptr;
    int nx_ = 0;

   public:
    GpuPotential_T() = default;
    GpuPotential_T(const GpuPotential_T&) = delete;
    GpuPotential_T& operator=(const GpuPotential_T&) = delete;
    ~GpuPotential_T() { tm_ctx_destroy(h_); }
    template <class Vec>
    void setGradientMatrices(const Sparse& tx, const Sparse& ty, const Sparse& tz, const Sparse& gx, const Sparse& gy,
                             const Sparse& gz, const Vec& Js) {
        nx_ = (int)gx.rows();
        tm_ctx_destroy(h_);
        h_ = nullptr;
        check(tm_ctx_create(nx_, &h_), "tm_ctx_create");
        const Sparse* ops[6] = {&tx, &ty, &tz, &gx, &gy, &gz};
        const int ids[6] = {TM_OP_TGRAD_X, TM_OP_TGRAD_Y, TM_OP_TGRAD_Z, TM_OP_GRAD_X, TM_OP_GRAD_Y, TM_OP_GRAD_Z};
        for (int k = 0; k < 6; ++k) {
            detail::CsrView<Sparse> c(*ops[k]);
            check(tm_ctx_set_op(h_, ids[k], c.rp.data(), c.ci.data(), c.val.data()), "tm_ctx_set_op");
        }
        check(tm_ctx_set_js(h_, Js.data()), "tm_ctx_set_js");
    }
    // sum_d tGrad_d (Js m_d), not divided by mu0 (GpuPotential.cu:86-106)
    Vector calcDivM(const Dense& mag) {
        Vector b(nx_);
        check(tm_div_m(h_, detail::soa(mag), TM_HOST, b.data(), TM_HOST), "tm_div_m");
        return b;
    }
    template <class Vec>
    Dense calcGradientField(const Vec& u) {
        Dense H(nx_, 3);
        check(tm_grad_field(h_, u.data(), TM_HOST, H.data(), TM_HOST), "tm_grad_field");
        return H;
    }
};

// ---- LLGWrapper (LLGWrapper.h:57-63) and the CVODE device right-hand side --
// The reference's device right-hand side for CVODE (TheLLG::rhs_d,
// TheLLG.h:109, LLGWrapper.cu:213-228) copies u into the field object,
// evaluates the fields, sums them on the host and returns the LLG right-hand
// side in a device vector.  Here one call evaluates the whole right-hand side
// on the device of the context that owns the operators (the EffFieldGPU's),
// including demag when the context has its boundary, BEM matrix and demag
// solvers set.
class LLGWrapper {
    int nx3 = 0;
    tm_ctx* ctx_ = nullptr;
    double alpha_ = 0.0;
    int variant_ = TM_LLG_CLASSIC;

   public:
    void init(int nx) { nx3 = 3 * nx; }
    int size() const { return nx3; }
    // selectLLGTypeGPU (LLGWrapper.cu:109-120): classic or no-precession form.
    void bind(tm_ctx* ctx, double alpha, bool removePrecession) {
        ctx_ = ctx;
        alpha_ = alpha;
        variant_ = removePrecession ? TM_LLG_NOPREC : TM_LLG_CLASSIC;
    }
    // u, udot: 3n device pointers (SoA x|y|z).  Synchronous like the
    // reference (udot is complete on return, whatever stream reads it).
    void rhs_d(const double* u, double* udot, tm_eval_stats* stats = nullptr) {
        if (!ctx_) throw Error(TM_ERR_NOT_READY, "LLGWrapper::rhs_d before bind");
        check(tm_eval_rhs(ctx_, u, TM_DEVICE, alpha_, variant_, udot, TM_DEVICE, nullptr, TM_DEVICE, stats),
              "tm_eval_rhs");
        check(tm_ctx_synchronize(ctx_), "tm_ctx_synchronize");
    }
    // operator()(dev_vec) (LLGWrapper.cu:104-107)
    void operator()(const DeviceVector& u, DeviceVector& udot) {
        if (udot.size() != u.size()) udot.resize(u.size());
        rhs_d(u.data(), udot.data());
    }
};

// user_data of the CVODE callback (TheLLG::alloc_user_data, TheLLG.cpp:657-663)
struct UserData {
    LLGWrapper* llg;
    int nx;
};

// TheLLG::rhs_d (TheLLG.h:109; LLGWrapper.cu:213-228) with its CVODE
// signature int(realtype t, N_Vector u, N_Vector u_dot, void* user_data).
// NVector is SUNDIALS' N_Vector and DevPtr the accessor of its device array
// (N_VGetDeviceArrayPointer_Hip for the HIP N_Vector).  Returns 0, or -1 (an
// unrecoverable error for CVODE) when the evaluation fails; never throws.
template <class NVector, double* (*DevPtr)(NVector)>
int rhs_d(double t, NVector u, NVector u_dot, void* user_data) {
    (void)t;
    UserData* d = static_cast<UserData*>(user_data);
    if (!d || !d->llg) return -1;
    try {
        d->llg->rhs_d(DevPtr(u), DevPtr(u_dot));
    } catch (const Error&) {
        return -1;
    }
    return 0;
}

}  // namespace tmg

#endif  // TETMAG_GPU_HPP_
