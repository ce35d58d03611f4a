/*
 * tetmag_hip.h -- C ABI of the MI355X (gfx950) effective-field + LLG-RHS path.
 *
 * This is the boundary between a tetmag-style host (TheLLG, DemagField,
 * TheSimulation) and the HIP kernels.  Every entry point:
 *   - takes plain pointers and sizes (no C++ or torch types),
 *   - returns an int status (TM_OK or one of TM_ERR_*), never throws,
 *   - names the reference interface it replaces (file:line in tetmag v1.7.7).
 *
 * Layout conventions (same as the reference):
 *   - a "3n" field or magnetisation is the column-major n x 3 matrix
 *     (x-block, y-block, z-block), i.e. the CVODE state vector layout
 *     (TheLLG.cpp:156, typedefs.h:39-41);
 *   - sparse operators are CSR with int32 indices and FP64 values, the arrays
 *     of an Eigen RowMajor SparseMatrix (typedefs.h:35);
 *   - the boundary-element matrix is dense bnx x bnx, column-major, rows and
 *     columns in the order of the ascending boundary-node list
 *     (Boundary.cpp:200-211, BEMprocessing.cpp:375-411);
 *   - mesh geometry is xyz (n x 3 row-major, mesh units) and fel (ntet x 4
 *     row-major).
 *
 * Pointer arguments carry a "where" flag: TM_HOST for host memory (the call
 * copies and synchronises), TM_DEVICE for device memory on the context's GPU
 * (the call is stream-ordered on the context stream and returns without
 * synchronising, except where a scalar result is returned).
 *
 * Thread safety: one context per host thread.  Calls on one context are not
 * reentrant (the reference GPU classes are not thread-safe either,
 * EffFieldGPU.h:54-127, LLGWrapper.cu:43-44).
 */
#ifndef TETMAG_HIP_H_
#define TETMAG_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TM_OK 0
#define TM_ERR_NO_DEVICE 1   /* no usable GPU (SetDevice.cu:43-49 falls back to CPU) */
#define TM_ERR_BAD_ARG 2     /* invalid size, pointer, layout or operator */
#define TM_ERR_NONFINITE 3   /* NaN/Inf met in a solve (GPUSolver.cu:95-97 checks isfinite) */
#define TM_ERR_MAXITER 4     /* solver did not reach the tolerance within maxit */
#define TM_ERR_HIP 5         /* a HIP runtime call failed */
#define TM_ERR_NOT_READY 6   /* a required operator/vector was not set */
#define TM_ERR_UNSUPPORTED 7 /* matrix form the solver cannot handle (e.g. non-symmetric free block) */

#define TM_HOST 0
#define TM_DEVICE 1

/* ---- sparse operator ids for tm_ctx_set_op ----------------------------- */
#define TM_OP_EXCHANGE 0 /* XC_field_OP = -2 diag(A/V) K  (TheLLG.cpp:88-93 -> EffFieldGPU::setExchangeMatOnDev) */
#define TM_OP_CURL_X 1   /* -tGradX as passed to EffFieldGPU::setGradientMatsOnDev (TheLLG.cpp:94-97) */
#define TM_OP_CURL_Y 2
#define TM_OP_CURL_Z 3
#define TM_OP_TGRAD_X 4  /* tGrad as passed to GpuPotential::setGradientMatrices (DemagField.cpp:248) */
#define TM_OP_TGRAD_Y 5
#define TM_OP_TGRAD_Z 6
#define TM_OP_GRAD_X 7   /* grad as passed to GpuPotential::setGradientMatrices */
#define TM_OP_GRAD_Y 8
#define TM_OP_GRAD_Z 9
#define TM_OP_STT_GRAD_X 10 /* Grad ops of EffFieldGPU::setSTTDataOnDevice (EffFieldGPU.cu:253-287) */
#define TM_OP_STT_GRAD_Y 11
#define TM_OP_STT_GRAD_Z 12
#define TM_NUM_OPS 13

/* ---- LLG variants (TheLLG.cpp:129-141, selectLLGType) ------------------ */
#define TM_LLG_NOPREC 0  /* dm/dtau = -alpha m x (m x H)           (EffFieldGPU.cu:341-362) */
#define TM_LLG_CLASSIC 1 /* dm/dtau = -m x H - alpha m x (m x H)   (EffFieldGPU.cu:316-339) */

/* ---- solid-angle form of the GPU boundary-element assembly -------------- */
#define TM_BEM_REFERENCE 0 /* 2 acos(.) of Lindholm.cpp:41-60, same operation order (bitwise-faithful) */
#define TM_BEM_ACCURATE 1  /* Van Oosterom-Strackee 2 atan2(.): full precision for distant triangles */

/* ---- preconditioners for tm_solver_create ------------------------------ */
#define TM_PRECOND_JACOBI 0 /* diagonal */
#define TM_PRECOND_AMG 1    /* smoothed-aggregation AMG V(1,1)-cycle (the reference uses AMGCL SA, GPUSolver.h:53-63) */

/* ---- device selection: SetDevice (SetDevice.h:36-44) ------------------- */

/* Number of visible GPUs; TM_ERR_NO_DEVICE when none (SetDevice.cu:41-50). */
int tm_device_count(int* ndev);
/* SetDevice::set (SetDevice.cu:41-59): selects `dev`, falling back to device
 * 0 when dev is out of range; *chosen receives the device actually used. */
int tm_set_device(int dev, int* chosen);

/* ---- generic sparse matrix: SpMatCUDA (SpMatCUDA.h:49-79) -------------- */
typedef struct tm_spmat tm_spmat;
/* SpMatCUDA(const SpMat&) (SpMatCUDA.cu:59-75): deep copy of a CSR matrix to
 * the device.  rows == cols is not required. */
int tm_spmat_create(int rows, int cols, const int* rowptr, const int* colidx, const double* vals, tm_spmat** out);
int tm_spmat_destroy(tm_spmat* A);
/* SpMatCUDA::mvp / setX / mvpResEig / mvpResDev (SpMatCUDA.cu:131-169): y = A x. */
int tm_spmat_mvp(tm_spmat* A, const double* x, int xwhere, double* y, int ywhere);

/* ---- linear solver: GPUSolver (GPUSolver.h:65-87) ----------------------- */
typedef struct tm_solver tm_solver;
/* SolverFactory::setMatrix + GPUSolver::compute/setup (SolverFactory.cpp:45-48,
 * GPUSolver.cu:57-79).  Accepts the matrices tetmag actually hands to its
 * solvers: symmetric stiffness matrices some of whose rows were replaced by
 * identity rows (Neumann: DemagField.cpp:204-220, Dirichlet:
 * FEMprocessing.cpp:329-369).  Identity rows are treated as fixed values and
 * the remaining free block, which must be symmetric, is solved by
 * preconditioned CG.  Returns TM_ERR_UNSUPPORTED if the free block is not
 * symmetric. */
int tm_solver_create(int n, const int* rowptr, const int* colidx, const double* vals, int precond, tm_solver** out);
int tm_solver_destroy(tm_solver* s);
/* GPUSolver::solve (GPUSolver.cu:82-86): solves A x = b to
 * ||b - A x|| <= tol ||b||.  x is the warm start on entry when warm == 1;
 * warm == 2 starts from the solver's own earlier solutions (the reference
 * keeps x_d between calls): the combination of the last six solutions
 * whose right-hand sides best match b in the least-squares sense, the last
 * solution alone being one of the candidates; warm == 0 starts from zero.
 * Returns TM_ERR_MAXITER when maxit is reached first (the reference caps at
 * 100 GMRES iterations silently, GPUSolver.cu:47).  iters/relres may be NULL. */
int tm_solver_solve(tm_solver* s, const double* b, int bwhere, double* x, int xwhere, double tol, int maxit,
                    int warm, int* iters, double* relres);

/* Solves A^T x = b for the solver's matrix A (symmetric apart from identity
 * rows): CG on the free block, then the identity rows' values from the
 * coupling.  Warm start: a separate solution vector from tm_solver_solve.
 * The adjoint of GPUSolver::solve (no counterpart in the reference). */
int tm_solver_solve_transposed(tm_solver* s, const double* b, int bwhere, double* x, int xwhere, double tol,
                               int maxit, int warm, int* iters, double* relres);
/* Preconditioner hierarchy of an AMG solver: *nlevels receives the number of
 * levels including the dense coarsest one, sizes[0..min(cap,nlevels)-1] their
 * unknowns (0 levels for a Jacobi solver). */
int tm_solver_amg_info(tm_solver* s, int* nlevels, int* sizes, int cap);

/* ---- field context: EffFieldGPU + GpuPotential + DemagField state ------- */
typedef struct tm_ctx tm_ctx;

typedef struct tm_eval_stats {
    int iters_neumann;      /* CG iterations of the u1 solve */
    int iters_dirichlet;    /* CG iterations of the u2 solve */
    double relres_neumann;  /* final relative residuals */
    double relres_dirichlet;
    double max_torque;      /* max_i |m_i x Heff_i| in A/m (TheLLG.cpp:426-439 multiplies by mu0) */
} tm_eval_stats;

/* EffFieldGPU(int) + init (EffFieldGPU.cu:50-53, 216-251) on the current device. */
int tm_ctx_create(int n, tm_ctx** out);
int tm_ctx_destroy(tm_ctx* c);
/* Run the context's work on a caller-provided hipStream_t (NULL = the
 * context's own stream). */
int tm_ctx_set_stream(tm_ctx* c, void* hip_stream);
/* Blocks until all work queued on the context stream has finished. */
int tm_ctx_synchronize(tm_ctx* c);

/* Sparse operators: EffFieldGPU::setExchangeMatOnDev (EffFieldGPU.cu:211-214),
 * setGradientMatsOnDev (EffFieldGPU.cu:205-209), setSTTDataOnDevice matrices
 * (EffFieldGPU.cu:253-287), GpuPotential::setGradientMatrices
 * (GpuPotential.cu:54-83).  n x n CSR, host arrays, deep-copied. */
int tm_ctx_set_op(tm_ctx* c, int op, const int* rowptr, const int* colidx, const double* vals);

/* GpuPotential::setGradientMatrices Js argument (GpuPotential.cu:80-82) and
 * TheLLG's invJs (TheLLG.cpp:110, 1/Js with non-finite -> 0).  n values, T. */
int tm_ctx_set_js(tm_ctx* c, const double* Js);
/* EffFieldGPU::setUniaxialAnisotropy (EffFieldGPU.cu:55-64): axis 3n, Ku n (J/m^3). */
int tm_ctx_set_uniaxial(tm_ctx* c, const double* axis3n, const double* Ku);
/* EffFieldGPU::setDMIdata (EffFieldGPU.cu:289-314): D (J/m^2 per mesh unit,
 * already scaled, SimulationData.cpp:45-49), invNodeVol, node normals 3n,
 * node surface area. */
int tm_ctx_set_dmi(tm_ctx* c, const double* D, const double* invNodeVol, const double* nv3n, const double* nodeArea);
/* EffFieldGPU::setSTTDataOnDevice vectors eta_j{x,y,z} (EffFieldGPU.cu:253-287). */
int tm_ctx_set_stt_eta(tm_ctx* c, const double* eta_x, const double* eta_y, const double* eta_z);
/* TheLLG::setZeemanField (TheLLG.cpp:477-483): applied field in A/m; either
 * a uniform vector (H3n == NULL) or a per-node 3n field. */
int tm_ctx_set_zeeman(tm_ctx* c, const double* H3n, double hx, double hy, double hz);

/* EffFieldGPU::setMagDev (EffFieldGPU.cu:164-203). */
int tm_ctx_set_mag(tm_ctx* c, const double* m3n, int where);
int tm_ctx_get_mag(tm_ctx* c, double* m3n, int where);
/* EffFieldGPU::NormalizeMag (EffFieldGPU.cu:625-663): normalises m in place
 * and stores it as the context magnetisation. */
int tm_ctx_normalize_mag(tm_ctx* c, double* m3n, int where);

/* Individual fields from the context magnetisation, each n x 3 column-major,
 * same units as the reference return values (J/m^3 before invJs for
 * exchange/anisotropy/DMI). */
int tm_exchange_field(tm_ctx* c, double* out3n, int where);  /* ExchangeFieldGPU (EffFieldGPU.cu:71-86) */
int tm_uniaxial_field(tm_ctx* c, double* out3n, int where);  /* UniaxialAnisotropyField (EffFieldGPU.cu:538-560) */
int tm_cubic_field(tm_ctx* c, double* out3n, int where);     /* CubicAnisotropyField (EffFieldGPU.cu:514-535): zeros, as the reference */
int tm_dmi_field(tm_ctx* c, double* out3n, int where);       /* DMIField (EffFieldGPU.cu:596-622) */
int tm_stt_uterm(tm_ctx* c, double* out3n, int where);       /* UTermSTT_GPU (EffFieldGPU.cu:107-148) */

/* ClassicLLG_dev/_hst and LLG_noPrec_dev/_hst (EffFieldGPU.cu:316-430):
 * dm/dtau from the context magnetisation and a given Heff. */
int tm_llg_rhs(tm_ctx* c, const double* Heff3n, int hwhere, double alpha, int variant, double* out3n, int owhere);
/* STT_term_LLG_dev (EffFieldGPU.cu:434-475): -(beta-alpha) m x u - (1+alpha beta) m x (m x u). */
int tm_stt_llg(tm_ctx* c, const double* U3n, int uwhere, double alpha, double beta, double* out3n, int owhere);
/* MaxTorque (EffFieldGPU.cu:484-512): max_i |m_i x H_i|. */
int tm_max_torque(tm_ctx* c, const double* Heff3n, int hwhere, double* out);

/* GpuPotential::calcDivM (GpuPotential.cu:86-106): sum_d tGrad_d (Js .* m_d)
 * from the given magnetisation (NOT divided by mu0, as the reference). */
int tm_div_m(tm_ctx* c, const double* m3n, int mwhere, double* b, int bwhere);
/* GpuPotential::calcGradientField (GpuPotential.cu:109-122): H = -grad u. */
int tm_grad_field(tm_ctx* c, const double* u, int uwhere, double* H3n, int hwhere);

/* Boundary nodes (DemagField ctor, DemagField.cpp:65-89): ascending global
 * node ids, bnx of them. */
int tm_ctx_set_boundary(tm_ctx* c, int bnx, const int* bnodes);
/* DemagField::dirichletBEM (DemagField.cpp:67): dense bnx x bnx column-major
 * matrix with leading dimension ld >= bnx. */
int tm_ctx_set_bem_dense(tm_ctx* c, const double* B, int ld, int where);
/* Dense boundary-element matrix assembled on the GPU (BEMprocessing.cpp:358-411,
 * Lindholm.cpp:41-128) from the boundary triangulation: bxyz (bnx x 3
 * row-major coordinates of the boundary nodes), faces (nbel x 3 boundary-local
 * node indices), face areas, outward unit face normals (nbel x 3 row-major)
 * and the interior solid angle at every boundary node.  `formula` selects
 * the solid-angle form of the Lindholm weights (TM_BEM_REFERENCE or
 * TM_BEM_ACCURATE). */
int tm_ctx_assemble_bem(tm_ctx* c, int nbel, const double* bxyz, const int* faces, const double* face_area,
                        const double* face_normal, const double* solid_angle, int formula);
/* Copies the context's dense matrix out (column-major, ld = bnx). */
int tm_ctx_get_bem_dense(tm_ctx* c, double* B, int where);
/* DemagField::denseMVP (DemagField.cpp:145-147): g = B u on boundary vectors. */
int tm_bem_apply(tm_ctx* c, const double* ub, int uwhere, double* gb, int gwhere);
/* No counterpart in the reference: tm_bem_apply for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  y = B x for a dense bnx x bnx
 * column-major B (leading dimension ld >= bnx, ld even, B 16-byte aligned),
 * x and y device pointers, enqueued on `stream` (a hipStream_t; NULL = the
 * default stream), no context needed.  Nothing is allocated: the partial
 * products of the column splits live in a static device array of 32 MiB
 * (4,194,304 doubles) that every GPU the library runs on holds until the
 * process ends; sizes whose partials do not fit (bnx above about 92,000)
 * return TM_ERR_BAD_ARG before any device work.  Same kernels and results
 * as tm_bem_apply.  Calls that overlap on one GPU share that array and race.
 * The array's address on a GPU is looked up when tm_set_device selects it,
 * else on the first call there (call tm_set_device before a capture). */
int tm_bem_apply_noalloc(int bnx, int ld, const double* B, const double* x, double* y, void* stream);

/* DemagField::initializeSolvers (DemagField.cpp:243-282): the Neumann matrix
 * (stiffness with row nx/2 replaced by the identity) and the Dirichlet matrix
 * (stiffness with boundary rows replaced by the identity), CSR host arrays. */
int tm_ctx_set_demag_solvers(tm_ctx* c, const int* rpN, const int* colN, const double* valN, const int* rpD,
                             const int* colD, const double* valD, int precond);
/* AMG hierarchy of the context's Neumann (which = 0) or Dirichlet (1) solver. */
int tm_ctx_solver_amg_info(tm_ctx* c, int which, int* nlevels, int* sizes, int cap);
/* Solver tolerance and cap (ProgramSpecs.cpp:70 "CG tolerance"). */
int tm_ctx_set_solver_params(tm_ctx* c, double tol, int maxit);
/* DemagField::calcPotential (DemagField.cpp:163-181): u = u1 + u2 for the
 * given magnetisation (m3n == NULL uses the context magnetisation).  Solves
 * are warm-started from the previous call (AMGCLSolver.cpp:97-103). */
int tm_demag_potential(tm_ctx* c, const double* m3n, int mwhere, double* u, int uwhere, tm_eval_stats* stats);
/* DemagField::calcField (DemagField.cpp:92-94): Hdem = -grad u. */
int tm_demag_field(tm_ctx* c, const double* m3n, int mwhere, double* H3n, int hwhere, tm_eval_stats* stats);
/* Last solutions of the two solves (u1, u2), n each. */
int tm_demag_last_potentials(tm_ctx* c, double* u1, double* u2, int where);

/* One full effective-field + LLG right-hand-side evaluation with the demag
 * field recomputed (the composition TheLLG::rhs_d -> classicVersion_GPU ->
 * evaluateAllEffectiveFields -> totalEffectiveField -> ClassicLLG_dev of
 * LLGWrapper.cu:78-82, 213-228 and TheLLG.cpp:238-262, with DemagField::calcField
 * as in TheLLG::calcDemagField, TheLLG.cpp:278-283):
 *   Heff = (Hexc + Hku + Hdmi) invJs + Hext + Hdem,  dm/dtau = LLG(m, Heff).
 * Terms without data set on the context are zero.  m3n == NULL uses the
 * context magnetisation.  Heff3n may be NULL. */
int tm_eval_rhs(tm_ctx* c, const double* m3n, int mwhere, double alpha, int variant, double* dmdt3n, int dwhere,
                double* Heff3n, int hwhere, tm_eval_stats* stats);

/* Energies as densities (J/m^3, divided by the total volume) and volume
 * averages, as TheSimulation::calculateEnergies (TheSimulation.cpp:384-396)
 * and getMeanM / TheLLG::getMeanH (TheSimulation.cpp:100-102,
 * TheLLG.cpp:421-423) report them; cubic and surface terms are zero, as on
 * the reference's GPU path. */
typedef struct tm_energies {
    double demag;     /* -1/2 sum Js V Hdem.m    (DemagField.cpp:97-99) */
    double exchange;  /* -1/2 sum V Hxc.m        (EffFieldCalc.cpp:65-69) */
    double uniaxial;  /* sum V Ku (1 - (k.m)^2)  (EffFieldCalc.cpp:86-95) */
    double zeeman;    /* -sum Js V Hext.m        (TheLLG.cpp:345-348) */
    double cubic;
    double surface;
    double dmi;       /* -1/2 sum V Hdmi.m       (EffFieldCalc.cpp:172-175) */
    double total;
    double mean_m[3];
    double mean_h[3]; /* applied field average, A/m */
    double volume;    /* sum of node volumes (mesh units^3) */
} tm_energies;

/* Interfacial DMI with interface normal z (no counterpart in the reference,
 * which has bulk DMI only): D per node (J/m^2 divided by the mesh scale), null
 * to switch it off.  The field is the gradient of the lumped P1 energy
 * sum_i D_i [m_z (Q_x m_x + Q_y m_y) - m_x Q_x m_z - m_y Q_y m_z]_i with
 * Q_d = tGrad_d^T; it needs the node volumes and the GRAD_X/Y, TGRAD_X/Y
 * operators.  Adds to Heff and to the DMI energy; works beside bulk DMI. */
int tm_ctx_set_dmi_interfacial(tm_ctx* c, const double* D);
int tm_idmi_field(tm_ctx* c, double* out3n, int where);

/* Node volumes (MeshData NodeVolume) used by the energies. */
int tm_ctx_set_node_volume(tm_ctx* c, const double* V);
/* Energies of m (null = the context's m).  recompute_demag = 0 reuses the
 * last potentials (TheSimulation::calculateEnergies with calcDemag false). */
int tm_energies_eval(tm_ctx* c, const double* m3n, int mwhere, int recompute_demag, tm_energies* out);
/* "freeze demag" (ProgramSpecs.cpp:81, TheSimulation.cpp:235-236): when on,
 * tm_eval_rhs reuses the potentials of the last demag solve, refreshed by
 * tm_ctx_refresh_demag (or by any call that solves). */
int tm_ctx_set_freeze_demag(tm_ctx* c, int on);
int tm_ctx_refresh_demag(tm_ctx* c, const double* m3n, int mwhere, tm_eval_stats* stats);

/* ---- adjoints (no counterpart in the reference; gradients of trajectory
 * objectives with respect to m, applied fields, Ku and D) -------------------
 * Vector-Jacobian product of the right-hand side f(m) = LLG(m, Heff(m)) of
 * tm_eval_rhs: g_m = (df/dm)^T lam, including the transposed demag chain
 * (transposed gradient, Dirichlet and Neumann solves with the transposed
 * row-replaced matrices, B^T) unless demag is frozen.  The forward is
 * re-evaluated at m, which becomes the context's magnetisation.  Optional
 * outputs (null to skip): g_hext3 (host, 3) = gradient w.r.t. a uniform
 * applied field; g_hfield3n = w.r.t. a per-node applied field; g_ku (n) =
 * w.r.t. per-node Ku; g_d (n) = w.r.t. per-node bulk D (needs node volumes);
 * g_di (n) = w.r.t. per-node interfacial D.
 * g_m, g_hfield3n, g_ku and g_d are written where gwhere says. */
int tm_eval_rhs_vjp(tm_ctx* c, const double* m3n, int mwhere, double alpha, int variant, const double* lam3n,
                    int lwhere, double* gm3n, int gwhere, double* g_hext3, double* g_hfield3n, double* g_ku,
                    double* g_d, double* g_di, tm_eval_stats* stats);
/* g = (dHdem/dm)^T v (Hdem is linear in m). */
int tm_demag_field_vjp(tm_ctx* c, const double* v3n, int vwhere, double* g3n, int gwhere, tm_eval_stats* stats);
/* x_b = B^T y_b on the boundary (the transpose of tm_bem_apply). */
int tm_bem_apply_transposed(tm_ctx* c, const double* yb, int ywhere, double* xb, int xwhere);

/* ---- on-device time stepping (the reference integrates with CVODE on the
 * host: LLGWrapper.cu:139-228, TheLLG.cpp:578-711) --------------------------- */
typedef struct tm_rk4_stats {
    int sync_steps;          /* steps run with host-polled solves (the first, or use_graph = 0) */
    int graph_steps;         /* steps replayed from the captured graph */
    int redone_steps;        /* replays in which a solve missed its budget, redone */
    int captures;            /* graph captures */
    int budget_neumann;      /* CG iteration budgets of the graph */
    int budget_dirichlet;
    int max_iters_neumann;   /* largest CG iteration counts seen */
    int max_iters_dirichlet;
    double ms;               /* device time of the call */
} tm_rk4_stats;
/* nsteps classical RK4 steps of dtau (tau units) on the context's
 * magnetisation, renormalised after each step; demag recomputed at every
 * stage unless frozen.  use_graph = 1 captures one step in a HIP graph (CG
 * with fixed iteration budgets that stop early on the device; every replay is
 * checked and a step whose solve missed the tolerance is redone with
 * host-polled solves and larger budgets).  The field parameters in effect at
 * the call are used for all steps. */
int tm_ctx_run_rk4(tm_ctx* c, double dtau, int nsteps, double alpha, int variant, int use_graph,
                   tm_rk4_stats* stats);

typedef struct tm_ode_stats {
    int accepted, rejected;  /* steps */
    int evaluations;         /* right-hand side evaluations (6 per step + 1) */
    int sync_steps;          /* steps run with host-polled solves */
    int graph_steps;         /* steps replayed from the captured graph */
    int redone_steps;        /* replays in which a solve missed its budget, redone */
    int captures;
    int budget_neumann, budget_dirichlet;
    int max_iters_neumann, max_iters_dirichlet;
    double tau;              /* time reached (tau units, from 0 at the call) */
    double h_next;           /* proposed next step */
    double last_err;         /* weighted RMS error estimate of the last step */
    double ms;               /* device time of the call */
} tm_ode_stats;
/* Adaptive integration of the context magnetisation over tau_span (tau
 * units) with the Dormand-Prince 5(4) pair, the error measured in the
 * weighted RMS norm of CVODE with weights 1/(rtol |m_i| + atol) -- the
 * integration TheLLG::gpuODE / integrateSUNDIALS run with CVODE's Adams
 * method and rtol = atol = 1e-6 (LLGWrapper.cu:139-211, TheLLG.cpp:578-654).
 * The right-hand side is that of tm_eval_rhs (demag recomputed at every
 * stage unless frozen), parameters as set at the call.  *h_io: first step on
 * entry (<= 0: chosen from the largest torque), proposed next step on exit
 * (pass it to the next window).  At most max_steps steps (accepted plus
 * rejected); TM_ERR_MAXITER if tau_span is not reached.  use_graph = 1
 * captures one step in a HIP graph (fixed CG budgets that stop early on the
 * device; a step whose solve missed its tolerance is redone with host-polled
 * solves).  normalize = 1 renormalises m at the end, as TheSimulation does
 * after each integration window (TheSimulation.cpp:293, 404-412). */
int tm_ctx_integrate(tm_ctx* c, double tau_span, double rtol, double atol, double* h_io, int max_steps, double alpha,
                     int variant, int use_graph, int normalize, tm_ode_stats* stats);

/* Explicit Euler advance of the context magnetisation with renormalisation:
 * m <- (m + dtau dm/dtau) / |m + dtau dm/dtau|  (the renormalisation of
 * EffFieldGPU::NormalizeMag, EffFieldGPU.cu:625-663, fused with the update). */
int tm_ctx_step_euler(tm_ctx* c, const double* dmdt3n, int where, double dtau);

/* Phase timers of tm_eval_rhs / tm_demag_field, measured with HIP events on the
 * context stream (the reference's u1 / h2 / u2 / total timers,
 * DemagField.cpp:155-160).  Milliseconds summed over `count` evaluations. */
typedef struct tm_phase_times {
    double div_ms;        /* Poisson right-hand side (k_div) */
    double neumann_ms;    /* u1 solve */
    double bem_ms;        /* dense B product + scatter */
    double dirichlet_ms;  /* u2 solve */
    double fields_ms;     /* fused Heff + LLG kernel */
    long long iters_neumann;
    long long iters_dirichlet;
    int count;
} tm_phase_times;
int tm_ctx_enable_timers(tm_ctx* c, int on);
int tm_ctx_get_timers(tm_ctx* c, tm_phase_times* out, int reset);

/* ---- batched evaluation: up to 64 magnetisation samples on one mesh ------ */
/* No reference counterpart as one call: the reference evaluates one sample at
 * a time (TheLLG::rhs_d, LLGWrapper.cu:213-228) and its hysteresis driver
 * loops over field values (Hysteresis.cpp:43-169).  Samples share the
 * context's operators, boundary matrix and AMG hierarchies; arrays are
 * [3][n][S] (component, node, sample) for fields and [n][S] for potentials. */
#define TM_BATCH_MAX 64
#define TM_PREC_FP64 0  /* all-FP64 batched solves */
#define TM_PREC_MIXED 1 /* FP32 inner CG iterations, FP64 iterative refinement to the tolerance */
#define TM_PREC_MIXED_BEM32 2 /* as MIXED, and the boundary product from an FP32 copy of B: FP32
                               * products accumulated in FP32 over 512-column chunks, chunk sums
                               * added in FP64 */
#define TM_PREC_MIXED_BEMI8 3 /* as MIXED, and the boundary product from int8 digits: every row of B
                               * and sample of U scaled by a power of two and split into base-2^7
                               * digits (five for B, six for U), digit products (p + q <= 5)
                               * accumulated exactly in int32 on the int8 matrix cores, combined in
                               * FP64 (~1e-8 relative product error, from B's rounding; B's digits
                               * take 5 bytes per entry besides B) */

typedef struct tm_batch tm_batch;
typedef struct tm_batch_stats {
    int nsamples;
    int outer_neumann, outer_dirichlet;  /* refinement steps */
    int inner_neumann, inner_dirichlet;  /* inner CG iterations (all samples advance together) */
    double max_relres_neumann;           /* worst final FP64 relative residual over the samples */
    double max_relres_dirichlet;
    double max_torque[TM_BATCH_MAX];     /* per sample, A/m */
} tm_batch_stats;

/* Needs a context with operators, boundary matrix and AMG demag solvers.  The
 * batch takes the context's operators, Js, solver hierarchies, solver
 * tolerance and uniform applied field as they are at creation (create a new
 * batch after changing them; per-sample fields: tm_batch_set_zeeman); the
 * context's per-node applied field is read at each evaluation.  Batched
 * vectors are addressed with 32-bit byte offsets: 3 n S 8 bytes must stay
 * below 2 GiB (n up to 1.39 million nodes at S = 64), else
 * TM_ERR_UNSUPPORTED. */
int tm_batch_create(tm_ctx* c, int nsamples, int precision, tm_batch** out);
int tm_batch_destroy(tm_batch* b);
/* Per-sample uniform applied field (A/m), [3][S]; default = the context's
 * uniform field.  A per-node field set on the context (tm_ctx_set_zeeman with
 * H3n, the carrier of the pulse / sweep / local fields of TheLLG.cpp:320-332)
 * is added to every sample, read at each evaluation. */
int tm_batch_set_zeeman(tm_batch* b, const double* h3s);
int tm_batch_set_mag(tm_batch* b, const double* m, int where);
int tm_batch_get_mag(tm_batch* b, double* m, int where);
/* One Heff + LLG RHS evaluation of every sample (demag recomputed; solves
 * warm-started per sample).  m == NULL uses the batch magnetisation. */
int tm_batch_eval_rhs(tm_batch* b, const double* m, int mwhere, double alpha, int variant, double* dmdt, int dwhere,
                      double* heff, int hwhere, tm_batch_stats* stats);
int tm_batch_get_potentials(tm_batch* b, double* u1, double* u2, int where);
/* m <- normalize(m + dtau dm/dtau) for every sample. */
int tm_batch_step_euler(tm_batch* b, const double* dmdt, int where, double dtau);
/* One explicit Euler step of the batch magnetisation: dm/dt at the current m
 * as tm_batch_eval_rhs (into dmdt, stats as there), then m <- normalize(m +
 * dtau dm/dt), the step written by the field kernel itself into a second
 * buffer that becomes the batch's m (the same numbers as tm_batch_eval_rhs
 * followed by tm_batch_step_euler, without the separate pass over m and
 * dm/dt).  No reference counterpart: the reference's integrator (CVODE,
 * LLGWrapper.cu:139-228) updates m itself. */
int tm_batch_eval_euler(tm_batch* b, double alpha, int variant, double dtau, double* dmdt, int dwhere,
                        tm_batch_stats* stats);
/* nsteps classical RK4 steps of every sample's magnetisation (renormalised
 * after each step), each stage a full batched evaluation.  The hysteresis
 * driver of the reference steps one field value at a time through CVODE
 * (Hysteresis.cpp:43-169, TheLLG.cpp:578-711).  use_graph = 1: after a first
 * host-polled step has set the solver budgets (refinement steps and inner
 * iterations per step, per solve), one whole step -- four evaluations with
 * fixed-budget solves whose kernels return at once past convergence -- is
 * captured in a HIP graph and replayed with one host synchronisation per
 * step; a step in which a sample missed the tolerance is restored and redone
 * with host-polled solves, and the budgets grow.  The captured steps run the
 * boundary product on the fastest of the library's own FP64 GEMM kernels that
 * agreed with rocBLAS at tm_batch_create (S = 64, FP64 boundary product);
 * other batches run every step host-polled.
 * `last` (may be null): statistics of the final evaluation when it was
 * host-polled (only nsamples is set after a replayed final step); `rk` (may
 * be null): step counts, budgets (summed inner iterations per solve) and the
 * device time. */
int tm_batch_run_rk4(tm_batch* b, double dtau, int nsteps, double alpha, int variant, int use_graph,
                     tm_batch_stats* last, tm_rk4_stats* rk);
/* The boundary GEMM chosen by timing every candidate for this shape at
 * tm_batch_create.  *own_kernel = 0: a rocBLAS GEMM, *solution its solution
 * index (0 = library default; rocBLAS indices may be negative).
 * *own_kernel = 1: the library's own FP64 MFMA kernel (only chosen when it
 * agrees with rocBLAS to 1e-12), *solution its tile shape, run with `ksplit`
 * partial products over K.  Also the number of candidates timed and the
 * timings of the library default and of the choice.  Any output may be NULL. */
int tm_batch_gemm_info(tm_batch* b, int* solution, int* own_kernel, int* candidates, double* ms_default,
                       double* ms_tuned, int* ksplit);
/* In-pipeline timing of the boundary GEMM (no reference counterpart): on = 1
 * starts recording a HIP event pair around the GEMM launch of each following
 * evaluation on the batch's stream (no host wait; the first 256 are kept);
 * on = 0 stops, waits for the last recorded event and returns the mean
 * duration in ms of the launches recorded since the last start and their
 * number (either output may be NULL). */
int tm_batch_gemm_live(tm_batch* b, int on, double* ms_mean, int* count);
/* y[j][s] = sum_k B[j][k] x[k][s] for every sample s: the batch's boundary
 * product (the GEMM tm_batch_eval_rhs runs, in the batch's precision) on
 * caller values; x, y: b x S, boundary-node major, the sample index fastest.
 * The batched form of DemagField::denseMVP (DemagField.cpp:145-147). */
int tm_batch_bem_apply(tm_batch* b, const double* x, int xwhere, double* y, int ywhere);
/* Kernel timing of a batch (no reference counterpart): the hot kernels of an
 * evaluation -- the divergence, the fused field + LLG kernel, the tuned
 * boundary GEMM (FP64 forms), and of the Neumann solve the level-0 kernels
 * of one inner CG iteration (SpMM, update, the pre-smoothing restriction,
 * the prolongation, the post-smoother) and the FP64 outer residual -- each
 * launched `reps` times alone on the batch's
 * stream (all samples active) and timed with HIP events.  For each: its
 * name, the mean ms per launch and its algorithmic bytes per launch (slot
 * offsets and weights read once, every vector read or written once, the
 * gathered rows counted once).  Leaves the batch's solver state undefined
 * until the next evaluation; *count receives the number written (<= cap). */
typedef struct tm_batch_kernel_time {
    char name[48];
    double ms;
    double bytes;
} tm_batch_kernel_time;
int tm_batch_time_kernels(tm_batch* b, int reps, tm_batch_kernel_time* out, int cap, int* count);
int tm_batch_enable_timers(tm_batch* b, int on);
int tm_batch_get_timers(tm_batch* b, tm_phase_times* out, int reset);

/* ---- host-side setup helpers (no GPU needed) ---------------------------- */
/* Kuhn-split cuboid: (nx+1)(ny+1)(nz+1) nodes, 6 nx ny nz tets, spacing h,
 * interior nodes jittered uniformly by +-jitter*h per coordinate (seeded). */
int tm_mesh_kuhn_cuboid_size(int nx, int ny, int nz, int* n_nodes, int* n_tets);
int tm_mesh_kuhn_cuboid(int nx, int ny, int nz, double h, double jitter, uint64_t seed, double* xyz, int* fel);
/* Node renumbering for gather locality (no counterpart in the reference; a
 * host applies it once to the points and cells it has read, after
 * MeshReader::read, io/MeshReaderInterface.cpp:68-82, before any operator is
 * built): perm[k] = the
 * old number of new node k, nodes ordered along a Z-order curve of their
 * coordinates (xyz: n x 3 row-major), the axis of smallest extent fastest.
 * The batched gather kernels take 7 % less time per evaluation on the
 * 282,897-node cuboid than with its row-by-row numbering. */
int tm_mesh_locality_order(int n, const double* xyz, int* perm);
/* P1 operator pattern (FEMprocessing.cpp:228-286 touch exactly the node pairs
 * sharing a tet) and values: stiffness, grad_d, tGrad_d, node volumes. */
int tm_fem_pattern(int n, int ntet, const int* fel, int* rowptr);
int tm_fem_assemble(int n, int ntet, const double* xyz, const int* fel, const int* rowptr, int* col, double* K,
                    double* gx, double* gy, double* gz, double* tgx, double* tgy, double* tgz, double* nodevol,
                    double* elvol);
/* Boundary extraction (Boundary.cpp:47-224). */
int tm_boundary_faces(int ntet, const int* fel, int* faces, int* face_tet, int* nbel_out);
int tm_boundary_props(int n, const double* xyz, const int* fel, int nbel, const int* faces, const int* face_tet,
                      int* bnodes, int* bindex, double* face_area, double* face_normal, double* node_area,
                      double* node_normal, int* bnx_out);
/* Interior solid angles at boundary nodes (BEMprocessing.cpp:358-372). */
int tm_boundary_solid_angles(int n, int ntet, const double* xyz, const int* fel, const int* bindex, int bnx,
                             double* omega);
/* Seeded inputs. */
int tm_init_random_unit(int n, uint64_t seed, double* m3n);
int tm_uniform(uint64_t seed, int count, double* out);

/* Library version string. */
const char* tm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TETMAG_HIP_H_ */
