/*
 * ifem_gpu.h -- C ABI of the MI355X (gfx950) isogeometric assembly and
 * linear-solve library (libifem_gpu.so).
 *
 * The entry points replace, for one tensor-product spline patch, the
 * reference's hot path
 *   ASMbase::integrate -> Integrand::evalInt -> GlobalIntegral::assemble
 *   -> SAM::assembleSystem -> SystemMatrix (+ LinSolParams) -> solve
 * Each declaration names the reference interface it stands in for.
 *
 * Conventions: plain pointers and sizes only; host arrays passed in are
 * copied, never retained.  Every call returns IFEM_GPU_OK (0) or an error
 * code, and ifem_gpu_last_error() then returns a thread-local message.
 * Nothing throws across this boundary.  One context per host thread; all GPU
 * work of a context runs on its stream.
 */
#ifndef IFEM_GPU_H
#define IFEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFEM_GPU_ABI_VERSION 2

enum ifem_gpu_status {
  IFEM_GPU_OK = 0,
  IFEM_GPU_EINVAL = 1,       /* bad argument or call order */
  IFEM_GPU_EUNSUPPORTED = 2, /* valid input the GPU path does not handle */
  IFEM_GPU_EDEVICE = 3,      /* HIP runtime error or no device */
  IFEM_GPU_ENOCONV = 4,      /* solver stopped without meeting tolerances */
  IFEM_GPU_ENOMEM = 5
};

typedef struct ifem_gpu_ctx ifem_gpu_ctx;

/* Thread-local text of the last error (empty string if none). */
const char* ifem_gpu_last_error(void);
int ifem_gpu_abi_version(void);

/* Context on HIP device `device`.  stream: a hipStream_t to run on, or NULL
 * for a new default-priority stream owned by the context.
 * Stands in for LinAlgInit::Init / AlgEqSystem construction
 * (src/LinAlg/LinAlgInit.C:25-43, src/ASM/AlgEqSystem.C:33). */
int ifem_gpu_ctx_create(int device, void* stream, ifem_gpu_ctx** ctx);
int ifem_gpu_ctx_destroy(ifem_gpu_ctx* ctx);
/* The hipStream_t the context runs on. */
void* ifem_gpu_ctx_stream(ifem_gpu_ctx* ctx);
int ifem_gpu_synchronize(ifem_gpu_ctx* ctx);

/* ---- patch geometry --------------------------------------------------------
 * Non-rational trivariate B-spline volume, GoTools conventions:
 * order[d] = p+1, ncoef[d] control points, knots[d] has ncoef+order entries
 * (open, simple interior knots), coefs xyz-interleaved with node (i,j,k) at
 * ((k*n2+j)*n1+i)*3.  Replaces ASMs3D::read/generateFEMTopology for the data
 * the integration loop reads (src/ASM/ASMs3D.C:117-160, 433-519; element
 * coordinates src/ASM/ASMs3D.C:1640-1678).  rational must be 0. */
int ifem_gpu_patch_create(ifem_gpu_ctx* ctx, const int order[3], const int ncoef[3],
                          const double* const knots[3], const double* coefs,
                          int rational, int* patch_id);

/* ---- SAM -------------------------------------------------------------------
 * Equation numbers per DOF, node-major / component-minor (DOF = node*ncomp+c),
 * 1-based; 0 = eliminated (homogeneous Dirichlet, msc = 0).  Free DOFs must be
 * numbered in increasing DOF order, as SAM::initSystemEquations does
 * (src/LinAlg/SAM.C:319-323); constrained (negative) equations need
 * ifem_gpu_sam_bind_constraints.  Replaces SAM::initSystemEquations/getElmEqns
 * (src/LinAlg/SAM.C:225-336, 655-696) for the patch. */
int ifem_gpu_sam_bind(ifem_gpu_ctx* ctx, int patch_id, int ncomp, int ndof,
                      const int* meqn, int neq);

/* Same with constraint equations in SAM's layout (src/LinAlg/SAM.h:330-350,
 * src/LinAlg/SAM.C:260-308): meqn[i] = -k marks DOF i as the slave of
 * constraint k (1-based); its entries are mpmceq[k-1] .. mpmceq[k]-1 (1-based)
 * of mmceq (DOF numbers, the slave first) and ttcc (the constant c0 first).
 * Constraints without masters (mpmceq[k]-mpmceq[k-1] == 1) are prescribed
 * values: the assembly moves the eliminated columns to the right-hand side,
 * b_i -= K_ij c0_j, as SparseMatrix::assemble does
 * (src/LinAlg/SparseMatrix.C:614-688).  A constraint with masters,
 * u_s = c0 + sum_m c_m u_m (periodicity, couplings), is eliminated as
 * assemSparse/SAM::assembleRHS do (src/LinAlg/SparseMatrix.C:612-688,
 * src/LinAlg/SAM.C:605-617): the GPU assembles an extended system with the
 * slaves as extra unknowns and condenses it to K = T^T K_x T,
 * b = T^T (b_x - K_x t0); the masters must be free DOFs (chains resolved, as
 * SAMpatch::initConstraintEqs leaves them, else IFEM_GPU_EINVAL), and the
 * pattern then holds the master-master couplings of SAM::getDofCouplings
 * (src/LinAlg/SAM.C:377-427).  ifem_gpu_expand_solution and
 * ifem_gpu_integrate_norms use u_s = c0 + sum c_m u_m at the slaves
 * (SAM::expandSolution, src/LinAlg/SAM.C:768-810).  The multigrid and learned
 * preconditioners refuse systems with master constraints
 * (IFEM_GPU_EUNSUPPORTED); a traction needs the FP64 system assembled
 * first. */
int ifem_gpu_sam_bind_constraints(ifem_gpu_ctx* ctx, int patch_id, int ncomp, int ndof,
                                  const int* meqn, int neq, int nceq, const int* mpmceq,
                                  const int* mmceq, const double* ttcc);

/* Several patches in one system, sharing the nodes of their common faces
 * (ASMs3D::connectPatch / SAM's global node numbering MLGN,
 * src/ASM/ASMs3D.C:702-815, src/ASM/SAMpatch.C:21-161): mlgn[k] maps local
 * node i of patch patch_ids[k] (GoTools order) to its global node (1-based,
 * distinct within the patch); meqn holds ncomp equation numbers per global
 * node, as in ifem_gpu_sam_bind; nceq > 0 adds prescribed values as in
 * ifem_gpu_sam_bind_constraints (constraints with masters:
 * IFEM_GPU_EUNSUPPORTED).  ifem_gpu_init_assembly then builds the
 * union pattern on the device, and ifem_gpu_assemble adds the patches in
 * order into the shared rows (the same integrand on every patch).  The
 * multigrid and learned preconditioners need a single patch. */
int ifem_gpu_sam_bind_patches(ifem_gpu_ctx* ctx, int npatch, const int* patch_ids,
                              const int* const* mlgn, int nnod, int ncomp, const int* meqn,
                              int neq, int nceq, const int* mpmceq, const int* mmceq,
                              const double* ttcc);

/* ---- symbolic assembly -----------------------------------------------------
 * Builds the CSR sparsity pattern (symmetric, full storage, sorted columns)
 * on the device.  Replaces SystemMatrix::initAssembly -> SparseMatrix::
 * preAssemble -> SAM::getDofCouplings -> optimiseCols
 * (src/LinAlg/SparseMatrix.C:717-762, src/LinAlg/SAM.C:377-427,
 * src/LinAlg/SparseMatrix.C:984-1023).  nnz may be NULL. */
int ifem_gpu_init_assembly(ifem_gpu_ctx* ctx, int64_t* nnz);

/* ---- integrands ------------------------------------------------------------
 * Recognised integrand kinds of the GPU path (the physics of
 * Integrand::evalInt, src/ASM/Integrand.h:206,344):
 *  IFEM_GPU_POISSON:    EK += kappa(X) detJxW dNdX dNdX^T, EV += f(X) detJxW N
 *                       (Apps/Utils/EqualOrderOperators.C:158-211,243-255)
 *  IFEM_GPU_ELASTICITY: EK += detJxW B^T C B (isotropic, E, nu),
 *                       EV += detJxW body N
 * Fields: ZERO (0), CONST scale, SINPROD scale*sin(pi x)sin(pi y)sin(pi z),
 * FOURIER mean + scale*sum a_m cos(k_m.x+phi_m), EXP_FOURIER exp(FOURIER).
 * modes: nmodes x {kx, ky, kz, phi, a}.  kappa of type ZERO means 1. */
enum { IFEM_GPU_FN_ZERO = 0, IFEM_GPU_FN_CONST = 1, IFEM_GPU_FN_SINPROD = 2,
       IFEM_GPU_FN_FOURIER = 3, IFEM_GPU_FN_EXP_FOURIER = 4 };
enum { IFEM_GPU_POISSON = 1, IFEM_GPU_ELASTICITY = 2 };

typedef struct {
  int type;
  int nmodes;
  double scale;
  double mean;
  const double* modes;
} ifem_gpu_function;

typedef struct {
  int kind;
  double E, nu;
  ifem_gpu_function kappa;
  ifem_gpu_function source;
  double body[3];
} ifem_gpu_integrand;

/* Precision of assembled values and of the solver's inner iteration. */
enum { IFEM_GPU_FP64 = 64, IFEM_GPU_FP32 = 32 };

/* ---- numeric assembly ------------------------------------------------------
 * Volume integration over all elements of the bound patch with ng[d] Gauss
 * points per direction, scattered into K (CSR values) and b.  Zeroes K and b
 * first (AlgEqSystem::initialize, src/ASM/AlgEqSystem.C:184-212).  Replaces
 * ASMs3D::integrate (src/ASM/ASMs3D.C:2071-2335) + AlgEqSystem::assemble
 * (src/ASM/AlgEqSystem.C:215-303) + SAM::assembleSystem
 * (src/LinAlg/SAM.C:430-467, 518-543).  precision: IFEM_GPU_FP64 writes the
 * FP64 system; IFEM_GPU_FP32 writes the FP32 copy of K used by the mixed
 * precision solver (b stays FP64 and is assembled in FP64). */
int ifem_gpu_assemble(ifem_gpu_ctx* ctx, const ifem_gpu_integrand* itg,
                      const int ng[3], int precision);

/* No counterpart in the reference (part of this library's autotuner): times
 * the line-assembly kernel of the bound system with each element-chunk
 * variant (tile shape of the (p, nen) kernel: LDS per block against blocks
 * per CU): 0 full chunks, 1 half chunks, 2 full chunks with the Gauss points
 * in two slabs (p <= 2 with at most 3 points per direction: three blocks per
 * CU), keeps the fastest for this (p, ng, ncomp, precision) in the context
 * and leaves the system assembled as ifem_gpu_assemble does (apply Neumann
 * loads afterwards).  *half (may be NULL) receives the choice. */
int ifem_gpu_autotune_assembly(ifem_gpu_ctx* ctx, const ifem_gpu_integrand* itg,
                               const int ng[3], int precision, int* half);

/* No counterpart in the reference: fix the element-chunk variant (as above;
 * -1 restores the built-in default) of the line kernel for (p, ng, ncomp,
 * precision) in this context.  EUNSUPPORTED for a variant that does not
 * exist for that degree and rule. */
int ifem_gpu_set_assembly_variant(ifem_gpu_ctx* ctx, int p, int ng, int ncomp, int precision,
                                  int variant);

/* Neumann traction t (constant vector, ncomp components) on patch face
 * face = 1..6 (IFEM numbering: 1 u=0, 2 u=1, 3 v=0, 4 v=1, 5 w=0, 6 w=1),
 * ng Gauss points per tangent direction, added into b.  Replaces
 * ASMs3D::integrate(Integrand&, int lIndex, ...) (src/ASM/ASMs3D.C:2547-2812). */
int ifem_gpu_assemble_neumann(ifem_gpu_ctx* ctx, int face, const double* traction, int ng);
/* Same on patch `index` (0-based, in the order given to
 * ifem_gpu_sam_bind_patches; 0 for a single-patch system). */
int ifem_gpu_assemble_neumann_patch(ifem_gpu_ctx* ctx, int index, int face,
                                    const double* traction, int ng);

/* ---- norms ------------------------------------------------------------------
 * Norms of the current solution u_h (after ifem_gpu_solve; eliminated DOFs
 * are 0), integrated with ng Gauss points per direction as NormBase /
 * ElmNorm / GlbNorm do with the SQRT final operation (src/ASM/IntegrandBase.h:
 * 378-530, src/ASM/GlbNorm.C:33-79): norms[0] = ||u_h||_L2, norms[1] =
 * ||u_h||_a (Poisson: sqrt(int kappa |grad u_h|^2); elasticity: sqrt(int
 * eps:C:eps)); with an exact scalar solution (Poisson only, NULL for none):
 * norms[2] = ||u||_L2, norms[3] = ||u||_a, norms[4] = ||u - u_h||_L2,
 * norms[5] = ||u - u_h||_a (else 0). */
int ifem_gpu_integrate_norms(ifem_gpu_ctx* ctx, const ifem_gpu_integrand* itg, const int ng[3],
                             const ifem_gpu_function* exact, double norms[6]);

/* ---- linear solve ----------------------------------------------------------
 * LinSolParams keys (src/LinAlg/LinSolParams.C:68-77,158-195) as fields:
 * type "cg" (CG only), pc "jacobi" or a multigrid ("gamg", "ml": PETSc's
 * algebraic multigrids, src/LinAlg/LinSolParams.C:60-155, here a spline
 * multigrid: nested knot vectors, knot-insertion prolongation, Galerkin
 * coarse operators, Chebyshev-Jacobi smoothing of mg_smooth steps
 * (multigrid_no_smooth), dense coarsest solve; applied in flexible CG, FP32
 * inside), rtol, atol, dtol, maxits.
 * Extensions: precision (FP64: plain FP64 PCG; FP32: FP32 inner PCG inside
 * FP64 iterative refinement), inner_rtol (FP32 inner residual reduction),
 * check_every (iterations between host convergence checks).
 * Convergence: ||r||_2 <= max(rtol*||b||_2, atol), confirmed on the true FP64
 * residual b - A x.  Replaces SystemMatrix::solve with PETSc KSPCG+PCJACOBI /
 * dune CGSolver+SeqJac (src/LinAlg/PETScMatrix.C:778-832,980-990;
 * src/LinAlg/ISTLMatrix.C:169-196). */
enum { IFEM_GPU_PC_JACOBI = 0, IFEM_GPU_PC_MG = 1 };
typedef struct {
  double rtol, atol, dtol;
  int maxits;
  int precision;
  double inner_rtol;
  int check_every;
  int verbosity;
  int pc;          /* IFEM_GPU_PC_JACOBI, or IFEM_GPU_PC_MG: spline multigrid */
  int mg_smooth;   /* Chebyshev steps per pre/post smoothing (multigrid) */
} ifem_gpu_solver_params;

typedef struct {
  int iterations;        /* total inner (Krylov) iterations */
  int outer_iterations;  /* refinement steps (1 for plain FP64) */
  int converged;
  double rel_residual;   /* true ||b-Ax||_2/||b||_2 in FP64 */
  double solve_ms;       /* device time of the solve on the ctx stream */
  /* Preconditioners other than Jacobi (multigrid, learned) run in flexible
   * CG.  pc_iterations: iterations that applied that preconditioner;
   * fallback: 1 if flexible CG broke down (<r, z> <= 0) and the solve went
   * on with Jacobi-PCG from the current iterate (iterations - pc_iterations
   * of them); pc_rejected: iterations in which the learned correction was
   * dropped for that one step because <r, M0 r + correction> fell below
   * 0.1 <r, M0 r> (the step then used M0 r alone). */
  int pc_iterations;
  int fallback;
  int pc_rejected;
} ifem_gpu_stats;

void ifem_gpu_solver_defaults(ifem_gpu_solver_params* p);
/* Solves K x = b; the solution stays on the device (ifem_gpu_get_solution). */
int ifem_gpu_solve(ifem_gpu_ctx* ctx, const ifem_gpu_solver_params* params,
                   ifem_gpu_stats* stats);

/* ---- data access -----------------------------------------------------------
 * Sizes of the bound system; any pointer may be NULL. */
int ifem_gpu_system_info(ifem_gpu_ctx* ctx, int* neq, int64_t* nnz, int* ncomp);
/* Copies of the device CSR (0-based) and RHS to host arrays (NULL skips). */
int ifem_gpu_get_system(ifem_gpu_ctx* ctx, int64_t* rowptr, int32_t* colind,
                        double* values, double* rhs);
/* FP32 copy of K values (after an IFEM_GPU_FP32 assembly). */
int ifem_gpu_get_values_fp32(ifem_gpu_ctx* ctx, float* values);
/* Solution in equation order (neq). */
int ifem_gpu_get_solution(ifem_gpu_ctx* ctx, double* x);
/* Solution expanded to DOF order, eliminated DOFs = 0 or their prescribed
 * value (SAM::expandSolution with scaleSD = 1, src/LinAlg/SAM.C:768-810). */
int ifem_gpu_expand_solution(ifem_gpu_ctx* ctx, double* xdof);
/* y = K x with host vectors (SparseMatrix::multiply,
 * src/LinAlg/SparseMatrix.C:578-605). */
int ifem_gpu_spmv(ifem_gpu_ctx* ctx, const double* x, double* y);
/* No counterpart in the reference: ifem_gpu_spmv for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same product as ifem_gpu_spmv with
 * host vectors, staged through a fixed device array of the library (64 MiB
 * per GPU, held until the process ends) instead of buffers allocated per
 * call; no hipMalloc/hipFree or allocating library call.  Systems above
 * 4,194,304 equations are refused (IFEM_GPU_EINVAL).  All calls on one GPU
 * share that array: calls that overlap on one GPU (other threads or
 * contexts) race. */
int ifem_gpu_spmv_noalloc(ifem_gpu_ctx* ctx, const double* x, double* y);
/* y = K x on device vectors (neq each, resident in HBM): precision
 * IFEM_GPU_FP64 uses the FP64 values with double x, y; IFEM_GPU_FP32 the FP32
 * values with float x, y.  Runs on the ctx stream, does not synchronize. */
int ifem_gpu_spmv_device(ifem_gpu_ctx* ctx, int precision, const void* x, void* y);
/* Overwrite the device RHS with a host vector (neq). */
int ifem_gpu_set_rhs(ifem_gpu_ctx* ctx, const double* b);

/* Device pointers of the resident system for callers that keep data in HBM
 * (rowptr int64, colind int32, values fp64, rhs fp64, solution fp64). */
int ifem_gpu_device_pointers(ifem_gpu_ctx* ctx, void** rowptr, void** colind,
                             void** values, void** rhs, void** solution);

/* Device time (ms, HIP events on the ctx stream) of the last numeric
 * assembly kernel and of the last solve. */
int ifem_gpu_last_timings(ifem_gpu_ctx* ctx, double* assemble_ms, double* solve_ms);
/* The two parts of the last numeric assembly (device ms): the Gauss-point
 * geometry pass (Jacobian, detJxW, material and source per element and
 * Gauss point) and the line kernel (contraction, scatter, RHS). */
int ifem_gpu_last_assembly_parts(ifem_gpu_ctx* ctx, double* geometry_ms, double* lines_ms);

/* ---- learned preconditioner -------------------------------------------------
 * No counterpart in the reference (an addition of this library): a 3D
 * convolutional U-net on the lattice of free DOFs maps a residual to an
 * error correction; it is trained on the device by the library's own forward,
 * backward and Adam kernels and applied inside flexible CG
 * (z = D^-1 r + scale * |r|_rms * U(r/|r|_rms, aux)).
 * Tensors are NCDHW float on the device; one network per context. */
typedef struct {
  int levels;          /* 1..4; grid sizes must divide by 2^(levels-1) */
  int channels[4];     /* channels per level */
  int in_channels;     /* residual components + auxiliary channels */
  int out_channels;    /* residual components */
  uint64_t seed;       /* He-normal initialisation seed */
  int activation;      /* 0: ReLU after every hidden convolution, 1: none (linear) */
} ifem_gpu_unet_config;

int ifem_gpu_unet_create(ifem_gpu_ctx* ctx, const ifem_gpu_unet_config* cfg, int64_t* nparams);
int ifem_gpu_unet_get_params(ifem_gpu_ctx* ctx, float* params);
int ifem_gpu_unet_set_params(ifem_gpu_ctx* ctx, const float* params);
/* y = U(x), N samples of D x H x W (device pointers). */
int ifem_gpu_unet_forward(ifem_gpu_ctx* ctx, const float* x, float* y, int N, int D, int H, int W);
/* Loss 0.5*sum_n w_n*sum((U(x_n)-t_n)^2)/N (w: N per-sample weights on the
 * device, NULL for all ones) and its gradient w.r.t. the parameters (copied
 * to grad if not NULL). */
int ifem_gpu_unet_gradient(ifem_gpu_ctx* ctx, const float* x, const float* t, const float* w,
                           int N, int D, int H, int W, float* grad, float* loss);
/* One Adam step on that loss (loss may be NULL to skip the read-back). */
int ifem_gpu_unet_train_step(ifem_gpu_ctx* ctx, const float* x, const float* t, const float* w,
                             int N, int D, int H, int W, float lr, float* loss);

/* Gradient of sum(dy * U(x)) w.r.t. the parameters for a caller-supplied
 * output gradient dy (device, shape of U(x)), e.g. the energy-norm loss
 * 0.5 z^T K z - z^T r of the preconditioned residual z, whose gradient needs
 * K (ifem_gpu_spmv_device).  Runs the forward pass on x first; grad (host,
 * may be NULL) receives the gradient; lr > 0 applies one Adam step. */
int ifem_gpu_unet_backward(ifem_gpu_ctx* ctx, const float* x, const float* dy, int N, int D, int H,
                           int W, float lr, float* grad);

/* Add the network's correction to the preconditioner of the next solves
 * (enable != 0) or remove it (enable == 0): flexible CG with
 * z = M0 r + scale * |r|_rms * U(r / |r|_rms, aux), M0 = D^-1 (pc Jacobi) or
 * one spline-multigrid V-cycle (pc = IFEM_GPU_PC_MG); an iteration whose
 * correction would make <r, z> < 0.1 <r, M0 r> uses M0 r alone (counted in
 * pc_rejected).  The nodes with free DOFs must form a box of
 * the control-point grid (all ncomp components free inside, none outside);
 * the box (node i fastest) is placed at the origin of a D x H x W lattice
 * with D, H, W >= the box extents in k, j, i (zero padding beyond it).
 * aux: naux auxiliary fields (D*H*W each, device pointer, copied).
 * features 0: network input [r^, aux] (ncomp + naux channels);
 * features 1: [r^, r^*aux_j per component] (ncomp*(1+naux) channels), which
 * keeps a linear network linear in the residual. */
int ifem_gpu_set_learned_preconditioner(ifem_gpu_ctx* ctx, int enable, int D, int H, int W,
                                        const float* aux, int naux, int features, double scale);

/* ---- multigrid inspection -----------------------------------------------
 * No counterpart in the reference (PETSc keeps its multigrid internal); for
 * tests and tuning.  ifem_gpu_mg_setup builds the spline multigrid hierarchy
 * of the bound system from the current FP64 matrix (ifem_gpu_solve with pc =
 * IFEM_GPU_PC_MG does this itself): level l+1 drops every second distinct
 * interior knot of level l per direction.  ifem_gpu_mg_level: control points
 * per direction, free DOFs and lambda_max(D^-1 A) of level l (0 on the
 * coarsest); with S non-NULL (levels >= 1) the Galerkin operator P^T A P in
 * stencil form, S[((node*noff + off)*nc + ci)*nc + cj], node = i + n0*(j +
 * n1*k), noff = (2p+1)^3, off = ((e3+p)*(2p+1) + e2+p)*(2p+1) + e1+p for the
 * neighbour (i+e1, j+e2, k+e3).  ifem_gpu_mg_apply: z = M^-1 r, one V-cycle, host
 * vectors in equation order. */
int ifem_gpu_mg_setup(ifem_gpu_ctx* ctx, int smooth, int* nlevels, double* setup_ms);
int ifem_gpu_mg_level(ifem_gpu_ctx* ctx, int level, int n[3], int* nfree, double* lmax,
                      double* S);
int ifem_gpu_mg_apply(ifem_gpu_ctx* ctx, const double* r, double* z);

/* ---- batch of independent patches ------------------------------------------
 * Many scalar (Poisson / diffusion) patches, each with its own SAM numbering
 * and integrand, assembled by one persistent launch into one block-diagonal
 * system and solved together with one Jacobi-PCG per patch (own scalars and
 * convergence test; the iteration loop is replayed from a captured hipGraph).
 * Per patch, assembly and solve are those of ifem_gpu_assemble and
 * ifem_gpu_solve; the batch stands in for the per-patch loop of
 * SIMbase::assembleSystem (src/SIM/SIMbase.C:1229-1400) over independent
 * models, each followed by SystemMatrix::solve
 * (src/LinAlg/PETScMatrix.C:778-832).  Patches need equal degrees p = 1..3
 * and ng = p+1 Gauss points per direction. */
int ifem_gpu_batch_clear(ifem_gpu_ctx* ctx);
/* Adds patch patch_id (from ifem_gpu_patch_create) with equation numbers
 * meqn per node (1-based, 0 = eliminated; same rules as ifem_gpu_sam_bind
 * with ncomp = 1) and integrand itg (kind IFEM_GPU_POISSON; copied).
 * *index receives the patch's position in the batch. */
int ifem_gpu_batch_add(ifem_gpu_ctx* ctx, int patch_id, const int* meqn, int neq,
                       const ifem_gpu_integrand* itg, int ng, int* index);
/* Basis tables, global sparsity pattern (SystemMatrix::initAssembly per
 * patch, src/LinAlg/SparseMatrix.C:717-762), work list and solver buffers.
 * neq, nnz (totals) may be NULL. */
int ifem_gpu_batch_init(ifem_gpu_ctx* ctx, int* neq, int64_t* nnz);
/* One persistent launch over all (patch, node line) work items, largest
 * first: IFEM_GPU_FP64 writes K and b, IFEM_GPU_FP32 the FP32 copy of K. */
int ifem_gpu_batch_assemble(ifem_gpu_ctx* ctx, int precision);
/* Tunes, for each degree present (nen = (p+1)^3 element nodes), the element
 * chunk of the assembly (full or half) and the SpMV lanes per row
 * (16/32/64) by timing the batch, then re-assembles with the choice.
 * table (may be NULL) receives 4 x {half, lanes} indexed by p. */
int ifem_gpu_batch_autotune(ifem_gpu_ctx* ctx, int precision, int* table);
/* Solves every patch to max(rtol*||b_q||, atol) on its true FP64 residual;
 * precision IFEM_GPU_FP32 runs FP32 inner CG inside FP64 refinement.
 * stats: iterations = max over patches, outer_iterations = refinement
 * passes, converged = all converged, rel_residual = max over patches. */
int ifem_gpu_batch_solve(ifem_gpu_ctx* ctx, const ifem_gpu_solver_params* params,
                         ifem_gpu_stats* stats);
/* Per-patch results of the last batch solve (arrays of the batch size, NULL
 * skips). */
int ifem_gpu_batch_results(ifem_gpu_ctx* ctx, int* iterations, int* converged,
                           double* rel_residual);
/* Size of patch `index`'s block and its first global equation. */
int ifem_gpu_batch_patch_info(ifem_gpu_ctx* ctx, int index, int* neq, int64_t* nnz, int* eq0);
/* Patch `index`'s block in local numbering: CSR (0-based, neq+1 / nnz),
 * FP64 values, FP32 values, RHS; NULL skips. */
int ifem_gpu_batch_get_system(ifem_gpu_ctx* ctx, int index, int64_t* rowptr, int32_t* colind,
                              double* values, float* values32, double* rhs);
/* Patch `index`'s solution in its equation order. */
int ifem_gpu_batch_get_solution(ifem_gpu_ctx* ctx, int index, double* x);
/* Device times (ms) of the last batch init, assembly and solve. */
int ifem_gpu_batch_timings(ifem_gpu_ctx* ctx, double* symbolic_ms, double* assemble_ms,
                           double* solve_ms);

#ifdef __cplusplus
}
#endif
#endif
