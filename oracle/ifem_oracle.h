/*
 * ifem_oracle.h -- CPU restatement of IFEM's 3D isogeometric assembly and
 * Jacobi-PCG hot path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.
 * The product (libifem_gpu.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Arrays are 0-based; equation numbers follow SAM (1-based, 0 = eliminated).
 */
#ifndef IFEM_ORACLE_H
#define IFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar field descriptors shared with the GPU path (same numbering). */
enum { ORC_FN_ZERO = 0, ORC_FN_CONST = 1, ORC_FN_SINPROD = 2,
       ORC_FN_FOURIER = 3, ORC_FN_EXP_FOURIER = 4 };

typedef struct {
  int type;            /* ORC_FN_* */
  int nmodes;          /* FOURIER kinds: number of modes */
  double scale;        /* CONST: value; SINPROD: amplitude; FOURIER: sigma */
  double mean;         /* FOURIER kinds: additive mean */
  const double* modes; /* nmodes x {kx, ky, kz, phase, amp} */
} orc_function;

enum { ORC_POISSON = 1, ORC_ELASTICITY = 2 };

typedef struct {
  int kind;            /* ORC_POISSON (1 comp) or ORC_ELASTICITY (3 comps) */
  double E, nu;        /* isotropic Hooke material (elasticity) */
  orc_function kappa;  /* Poisson conductivity (ZERO type means kappa = 1) */
  orc_function source; /* Poisson source f */
  double body[3];      /* elasticity body force per unit volume */
} orc_integrand;

typedef struct {
  int p[3];              /* polynomial degree per direction (GoTools order-1) */
  int n[3];              /* control points per direction */
  const double* knots[3];/* n+p+1 knots per direction, open, simple interior */
  const double* coefs;   /* control points, node (i,j,k) at ((k*n2+j)*n1+i)*3 */
} orc_patch;

/* a1: n-point Gauss-Legendre rule on [-1,1] (n <= 10). */
int orc_gauss(int n, double* xi, double* w);
/* a2: Gauss point count rule of ASMbase::getNoGaussPt. */
int orc_num_gauss(int order, int nGauss, int neumann);
/* a3: Gauss point parameters of span [u0,u1]. */
double orc_gauss_param(double u0, double u1, double xi);
/* a4: knot span and B-spline values/first derivatives at u.
 * Returns the span index s (u in [t_s, t_{s+1})); the nonzero functions are
 * s-p..s.  N and dN hold p+1 values each. */
int orc_find_span(int p, int n, const double* t, double u);
void orc_basis_funs(int p, const double* t, int s, double u, double* N, double* dN);

/* a5: element nodes, local node a = a1 + (p1+1)*(a2 + (p2+1)*a3). */
int orc_num_elements(const orc_patch* P, int nel[3]);
void orc_element_nodes(const orc_patch* P, int e1, int e2, int e3, int* mnpc);

/* a8: utl::Jacobian.  X is 3 x nen (row-major X[i*nen+b]), dNdu nen x 3
 * (dNdu[b*3+k]); outputs J (3x3 inverse on return, row-major), dNdX nen x 3.
 * Returns detJ (0 when |detJ| <= 1e-12). */
double orc_jacobian(int nen, const double* X, const double* dNdu,
                    double* Jinv, double* dNdX);

/* a9: element integrals of one element.  EK is nedof x nedof row-major,
 * EV nedof; nedof = ncomp*nen, DOF index = ncomp*a + i. */
int orc_element_matrices(const orc_patch* P, const orc_integrand* itg,
                         const int ng[3], int e1, int e2, int e3,
                         double* EK, double* EV);

/* a17: equation numbers from status codes (msc 1 = free, 0 = fixed).
 * meqn[dof] = 1-based equation or 0.  Returns neq. */
int orc_sam_meqn(int ndof, const int* msc, int* meqn);
/* Homogeneous Dirichlet on patch faces: fixmask[face 0..5] bit c fixes comp c.
 * Faces: 0 u=0, 1 u=1, 2 v=0, 3 v=1, 4 w=0, 5 w=1 (IFEM face 1..6). */
void orc_face_msc(const orc_patch* P, int ncomp, const int fixmask[6], int* msc);

/* a16: sparsity pattern (SAM::getDofCouplings + SparseMatrix::optimiseCols).
 * Fills rowptr[neq+1]; if colind != NULL also the sorted column indices. */
int64_t orc_pattern(const orc_patch* P, int ncomp, const int* meqn, int neq,
                    int64_t* rowptr, int32_t* colind);

/* a11-a15: numeric assembly of K (CSR values) and b over all elements, using
 * the two-colour stripe partition; nthreads <= 0 uses all threads. */
int orc_assemble(const orc_patch* P, const orc_integrand* itg, const int ng[3],
                 int ncomp, const int* meqn, int neq, const int64_t* rowptr,
                 const int32_t* colind, double* A, double* b, int nthreads);
/* Same, restricted to elements with e3 in [e3_begin, e3_end). */
int orc_assemble_range(const orc_patch* P, const orc_integrand* itg,
                       const int ng[3], int ncomp, const int* meqn, int neq,
                       const int64_t* rowptr, const int32_t* colind,
                       double* A, double* b, int nthreads,
                       int e3_begin, int e3_end);

/* Same with prescribed values gdof (per DOF; used where meqn <= 0): the
 * constrained columns go to the right-hand side, b_i -= g_j K_ij
 * (src/LinAlg/SparseMatrix.C:639-661).  gdof NULL: homogeneous. */
int orc_assemble_range_g(const orc_patch* P, const orc_integrand* itg,
                         const int ng[3], int ncomp, const int* meqn, int neq,
                         const int64_t* rowptr, const int32_t* colind,
                         double* A, double* b, int nthreads,
                         int e3_begin, int e3_end, const double* gdof);

/* SAM arrays with constraint equations (1-based, SAM layout): meqn per DOF
 * (> 0 equation, 0 fixed, -k slave of constraint k); constraint k holds the
 * entries mpmceq[k-1] .. mpmceq[k]-1 of mmceq (DOF numbers) and ttcc (first
 * the slave and c0, then masters and coefficients). */
typedef struct {
  const int* meqn;
  int ndof;
  int nceq;
  const int* mpmceq;
  const int* mmceq;
  const double* ttcc;
} orc_sam;

/* SAM::getDofCouplings with constraints (src/LinAlg/SAM.C:377-427): nnz, or
 * -1; colind may be NULL (count only). */
int64_t orc_pattern_sam(const orc_patch* P, int ncomp, const orc_sam* s, int neq,
                        int64_t* rowptr, int32_t* colind);
/* assemSparse + SAM::assembleRHS with constraint equations, masters included
 * (src/LinAlg/SparseMatrix.C:612-688, src/LinAlg/SAM.C:605-617); adds into
 * A and b. */
int orc_assemble_sam(const orc_patch* P, const orc_integrand* itg, const int ng[3], int ncomp,
                     const orc_sam* s, int neq, const int64_t* rowptr, const int32_t* colind,
                     double* A, double* b);
/* SAM::expandVector, scaleSD = 1 (src/LinAlg/SAM.C:783-807). */
void orc_expand_sam(const orc_sam* s, const double* xeq, double* xdof);

/* a21: Neumann traction t on face (0..5 as above), ng points per tangent
 * direction, added into b. */
int orc_assemble_traction(const orc_patch* P, int face, const double t[3],
                          int ng, int ncomp, const int* meqn, double* b);

/* a19: y = A x (CSR). */
void orc_spmv(int n, const int64_t* rowptr, const int32_t* colind,
              const double* A, const double* x, double* y);

/* a18: Jacobi-preconditioned CG from x = 0.  Stops when
 * ||r||_2 <= max(rtol*||b||_2, atol); then re-checks the true residual and
 * continues from it if needed.  Returns iterations (negative if maxits hit).
 * resid_out (optional) gets the final true relative residual. */
int orc_pcg_jacobi(int n, const int64_t* rowptr, const int32_t* colind,
                   const double* A, const double* b, double* x,
                   double rtol, double atol, int maxits, double* resid_out);
/* Fixed number of plain PCG iterations (no stopping test), for iterate parity. */
void orc_pcg_jacobi_fixed(int n, const int64_t* rowptr, const int32_t* colind,
                          const double* A, const double* b, double* x, int iters);

/* a20: equation vector -> DOF vector (fixed DOFs = 0; with gdof the
 * prescribed values). */
void orc_expand_solution(int ndof, const int* meqn, const double* xeq, double* xdof);
void orc_expand_solution_g(int ndof, const int* meqn, const double* xeq, const double* gdof,
                           double* xdof);

/* Field evaluation shared by all integrands, and its gradient. */
double orc_eval_function(const orc_function* f, const double X[3]);
void orc_eval_function_grad(const orc_function* f, const double X[3], double g[3]);

/* Norms of the field u_h = sum_a N_a x_a (xdof: DOF vector, eliminated DOFs
 * included), NormBase/GlbNorm with the SQRT final operation: out[0] =
 * ||u_h||_L2, out[1] = ||u_h||_a (Poisson: int kappa |grad u_h|^2,
 * elasticity: int eps:C:eps); with an exact scalar solution (Poisson):
 * out[2] = ||u||_L2, out[3] = ||u||_a, out[4] = ||u-u_h||_L2,
 * out[5] = ||u-u_h||_a (else 0). */
int orc_norms(const orc_patch* P, const orc_integrand* itg, const int ng[3],
              const double* xdof, const orc_function* exact, double* out);

#ifdef __cplusplus
}
#endif
#endif
