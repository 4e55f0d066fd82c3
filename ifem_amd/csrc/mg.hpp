// mg.hpp -- spline multigrid preconditioner (not installed).
//
// Nested spline spaces of one patch: every coarser level drops every second
// distinct interior knot of each direction, so the coarse B-splines are
// linear combinations of the fine ones (knot insertion gives the exact
// prolongation P, a tensor product of 1D matrices).  Coarse operators are
// Galerkin products P^T A P, formed direction by direction on a stencil
// layout; smoothing is Chebyshev on D^-1 A; the coarsest level is solved by
// a dense inverse.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ifem {

struct Context;

constexpr int kMgKC = 8; // fine functions per coarse function, at most
constexpr int kMgKR = 4; // coarse functions per fine function, at most

// Transfer in one direction between a level and the next coarser one:
// column form (coarse I: fine indices clo[I] .. clo[I]+ccnt[I]-1, weights
// cw[I*kMgKC + k]) and row form (fine i: coarse rlo[i] .., weights rw).
struct MgDir {
  int nf = 0, nc = 0;
  int *clo = nullptr, *ccnt = nullptr;
  double* cw = nullptr;
  int *rlo = nullptr, *rcnt = nullptr;
  double* rw = nullptr;
};

struct MgLevel {
  int n[3] = {0, 0, 0};     // control points per direction
  int nnode = 0, ndof = 0;  // ndof = nnode * ncomp
  int nvec = 0;             // vector length: neq on level 0, ndof below
  int nfree = 0;
  MgDir dir[3];             // from the finer level (unused on level 0)
  double* S64 = nullptr;    // stencil [nnode][noff][nc][nc] (coarse levels)
  float* S32 = nullptr;
  unsigned char* freem = nullptr; // per DOF, coarse levels
  float* dinv = nullptr;    // per vector entry, 0 at eliminated DOFs
  float *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr, *q = nullptr;
  double lmax = 0.0;        // estimate of the largest eigenvalue of D^-1 A
  std::vector<unsigned char> h_free;
  std::vector<double> knots[3];
};

struct Multigrid {
  int p = 0, nc = 0, noff = 0;
  int degree = 2;           // Chebyshev steps per smoothing
  int patch = -1, neq = 0;  // what the hierarchy was built for
  int64_t nnz = 0;
  std::vector<MgLevel> lev;
  double* S0 = nullptr;     // fine operator, node-major stencil (setup only)
  int* dof_of_eq = nullptr; // fine equation -> DOF
  double *T1 = nullptr, *T2 = nullptr;
  size_t s0cap = 0, t1cap = 0, t2cap = 0;
  __half* val16 = nullptr;  // K / kscale in FP16 (fine-level smoother), CSR order
  uint2* val16n = nullptr;  // the same in node blocks (3-component node-uniform systems)
  double kscale = 1.0;      // largest diagonal entry of K
  int ncf = 0;              // free DOFs of the coarsest level
  int* cdofs = nullptr;
  float* cinv = nullptr;    // dense inverse of the coarsest operator
  double *part = nullptr, *part2 = nullptr; // partial sums (eigenvalue estimate)
  double* lam = nullptr;
  double* lzh = nullptr;    // Lanczos coefficients per level (device)
  double* h_lzh = nullptr;  // pinned host copies
  double* h_coarse = nullptr;
  int* h_err = nullptr;
  int* err = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr, e_copy = nullptr, e_gal = nullptr, e_side = nullptr;
  hipStream_t side = nullptr; // coarse-level eigenvalue estimates beside the fine one
  double *part_s = nullptr, *part_s2 = nullptr;
  double setup_ms = 0.0;
  ~Multigrid();
};

// Builds (first call for this system) and refreshes the hierarchy from the
// current FP64 matrix; degree = Chebyshev steps per smoothing.
int mg_setup(Context& c, int degree);
// z = M^-1 r on the context stream (r, z: equation order).  With scal, the
// kernels skip their work once scal[S_DONE] is set.
int mg_apply(Context& c, const double* r, double* z, const double* scal);
int mg_apply(Context& c, const float* r, float* z, const double* scal);

} // namespace ifem
