// internal.hpp -- shared state and helpers of libifem_gpu.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cmath>
#include <string>
#include <vector>

#include "batch.hpp"
#include "ifem_gpu.h"
#include "mg.hpp"
#include "unet.hpp"

#include <memory>

namespace ifem {

constexpr int kMaxP = 3;
constexpr int kMaxNG = 5;

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define IFEM_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::ifem::fail(IFEM_GPU_EDEVICE, std::string(#call) + ": " +       \
                                                hipGetErrorString(e_));       \
  } while (0)

// Device-side description of one scalar field (matches ifem_gpu_function).
struct DevFunction {
  int type = 0;
  int nmodes = 0;
  double scale = 0.0;
  double mean = 0.0;
  const double* modes = nullptr; // device, nmodes*5
  // bounds over the modes of |k_x| + |k_y| + |k_z| and of |phase| (host
  // computed): k.x + phase stays below kbound*max|x_i| + pbound
  double kbound = 0.0, pbound = 0.0;
};

// kbound and pbound of a mode table (nmodes rows of kx ky kz phase amp).
inline void mode_bounds(const double* modes, int nmodes, double& kbound, double& pbound)
{
  kbound = pbound = 0.0;
  for (int m = 0; m < nmodes; m++) {
    const double* md = modes + 5 * m;
    const double k1 = std::fabs(md[0]) + std::fabs(md[1]) + std::fabs(md[2]);
    kbound = k1 > kbound ? k1 : kbound;
    pbound = std::fabs(md[3]) > pbound ? std::fabs(md[3]) : pbound;
  }
}

struct PatchRec {
  int p[3] = {0, 0, 0};
  int n[3] = {0, 0, 0};
  int nel[3] = {0, 0, 0};
  double* d_knots[3] = {nullptr, nullptr, nullptr};
  double* d_coefs = nullptr;
  std::vector<double> knots[3]; // host copies
};

// Solver work vectors (equation order).
struct SolverWork {
  int n = 0;
  double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double* dinv = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
  float* dinv32 = nullptr;
  double* partial = nullptr; // per-block partial sums (3 x grid)
  double* scal = nullptr;    // device scalars
  double* h_scal = nullptr;  // pinned host mirror
  int nblocks = 0;
};

// One patch of a system with several patches sharing nodes.
struct SubPatch {
  int patch = -1;
  int* d_leq = nullptr;       // local DOF -> global equation (0-based) or -1
  double* d_lgval = nullptr;  // local DOF -> prescribed value (nullptr: none)
  // position of each (local row, valid window column) in the shared CSR:
  // d_pmap[d_pptr[d] + k] for the k-th valid column of local DOF d in the
  // (2p+1)^3 window order (nullptr: search the row instead)
  int64_t* d_pmap = nullptr;
  int64_t* d_pptr = nullptr;
  std::vector<int> h_leq;
};

// Target system of an assembly pass.
struct SysView {
  int64_t* rowptr;
  int32_t* colind;
  double* val;
  float* val32;
  double* rhs;
  int64_t nnz;
  int neq;
};

// Constraint equations with master DOFs: the slaves with masters become
// extra unknowns ne = neq .. neq+ns-1 of an extended system that the line
// kernels assemble; it is condensed to the free equations, K = T^T K_x T,
// b = T^T (b_x - K_x t0) with u_x = T u + t0.
struct MpcSystem {
  int ne = 0, ns = 0;
  int64_t nnz = 0;
  int64_t* d_rowptr = nullptr;
  int32_t* d_colind = nullptr;
  double* d_val = nullptr;
  float* d_val32 = nullptr;
  double* d_rhs = nullptr;
  SubPatch ext;                 // local DOF -> extended equation; prescribed values
  // per reduced row: contributing extended rows and weights
  int* d_cptr = nullptr;
  int* d_cext = nullptr;
  double* d_cw = nullptr;
  // per slave k (extended equation neq+k): masters (reduced equations),
  // coefficients and constant
  int* d_sptr = nullptr;
  int* d_smaster = nullptr;
  double* d_scoef = nullptr;
  double* d_sc0 = nullptr;
  int* d_slave_of_dof = nullptr; // per DOF: k or -1
  double* d_xdof = nullptr;      // expanded solution (norm integration)
  std::vector<int> h_slave_of_dof, h_sptr, h_smaster;
  std::vector<double> h_scoef, h_sc0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<PatchRec> patches;

  // bound system
  int patch = -1;
  int ncomp = 0, ndof = 0, neq = 0;
  int* d_eq = nullptr;          // per DOF: 0-based equation or -1
  std::vector<int> h_eq;
  double* d_gval = nullptr;     // per DOF prescribed value (nullptr: all 0)
  std::vector<double> h_gval;
  std::vector<SubPatch> multi;  // non-empty: the system holds several patches
  std::unique_ptr<MpcSystem> mpc; // constraint equations with masters
  int64_t nnz = 0;
  int64_t* d_rowptr = nullptr;
  int32_t* d_colind = nullptr;
  int32_t* d_bcol = nullptr;    // node-triple column index (3-comp, node-uniform)
  double* d_val = nullptr;
  float* d_val32 = nullptr;
  double* d_rhs = nullptr;
  bool pattern_ready = false;
  bool fp64_ready = false, fp32_ready = false;

  SolverWork w;
  double* noalloc_scratch = nullptr; // this device's staging array (ifem_gpu_spmv_noalloc)
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_mid = nullptr;
  double last_assemble_ms = 0.0, last_solve_ms = 0.0;
  double last_geometry_ms = 0.0, last_lines_ms = 0.0; // parts of the last assembly

  // owned device buffers for integrand mode tables
  // basis tables (values, derivatives, element spans per direction) of a
  // patch for one Gauss count: built on first use, kept until the context
  // ends (patches never change once created)
  struct BasisTables {
    const double* knots0 = nullptr; // the patch's device knot vector 1 (identifies it)
    int ng = 0;
    double *tabN[3] = {nullptr, nullptr, nullptr}, *tabD[3] = {nullptr, nullptr, nullptr};
    double* span[3] = {nullptr, nullptr, nullptr};
  };
  std::vector<BasisTables> basis_cache;
  // device copies of the Fourier mode tables of the fields of the current
  // call (conductivity, source, exact solution), grown as needed
  double* d_modes[3] = {nullptr, nullptr, nullptr};
  size_t modes_cap[3] = {0, 0, 0};
  void* d_asmargs = nullptr; // assembly kernel arguments (device copy)
  double* d_gpgeo = nullptr; // geometry records per (element, Gauss point)
  size_t gpgeo_cap = 0;      // doubles
  // element-chunk variant of the line kernel per (p, ng-p-1, ncomp==3, fp32):
  // -1 built-in default, 0 full, 1 half
  int asm_half[4][2][2][2] = {{{{-1, -1}, {-1, -1}}, {{-1, -1}, {-1, -1}}},
                              {{{-1, -1}, {-1, -1}}, {{-1, -1}, {-1, -1}}},
                              {{{-1, -1}, {-1, -1}}, {{-1, -1}, {-1, -1}}},
                              {{{-1, -1}, {-1, -1}}, {{-1, -1}, {-1, -1}}}};

  // learned preconditioner
  std::unique_ptr<UNet> net;
  bool learned = false;
  int lat[3] = {0, 0, 0}; // D, H, W of the (padded) network lattice
  int box[3] = {0, 0, 0}; // k, j, i extents of the free-node box
  int naux = 0;
  int features = 0;          // 0: [r, aux], 1: [r, r*aux]
  float* d_aux = nullptr;   // naux * D*H*W
  float* d_netin = nullptr; // (ncomp + naux) * D*H*W
  float* d_netout = nullptr;
  double* d_rold = nullptr; // previous residual (flexible CG)
  float* d_rold32 = nullptr;
  double* d_zn = nullptr;   // learned correction of the current iteration
  float* d_zn32 = nullptr;
  double learned_scale = 1.0;
  bool last_fallback = false; // the last solve fell back to Jacobi-PCG

  // batch of independent patches (batch.hip)
  std::unique_ptr<Batch> batch;
  // spline multigrid hierarchy (pc = IFEM_GPU_PC_MG)
  std::unique_ptr<Multigrid> mg;
  // face Gauss-point weights of the traction pass
  double* d_facew = nullptr;
  size_t facew_cap = 0;
};

// assemble.hip
int build_pattern(Context& c);
int assemble_volume(Context& c, const ifem_gpu_integrand& itg, const int ng[3],
                    int precision);
int assemble_traction(Context& c, int index, int face0, const double* t, int ng);
// constraint equations with masters (mpc.hip)
int mpc_bind(Context& c, int nceq, const int* mpmceq, const int* mmceq, const double* ttcc);
int mpc_build_pattern(Context& c);
int mpc_condense(Context& c, bool fp32, bool rhs_only);
int mpc_expand_device(Context& c);
void mpc_free(Context& c);
void mpc_expand_host(const Context& c, const std::vector<double>& x, double* xdof);
int autotune_assembly(Context& c, const ifem_gpu_integrand& itg, const int ng[3], int precision,
                      int* half);
int set_assembly_variant(Context& c, int p, int ng, int nc, bool fp32, int variant);
int integrate_norms(Context& c, const ifem_gpu_integrand& itg, const int ng[3],
                    const ifem_gpu_function* exact, double* out);
int batch_symbolic(Context& c);
int batch_upload_args(Context& c);
int batch_assemble(Context& c, int precision, double* ms);

// batch.hip
int batch_init(Context& c);
int batch_solve(Context& c, const ifem_gpu_solver_params& prm, ifem_gpu_stats* st);
int batch_autotune(Context& c, int precision, int* table);

// solver.hip
int solver_setup(Context& c);
int solve(Context& c, const ifem_gpu_solver_params& prm, ifem_gpu_stats* st);
int spmv_device(Context& c, const double* d_x, double* d_y);
int spmv_device32(Context& c, const float* d_x, float* d_y);
void solver_free(Context& c);
int set_learned(Context& c, int enable, int D, int H, int W, const float* aux, int naux,
                int features, double scale);

} // namespace ifem

struct ifem_gpu_ctx {
  ifem::Context c;
};
