// batch.hpp -- state of a batch of independent scalar patches assembled by
// one persistent launch and solved together (block-diagonal global system,
// one CG per patch with its own scalars).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ifem_gpu.h"

namespace ifem {

struct BatchPatch {
  int patch = -1;                // index into Context::patches
  int p = 0, ng = 0;             // degree, Gauss points per direction
  int nnod = 0, neq = 0, eq0 = 0; // nodes, equations, first global equation
  int chunk0 = 0, nchunk = 0;    // solver row chunks of this patch
  ifem_gpu_integrand itg{};      // mode pointers replaced by the copies below
  std::vector<double> kappa_modes, source_modes;
  int* d_eq = nullptr;           // per node: global 0-based equation or -1
  double *d_kmodes = nullptr, *d_smodes = nullptr;
  double *tabN[3] = {nullptr, nullptr, nullptr}, *tabD[3] = {nullptr, nullptr, nullptr};
  double* span[3] = {nullptr, nullptr, nullptr};
};

// Per-degree tuning: element-chunk variant of the assembly and lanes per row
// of the SpMV.
struct BatchTune {
  int half = 0;
  int lanes = 0; // 0: by row length
};

struct Batch {
  std::vector<BatchPatch> pp;
  bool ready = false, fp64_ready = false, fp32_ready = false;
  // the Gauss-point geometry records hold the last assembly's pass
  // (Jacobians and material of every patch; the source term only after an
  // FP64 pass, which always recomputes them): an FP32 assembly of the same
  // batch reuses them instead of recomputing them
  bool geo_valid = false;
  int neq = 0;
  int64_t nnz = 0;
  int nchunk = 0;
  int64_t* d_rowptr = nullptr;
  int32_t* d_colind = nullptr;
  double* d_val = nullptr;
  float* d_val32 = nullptr;
  double* d_rhs = nullptr;

  // assembly work list and per-patch arguments (AsmArgs, FP64 and FP32)
  void* d_args64 = nullptr;
  void* d_args32 = nullptr;
  int2* d_items = nullptr;
  int nitems = 0;
  int item0[4] = {0, 0, 0, 0}, item1[4] = {0, 0, 0, 0}; // items of degree p: [item0[p], item1[p])
  int* d_queue = nullptr;   // one queue counter per degree
  int ncu = 0;              // compute units (persistent grid size)
  // geometry records per (element, Gauss point) (patch q's records start at
  // h_gp0[q]); ngp records in all.  Per degree p: its patches
  // d_glist[gdeg_off[p] ..+gdeg_n[p]), the first point of each among that
  // degree's gdeg_total[p] points in d_gq0 (same offsets)
  double* d_geo = nullptr;
  std::vector<int64_t> h_gp0;
  int* d_glist = nullptr;
  int64_t* d_gq0 = nullptr;
  int gdeg_off[4] = {0, 0, 0, 0}, gdeg_n[4] = {0, 0, 0, 0};
  int64_t gdeg_total[4] = {0, 0, 0, 0};
  int64_t ngp = 0;
  BatchTune tune[4];
  bool tuned = false;

  // solver chunks: chunk c covers rows [crow0[c], crow0[c+1]) of patch cpatch[c]
  int* d_crow0 = nullptr;
  int* d_cpatch = nullptr;
  int* d_clanes = nullptr;
  int* d_pchunk = nullptr; // per patch: first chunk, number of chunks
  std::vector<int> h_cpatch;

  // solver vectors (global equation order)
  double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr,
        *dinv32 = nullptr;
  double* part = nullptr;   // 3 x nchunk partial sums
  double* pscal = nullptr;  // npatch x 8 per-patch scalars
  double* h_pscal = nullptr; // pinned mirror
  double* pnorm = nullptr;  // per patch: scale of the FP32 inner right-hand side

  // captured iteration loops (FP64, FP32): check iterations each
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_check[2] = {0, 0};

  // results of the last solve (per patch)
  std::vector<int> iters, conv;
  std::vector<double> relres;
  double symbolic_ms = 0.0, assemble_ms = 0.0, solve_ms = 0.0;

  ~Batch();
  void release();
};

} // namespace ifem
