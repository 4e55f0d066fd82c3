// batch.hip -- many independent scalar patches assembled by one persistent
// launch and solved together, each with its own Jacobi-PCG.
//
// The global system is block diagonal: patch q owns equations
// [eq0_q, eq0_q + neq_q).  The solver splits the rows into chunks of at most
// 256 rows that never straddle two patches; one block handles one chunk.
// Dot products are per-chunk partials that every consumer block of patch q
// re-reduces over q's chunk range in index order, so all blocks of a patch
// agree bitwise on alpha and beta without grid-wide fences.  Each patch has
// its own scalars (rz, rr, tolerance, iteration count, done flag); the first
// chunk's block of a patch records them.  Patches that have converged skip
// their chunks.  A fixed number of iterations (3 launches each) is captured
// once in a hipGraph and replayed between host convergence checks.
//
// Per patch this is the solve of SystemMatrix::solve with KSPCG + PCJACOBI
// (src/LinAlg/PETScMatrix.C:778-832, 980-990) under the LinSolParams
// contract (src/LinAlg/LinSolParams.C:68-77); the FP32 mode runs the inner
// CG in FP32 inside FP64 iterative refinement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "devutil.hpp"
#include "internal.hpp"

namespace ifem {

constexpr int kChunkRows = kThreads; // rows per solver chunk
constexpr int kBS = 8;               // per-patch scalar stride (Scal slots)

// ---------------------------------------------------------------------------
template <typename V, typename ACC, int L, bool DOT>
__device__ inline double bspmv_rows(int r0, int r1, const int64_t* __restrict__ rowptr,
                                    const int32_t* __restrict__ colind, const V* __restrict__ val,
                                    const V* __restrict__ p, V* __restrict__ q)
{
  const int g = threadIdx.x / L, l = threadIdx.x % L;
  double pq = 0.0;
  for (int row = r0 + g; row < r1; row += kThreads / L) {
    const int64_t b = rowptr[row], e = rowptr[row + 1];
    ACC s = 0;
    for (int64_t k = b + l; k < e; k += L) s += (ACC)val[k] * (ACC)p[colind[k]];
#pragma unroll
    for (int off = L / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, L);
    if (l == 0) {
      q[row] = (V)s;
      if (DOT) pq += (double)((ACC)p[row] * s);
    }
  }
  return pq;
}

// q = A p over one chunk per block (lanes per row chosen per chunk); with
// DOT, per-chunk partial <p,q> and converged patches are skipped.
template <typename V, typename ACC, bool DOT>
__global__ __launch_bounds__(kThreads) void k_bspmv(
    const int* __restrict__ crow0, const int* __restrict__ cpatch, const int* __restrict__ clanes,
    const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const V* __restrict__ val, const V* __restrict__ p, V* __restrict__ q,
    double* __restrict__ part_pq, const double* __restrict__ pscal)
{
  const int c = blockIdx.x;
  const int pt = cpatch[c];
  if (DOT && pscal[pt * kBS + S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const int r0 = crow0[c], r1 = crow0[c + 1];
  double pq;
  switch (clanes[c]) {
  case 16: pq = bspmv_rows<V, ACC, 16, DOT>(r0, r1, rowptr, colind, val, p, q); break;
  case 32: pq = bspmv_rows<V, ACC, 32, DOT>(r0, r1, rowptr, colind, val, p, q); break;
  default: pq = bspmv_rows<V, ACC, 64, DOT>(r0, r1, rowptr, colind, val, p, q); break;
  }
  if (!DOT) return;
  const double bs = block_sum<double>(pq, sh);
  if (threadIdx.x == 0) part_pq[c] = bs;
}

// alpha_q = rz_q / <p,q>_q; x += alpha p, r -= alpha q, z = D^-1 r;
// per-chunk partials <r,z>, <r,r>.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_bupdate(
    const int* __restrict__ crow0, const int* __restrict__ cpatch, const int* __restrict__ pchunk,
    V* __restrict__ x, V* __restrict__ r, V* __restrict__ z, const V* __restrict__ p,
    const V* __restrict__ q, const V* __restrict__ dinv, const double* __restrict__ part_pq,
    double* __restrict__ part_rz, double* __restrict__ part_rr, double* __restrict__ pscal,
    int parity)
{
  const int c = blockIdx.x;
  const int pt = cpatch[c];
  double* sc = pscal + pt * kBS;
  if (sc[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const int c0 = pchunk[2 * pt], nc = pchunk[2 * pt + 1];
  const double pq = sum_parts(part_pq + c0, nc, sh);
  double a = 0.0;
  if (pq != 0.0 && __builtin_isfinite(pq)) a = sc[S_RZ + parity] / pq;
  else if (c == c0 && threadIdx.x == 0) sc[S_DONE] = 2.0; // breakdown
  const V alpha = (V)a;
  double rz = 0.0, rr = 0.0;
  const int i = crow0[c] + threadIdx.x;
  if (i < crow0[c + 1]) {
    const V xi = x[i] + alpha * p[i];
    const V ri = r[i] - alpha * q[i];
    const V zi = dinv[i] * ri;
    x[i] = xi;
    r[i] = ri;
    z[i] = zi;
    rz = (double)(ri * zi);
    rr = (double)(ri * ri);
  }
  const double brz = block_sum<double>(rz, sh);
  const double brr = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) {
    part_rz[c] = brz;
    part_rr[c] = brr;
  }
}

// beta_q = <r,z>_new / <r,z>_old; p = z + beta p; the patch's first chunk
// records rz, rr, the iteration count and the done flag.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_bpupdate(
    const int* __restrict__ crow0, const int* __restrict__ cpatch, const int* __restrict__ pchunk,
    V* __restrict__ p, const V* __restrict__ z, const double* __restrict__ part_rz,
    const double* __restrict__ part_rr, double* __restrict__ pscal, int parity)
{
  const int c = blockIdx.x;
  const int pt = cpatch[c];
  double* sc = pscal + pt * kBS;
  if (sc[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const int c0 = pchunk[2 * pt], nc = pchunk[2 * pt + 1];
  const double rzn = sum_parts(part_rz + c0, nc, sh);
  const double rr = sum_parts(part_rr + c0, nc, sh);
  const double rzo = sc[S_RZ + parity];
  const bool conv = !(rr > sc[S_TOL2]);
  if (c == c0 && threadIdx.x == 0) {
    sc[S_RZ + 1 - parity] = rzn;
    sc[S_RR] = rr;
    sc[S_ITER] += 1.0;
    if (conv) sc[S_DONE] = 1.0;
    else if (sc[S_ITER] >= sc[S_MAXIT]) sc[S_DONE] = 3.0;
  }
  if (conv) return;
  const V beta = (V)(rzo != 0.0 ? rzn / rzo : 0.0);
  const int i = crow0[c] + threadIdx.x;
  if (i < crow0[c + 1]) p[i] = z[i] + beta * p[i];
}

// CG start from r: z = D^-1 r, p = z, per-chunk partials <r,z>, <r,r>.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_bstart(const int* __restrict__ crow0,
                                                     const V* __restrict__ r, V* __restrict__ z,
                                                     V* __restrict__ p, const V* __restrict__ dinv,
                                                     double* __restrict__ part_rz,
                                                     double* __restrict__ part_rr)
{
  __shared__ double sh[kThreads / 64];
  const int c = blockIdx.x;
  const int i = crow0[c] + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < crow0[c + 1]) {
    const V ri = r[i];
    const V zi = dinv[i] * ri;
    z[i] = zi;
    p[i] = zi;
    rz = (double)(ri * zi);
    rr = (double)(ri * ri);
  }
  const double brz = block_sum<double>(rz, sh);
  const double brr = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) {
    part_rz[c] = brz;
    part_rr[c] = brr;
  }
}

// One block per patch: initial scalars; patches flagged done stay done.
__global__ __launch_bounds__(kThreads) void k_bstart_check(const int* __restrict__ pchunk,
                                                           const double* __restrict__ part_rz,
                                                           const double* __restrict__ part_rr,
                                                           double* __restrict__ pscal)
{
  __shared__ double sh[kThreads / 64];
  const int pt = blockIdx.x;
  double* sc = pscal + pt * kBS;
  if (sc[S_DONE] != 0.0) return;
  const int c0 = pchunk[2 * pt], nc = pchunk[2 * pt + 1];
  const double rz = sum_parts(part_rz + c0, nc, sh);
  const double rr = sum_parts(part_rr + c0, nc, sh);
  if (threadIdx.x == 0) {
    sc[S_RZ] = rz;
    sc[S_RR] = rr;
    sc[S_DONE] = (rr > sc[S_TOL2]) ? 0.0 : 1.0;
    if (sc[S_ITER] >= sc[S_MAXIT]) sc[S_DONE] = 3.0;
  }
}

// r = b - y (y == nullptr: r = b), per-chunk partials of ||r||^2.
__global__ __launch_bounds__(kThreads) void k_bresidual(const int* __restrict__ crow0,
                                                        const double* __restrict__ b,
                                                        const double* __restrict__ y,
                                                        double* __restrict__ r,
                                                        double* __restrict__ part)
{
  __shared__ double sh[kThreads / 64];
  const int c = blockIdx.x;
  const int i = crow0[c] + threadIdx.x;
  double rr = 0.0;
  if (i < crow0[c + 1]) {
    const double ri = y ? b[i] - y[i] : b[i];
    r[i] = ri;
    rr = ri * ri;
  }
  const double s = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) part[c] = s;
}

// One block per patch: pscal[slot] = sum of the patch's chunk partials.
__global__ __launch_bounds__(kThreads) void k_bpatch_sum(const int* __restrict__ pchunk,
                                                         const double* __restrict__ part,
                                                         double* __restrict__ pscal, int slot)
{
  __shared__ double sh[kThreads / 64];
  const int pt = blockIdx.x;
  const double s = sum_parts(part + pchunk[2 * pt], pchunk[2 * pt + 1], sh);
  if (threadIdx.x == 0) pscal[pt * kBS + slot] = s;
}

// r32 = r / s_q (s_q = 0: r32 = 0), x32 = 0
__global__ __launch_bounds__(kThreads) void k_bdemote(const int* __restrict__ crow0,
                                                      const int* __restrict__ cpatch,
                                                      const double* __restrict__ r,
                                                      const double* __restrict__ s,
                                                      float* __restrict__ r32,
                                                      float* __restrict__ x32)
{
  const int c = blockIdx.x;
  const int i = crow0[c] + threadIdx.x;
  if (i >= crow0[c + 1]) return;
  const double sq = s[cpatch[c]];
  r32[i] = sq > 0.0 ? (float)(r[i] * (1.0 / sq)) : 0.0f;
  x32[i] = 0.0f;
}

// x += s_q * x32
__global__ __launch_bounds__(kThreads) void k_baxpy(const int* __restrict__ crow0,
                                                    const int* __restrict__ cpatch,
                                                    double* __restrict__ x,
                                                    const float* __restrict__ x32,
                                                    const double* __restrict__ s)
{
  const int c = blockIdx.x;
  const int i = crow0[c] + threadIdx.x;
  if (i >= crow0[c + 1]) return;
  x[i] += s[cpatch[c]] * (double)x32[i];
}

// ---------------------------------------------------------------------------
Batch::~Batch() { release(); }

void Batch::release()
{
  for (BatchPatch& bp : pp) {
    hipFree(bp.d_eq);
    hipFree(bp.d_kmodes);
    hipFree(bp.d_smodes);
    for (int d = 0; d < 3; d++) {
      hipFree(bp.tabN[d]);
      hipFree(bp.tabD[d]);
      hipFree(bp.span[d]);
    }
  }
  pp.clear();
  void** d[] = {(void**)&d_rowptr, (void**)&d_colind, (void**)&d_val, (void**)&d_val32,
                (void**)&d_rhs, &d_args64, &d_args32, (void**)&d_items, (void**)&d_queue,
                (void**)&d_crow0, (void**)&d_cpatch, (void**)&d_clanes, (void**)&d_pchunk,
                (void**)&x, (void**)&r, (void**)&z, (void**)&p, (void**)&q, (void**)&dinv,
                (void**)&r32, (void**)&z32, (void**)&p32, (void**)&q32, (void**)&x32,
                (void**)&dinv32, (void**)&part, (void**)&pscal, (void**)&pnorm, (void**)&d_geo,
                (void**)&d_glist, (void**)&d_gq0};
  for (void** v : d) {
    hipFree(*v);
    *v = nullptr;
  }
  if (h_pscal) hipHostFree(h_pscal);
  h_pscal = nullptr;
  for (int g = 0; g < 2; g++) {
    if (graph[g]) hipGraphExecDestroy(graph[g]);
    graph[g] = nullptr;
    graph_check[g] = 0;
  }
  ready = fp64_ready = fp32_ready = tuned = geo_valid = false;
  neq = nitems = nchunk = 0;
  nnz = ngp = 0;
  h_gp0.clear();
  for (int p = 0; p < 4; p++) {
    gdeg_off[p] = gdeg_n[p] = 0;
    gdeg_total[p] = 0;
  }
}

static int default_lanes(int p)
{
  // interior row lengths (2p+1)^3: 27, 125, 343
  return p <= 1 ? 16 : 64;
}

static int upload_lanes(Context& c)
{
  Batch& B = *c.batch;
  std::vector<int> lanes(B.nchunk > 0 ? B.nchunk : 1, 64);
  for (int ch = 0; ch < B.nchunk; ch++) {
    const int p = B.pp[B.h_cpatch[ch]].p;
    lanes[ch] = B.tune[p].lanes > 0 ? B.tune[p].lanes : default_lanes(p);
  }
  IFEM_HIP(hipMemcpy(B.d_clanes, lanes.data(), sizeof(int) * lanes.size(), hipMemcpyHostToDevice));
  return IFEM_GPU_OK;
}

// Solver chunks and work vectors of the batch.
static int batch_solver_setup(Context& c)
{
  Batch& B = *c.batch;
  const int np = (int)B.pp.size();
  std::vector<int> crow0, cpatch, pchunk(2 * (np > 0 ? np : 1), 0);
  for (int q = 0; q < np; q++) {
    BatchPatch& bp = B.pp[q];
    bp.chunk0 = (int)cpatch.size();
    for (int r = 0; r < bp.neq; r += kChunkRows) {
      crow0.push_back(bp.eq0 + r);
      cpatch.push_back(q);
    }
    bp.nchunk = (int)cpatch.size() - bp.chunk0;
    pchunk[2 * q] = bp.chunk0;
    pchunk[2 * q + 1] = bp.nchunk;
  }
  B.nchunk = (int)cpatch.size();
  crow0.push_back(B.neq);
  B.h_cpatch = cpatch;
  const size_t nch = B.nchunk > 0 ? B.nchunk : 1;
  IFEM_HIP(hipMalloc(&B.d_crow0, sizeof(int) * crow0.size()));
  IFEM_HIP(hipMalloc(&B.d_cpatch, sizeof(int) * nch));
  IFEM_HIP(hipMalloc(&B.d_clanes, sizeof(int) * nch));
  IFEM_HIP(hipMalloc(&B.d_pchunk, sizeof(int) * pchunk.size()));
  IFEM_HIP(hipMemcpy(B.d_crow0, crow0.data(), sizeof(int) * crow0.size(), hipMemcpyHostToDevice));
  if (B.nchunk > 0)
    IFEM_HIP(hipMemcpy(B.d_cpatch, cpatch.data(), sizeof(int) * cpatch.size(), hipMemcpyHostToDevice));
  IFEM_HIP(hipMemcpy(B.d_pchunk, pchunk.data(), sizeof(int) * pchunk.size(), hipMemcpyHostToDevice));
  int rc = upload_lanes(c);
  if (rc) return rc;
  const size_t n = B.neq > 0 ? B.neq : 1;
  double** d[] = {&B.x, &B.r, &B.z, &B.p, &B.q, &B.dinv};
  for (double** v : d) IFEM_HIP(hipMalloc(v, sizeof(double) * n));
  float** f[] = {&B.r32, &B.z32, &B.p32, &B.q32, &B.x32, &B.dinv32};
  for (float** v : f) IFEM_HIP(hipMalloc(v, sizeof(float) * n));
  IFEM_HIP(hipMemsetAsync(B.x, 0, sizeof(double) * n, c.stream));
  IFEM_HIP(hipMalloc(&B.part, sizeof(double) * 3 * nch));
  const size_t ns = (size_t)kBS * (np > 0 ? np : 1);
  IFEM_HIP(hipMalloc(&B.pscal, sizeof(double) * ns));
  IFEM_HIP(hipMalloc(&B.pnorm, sizeof(double) * (np > 0 ? np : 1)));
  IFEM_HIP(hipHostMalloc(&B.h_pscal, sizeof(double) * ns, hipHostMallocDefault));
  B.iters.assign(np, 0);
  B.conv.assign(np, 0);
  B.relres.assign(np, 0.0);
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

template <typename V, typename ACC, bool DOT>
static void launch_bspmv(Context& c, const V* val, const V* p, V* q)
{
  Batch& B = *c.batch;
  if (B.nchunk == 0) return;
  k_bspmv<V, ACC, DOT><<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, B.d_cpatch, B.d_clanes,
                                                            B.d_rowptr, B.d_colind, val, p, q,
                                                            B.part, B.pscal);
}

static int read_pscal(Context& c)
{
  Batch& B = *c.batch;
  const size_t np = B.pp.size();
  if (np == 0) return IFEM_GPU_OK;
  IFEM_HIP(hipMemcpyAsync(B.h_pscal, B.pscal, sizeof(double) * kBS * np, hipMemcpyDeviceToHost,
                          c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

static int write_pscal(Context& c)
{
  Batch& B = *c.batch;
  const size_t np = B.pp.size();
  if (np == 0) return IFEM_GPU_OK;
  IFEM_HIP(hipMemcpyAsync(B.pscal, B.h_pscal, sizeof(double) * kBS * np, hipMemcpyHostToDevice,
                          c.stream));
  return IFEM_GPU_OK;
}

// Per-patch sums of squares of r = b - y into out[q] (y == nullptr: of b).
static int batch_norms(Context& c, const double* y, double* r, std::vector<double>& out)
{
  Batch& B = *c.batch;
  const int np = (int)B.pp.size();
  if (B.nchunk > 0)
    k_bresidual<<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, B.d_rhs, y, r, B.part);
  if (np > 0) k_bpatch_sum<<<np, kThreads, 0, c.stream>>>(B.d_pchunk, B.part, B.pscal, S_BB);
  IFEM_HIP(hipGetLastError());
  int rc = read_pscal(c);
  if (rc) return rc;
  out.resize(np);
  for (int q = 0; q < np; q++) out[q] = B.h_pscal[q * kBS + S_BB];
  return IFEM_GPU_OK;
}

// The iteration loop (check iterations, 3 launches each) as a graph.
template <typename V, typename ACC>
static int get_graph(Context& c, int gi, const V* val, V* x, V* r, V* z, V* p, V* q, const V* dinv,
                     int check)
{
  Batch& B = *c.batch;
  if (B.graph[gi] && B.graph_check[gi] == check) return IFEM_GPU_OK;
  if (B.graph[gi]) {
    hipGraphExecDestroy(B.graph[gi]);
    B.graph[gi] = nullptr;
  }
  double* part_pq = B.part;
  double* part_rz = B.part + B.nchunk;
  double* part_rr = B.part + 2 * B.nchunk;
  hipGraph_t g = nullptr;
  IFEM_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < check; k++) {
    const int parity = k & 1;
    launch_bspmv<V, ACC, true>(c, val, p, q);
    k_bupdate<V><<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, B.d_cpatch, B.d_pchunk, x, r, z,
                                                      p, q, dinv, part_pq, part_rz, part_rr,
                                                      B.pscal, parity);
    k_bpupdate<V><<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, B.d_cpatch, B.d_pchunk, p, z,
                                                       part_rz, part_rr, B.pscal, parity);
  }
  const hipError_t e1 = hipStreamEndCapture(c.stream, &g);
  if (e1 != hipSuccess) return fail(IFEM_GPU_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e1));
  const hipError_t e2 = hipGraphInstantiate(&B.graph[gi], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    B.graph[gi] = nullptr;
    return fail(IFEM_GPU_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  }
  B.graph_check[gi] = check;
  return IFEM_GPU_OK;
}

// One PCG pass for every patch whose done flag the host left at 0 (the
// scalars TOL2, MAXIT are set by the caller).  Returns per-patch iterations
// and status in the pinned scalar mirror.
template <typename V, typename ACC>
static int batch_pcg(Context& c, int gi, const V* val, V* x, V* r, V* z, V* p, V* q,
                     const V* dinv, int maxits, int check)
{
  Batch& B = *c.batch;
  const int np = (int)B.pp.size();
  int rc = write_pscal(c);
  if (rc) return rc;
  if (B.nchunk == 0) return read_pscal(c);
  double* part_rz = B.part + B.nchunk;
  double* part_rr = B.part + 2 * B.nchunk;
  k_bstart<V><<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, r, z, p, dinv, part_rz, part_rr);
  k_bstart_check<<<np, kThreads, 0, c.stream>>>(B.d_pchunk, part_rz, part_rr, B.pscal);
  IFEM_HIP(hipGetLastError());
  rc = get_graph<V, ACC>(c, gi, val, x, r, z, p, q, dinv, check);
  if (rc) return rc;
  int launched = 0;
  while (true) {
    IFEM_HIP(hipGraphLaunch(B.graph[gi], c.stream));
    launched += check;
    rc = read_pscal(c);
    if (rc) return rc;
    bool all = true;
    for (int pt = 0; pt < np && all; pt++) all = B.h_pscal[pt * kBS + S_DONE] != 0.0;
    if (all || launched >= maxits + check) break;
  }
  return IFEM_GPU_OK;
}

int batch_solve(Context& c, const ifem_gpu_solver_params& prm, ifem_gpu_stats* st)
{
  Batch& B = *c.batch;
  const int np = (int)B.pp.size();
  if (prm.maxits < 1) return fail(IFEM_GPU_EINVAL, "batch_solve: maxits must be positive");
  if (!(prm.rtol >= 0.0) || !(prm.atol >= 0.0))
    return fail(IFEM_GPU_EINVAL, "batch_solve: tolerances must be non-negative");
  const bool mixed = prm.precision == IFEM_GPU_FP32;
  if (!B.fp64_ready) return fail(IFEM_GPU_EINVAL, "batch_solve: assemble the FP64 system first");
  if (mixed && !B.fp32_ready) return fail(IFEM_GPU_EINVAL, "batch_solve: no FP32 values");
  int check = prm.check_every > 0 ? prm.check_every : 32;
  check += check & 1; // even: every graph replay starts at parity 0
  const int n = B.neq;
  const int nb = (n + 255) / 256;
  IFEM_HIP(hipEventRecord(c.ev0, c.stream));
  std::vector<double> bb;
  int rc = batch_norms(c, nullptr, B.r, bb); // r = b
  if (rc) return rc;
  std::vector<double> bnorm(np), tol(np), rr(np);
  std::vector<int> state(np, 0); // 0 iterating, 1 converged, 2 stopped
  B.iters.assign(np, 0);
  B.relres.assign(np, 0.0);
  for (int q = 0; q < np; q++) {
    bnorm[q] = std::sqrt(bb[q]);
    tol[q] = std::fmax(prm.rtol * bnorm[q], prm.atol);
    rr[q] = bb[q];
    if (bnorm[q] == 0.0) state[q] = 1;
    B.relres[q] = bnorm[q] > 0 ? 1.0 : 0.0;
  }
  if (n > 0) {
    k_diag_inv<double><<<nb, 256, 0, c.stream>>>(n, B.d_rowptr, B.d_colind, B.d_val, B.dinv);
    k_fill<double><<<nb, 256, 0, c.stream>>>(n, B.x, 0.0);
    if (mixed)
      k_diag_inv<float><<<nb, 256, 0, c.stream>>>(n, B.d_rowptr, B.d_colind, B.d_val32, B.dinv32);
  }
  IFEM_HIP(hipGetLastError());
  int outer = 0;
  bool diverged = false;
  const double inner = prm.inner_rtol > 0 ? prm.inner_rtol : 1e-4;
  std::vector<double> snorm(np, 0.0);
  while (outer < 100) {
    bool any = false;
    for (int q = 0; q < np; q++) any |= state[q] == 0;
    if (!any) break;
    // per-patch scalars of this pass
    for (int q = 0; q < np; q++) {
      double* sc = B.h_pscal + q * kBS;
      std::memset(sc, 0, sizeof(double) * kBS);
      const double rn = std::sqrt(rr[q]);
      double t2;
      if (mixed) {
        // unit-norm inner right-hand side; never ask for more than the
        // outer tolerance needs
        const double target = rn > 0 ? std::fmax(inner, 0.5 * tol[q] / rn) : 1.0;
        t2 = target * target;
        snorm[q] = state[q] == 0 ? rn : 0.0;
      } else {
        t2 = tol[q] * tol[q];
      }
      sc[S_TOL2] = t2;
      sc[S_MAXIT] = std::max(1, prm.maxits - B.iters[q]);
      sc[S_DONE] = state[q] == 0 ? 0.0 : 1.0;
    }
    if (mixed) {
      IFEM_HIP(hipMemcpyAsync(B.pnorm, snorm.data(), sizeof(double) * np, hipMemcpyHostToDevice,
                              c.stream));
      if (B.nchunk > 0)
        k_bdemote<<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, B.d_cpatch, B.r, B.pnorm, B.r32,
                                                       B.x32);
      rc = batch_pcg<float, float>(c, 1, B.d_val32, B.x32, B.r32, B.z32, B.p32, B.q32, B.dinv32,
                                   prm.maxits, check);
      if (rc) return rc;
      if (B.nchunk > 0)
        k_baxpy<<<B.nchunk, kThreads, 0, c.stream>>>(B.d_crow0, B.d_cpatch, B.x, B.x32, B.pnorm);
    } else {
      rc = batch_pcg<double, double>(c, 0, B.d_val, B.x, B.r, B.z, B.p, B.q, B.dinv, prm.maxits,
                                     check);
      if (rc) return rc;
    }
    IFEM_HIP(hipGetLastError());
    outer++;
    std::vector<int> it(np), status(np);
    for (int q = 0; q < np; q++) {
      it[q] = (int)B.h_pscal[q * kBS + S_ITER];
      status[q] = (int)B.h_pscal[q * kBS + S_DONE];
      if (state[q] == 0) B.iters[q] += it[q];
    }
    // true FP64 residual r = b - A x (also the restart residual)
    launch_bspmv<double, double, false>(c, B.d_val, B.x, B.q);
    rc = batch_norms(c, B.q, B.r, rr);
    if (rc) return rc;
    for (int q = 0; q < np; q++) {
      if (state[q] != 0) continue;
      const double tr = std::sqrt(rr[q]);
      B.relres[q] = bnorm[q] > 0 ? tr / bnorm[q] : 0.0;
      if (tr <= tol[q]) state[q] = 1;
      else if (!(tr <= prm.dtol * bnorm[q])) { state[q] = 2; diverged = true; }
      else if (B.iters[q] >= prm.maxits || status[q] == 2 || it[q] == 0) state[q] = 2;
    }
  }
  IFEM_HIP(hipEventRecord(c.ev1, c.stream));
  IFEM_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  IFEM_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  B.solve_ms = ms;
  int nconv = 0, maxit = 0;
  double maxrel = 0.0;
  for (int q = 0; q < np; q++) {
    B.conv[q] = state[q] == 1;
    nconv += B.conv[q];
    maxit = std::max(maxit, B.iters[q]);
    maxrel = std::fmax(maxrel, B.relres[q]);
  }
  if (st) {
    st->iterations = maxit;
    st->outer_iterations = outer;
    st->converged = nconv == np ? 1 : 0;
    st->rel_residual = maxrel;
    st->solve_ms = ms;
    st->pc_iterations = 0; // Jacobi only
    st->fallback = 0;
    st->pc_rejected = 0;
  }
  if (diverged) return fail(IFEM_GPU_ENOCONV, "batch_solve: a patch diverged (residual above dtol*||b||)");
  if (nconv != np)
    return fail(IFEM_GPU_ENOCONV, "batch_solve: " + std::to_string(np - nconv) + " of " +
                                      std::to_string(np) + " patches not converged");
  return IFEM_GPU_OK;
}

// Autotuner: for each degree present, the element-chunk variant of the
// assembly (full or half chunk) and the SpMV lanes per row (16, 32, 64),
// each chosen by timing the whole batch with the other degrees fixed.
int batch_autotune(Context& c, int precision, int* table)
{
  Batch& B = *c.batch;
  bool present[4] = {false, false, false, false};
  for (const BatchPatch& bp : B.pp) present[bp.p] = true;
  int rc;
  for (int p = 1; p <= 3; p++) {
    if (!present[p]) continue;
    double best = 1e300;
    int best_half = 0;
    for (int half = 0; half <= 1; half++) {
      B.tune[p].half = half;
      rc = batch_upload_args(c);
      if (rc) return rc;
      double t = 1e300;
      for (int rep = 0; rep < 2; rep++) {
        double ms = 0.0;
        rc = batch_assemble(c, precision, &ms);
        if (rc) return rc;
        t = std::fmin(t, ms);
      }
      if (t < best) {
        best = t;
        best_half = half;
      }
    }
    B.tune[p].half = best_half;
  }
  rc = batch_upload_args(c);
  if (rc) return rc;
  // re-assemble with the chosen variants (their summation orders differ)
  if (B.fp64_ready || precision == IFEM_GPU_FP64) {
    rc = batch_assemble(c, IFEM_GPU_FP64, nullptr);
    if (rc) return rc;
  }
  if (B.fp32_ready || precision == IFEM_GPU_FP32) {
    rc = batch_assemble(c, IFEM_GPU_FP32, nullptr);
    if (rc) return rc;
  }
  const int lanes_c[3] = {16, 32, 64};
  for (int p = 1; p <= 3; p++) {
    if (!present[p]) continue;
    double best = 1e300;
    int best_l = default_lanes(p);
    for (int L : lanes_c) {
      B.tune[p].lanes = L;
      rc = upload_lanes(c);
      if (rc) return rc;
      IFEM_HIP(hipEventRecord(c.ev0, c.stream));
      for (int rep = 0; rep < 5; rep++) {
        if (precision == IFEM_GPU_FP32)
          launch_bspmv<float, float, false>(c, B.d_val32, B.x32, B.q32);
        else
          launch_bspmv<double, double, false>(c, B.d_val, B.x, B.q);
      }
      IFEM_HIP(hipGetLastError());
      IFEM_HIP(hipEventRecord(c.ev1, c.stream));
      IFEM_HIP(hipEventSynchronize(c.ev1));
      float ms = 0.f;
      IFEM_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
      if (ms < best) {
        best = ms;
        best_l = L;
      }
    }
    B.tune[p].lanes = best_l;
  }
  rc = upload_lanes(c);
  if (rc) return rc;
  B.tuned = true;
  if (table)
    for (int p = 0; p < 4; p++) {
      table[2 * p] = B.tune[p].half;
      table[2 * p + 1] = B.tune[p].lanes > 0 ? B.tune[p].lanes : (p > 0 ? default_lanes(p) : 0);
    }
  return IFEM_GPU_OK;
}

int batch_init(Context& c)
{
  int rc = batch_symbolic(c);
  if (rc) return rc;
  rc = batch_solver_setup(c);
  if (rc) return rc;
  c.batch->ready = true;
  return IFEM_GPU_OK;
}

} // namespace ifem
