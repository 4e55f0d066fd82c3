// solver.hip -- Jacobi-preconditioned CG on the device, FP64 and FP32-inner
// with FP64 iterative refinement.
//
// Stands in for SystemMatrix::solve with PETSc KSPCG + PCJACOBI or dune
// CGSolver + SeqJac (src/LinAlg/PETScMatrix.C:778-832, 980-990;
// src/LinAlg/ISTLMatrix.C:169-196), with the LinSolParams convergence
// contract (rtol, atol, maxits; src/LinAlg/LinSolParams.C:68-77).
//
// One iteration is three launches on a fixed grid:
//   k_spmv_dot : q = A p, per-block partials of <p,q>
//   k_update   : alpha = rz / sum(partials <p,q>); x += alpha p,
//                r -= alpha q, z = D^-1 r; per-block partials <r,z>, <r,r>
//   k_pupdate  : beta = sum(<r,z>) / rz; p = z + beta p; block 0 records the
//                scalars (rz, rr, iteration count, convergence flag)
// Every block that needs a global sum reduces the 2048 per-block partials
// itself, in index order, so all blocks agree bitwise and no grid-wide fence
// or atomic is needed (a device-scope fence per block forces L2 write-backs
// across the XCDs).  Scalars stay on the device; the host reads them every
// check_every iterations.  Results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include <hip/hip_fp16.h>

#include <type_traits>

#include "devutil.hpp"
#include "internal.hpp"

namespace ifem {

// ---------------------------------------------------------------------------
// SpMV q = A p (CSR), LANES lanes per row; with DOT, per-block partial <p,q>.
// SKIP: return at once when scal[S_DONE] is set (work queued past convergence).
// VV: type of the stored values (V, or FP16 values scaled by 1/scale).
template <typename V, typename ACC, int LANES, bool DOT, bool SKIP = DOT, typename VV = V>
__global__ __launch_bounds__(kThreads) void k_spmv_dot(
    int n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const VV* __restrict__ val, const V* __restrict__ p, V* __restrict__ q,
    double* __restrict__ part_pq, const double* __restrict__ scal, ACC scale)
{
  if (SKIP && scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const int groups_per_block = kThreads / LANES;
  const int g = threadIdx.x / LANES, l = threadIdx.x % LANES;
  double pq = 0.0;
  for (int row = blockIdx.x * groups_per_block + g; row < n; row += gridDim.x * groups_per_block) {
    const int64_t b = rowptr[row], e = rowptr[row + 1];
    ACC s = 0;
    for (int64_t k = b + l; k < e; k += LANES) s += (ACC)val[k] * (ACC)p[colind[k]];
#pragma unroll
    for (int off = LANES / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, LANES);
    if (!std::is_same<VV, V>::value) s *= scale;
    if (l == 0) {
      q[row] = (V)s;
      if (DOT) pq += (double)((ACC)p[row] * s);
    }
  }
  if (!DOT) return;
  const double bs = block_sum<double>(pq, sh);
  if (threadIdx.x == 0) part_pq[blockIdx.x] = bs;
}

// SpMV for 3-component systems whose DOFs are free or fixed per node: the
// three rows of a node share one column pattern made of node triples, so the
// gather reads one node index (bcol) per 9 values.  One wave per node; the
// CSR values are read in place.  Blocks are mapped to the 8 XCDs in
// contiguous node ranges so each XCD's L2 holds its slice of p.
template <typename V, typename ACC, bool DOT, bool SKIP = DOT, typename VV = V>
__global__ __launch_bounds__(kThreads) void k_spmv_node3(
    int nnode, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ bcol,
    const VV* __restrict__ val, const V* __restrict__ p, V* __restrict__ q,
    double* __restrict__ part_pq, const double* __restrict__ scal, ACC scale)
{
  if (SKIP && scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  constexpr int WPB = kThreads / 64;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nxcd = 8;
  const int xcd = blockIdx.x % nxcd, bl = blockIdx.x / nxcd, nbl = gridDim.x / nxcd;
  const int lo = (int)((int64_t)nnode * xcd / nxcd), hi = (int)((int64_t)nnode * (xcd + 1) / nxcd);
  double pq = 0.0;
  for (int nd = lo + bl * WPB + wave; nd < hi; nd += nbl * WPB) {
    const int64_t r0 = rowptr[3 * nd];
    const int cnt = (int)((rowptr[3 * nd + 1] - r0) / 3);
    const int64_t nb = r0 / 9;
    const VV* v0 = val + r0;
    const VV* v1 = v0 + 3 * cnt;
    const VV* v2 = v1 + 3 * cnt;
    ACC s0 = 0, s1 = 0, s2 = 0;
    for (int k = lane; k < cnt; k += 64) {
      const int64_t c = 3 * (int64_t)bcol[nb + k];
      const ACC x0 = (ACC)p[c], x1 = (ACC)p[c + 1], x2 = (ACC)p[c + 2];
      s0 += (ACC)v0[3 * k] * x0 + (ACC)v0[3 * k + 1] * x1 + (ACC)v0[3 * k + 2] * x2;
      s1 += (ACC)v1[3 * k] * x0 + (ACC)v1[3 * k + 1] * x1 + (ACC)v1[3 * k + 2] * x2;
      s2 += (ACC)v2[3 * k] * x0 + (ACC)v2[3 * k + 1] * x1 + (ACC)v2[3 * k + 2] * x2;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
    }
    if (!std::is_same<VV, V>::value) {
      s0 *= scale;
      s1 *= scale;
      s2 *= scale;
    }
    if (lane == 0) {
      q[3 * nd] = (V)s0;
      q[3 * nd + 1] = (V)s1;
      q[3 * nd + 2] = (V)s2;
      if (DOT)
        pq += (double)((ACC)p[3 * nd] * s0) + (double)((ACC)p[3 * nd + 1] * s1) +
              (double)((ACC)p[3 * nd + 2] * s2);
    }
  }
  if (!DOT) return;
  const double bs = block_sum<double>(pq, sh);
  if (threadIdx.x == 0) part_pq[blockIdx.x] = bs;
}

// FP16 values of a 3-component node-uniform system in node blocks: for node
// nd and its k-th neighbour node (k < cnt, the order of bcol), 12 halfs at
// ((rowptr[3nd]/9) + k) * 12: [row r of the node][column component j,
// padded with a zero to 4], so that each row's three values of a neighbour
// are one 8-byte load (instead of three 2-byte loads).  Built from the FP64
// CSR values times inv_scale, one wave per node.
__global__ void k_demote16_node3(int nnode, const int64_t* __restrict__ rowptr,
                                 const double* __restrict__ a, uint2* __restrict__ b,
                                 double inv_scale)
{
  const int lane = threadIdx.x & 63;
  const int nw = (int)(((int64_t)gridDim.x * blockDim.x) >> 6);
  for (int nd = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6); nd < nnode;
       nd += nw) {
    const int64_t r0 = rowptr[3 * nd];
    const int cnt = (int)((rowptr[3 * nd + 1] - r0) / 3);
    const int64_t nb = r0 / 9;
    for (int k = lane; k < cnt; k += 64) {
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const double* src = a + r0 + (int64_t)r * 3 * cnt + 3 * k;
        const uint32_t h0 = __half_as_ushort(__float2half((float)(src[0] * inv_scale)));
        const uint32_t h1 = __half_as_ushort(__float2half((float)(src[1] * inv_scale)));
        const uint32_t h2 = __half_as_ushort(__float2half((float)(src[2] * inv_scale)));
        b[(nb + k) * 3 + r] = make_uint2(h0 | (h1 << 16), h2);
      }
    }
  }
}

__device__ inline float lo_half(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
__device__ inline float hi_half(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

// q = scale * (K16 p) with the node-block FP16 values (k_demote16_node3), FP32
// sums: the fine-level smoother's SpMV.  Same wave-per-node and XCD mapping
// as k_spmv_node3.
template <bool SKIP>
__global__ __launch_bounds__(kThreads) void k_spmv_node3_h(
    int nnode, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ bcol,
    const uint2* __restrict__ val, const float* __restrict__ p, float* __restrict__ q,
    const double* __restrict__ scal, float scale)
{
  if (SKIP && scal[S_DONE] != 0.0) return;
  constexpr int WPB = kThreads / 64;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nxcd = 8;
  const int xcd = blockIdx.x % nxcd, bl = blockIdx.x / nxcd, nbl = gridDim.x / nxcd;
  const int lo = (int)((int64_t)nnode * xcd / nxcd), hi = (int)((int64_t)nnode * (xcd + 1) / nxcd);
  for (int nd = lo + bl * WPB + wave; nd < hi; nd += nbl * WPB) {
    const int64_t r0 = rowptr[3 * nd];
    const int cnt = (int)((rowptr[3 * nd + 1] - r0) / 3);
    const int64_t nb = r0 / 9;
    const uint2* vb = val + nb * 3;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int k = lane; k < cnt; k += 64) {
      const int64_t c = 3 * (int64_t)bcol[nb + k];
      const float x0 = p[c], x1 = p[c + 1], x2 = p[c + 2];
      const uint2 w0 = vb[3 * k], w1 = vb[3 * k + 1], w2 = vb[3 * k + 2];
      s0 += lo_half(w0.x) * x0 + hi_half(w0.x) * x1 + lo_half(w0.y) * x2;
      s1 += lo_half(w1.x) * x0 + hi_half(w1.x) * x1 + lo_half(w1.y) * x2;
      s2 += lo_half(w2.x) * x0 + hi_half(w2.x) * x1 + lo_half(w2.y) * x2;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
    }
    if (lane == 0) {
      q[3 * nd] = s0 * scale;
      q[3 * nd + 1] = s1 * scale;
      q[3 * nd + 2] = s2 * scale;
    }
  }
}

// bcol[r0/9 + k] = colind[r0 + 3k] / 3 for each node's first row.
__global__ void k_build_bcol(int nnode, const int64_t* __restrict__ rowptr,
                             const int32_t* __restrict__ colind, int32_t* __restrict__ bcol)
{
  const int nd = blockIdx.x;
  const int64_t r0 = rowptr[3 * nd];
  const int cnt = (int)((rowptr[3 * nd + 1] - r0) / 3);
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) bcol[r0 / 9 + k] = colind[r0 + 3 * k] / 3;
}

template <typename V>
__global__ __launch_bounds__(kThreads) void k_update(int n, V* __restrict__ x, V* __restrict__ r,
                                                     V* __restrict__ z, const V* __restrict__ p,
                                                     const V* __restrict__ q,
                                                     const V* __restrict__ dinv,
                                                     const double* __restrict__ part_pq,
                                                     double* __restrict__ part_rz,
                                                     double* __restrict__ part_rr,
                                                     double* __restrict__ scal, int parity)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double pq = sum_parts(part_pq, gridDim.x, sh);
  double a = 0.0;
  if (pq != 0.0 && __builtin_isfinite(pq)) a = scal[S_RZ + parity] / pq;
  else if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_DONE] = 2.0; // breakdown
  const V alpha = (V)a;
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const V xi = x[i] + alpha * p[i];
    const V ri = r[i] - alpha * q[i];
    const V zi = dinv[i] * ri;
    x[i] = xi;
    r[i] = ri;
    z[i] = zi;
    rz += (double)(ri * zi);
    rr += (double)(ri * ri);
  }
  const double brz = block_sum<double>(rz, sh);
  const double brr = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = brz;
    part_rr[blockIdx.x] = brr;
  }
}

template <typename V>
__global__ __launch_bounds__(kThreads) void k_pupdate(int n, V* __restrict__ p,
                                                      const V* __restrict__ z,
                                                      const double* __restrict__ part_rz,
                                                      const double* __restrict__ part_rr,
                                                      double* __restrict__ scal, int parity)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double rzn = sum_parts(part_rz, gridDim.x, sh);
  const double rr = sum_parts(part_rr, gridDim.x, sh);
  const double rzo = scal[S_RZ + parity];
  const bool conv = !(rr > scal[S_TOL2]); // converged (or NaN)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_RZ + 1 - parity] = rzn;
    scal[S_RR] = rr;
    scal[S_ITER] += 1.0;
    if (conv) scal[S_DONE] = 1.0;
    else if (scal[S_ITER] >= scal[S_MAXIT]) scal[S_DONE] = 3.0;
  }
  if (conv) return;
  const V beta = (V)(rzo != 0.0 ? rzn / rzo : 0.0);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = z[i] + beta * p[i];
}

// CG start from residual r (already in place): z = D^-1 r, p = z, partials.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_start(int n, const V* __restrict__ r,
                                                    V* __restrict__ z, V* __restrict__ p,
                                                    const V* __restrict__ dinv,
                                                    double* __restrict__ part_rz,
                                                    double* __restrict__ part_rr)
{
  __shared__ double sh[kThreads / 64];
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const V ri = r[i];
    const V zi = dinv[i] * ri;
    z[i] = zi;
    p[i] = zi;
    rz += (double)(ri * zi);
    rr += (double)(ri * ri);
  }
  const double brz = block_sum<double>(rz, sh);
  const double brr = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = brz;
    part_rr[blockIdx.x] = brr;
  }
}

// One block: scalars after k_start.
__global__ __launch_bounds__(kThreads) void k_start_check(const double* __restrict__ part_rz,
                                                          const double* __restrict__ part_rr,
                                                          int nparts, double* __restrict__ scal)
{
  __shared__ double sh[kThreads / 64];
  const double rz = sum_parts(part_rz, nparts, sh);
  const double rr = sum_parts(part_rr, nparts, sh);
  if (threadIdx.x == 0) {
    scal[S_RZ] = rz;
    scal[S_RR] = rr;
    scal[S_DONE] = (rr > scal[S_TOL2]) ? 0.0 : 1.0;
    if (scal[S_ITER] >= scal[S_MAXIT]) scal[S_DONE] = 3.0;
  }
}

// r = b - y (FP64), per-block partials of ||r||^2.
__global__ __launch_bounds__(kThreads) void k_residual(int n, const double* __restrict__ b,
                                                       const double* __restrict__ y,
                                                       double* __restrict__ r,
                                                       double* __restrict__ part)
{
  __shared__ double sh[kThreads / 64];
  double rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double ri = b[i] - y[i];
    r[i] = ri;
    rr += ri * ri;
  }
  const double brr = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = brr;
}

// One block: scal[slot] = sum of partials.
__global__ __launch_bounds__(kThreads) void k_sum_to(const double* __restrict__ part, int nparts,
                                                     double* __restrict__ scal, int slot)
{
  __shared__ double sh[kThreads / 64];
  const double t = sum_parts(part, nparts, sh);
  if (threadIdx.x == 0) scal[slot] = t;
}

// x64 += s * d32 ; r32 = (float)(r64 * inv)
__global__ void k_axpy_promote(int n, double* __restrict__ x, const float* __restrict__ d, double s)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * (double)d[i];
}

__global__ void k_demote_scaled(int n, const double* __restrict__ r, float* __restrict__ r32,
                                double inv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r32[i] = (float)(r[i] * inv);
}

// ---------------------------------------------------------------------------
static int lanes_for(const Context& c)
{
  const double avg = c.neq > 0 ? (double)c.nnz / c.neq : 0.0;
  if (avg > 96) return 64;
  if (avg > 40) return 32;
  return 16;
}

template <typename V, typename ACC, bool DOT, bool SKIP = DOT, typename VV = V>
static void launch_spmv(Context& c, const VV* val, const V* p, V* q, ACC scale = ACC(1),
                        hipStream_t st = nullptr)
{
  const int L = lanes_for(c);
  SolverWork& w = c.w;
  const int grid = kGrid;
  if (!st) st = c.stream;
  if (c.d_bcol) {
    k_spmv_node3<V, ACC, DOT, SKIP, VV><<<grid, kThreads, 0, st>>>(
        c.neq / 3, c.d_rowptr, c.d_bcol, val, p, q, w.partial, w.scal, scale);
    return;
  }
  if (L == 64)
    k_spmv_dot<V, ACC, 64, DOT, SKIP, VV><<<grid, kThreads, 0, st>>>(
        c.neq, c.d_rowptr, c.d_colind, val, p, q, w.partial, w.scal, scale);
  else if (L == 32)
    k_spmv_dot<V, ACC, 32, DOT, SKIP, VV><<<grid, kThreads, 0, st>>>(
        c.neq, c.d_rowptr, c.d_colind, val, p, q, w.partial, w.scal, scale);
  else
    k_spmv_dot<V, ACC, 16, DOT, SKIP, VV><<<grid, kThreads, 0, st>>>(
        c.neq, c.d_rowptr, c.d_colind, val, p, q, w.partial, w.scal, scale);
}

// FP32 SpMV for the multigrid smoother on the finest level.
void launch_spmv32(Context& c, const float* val, const float* p, float* q, bool skip_done)
{
  if (skip_done) launch_spmv<float, float, false, true>(c, val, p, q);
  else launch_spmv<float, float, false, false>(c, val, p, q);
}

// Same with FP16 values holding K / scale (y = scale * (val x), FP32 sums).
void launch_spmv16(Context& c, const __half* val, float scale, const float* p, float* q,
                   bool skip_done, hipStream_t st)
{
  if (skip_done) launch_spmv<float, float, false, true, __half>(c, val, p, q, scale, st);
  else launch_spmv<float, float, false, false, __half>(c, val, p, q, scale, st);
}

// The node-block FP16 copy of K / scale (3-component node-uniform systems).
int demote16_node3(Context& c, double inv_scale, uint2* out)
{
  k_demote16_node3<<<4096, 256, 0, c.stream>>>(c.neq / 3, c.d_rowptr, c.d_val, out, inv_scale);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

void launch_spmv16_node3(Context& c, const uint2* val, float scale, const float* p, float* q,
                         bool skip_done, hipStream_t st)
{
  if (!st) st = c.stream;
  if (skip_done)
    k_spmv_node3_h<true><<<kGrid, kThreads, 0, st>>>(c.neq / 3, c.d_rowptr, c.d_bcol, val, p, q,
                                                      c.w.scal, scale);
  else
    k_spmv_node3_h<false><<<kGrid, kThreads, 0, st>>>(c.neq / 3, c.d_rowptr, c.d_bcol, val, p, q,
                                                       c.w.scal, scale);
}

int spmv_device(Context& c, const double* d_x, double* d_y)
{
  if (c.neq == 0) return IFEM_GPU_OK;
  launch_spmv<double, double, false>(c, c.d_val, d_x, d_y);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

int spmv_device32(Context& c, const float* d_x, float* d_y)
{
  if (c.neq == 0) return IFEM_GPU_OK;
  launch_spmv<float, float, false>(c, c.d_val32, d_x, d_y);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

void solver_free(Context& c)
{
  SolverWork& w = c.w;
  double* d[] = {w.x, w.r, w.z, w.p, w.q, w.dinv, w.partial, w.scal};
  for (double* p : d) hipFree(p);
  float* f[] = {w.r32, w.z32, w.p32, w.q32, w.x32, w.dinv32};
  for (float* p : f) hipFree(p);
  if (w.h_scal) hipHostFree(w.h_scal);
  w = SolverWork();
}

int solver_setup(Context& c)
{
  solver_free(c);
  SolverWork& w = c.w;
  const size_t n = c.neq > 0 ? c.neq : 1;
  w.n = c.neq;
  double** d[] = {&w.x, &w.r, &w.z, &w.p, &w.q, &w.dinv};
  for (double** p : d) IFEM_HIP(hipMalloc(p, sizeof(double) * n));
  float** f[] = {&w.r32, &w.z32, &w.p32, &w.q32, &w.x32, &w.dinv32};
  for (float** p : f) IFEM_HIP(hipMalloc(p, sizeof(float) * n));
  IFEM_HIP(hipMalloc(&w.partial, sizeof(double) * 8 * kGrid));
  IFEM_HIP(hipMalloc(&w.scal, sizeof(double) * 16));
  IFEM_HIP(hipMemsetAsync(w.x, 0, sizeof(double) * n, c.stream));
  IFEM_HIP(hipHostMalloc(&w.h_scal, sizeof(double) * 16, hipHostMallocDefault));
  // node-triple column index for node-uniform 3-component systems
  hipFree(c.d_bcol);
  c.d_bcol = nullptr;
  bool uniform = c.ncomp == 3 && c.neq % 3 == 0 && c.neq > 0;
  for (int nd = 0; uniform && nd < c.ndof / 3; nd++) {
    const int e0 = c.h_eq[3 * nd];
    if (e0 >= 0) uniform = e0 % 3 == 0 && c.h_eq[3 * nd + 1] == e0 + 1 && c.h_eq[3 * nd + 2] == e0 + 2;
    else uniform = c.h_eq[3 * nd + 1] < 0 && c.h_eq[3 * nd + 2] < 0;
  }
  if (uniform) {
    IFEM_HIP(hipMalloc(&c.d_bcol, sizeof(int32_t) * (c.nnz / 9 + 1)));
    k_build_bcol<<<c.neq / 3, 128, 0, c.stream>>>(c.neq / 3, c.d_rowptr, c.d_colind, c.d_bcol);
    IFEM_HIP(hipGetLastError());
  }
  w.nblocks = kGrid;
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

static int read_scalars(Context& c)
{
  IFEM_HIP(hipMemcpyAsync(c.w.h_scal, c.w.scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost,
                          c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

static int write_scalars(Context& c, const double* v)
{
  std::memcpy(c.w.h_scal, v, sizeof(double) * S_NSCAL);
  IFEM_HIP(hipMemcpyAsync(c.w.scal, c.w.h_scal, sizeof(double) * S_NSCAL, hipMemcpyHostToDevice,
                          c.stream));
  return IFEM_GPU_OK;
}

// Preconditioned CG on A (values val, vector type V) for A x = r0 where r is
// preloaded with the right-hand side and x with zero.  Stops when
// ||r||^2 <= tol2 or after maxits iterations.  Returns iterations done.
template <typename V, typename ACC>
static int pcg_run(Context& c, const V* val, V* x, V* r, V* z, V* p, V* q, const V* dinv,
                   double tol2, int maxits, int check_every, int* iters, double* rr_out,
                   int* status)
{
  SolverWork& w = c.w;
  const int n = c.neq;
  double s[S_NSCAL] = {0};
  s[S_TOL2] = tol2;
  s[S_MAXIT] = maxits;
  int rc = write_scalars(c, s);
  if (rc) return rc;
  double* part_pq = w.partial;
  double* part_rz = w.partial + kGrid;
  double* part_rr = w.partial + 2 * kGrid;
  k_start<V><<<kGrid, kThreads, 0, c.stream>>>(n, r, z, p, dinv, part_rz, part_rr);
  k_start_check<<<1, kThreads, 0, c.stream>>>(part_rz, part_rr, kGrid, w.scal);
  IFEM_HIP(hipGetLastError());
  int launched = 0;
  while (true) {
    for (int k = 0; k < check_every; k++) {
      const int parity = (launched + k) & 1;
      launch_spmv<V, ACC, true>(c, val, p, q);
      k_update<V><<<kGrid, kThreads, 0, c.stream>>>(n, x, r, z, p, q, dinv, part_pq, part_rz,
                                                    part_rr, w.scal, parity);
      k_pupdate<V><<<kGrid, kThreads, 0, c.stream>>>(n, p, z, part_rz, part_rr, w.scal, parity);
    }
    IFEM_HIP(hipGetLastError());
    launched += check_every;
    rc = read_scalars(c);
    if (rc) return rc;
    if (w.h_scal[S_DONE] != 0.0 || launched >= maxits + check_every) break;
  }
  *iters = (int)w.h_scal[S_ITER];
  *rr_out = w.h_scal[S_RR];
  *status = (int)w.h_scal[S_DONE];
  return IFEM_GPU_OK;
}

// ---------------------------------------------------------------------------
// Flexible CG with the learned preconditioner z = D^-1 r + s*|r|_rms*U(r/|r|_rms)
// (Polak-Ribiere beta = <r_new - r_old, z_new> / <r_old, z_old>).

// x += alpha p; r_old = r; r -= alpha q; partials ||r||^2
template <typename V>
__global__ __launch_bounds__(kThreads) void k_update_f(int n, V* __restrict__ x, V* __restrict__ r,
                                                       V* __restrict__ rold,
                                                       const V* __restrict__ p,
                                                       const V* __restrict__ q,
                                                       const double* __restrict__ part_pq,
                                                       double* __restrict__ part_rr,
                                                       double* __restrict__ scal, int parity)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double pq = sum_parts(part_pq, gridDim.x, sh);
  double a = 0.0;
  if (pq != 0.0 && __builtin_isfinite(pq)) a = scal[S_RZ + parity] / pq;
  else if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_DONE] = 2.0;
  const V alpha = (V)a;
  double rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const V ro = r[i];
    const V rn = ro - alpha * q[i];
    rold[i] = ro;
    r[i] = rn;
    x[i] += alpha * p[i];
    rr += (double)(rn * rn);
  }
  const double brr = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = brr;
}

// partials of ||r||^2 (start of flexible CG)
template <typename V>
__global__ __launch_bounds__(kThreads) void k_norm_parts(int n, const V* __restrict__ r,
                                                         double* __restrict__ part_rr)
{
  __shared__ double sh[kThreads / 64];
  double rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    rr += (double)(r[i] * r[i]);
  const double b = block_sum<double>(rr, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = b;
}

// network input: channel c of node k = r[k*nc + c] / |r|_rms; with feat 1
// also r^ * aux_j in channel nc + j*nc + c
// Free-DOF box (bw x bh x bd nodes, i fastest) inside the zero-padded
// D x H x W lattice: node index -> lattice index.
struct LatMap {
  int bw, bh, W, H;
  int64_t vol; // D*H*W
  __device__ int64_t at(int node) const
  {
    const int ii = node % bw, r = node / bw;
    const int jj = r % bh, kk = r / bh;
    return ((int64_t)kk * H + jj) * W + ii;
  }
};

template <typename V>
__global__ __launch_bounds__(kThreads) void k_to_lattice(int n, int nc, const V* __restrict__ r,
                                                         float* __restrict__ netin,
                                                         const float* __restrict__ aux, int naux,
                                                         int feat, LatMap lm,
                                                         const double* __restrict__ part_rr,
                                                         int nparts, const double* __restrict__ scal)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double rr = sum_parts(part_rr, nparts, sh);
  const double inv = rr > 0.0 ? 1.0 / sqrt(rr / n) : 0.0;
  const int64_t vol = lm.vol;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int node = i / nc, comp = i - node * nc;
    const int64_t li = lm.at(node);
    const float v = (float)((double)r[i] * inv);
    netin[(int64_t)comp * vol + li] = v;
    if (feat)
      for (int j = 0; j < naux; j++)
        netin[(int64_t)(nc + j * nc + comp) * vol + li] = v * aux[(int64_t)j * vol + li];
  }
}

// z = D^-1 r + s*|r|_rms*out ; partials <r,z> and <r_old,z>
template <typename V>
__global__ __launch_bounds__(kThreads) void k_from_lattice(
    int n, int nc, const V* __restrict__ r, const V* __restrict__ rold, const V* __restrict__ dinv,
    const float* __restrict__ out, V* __restrict__ z, double scale, LatMap lm,
    const double* __restrict__ part_rr, int nparts, double* __restrict__ part_rz,
    double* __restrict__ part_roz, const double* __restrict__ scal)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double rr = sum_parts(part_rr, nparts, sh);
  const double rms = sqrt(rr / n);
  const int64_t vol = lm.vol;
  double rz = 0.0, roz = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int node = i / nc, comp = i - node * nc;
    const V zi = dinv[i] * r[i] + (V)(scale * rms * (double)out[(int64_t)comp * vol + lm.at(node)]);
    z[i] = zi;
    rz += (double)(r[i] * zi);
    if (rold) roz += (double)(rold[i] * zi);
  }
  const double b1 = block_sum<double>(rz, sh);
  const double b2 = block_sum<double>(roz, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = b1;
    part_roz[blockIdx.x] = b2;
  }
}

// one block: scalars at the start of flexible CG
__global__ __launch_bounds__(kThreads) void k_start_check_f(const double* __restrict__ part_rz,
                                                            const double* __restrict__ part_rr,
                                                            int nparts, double* __restrict__ scal)
{
  __shared__ double sh[kThreads / 64];
  const double rz = sum_parts(part_rz, nparts, sh);
  const double rr = sum_parts(part_rr, nparts, sh);
  if (threadIdx.x == 0) {
    scal[S_RZ] = rz;
    scal[S_RR] = rr;
    scal[S_DONE] = (rr > scal[S_TOL2]) ? 0.0 : 1.0;
    if (!(rz > 0.0)) scal[S_DONE] = 2.0; // preconditioner not positive on r
    if (scal[S_ITER] >= scal[S_MAXIT]) scal[S_DONE] = 3.0;
  }
}

template <typename V>
__global__ void k_copy(int n, const V* __restrict__ a, V* __restrict__ b)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

template <typename V>
__global__ __launch_bounds__(kThreads) void k_pupdate_f(int n, V* __restrict__ p,
                                                        const V* __restrict__ z,
                                                        const double* __restrict__ part_rz,
                                                        const double* __restrict__ part_roz,
                                                        const double* __restrict__ part_rr,
                                                        double* __restrict__ scal, int parity)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double rzn = sum_parts(part_rz, gridDim.x, sh);
  const double roz = sum_parts(part_roz, gridDim.x, sh);
  const double rr = sum_parts(part_rr, gridDim.x, sh);
  const double rzo = scal[S_RZ + parity];
  const bool conv = !(rr > scal[S_TOL2]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_RZ + 1 - parity] = rzn;
    scal[S_RR] = rr;
    scal[S_ITER] += 1.0;
    if (conv) scal[S_DONE] = 1.0;
    else if (!(rzn > 0.0)) scal[S_DONE] = 2.0;
    else if (scal[S_ITER] >= scal[S_MAXIT]) scal[S_DONE] = 3.0;
  }
  if (conv) return;
  const V beta = (V)(rzo != 0.0 ? (rzn - roz) / rzo : 0.0);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = z[i] + beta * p[i];
}

// Learned correction zn = s*|r|_rms*U(r/|r|_rms, aux), mapped back from the
// lattice; with a Jacobi base (dinv != null) also z0 = D^-1 r, otherwise z0
// holds the base preconditioner's output already.  Per-block partials (part4,
// kGrid each): <r,z0>, <r_old,z0>, <r,zn>, <r_old,zn>.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_from_lattice_l(
    int n, int nc, const V* __restrict__ r, const V* __restrict__ rold, const V* __restrict__ dinv,
    V* __restrict__ z0, const float* __restrict__ out, V* __restrict__ zn, double scale, LatMap lm,
    const double* __restrict__ part_rr, double* __restrict__ part4, const double* __restrict__ scal)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double rr = sum_parts(part_rr, kGrid, sh);
  const double rms = sqrt(rr / n);
  const int64_t vol = lm.vol;
  double rz0 = 0.0, roz0 = 0.0, rzn = 0.0, rozn = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int node = i / nc, comp = i - node * nc;
    V a;
    if (dinv) {
      a = dinv[i] * r[i];
      z0[i] = a;
    } else {
      a = z0[i];
    }
    const V b = (V)(scale * rms * (double)out[(int64_t)comp * vol + lm.at(node)]);
    zn[i] = b;
    rz0 += (double)(r[i] * a);
    rzn += (double)(r[i] * b);
    if (rold) {
      roz0 += (double)(rold[i] * a);
      rozn += (double)(rold[i] * b);
    }
  }
  const double b1 = block_sum<double>(rz0, sh);
  const double b2 = block_sum<double>(roz0, sh);
  const double b3 = block_sum<double>(rzn, sh);
  const double b4 = block_sum<double>(rozn, sh);
  if (threadIdx.x == 0) {
    part4[blockIdx.x] = b1;
    part4[kGrid + blockIdx.x] = b2;
    part4[2 * kGrid + blockIdx.x] = b3;
    part4[3 * kGrid + blockIdx.x] = b4;
  }
}

// The learned correction is kept in an iteration when the combined
// preconditioner stays clearly positive on r: <r, z0 + zn> >= theta <r, z0>.
// Otherwise that iteration uses z0 alone (counted in scal[S_REJ]); flexible
// CG then never breaks down because of the network.
constexpr double kLearnedTheta = 0.1;

// p update of flexible CG with the learned correction (first: p = z at the
// start, no beta); block 0 records the scalars.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_pupdate_l(int n, V* __restrict__ p,
                                                        const V* __restrict__ z0,
                                                        const V* __restrict__ zn,
                                                        const double* __restrict__ part4,
                                                        const double* __restrict__ part_rr,
                                                        double* __restrict__ scal, int parity,
                                                        int first)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  const double rz0 = sum_parts(part4, kGrid, sh);
  const double roz0 = sum_parts(part4 + kGrid, kGrid, sh);
  const double rzn = sum_parts(part4 + 2 * kGrid, kGrid, sh);
  const double rozn = sum_parts(part4 + 3 * kGrid, kGrid, sh);
  const double rr = sum_parts(part_rr, kGrid, sh);
  const bool use = __builtin_isfinite(rzn) && __builtin_isfinite(rozn) &&
                   rz0 + rzn >= kLearnedTheta * rz0;
  const double rz = use ? rz0 + rzn : rz0;
  const double roz = use ? roz0 + rozn : roz0;
  const bool conv = !(rr > scal[S_TOL2]);
  if (first) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      scal[S_RZ] = rz;
      scal[S_RR] = rr;
      scal[S_DONE] = conv ? 1.0 : 0.0;
      if (!conv && !(rz > 0.0)) scal[S_DONE] = 2.0; // base preconditioner not positive on r
      if (scal[S_ITER] >= scal[S_MAXIT]) scal[S_DONE] = 3.0;
      if (!use) scal[S_REJ] += 1.0;
    }
    if (conv || !(rz > 0.0)) return;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
      p[i] = use ? z0[i] + zn[i] : z0[i];
    return;
  }
  const double rzo = scal[S_RZ + parity];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_RZ + 1 - parity] = rz;
    scal[S_RR] = rr;
    scal[S_ITER] += 1.0;
    if (conv) scal[S_DONE] = 1.0;
    else if (!(rz > 0.0)) scal[S_DONE] = 2.0;
    else if (scal[S_ITER] >= scal[S_MAXIT]) scal[S_DONE] = 3.0;
    if (!use) scal[S_REJ] += 1.0;
  }
  if (conv) return;
  const V beta = (V)(rzo != 0.0 ? (rz - roz) / rzo : 0.0);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = (use ? z0[i] + zn[i] : z0[i]) + beta * p[i];
}

// z0 (base: Jacobi inside k_from_lattice_l, or one multigrid V-cycle) and
// the learned correction zn, with the partials of k_pupdate_l
template <typename V>
static int apply_learned(Context& c, bool mg_base, const V* r, const V* rold, const V* dinv, V* z0,
                         V* zn, double* part_rr, double* part4)
{
  const int n = c.neq;
  if (mg_base) {
    int rc = mg_apply(c, r, z0, c.w.scal);
    if (rc) return rc;
  }
  LatMap lm;
  lm.bw = c.box[2];
  lm.bh = c.box[1];
  lm.W = c.lat[2];
  lm.H = c.lat[1];
  lm.vol = (int64_t)c.lat[0] * c.lat[1] * c.lat[2];
  k_to_lattice<V><<<kGrid, kThreads, 0, c.stream>>>(n, c.ncomp, r, c.d_netin, c.d_aux, c.naux,
                                                    c.features, lm, part_rr, kGrid, c.w.scal);
  const float* out = c.net->forward(c.d_netin, 1, c.lat[0], c.lat[1], c.lat[2],
                                   c.w.scal + S_DONE);
  k_from_lattice_l<V><<<kGrid, kThreads, 0, c.stream>>>(n, c.ncomp, r, rold,
                                                        mg_base ? (const V*)nullptr : dinv, z0, out,
                                                        zn, c.learned_scale, lm, part_rr, part4,
                                                        c.w.scal);
  return IFEM_GPU_OK;
}

// partials <r,z> and <r_old,z> after a preconditioner that does not form them
template <typename V>
__global__ __launch_bounds__(kThreads) void k_pc_dots(int n, const V* __restrict__ r,
                                                      const V* __restrict__ rold,
                                                      const V* __restrict__ z,
                                                      double* __restrict__ part_rz,
                                                      double* __restrict__ part_roz,
                                                      const double* __restrict__ scal)
{
  if (scal[S_DONE] != 0.0) return;
  __shared__ double sh[kThreads / 64];
  double rz = 0.0, roz = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double zi = (double)z[i];
    rz += (double)r[i] * zi;
    if (rold) roz += (double)rold[i] * zi;
  }
  const double b1 = block_sum<double>(rz, sh);
  const double b2 = block_sum<double>(roz, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = b1;
    part_roz[blockIdx.x] = b2;
  }
}

enum PcKind { PC_LEARNED = 0, PC_MULTIGRID = 1, PC_LEARNED_MG = 2 };

// Flexible CG with the spline multigrid (PC_MULTIGRID) or with a learned
// correction on a Jacobi (PC_LEARNED) or multigrid (PC_LEARNED_MG) base.
// zn: the learned correction (learned kinds only).  *rejected: iterations
// whose learned correction the safeguard dropped.
template <typename V, typename ACC>
static int fcg_run(Context& c, int kind, const V* val, V* x, V* r, V* z, V* zn, V* p, V* q,
                   V* rold, const V* dinv, double tol2, int maxits, int check_every, int* iters,
                   double* rr_out, int* status, int* rejected)
{
  SolverWork& w = c.w;
  const int n = c.neq;
  double s[S_NSCAL] = {0};
  s[S_TOL2] = tol2;
  s[S_MAXIT] = maxits;
  int rc = write_scalars(c, s);
  if (rc) return rc;
  const bool learned = kind != PC_MULTIGRID;
  const bool mg_base = kind != PC_LEARNED;
  double* part_pq = w.partial;
  double* part_rz = w.partial + kGrid;
  double* part_rr = w.partial + 2 * kGrid;
  double* part_roz = w.partial + 3 * kGrid;
  double* part4 = w.partial + 4 * kGrid;
  k_norm_parts<V><<<kGrid, kThreads, 0, c.stream>>>(n, r, part_rr);
  if (learned) {
    rc = apply_learned<V>(c, mg_base, r, (const V*)nullptr, dinv, z, zn, part_rr, part4);
    if (rc) return rc;
    k_pupdate_l<V><<<kGrid, kThreads, 0, c.stream>>>(n, p, z, zn, part4, part_rr, w.scal, 0, 1);
  } else {
    rc = mg_apply(c, r, z, w.scal);
    if (rc) return rc;
    k_pc_dots<V><<<kGrid, kThreads, 0, c.stream>>>(n, r, (const V*)nullptr, z, part_rz, part_roz,
                                                   w.scal);
    k_start_check_f<<<1, kThreads, 0, c.stream>>>(part_rz, part_rr, kGrid, w.scal);
    k_copy<V><<<(n + 255) / 256, 256, 0, c.stream>>>(n, z, p);
  }
  IFEM_HIP(hipGetLastError());
  int launched = 0;
  while (true) {
    for (int k = 0; k < check_every; k++) {
      const int parity = (launched + k) & 1;
      launch_spmv<V, ACC, true>(c, val, p, q);
      k_update_f<V><<<kGrid, kThreads, 0, c.stream>>>(n, x, r, rold, p, q, part_pq, part_rr,
                                                      w.scal, parity);
      if (learned) {
        rc = apply_learned<V>(c, mg_base, r, rold, dinv, z, zn, part_rr, part4);
        if (rc) return rc;
        k_pupdate_l<V><<<kGrid, kThreads, 0, c.stream>>>(n, p, z, zn, part4, part_rr, w.scal,
                                                         parity, 0);
      } else {
        rc = mg_apply(c, r, z, w.scal);
        if (rc) return rc;
        k_pc_dots<V><<<kGrid, kThreads, 0, c.stream>>>(n, r, rold, z, part_rz, part_roz, w.scal);
        k_pupdate_f<V><<<kGrid, kThreads, 0, c.stream>>>(n, p, z, part_rz, part_roz, part_rr,
                                                         w.scal, parity);
      }
    }
    IFEM_HIP(hipGetLastError());
    launched += check_every;
    rc = read_scalars(c);
    if (rc) return rc;
    if (w.h_scal[S_DONE] != 0.0 || launched >= maxits + check_every) break;
  }
  *iters = (int)w.h_scal[S_ITER];
  *rr_out = w.h_scal[S_RR];
  *status = (int)w.h_scal[S_DONE];
  *rejected = (int)w.h_scal[S_REJ];
  return IFEM_GPU_OK;
}

static int true_residual(Context& c, const double* x, double* r, double* rr)
{
  SolverWork& w = c.w;
  launch_spmv<double, double, false>(c, c.d_val, x, w.q);
  k_residual<<<kGrid, kThreads, 0, c.stream>>>(c.neq, c.d_rhs, w.q, r, w.partial);
  k_sum_to<<<1, kThreads, 0, c.stream>>>(w.partial, kGrid, w.scal, S_BB);
  IFEM_HIP(hipGetLastError());
  int rc = read_scalars(c);
  if (rc) return rc;
  *rr = w.h_scal[S_BB];
  return IFEM_GPU_OK;
}

static int norm2_rhs(Context& c, double* bb)
{
  SolverWork& w = c.w;
  k_fill<double><<<(c.neq + 255) / 256, 256, 0, c.stream>>>(c.neq, w.q, 0.0);
  k_residual<<<kGrid, kThreads, 0, c.stream>>>(c.neq, c.d_rhs, w.q, w.z, w.partial);
  k_sum_to<<<1, kThreads, 0, c.stream>>>(w.partial, kGrid, w.scal, S_BB);
  IFEM_HIP(hipGetLastError());
  int rc = read_scalars(c);
  if (rc) return rc;
  *bb = w.h_scal[S_BB];
  return IFEM_GPU_OK;
}

int solve(Context& c, const ifem_gpu_solver_params& prm, ifem_gpu_stats* st)
{
  SolverWork& w = c.w;
  const int n = c.neq;
  if (prm.maxits < 1) return fail(IFEM_GPU_EINVAL, "solve: maxits must be positive");
  if (!(prm.rtol >= 0.0) || !(prm.atol >= 0.0))
    return fail(IFEM_GPU_EINVAL, "solve: tolerances must be non-negative");
  const int check = prm.check_every > 0 ? prm.check_every : 32;
  int rc;
  IFEM_HIP(hipEventRecord(c.ev0, c.stream));
  double bb = 0.0;
  rc = norm2_rhs(c, &bb);
  if (rc) return rc;
  const double bnorm = std::sqrt(bb);
  const double tol = std::fmax(prm.rtol * bnorm, prm.atol);
  const int nb = (n + 255) / 256;
  k_diag_inv<double><<<nb, 256, 0, c.stream>>>(n, c.d_rowptr, c.d_colind, c.d_val, w.dinv);
  k_fill<double><<<nb, 256, 0, c.stream>>>(n, w.x, 0.0);
  IFEM_HIP(hipGetLastError());

  const bool use_mg = prm.pc == IFEM_GPU_PC_MG;
  if (prm.pc != IFEM_GPU_PC_JACOBI && !use_mg) return fail(IFEM_GPU_EINVAL, "solve: unknown pc");
  int check_pc = check;
  if (use_mg) {
    rc = mg_setup(c, prm.mg_smooth > 0 ? prm.mg_smooth : 2);
    if (rc) return rc;
    if (!c.d_rold) IFEM_HIP(hipMalloc(&c.d_rold, sizeof(double) * (n > 0 ? n : 1)));
    if (!c.d_rold32) IFEM_HIP(hipMalloc(&c.d_rold32, sizeof(float) * (n > 0 ? n : 1)));
    check_pc = std::min(check, 8); // few, expensive iterations
  }
  const int pc_kind = c.learned ? (use_mg ? PC_LEARNED_MG : PC_LEARNED) : PC_MULTIGRID;
  if (c.learned) {
    if (!c.d_zn) IFEM_HIP(hipMalloc(&c.d_zn, sizeof(double) * (n > 0 ? n : 1)));
    if (!c.d_zn32) IFEM_HIP(hipMalloc(&c.d_zn32, sizeof(float) * (n > 0 ? n : 1)));
    if (!c.d_rold) IFEM_HIP(hipMalloc(&c.d_rold, sizeof(double) * (n > 0 ? n : 1)));
    if (!c.d_rold32) IFEM_HIP(hipMalloc(&c.d_rold32, sizeof(float) * (n > 0 ? n : 1)));
  }
  int total_iters = 0, outer = 0, pc_iters = 0, rejected = 0;
  bool use_learned = c.learned || use_mg, fell_back = false;
  double rel = bnorm > 0 ? 1.0 : 0.0;
  bool converged = bnorm == 0.0;
  bool diverged = false;
  if (!converged && prm.precision == IFEM_GPU_FP64) {
    // plain FP64 Jacobi-PCG with true-residual confirmation / replacement
    IFEM_HIP(hipMemcpyAsync(w.r, c.d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
    while (true) {
      int it = 0, status = 0;
      double rr = 0.0;
      int rej = 0;
      if (use_learned)
        rc = fcg_run<double, double>(c, pc_kind, c.d_val, w.x, w.r, w.z, c.d_zn, w.p, w.q,
                                     c.d_rold, w.dinv, tol * tol, prm.maxits - total_iters,
                                     check_pc, &it, &rr, &status, &rej);
      else
        rc = pcg_run<double, double>(c, c.d_val, w.x, w.r, w.z, w.p, w.q, w.dinv, tol * tol,
                                     prm.maxits - total_iters, check, &it, &rr, &status);
      if (rc) return rc;
      total_iters += it;
      if (use_learned) pc_iters += it;
      rejected += rej;
      outer++;
      double trr = 0.0;
      rc = true_residual(c, w.x, w.r, &trr); // w.r <- b - A x (replacement)
      if (rc) return rc;
      rel = std::sqrt(trr) / bnorm;
      if (std::sqrt(trr) <= tol) { converged = true; break; }
      if (!(std::sqrt(trr) <= prm.dtol * bnorm)) { diverged = true; break; }
      if (total_iters >= prm.maxits) break;
      if (status == 2 || it == 0) {
        // breakdown: the learned preconditioner falls back to Jacobi-PCG,
        // restarted from the current iterate's true residual
        if (!use_learned) break;
        use_learned = false;
        fell_back = true;
      }
    }
  } else if (!converged) {
    // FP32 inner PCG inside FP64 iterative refinement
    if (!c.fp32_ready) return fail(IFEM_GPU_EINVAL, "solve: no FP32 values");
    k_diag_inv<float><<<nb, 256, 0, c.stream>>>(n, c.d_rowptr, c.d_colind, c.d_val32, w.dinv32);
    IFEM_HIP(hipMemcpyAsync(w.r, c.d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
    double rr64 = bb;
    const double inner = prm.inner_rtol > 0 ? prm.inner_rtol : 1e-4;
    while (total_iters < prm.maxits && outer < 100) {
      const double rn = std::sqrt(rr64);
      k_demote_scaled<<<nb, 256, 0, c.stream>>>(n, w.r, w.r32, 1.0 / rn);
      k_fill<float><<<nb, 256, 0, c.stream>>>(n, w.x32, 0.0f);
      int it = 0, status = 0;
      double rr = 0.0;
      // residual target relative to the scaled (unit-norm) residual; never
      // ask the inner solve for more than the outer tolerance needs
      double target = std::fmax(inner, 0.5 * tol / rn);
      int rej = 0;
      if (use_learned)
        rc = fcg_run<float, float>(c, pc_kind, c.d_val32, w.x32, w.r32, w.z32, c.d_zn32, w.p32,
                                   w.q32, c.d_rold32, w.dinv32, target * target,
                                   prm.maxits - total_iters, check_pc, &it, &rr, &status, &rej);
      else
        rc = pcg_run<float, float>(c, c.d_val32, w.x32, w.r32, w.z32, w.p32, w.q32, w.dinv32,
                                   target * target, prm.maxits - total_iters, check, &it, &rr,
                                   &status);
      if (rc) return rc;
      total_iters += it;
      if (use_learned) pc_iters += it;
      rejected += rej;
      outer++;
      k_axpy_promote<<<nb, 256, 0, c.stream>>>(n, w.x, w.x32, rn);
      rc = true_residual(c, w.x, w.r, &rr64);
      if (rc) return rc;
      rel = std::sqrt(rr64) / bnorm;
      if (std::sqrt(rr64) <= tol) { converged = true; break; }
      if (!(std::sqrt(rr64) <= prm.dtol * bnorm)) { diverged = true; break; }
      if (it == 0 || status == 2) {
        if (!use_learned) break;
        use_learned = false;
        fell_back = true;
      }
    }
  }
  IFEM_HIP(hipEventRecord(c.ev1, c.stream));
  IFEM_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  IFEM_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  c.last_solve_ms = ms;
  if (st) {
    st->iterations = total_iters;
    st->outer_iterations = outer;
    st->converged = converged ? 1 : 0;
    st->rel_residual = rel;
    st->solve_ms = ms;
    st->pc_iterations = pc_iters;
    st->fallback = fell_back ? 1 : 0;
    st->pc_rejected = rejected;
  }
  c.last_fallback = fell_back;
  if (diverged) return fail(IFEM_GPU_ENOCONV, "solve: diverged (residual above dtol*||b||)");
  if (!converged)
    return fail(IFEM_GPU_ENOCONV, "solve: not converged after " + std::to_string(total_iters) +
                                      " iterations, relative residual " + std::to_string(rel));
  return IFEM_GPU_OK;
}

int set_learned(Context& c, int enable, int D, int H, int W, const float* aux, int naux,
                int features, double scale)
{
  if (!enable) {
    c.learned = false;
    return IFEM_GPU_OK;
  }
  if (!c.multi.empty() || c.mpc)
    return fail(IFEM_GPU_EUNSUPPORTED,
                "set_learned_preconditioner: one patch without master constraints only");
  if (!c.net) return fail(IFEM_GPU_EINVAL, "set_learned_preconditioner: create the network first");
  if (!c.pattern_ready) return fail(IFEM_GPU_EINVAL, "set_learned_preconditioner: no system");
  // box of the nodes with free DOFs; all components free inside, none outside
  const PatchRec& Pt = c.patches[c.patch];
  const int n1 = Pt.n[0], n2 = Pt.n[1], n3 = Pt.n[2], nc = c.ncomp;
  int lo[3] = {n1, n2, n3}, hi[3] = {-1, -1, -1};
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        const int64_t node = (int64_t)i + (int64_t)n1 * (j + (int64_t)n2 * k);
        if (c.h_eq[node * nc] < 0) continue;
        const int idx[3] = {i, j, k};
        for (int d = 0; d < 3; d++) {
          lo[d] = std::min(lo[d], idx[d]);
          hi[d] = std::max(hi[d], idx[d]);
        }
      }
  const int bw = hi[0] - lo[0] + 1, bh = hi[1] - lo[1] + 1, bd = hi[2] - lo[2] + 1;
  if (bw < 1 || (int64_t)bw * bh * bd * nc != c.neq)
    return fail(IFEM_GPU_EUNSUPPORTED, "set_learned_preconditioner: the free DOFs are not a box");
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        const int64_t node = (int64_t)i + (int64_t)n1 * (j + (int64_t)n2 * k);
        for (int cc = 0; cc < nc; cc++)
          if (c.h_eq[node * nc + cc] < 0)
            return fail(IFEM_GPU_EUNSUPPORTED,
                        "set_learned_preconditioner: the free DOFs are not a box");
      }
  if (D < bd || H < bh || W < bw)
    return fail(IFEM_GPU_EINVAL, "set_learned_preconditioner: lattice smaller than the free box");
  if (naux < 0 || (naux > 0 && !aux)) return fail(IFEM_GPU_EINVAL, "set_learned_preconditioner: aux");
  if (features != 0 && features != 1)
    return fail(IFEM_GPU_EINVAL, "set_learned_preconditioner: features must be 0 or 1");
  const int cin = features ? c.ncomp * (1 + naux) : c.ncomp + naux;
  if (c.net->cfg.in_channels != cin || c.net->cfg.out_channels != c.ncomp)
    return fail(IFEM_GPU_EINVAL, "set_learned_preconditioner: network channels do not match");
  int rc = c.net->ensure(1, D, H, W);
  if (rc) return rc;
  const int64_t vol = (int64_t)D * H * W;
  hipFree(c.d_netin);
  hipFree(c.d_rold);
  hipFree(c.d_rold32);
  hipFree(c.d_aux);
  c.d_netin = nullptr;
  c.d_rold = nullptr;
  c.d_rold32 = nullptr;
  c.d_aux = nullptr;
  IFEM_HIP(hipMalloc(&c.d_netin, sizeof(float) * cin * vol));
  IFEM_HIP(hipMemsetAsync(c.d_netin, 0, sizeof(float) * cin * vol, c.stream)); // padding stays 0
  if (naux > 0) {
    IFEM_HIP(hipMalloc(&c.d_aux, sizeof(float) * naux * vol));
    IFEM_HIP(hipMemcpyAsync(c.d_aux, aux, sizeof(float) * naux * vol, hipMemcpyDeviceToDevice,
                            c.stream));
  }
  IFEM_HIP(hipMalloc(&c.d_rold, sizeof(double) * c.neq));
  IFEM_HIP(hipMalloc(&c.d_rold32, sizeof(float) * c.neq));
  if (naux > 0 && !features)
    IFEM_HIP(hipMemcpyAsync(c.d_netin + c.ncomp * vol, aux, sizeof(float) * naux * vol,
                            hipMemcpyDeviceToDevice, c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  c.lat[0] = D;
  c.lat[1] = H;
  c.lat[2] = W;
  c.box[0] = bd;
  c.box[1] = bh;
  c.box[2] = bw;
  c.naux = naux;
  c.features = features;
  c.learned_scale = scale;
  c.learned = true;
  return IFEM_GPU_OK;
}

} // namespace ifem
