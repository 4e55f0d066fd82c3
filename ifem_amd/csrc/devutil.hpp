// devutil.hpp -- device helpers shared by the solver kernels (solver.hip,
// batch.hip): fixed-order block reductions, fills, Jacobi diagonal.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ifem {

constexpr int kThreads = 256;
constexpr int kGrid = 2048; // fixed grid: partial sums have a fixed layout

enum Scal {
  S_RZ = 0, S_RZ1, S_RR, S_DONE, S_TOL2, S_ITER, S_MAXIT, S_BB,
  S_REJ, // iterations whose learned correction was dropped (flexible CG)
  S_NSCAL
};

template <typename A>
__device__ inline A block_sum(A v, A* sh)
{
  // fixed-shape reduction: wave butterflies, then wave totals in order;
  // the result is returned to every thread
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  A s = 0;
  for (int w = 0; w < kThreads / 64; w++) s += sh[w];
  return s;
}

// Sum of nparts partials in index order, computed by the whole block.
__device__ inline double sum_parts(const double* __restrict__ part, int nparts, double* sh)
{
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) s += part[i];
  return block_sum<double>(s, sh);
}

// Jacobi: inverse of the diagonal of A (CSR rows sorted: binary search).
template <typename V>
__global__ void k_diag_inv(int n, const int64_t* __restrict__ rowptr,
                           const int32_t* __restrict__ colind, const V* __restrict__ val,
                           V* __restrict__ dinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t lo = rowptr[i], hi = rowptr[i + 1] - 1;
  V d = 0;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int c = colind[mid];
    if (c == i) { d = val[mid]; break; }
    if (c < i) lo = mid + 1; else hi = mid - 1;
  }
  dinv[i] = d != V(0) ? V(1) / d : V(1);
}

template <typename V>
__global__ void k_fill(int n, V* __restrict__ v, V val)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = val;
}

} // namespace ifem
