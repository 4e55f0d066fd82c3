// pattern.hip -- sparsity pattern of several patches sharing nodes, built on
// the device: the union over the patches of each free row's window columns
// (SAM::getDofCouplings over the elements of all patches,
// src/LinAlg/SAM.C:377-427, then SparseMatrix::optimiseCols,
// src/LinAlg/SparseMatrix.C:984-1023), and each patch's map from (local row,
// window column) to the position in the shared CSR that the line kernels
// write through in accumulate mode.
//
// 1. per patch, per local DOF with a free row: the count of its valid window
//    columns, scanned into the patch's map offsets (SubPatch::d_pptr);
// 2. every (row, column) pair of every patch as one 64-bit key
//    row << 32 | column; radix sort; unique -> the CSR in row-major order
//    (sorted columns), rowptr from the row changes;
// 3. per patch, per (local row, window column): binary search of the column
//    in its row -> SubPatch::d_pmap.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <cstdint>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ifem {

struct PatchGrid {
  int n1, n2, n3, p, nc;
  const int* leq; // local DOF -> global equation (0-based) or -1
};

// Calls f(k, e) for the valid window columns e of local DOF d in the order
// (dk, dj, di) ascending, component minor, k counting from 0; returns the
// count.  The order is the one the line kernels walk a row's window in.
template <typename F>
__device__ inline int for_window(const PatchGrid& g, int64_t d, F f)
{
  const int nc = g.nc;
  const int64_t node = d / nc;
  const int i = (int)(node % g.n1), j = (int)((node / g.n1) % g.n2),
            k = (int)(node / ((int64_t)g.n1 * g.n2));
  int cnt = 0;
  for (int kk = k - g.p; kk <= k + g.p; kk++) {
    if (kk < 0 || kk >= g.n3) continue;
    for (int jj = j - g.p; jj <= j + g.p; jj++) {
      if (jj < 0 || jj >= g.n2) continue;
      for (int ii = i - g.p; ii <= i + g.p; ii++) {
        if (ii < 0 || ii >= g.n1) continue;
        const int64_t nb = (int64_t)ii + (int64_t)g.n1 * (jj + (int64_t)g.n2 * kk);
        for (int cj = 0; cj < nc; cj++) {
          const int e = g.leq[nb * nc + cj];
          if (e >= 0) f(cnt++, e);
        }
      }
    }
  }
  return cnt;
}

__global__ void k_pat_count(PatchGrid g, int64_t nd, int64_t* __restrict__ cnt)
{
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d > nd) return;
  if (d == nd || g.leq[d] < 0) {
    cnt[d] = 0; // cnt[nd] = 0: the exclusive scan's last entry is the total
    return;
  }
  cnt[d] = for_window(g, d, [](int, int) {});
}

__global__ void k_pat_emit(PatchGrid g, int64_t nd, const int64_t* __restrict__ pptr,
                           uint64_t* __restrict__ keys)
{
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const int r = g.leq[d];
  if (r < 0) return;
  uint64_t* out = keys + pptr[d];
  for_window(g, d, [&](int k, int e) { out[k] = ((uint64_t)(uint32_t)r << 32) | (uint32_t)e; });
}

// rowptr from the sorted unique keys (every free row has at least its
// diagonal), colind = the low halves
__global__ void k_pat_csr(const uint64_t* __restrict__ u, int64_t nnz, int neq,
                          int64_t* __restrict__ rowptr, int32_t* __restrict__ colind)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int r = (int)(u[i] >> 32);
  colind[i] = (int32_t)(uint32_t)(u[i] & 0xffffffffu);
  const int rprev = i == 0 ? -1 : (int)(u[i - 1] >> 32);
  for (int rr = rprev + 1; rr <= r; rr++) rowptr[rr] = i;
  if (i == nnz - 1)
    for (int rr = r + 1; rr <= neq; rr++) rowptr[rr] = nnz;
}

__global__ void k_pat_map(PatchGrid g, int64_t nd, const int64_t* __restrict__ pptr,
                          const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                          int64_t* __restrict__ pmap, int* __restrict__ missing)
{
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const int r = g.leq[d];
  if (r < 0) return;
  const int64_t rb = rowptr[r], re = rowptr[r + 1];
  int64_t* out = pmap + pptr[d];
  for_window(g, d, [&](int k, int e) {
    int64_t lo = rb, hi = re;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (colind[mid] < e) lo = mid + 1;
      else hi = mid;
    }
    if (lo == re || colind[lo] != e) *missing = 1;
    out[k] = lo;
  });
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

int device_union_pattern(Context& c, std::vector<SubPatch>& subs, int neq, int64_t** rowptr_out,
                         int32_t** colind_out, int64_t* nnz_out)
{
  hipStream_t s = c.stream;
  const int nc = c.ncomp;
  std::vector<PatchGrid> grids;
  std::vector<int64_t> nds, base;
  int64_t total = 0;
  // 1. counts and map offsets per patch
  for (SubPatch& sp : subs) {
    const PatchRec& P = c.patches[sp.patch];
    PatchGrid g{P.n[0], P.n[1], P.n[2], P.p[0], nc, sp.d_leq};
    const int64_t nd = (int64_t)P.n[0] * P.n[1] * P.n[2] * nc;
    hipFree(sp.d_pptr);
    hipFree(sp.d_pmap);
    sp.d_pptr = sp.d_pmap = nullptr;
    int64_t* cnt = nullptr;
    IFEM_HIP(hipMalloc(&cnt, sizeof(int64_t) * (nd + 1)));
    IFEM_HIP(hipMalloc(&sp.d_pptr, sizeof(int64_t) * (nd + 1)));
    k_pat_count<<<nblk(nd + 1), 256, 0, s>>>(g, nd, cnt);
    size_t tb = 0;
    IFEM_HIP(rocprim::exclusive_scan(nullptr, tb, cnt, sp.d_pptr, (int64_t)0, (size_t)(nd + 1),
                                     rocprim::plus<int64_t>(), s));
    void* tmp = nullptr;
    IFEM_HIP(hipMalloc(&tmp, tb > 0 ? tb : 1));
    IFEM_HIP(rocprim::exclusive_scan(tmp, tb, cnt, sp.d_pptr, (int64_t)0, (size_t)(nd + 1),
                                     rocprim::plus<int64_t>(), s));
    int64_t m = 0;
    IFEM_HIP(hipMemcpyAsync(&m, sp.d_pptr + nd, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    IFEM_HIP(hipStreamSynchronize(s));
    hipFree(tmp);
    hipFree(cnt);
    grids.push_back(g);
    nds.push_back(nd);
    base.push_back(total);
    total += m;
  }
  // 2. keys, sort, unique
  uint64_t *keys = nullptr, *sorted = nullptr;
  IFEM_HIP(hipMalloc(&keys, sizeof(uint64_t) * (total > 0 ? total : 1)));
  IFEM_HIP(hipMalloc(&sorted, sizeof(uint64_t) * (total > 0 ? total : 1)));
  for (size_t q = 0; q < subs.size(); q++)
    k_pat_emit<<<nblk(nds[q]), 256, 0, s>>>(grids[q], nds[q], subs[q].d_pptr, keys + base[q]);
  IFEM_HIP(hipGetLastError());
  int rbits = 1;
  while ((1ll << rbits) < (int64_t)neq + 1) rbits++;
  size_t tb = 0;
  IFEM_HIP(rocprim::radix_sort_keys(nullptr, tb, keys, sorted, (size_t)total, 0, 32 + rbits, s));
  void* tmp = nullptr;
  IFEM_HIP(hipMalloc(&tmp, tb > 0 ? tb : 1));
  IFEM_HIP(rocprim::radix_sort_keys(tmp, tb, keys, sorted, (size_t)total, 0, 32 + rbits, s));
  hipFree(tmp);
  size_t* d_nu = nullptr;
  IFEM_HIP(hipMalloc(&d_nu, sizeof(size_t)));
  tb = 0;
  IFEM_HIP(rocprim::unique(nullptr, tb, sorted, keys, d_nu, (size_t)total,
                           rocprim::equal_to<uint64_t>(), s));
  IFEM_HIP(hipMalloc(&tmp, tb > 0 ? tb : 1));
  IFEM_HIP(rocprim::unique(tmp, tb, sorted, keys, d_nu, (size_t)total,
                           rocprim::equal_to<uint64_t>(), s));
  size_t nu = 0;
  IFEM_HIP(hipMemcpyAsync(&nu, d_nu, sizeof(size_t), hipMemcpyDeviceToHost, s));
  IFEM_HIP(hipStreamSynchronize(s));
  hipFree(tmp);
  hipFree(d_nu);
  hipFree(sorted);
  // 3. CSR
  const int64_t nnz = (int64_t)nu;
  int64_t* rowptr = nullptr;
  int32_t* colind = nullptr;
  IFEM_HIP(hipMalloc(&rowptr, sizeof(int64_t) * (neq + 1)));
  IFEM_HIP(hipMalloc(&colind, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
  if (nnz > 0) k_pat_csr<<<nblk(nnz), 256, 0, s>>>(keys, nnz, neq, rowptr, colind);
  else IFEM_HIP(hipMemsetAsync(rowptr, 0, sizeof(int64_t) * (neq + 1), s));
  hipFree(keys);
  // 4. position maps
  int* d_missing = nullptr;
  IFEM_HIP(hipMalloc(&d_missing, sizeof(int)));
  IFEM_HIP(hipMemsetAsync(d_missing, 0, sizeof(int), s));
  for (size_t q = 0; q < subs.size(); q++) {
    int64_t m = 0;
    IFEM_HIP(hipMemcpyAsync(&m, subs[q].d_pptr + nds[q], sizeof(int64_t), hipMemcpyDeviceToHost, s));
    IFEM_HIP(hipStreamSynchronize(s));
    IFEM_HIP(hipMalloc(&subs[q].d_pmap, sizeof(int64_t) * (m > 0 ? m : 1)));
    k_pat_map<<<nblk(nds[q]), 256, 0, s>>>(grids[q], nds[q], subs[q].d_pptr, rowptr, colind,
                                            subs[q].d_pmap, d_missing);
  }
  IFEM_HIP(hipGetLastError());
  int missing = 0;
  IFEM_HIP(hipMemcpyAsync(&missing, d_missing, sizeof(int), hipMemcpyDeviceToHost, s));
  IFEM_HIP(hipStreamSynchronize(s));
  hipFree(d_missing);
  if (missing) {
    hipFree(rowptr);
    hipFree(colind);
    return fail(IFEM_GPU_EINVAL, "pattern: column missing from the union row");
  }
  *rowptr_out = rowptr;
  *colind_out = colind;
  *nnz_out = nnz;
  return IFEM_GPU_OK;
}

} // namespace ifem
