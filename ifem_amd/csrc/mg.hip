// mg.hip -- spline multigrid preconditioner for the PCG of one patch.
//
// The reference reaches multigrid through LinSolParams (pc "gamg"/"ml" and
// the multigrid_* keys handed to PETSc, src/LinAlg/LinSolParams.C:60-155);
// here the hierarchy comes from the spline spaces themselves:
//   * level l+1 keeps every second distinct interior knot of level l, so
//     its B-splines are exact combinations of level l's (knot insertion,
//     Boehm's algorithm, gives the prolongation P = P3 x P2 x P1);
//   * A_{l+1} = P^T A_l P (Galerkin), formed one direction at a time on a
//     node-major stencil layout [node][offset][ci][cj] (threads over
//     offsets read and write contiguous memory);
//   * a coarse DOF is kept iff all fine DOFs it is made of are free, so
//     homogeneous and prescribed Dirichlet faces carry over;
//   * smoothing: Chebyshev iteration on D^-1 A over [0.1, 1.05] x lambda_max
//     (lambda_max from 10 Lanczos steps), symmetric pre/post; the coarsest
//     level is solved with a dense inverse (Cholesky on the host, small);
//   * level 0 smooths with an FP16 copy of the CSR matrix (scaled by its
//     largest diagonal entry, FP32 sums) in equation order; coarse levels
//     use FP32 stencils in DOF order with zeros at eliminated DOFs.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "devutil.hpp"
#include "internal.hpp"
#include "mg.hpp"

namespace ifem {

void launch_spmv16(Context& c, const __half* val, float scale, const float* p, float* q,
                   bool skip_done, hipStream_t st);
void launch_spmv16_node3(Context& c, const uint2* val, float scale, const float* p, float* q,
                         bool skip_done, hipStream_t st);
int demote16_node3(Context& c, double inv_scale, uint2* out);

Multigrid::~Multigrid()
{
  if (side) hipStreamSynchronize(side);
  for (MgLevel& L : lev) {
    for (int d = 0; d < 3; d++) {
      MgDir& D = L.dir[d];
      hipFree(D.clo); hipFree(D.ccnt); hipFree(D.cw);
      hipFree(D.rlo); hipFree(D.rcnt); hipFree(D.rw);
    }
    hipFree(L.S64); hipFree(L.S32); hipFree(L.freem); hipFree(L.dinv);
    hipFree(L.x); hipFree(L.b); hipFree(L.r); hipFree(L.d); hipFree(L.q);
  }
  hipFree(S0); hipFree(T1); hipFree(T2);
  hipFree(val16); hipFree(val16n); hipFree(cdofs); hipFree(cinv); hipFree(part); hipFree(lam); hipFree(err);
  hipFree(dof_of_eq); hipFree(part2); hipFree(lzh);
  if (h_lzh) hipHostFree(h_lzh);
  if (h_coarse) hipHostFree(h_coarse);
  if (h_err) hipHostFree(h_err);
  if (e_copy) hipEventDestroy(e_copy);
  if (e_gal) hipEventDestroy(e_gal);
  if (e_side) hipEventDestroy(e_side);
  if (side) hipStreamDestroy(side);
  hipFree(part_s); hipFree(part_s2);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
}

// ---------------------------------------------------------------------------
// host: knots and prolongation

// Every second distinct interior breakpoint of t (degree p) is dropped.
static std::vector<double> coarsen_knots(const std::vector<double>& t, int p)
{
  const int m = (int)t.size();
  std::vector<double> br;
  std::vector<int> mult;
  for (int i = p + 1; i < m - p - 1; i++) {
    if (!br.empty() && t[i] == br.back()) mult.back()++;
    else { br.push_back(t[i]); mult.push_back(1); }
  }
  std::vector<double> out(t.begin(), t.begin() + p + 1);
  for (size_t k = 1; k < br.size(); k += 2)
    for (int r = 0; r < mult[k]; r++) out.push_back(br[k]);
  out.insert(out.end(), t.end() - (p + 1), t.end());
  return out;
}

// P (nf x nc, row-major) with fine coefficients = P * coarse coefficients:
// the knots of tf missing from tc are inserted one at a time (Boehm).
static bool prolongation(const std::vector<double>& tc, const std::vector<double>& tf, int p,
                         std::vector<double>& P, int& nf, int& nc)
{
  nc = (int)tc.size() - p - 1;
  nf = (int)tf.size() - p - 1;
  std::vector<double> ins;
  size_t a = 0;
  for (size_t b = 0; b < tf.size(); b++) {
    if (a < tc.size() && tc[a] == tf[b]) a++;
    else ins.push_back(tf[b]);
  }
  if (a != tc.size() || (int)ins.size() != nf - nc) return false;
  // columns: coefficient vectors of the coarse functions
  std::vector<std::vector<double>> col(nc, std::vector<double>(nc, 0.0));
  for (int j = 0; j < nc; j++) col[j][j] = 1.0;
  std::vector<double> t = tc;
  for (double u : ins) {
    const int n = (int)t.size() - p - 1;
    int k = p;
    while (k + 1 < n && t[k + 1] <= u) k++;
    for (int j = 0; j < nc; j++) {
      std::vector<double>& cv = col[j];
      std::vector<double> nv(n + 1);
      for (int i = 0; i <= n; i++) {
        if (i <= k - p) nv[i] = cv[i];
        else if (i >= k + 1) nv[i] = cv[i - 1];
        else {
          const double al = (u - t[i]) / (t[i + p] - t[i]);
          nv[i] = al * cv[i] + (1.0 - al) * cv[i - 1];
        }
      }
      cv.swap(nv);
    }
    t.insert(t.begin() + k + 1, u);
  }
  if (t != tf) return false;
  P.assign((size_t)nf * nc, 0.0);
  for (int j = 0; j < nc; j++)
    for (int i = 0; i < nf; i++) P[(size_t)i * nc + j] = std::fabs(col[j][i]) > 1e-15 ? col[j][i] : 0.0;
  return true;
}

template <typename T>
static int upload(T*& d, const std::vector<T>& h)
{
  IFEM_HIP(hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size())));
  if (!h.empty()) IFEM_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return IFEM_GPU_OK;
}

static int build_dir(const std::vector<double>& P, int nf, int nc, MgDir& D)
{
  std::vector<int> clo(nc), ccnt(nc), rlo(nf), rcnt(nf);
  std::vector<double> cw((size_t)nc * kMgKC, 0.0), rw((size_t)nf * kMgKR, 0.0);
  for (int j = 0; j < nc; j++) {
    int lo = -1, hi = -1;
    for (int i = 0; i < nf; i++)
      if (P[(size_t)i * nc + j] != 0.0) { if (lo < 0) lo = i; hi = i; }
    if (lo < 0 || hi - lo + 1 > kMgKC) return fail(IFEM_GPU_EUNSUPPORTED, "mg: prolongation too wide");
    clo[j] = lo;
    ccnt[j] = hi - lo + 1;
    for (int i = lo; i <= hi; i++) cw[(size_t)j * kMgKC + (i - lo)] = P[(size_t)i * nc + j];
  }
  for (int i = 0; i < nf; i++) {
    int lo = -1, hi = -1;
    for (int j = 0; j < nc; j++)
      if (P[(size_t)i * nc + j] != 0.0) { if (lo < 0) lo = j; hi = j; }
    if (lo < 0 || hi - lo + 1 > kMgKR) return fail(IFEM_GPU_EUNSUPPORTED, "mg: prolongation too wide");
    rlo[i] = lo;
    rcnt[i] = hi - lo + 1;
    for (int j = lo; j <= hi; j++) rw[(size_t)i * kMgKR + (j - lo)] = P[(size_t)i * nc + j];
  }
  D.nf = nf;
  D.nc = nc;
  int rc = upload(D.clo, clo);
  if (!rc) rc = upload(D.ccnt, ccnt);
  if (!rc) rc = upload(D.cw, cw);
  if (!rc) rc = upload(D.rlo, rlo);
  if (!rc) rc = upload(D.rcnt, rcnt);
  if (!rc) rc = upload(D.rw, rw);
  return rc;
}

// ---------------------------------------------------------------------------
// kernels

struct Grid3 {
  int n0, n1, n2;
};

// x / n and x % n for 0 <= x < 2^24 through a float reciprocal, corrected
__device__ inline int divmod_small(int x, int n, float inv, int& r)
{
  int q = (int)((float)x * inv);
  r = x - q * n;
  if (r < 0) { q--; r += n; }
  else if (r >= n) { q++; r -= n; }
  return q;
}

// Fine CSR (equation order) -> node-major stencil S0[node][off][ci][cj]
// (zero-filled beforehand).  One wave per row: lanes read the row's values
// and column indices contiguously and write into the row node's stencil
// block.  err is set when a column lies outside the +-p box.
template <int P, int NC>
__global__ __launch_bounds__(256) void k_mg_csr_scatter(
    int neq, Grid3 g, const int* __restrict__ dof_of_eq, const int64_t* __restrict__ rowptr,
    const int32_t* __restrict__ colind, const double* __restrict__ val, double* __restrict__ S,
    int* __restrict__ err)
{
  constexpr int W1 = 2 * P + 1, NOFF = W1 * W1 * W1, NCC = NC * NC;
  const int wave = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= neq) return;
  const int dof = dof_of_eq[wave];
  const int node = dof / NC, ci = dof - node * NC;
  const int i = node % g.n0, j = (node / g.n0) % g.n1, k = node / (g.n0 * g.n1);
  double* out = S + (int64_t)node * NOFF * NCC + ci * NC;
  const int64_t b = rowptr[wave], e = rowptr[wave + 1];
  const float inv0 = 1.0f / (float)g.n0, inv1 = 1.0f / (float)g.n1;
  for (int64_t q = b + lane; q < e; q += 64) {
    const int cd = dof_of_eq[colind[q]];
    const int cn = cd / NC, cj = cd - cn * NC;
    int r0, r1;
    const int q0 = divmod_small(cn, g.n0, inv0, r0);
    const int q1 = divmod_small(q0, g.n1, inv1, r1);
    const int e0 = r0 - i, e1 = r1 - j, e2 = q1 - k;
    if (e0 < -P || e0 > P || e1 < -P || e1 > P || e2 < -P || e2 > P) {
      err[0] = 1;
      continue;
    }
    const int off = ((e2 + P) * W1 + (e1 + P)) * W1 + (e0 + P);
    out[off * NCC + cj] = val[q];
  }
}

// The same scatter for 3-component node-uniform systems through the node
// index (bcol): one wave per node, a lane per neighbour node, writing its
// whole 3x3 block S[node][off][0..2][0..2] (72 contiguous bytes) from the
// node's three rows.
template <int P>
__global__ __launch_bounds__(256) void k_mg_csr_scatter3(
    int nnode, Grid3 g, const int* __restrict__ dof_of_eq, const int64_t* __restrict__ rowptr,
    const int32_t* __restrict__ bcol, const double* __restrict__ val, double* __restrict__ S,
    int* __restrict__ err)
{
  constexpr int W1 = 2 * P + 1, NOFF = W1 * W1 * W1;
  const int wave = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= nnode) return;
  const int node = dof_of_eq[3 * wave] / 3;
  const int i = node % g.n0, j = (node / g.n0) % g.n1, k = node / (g.n0 * g.n1);
  const int64_t r0 = rowptr[3 * wave];
  const int cnt = (int)((rowptr[3 * wave + 1] - r0) / 3);
  const int64_t nb = r0 / 9;
  double* out = S + (int64_t)node * NOFF * 9;
  const float inv0 = 1.0f / (float)g.n0, inv1 = 1.0f / (float)g.n1;
  for (int q = lane; q < cnt; q += 64) {
    const int cn = dof_of_eq[3 * bcol[nb + q]] / 3;
    int rr0, rr1;
    const int q0 = divmod_small(cn, g.n0, inv0, rr0);
    const int q1 = divmod_small(q0, g.n1, inv1, rr1);
    const int e0 = rr0 - i, e1 = rr1 - j, e2 = q1 - k;
    if (e0 < -P || e0 > P || e1 < -P || e1 > P || e2 < -P || e2 > P) {
      err[0] = 1;
      continue;
    }
    const int off = ((e2 + P) * W1 + (e1 + P)) * W1 + (e0 + P);
    double* o = out + off * 9;
#pragma unroll
    for (int ci = 0; ci < 3; ci++) {
      const double* v = val + r0 + (int64_t)ci * 3 * cnt + 3 * q;
      o[3 * ci] = v[0];
      o[3 * ci + 1] = v[1];
      o[3 * ci + 2] = v[2];
    }
  }
}

// One direction of the Galerkin product: Sout = P_d^T Sin P_d, where
// direction d of the grid goes from gin.n_d fine to ncd coarse points.  All
// stencils are node-major, S[node][off][ci][cj]; one thread per (node,
// offset) with the offset fastest, so reads and writes are contiguous.
// With freem (the last pass of a level): rows and columns of eliminated
// coarse DOFs are zeroed, the FP32 copy S32 and the inverse diagonal dinv
// are written as well.
template <int P, int NC>
__global__ __launch_bounds__(256) void k_mg_galerkin(
    const double* __restrict__ Sin, double* __restrict__ Sout, Grid3 gin, int d, int ncd,
    const int* __restrict__ clo, const int* __restrict__ ccnt, const double* __restrict__ cw,
    const unsigned char* __restrict__ freem, float* __restrict__ S32, float* __restrict__ dinv)
{
  constexpr int W1 = 2 * P + 1, NOFF = W1 * W1 * W1, NCC = NC * NC;
  int nin[3] = {gin.n0, gin.n1, gin.n2};
  int nout[3] = {gin.n0, gin.n1, gin.n2};
  nout[d] = ncd;
  const int64_t nnout = (int64_t)nout[0] * nout[1] * nout[2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnout * NOFF) return;
  const int64_t node = t / NOFF;
  const int off = (int)(t - node * NOFF);
  int o[3];
  o[0] = (int)(node % nout[0]);
  o[1] = (int)((node / nout[0]) % nout[1]);
  o[2] = (int)(node / ((int64_t)nout[0] * nout[1]));
  int e[3] = {off % W1 - P, (off / W1) % W1 - P, off / (W1 * W1) - P};
  double acc[NCC];
#pragma unroll
  for (int m = 0; m < NCC; m++) acc[m] = 0.0;
  const int I = o[d], J = I + e[d];
  bool ok = J >= 0 && J < ncd;
#pragma unroll
  for (int q = 0; q < 3; q++)
    if (q != d) {
      const int v = o[q] + e[q];
      if (v < 0 || v >= nout[q]) ok = false;
    }
  if (ok) {
    const int na = ccnt[I], nb = ccnt[J], a0 = clo[I], b0 = clo[J];
    // strides of direction d in the input grid (nodes) and offsets
    const int64_t sn = d == 0 ? 1 : (d == 1 ? (int64_t)nin[0] : (int64_t)nin[0] * nin[1]);
    const int so = d == 0 ? 1 : (d == 1 ? W1 : W1 * W1);
    int oi[3] = {o[0], o[1], o[2]};
    oi[d] = 0;
    const int64_t nbase = (int64_t)oi[0] + (int64_t)nin[0] * (oi[1] + (int64_t)nin[1] * oi[2]);
    const int obase = off - (e[d] + P) * so;
    for (int ka = 0; ka < na; ka++) {
      const int a = a0 + ka;
      const double wa = cw[(int64_t)I * kMgKC + ka];
      const int64_t ni = nbase + a * sn;
      for (int kb = 0; kb < nb; kb++) {
        const int db = b0 + kb - a;
        if (db < -P || db > P) continue;
        const double w = wa * cw[(int64_t)J * kMgKC + kb];
        const double* src = Sin + (ni * NOFF + obase + (db + P) * so) * NCC;
#pragma unroll
        for (int m = 0; m < NCC; m++) acc[m] += w * src[m];
      }
    }
  }
  double* dst = Sout + t * NCC;
  if (freem) {
    constexpr int center = (P * W1 + P) * W1 + P;
    const int64_t nb = ok ? node + e[0] + (int64_t)nout[0] * (e[1] + (int64_t)nout[1] * e[2]) : 0;
#pragma unroll
    for (int ci = 0; ci < NC; ci++)
#pragma unroll
      for (int cj = 0; cj < NC; cj++) {
        double v = acc[ci * NC + cj];
        if (!ok || !freem[node * NC + ci] || !freem[nb * NC + cj]) v = 0.0;
        dst[ci * NC + cj] = v;
        S32[t * NCC + ci * NC + cj] = (float)v;
        if (off == center && ci == cj)
          dinv[node * NC + ci] = (freem[node * NC + ci] && v > 0.0) ? (float)(1.0 / v) : 0.0f;
      }
    return;
  }
#pragma unroll
  for (int m = 0; m < NCC; m++) dst[m] = acc[m];
}

// y = A x on a coarse level (node-major stencil, DOF order): one wave per
// node, lanes over the offsets (contiguous 4*NC*NC-byte blocks), then a
// fixed-order wave reduction.
template <int P, int NC>
__global__ __launch_bounds__(256) void k_mg_spmv(const float* __restrict__ S,
                                                 const float* __restrict__ x,
                                                 float* __restrict__ y, Grid3 g,
                                                 const double* __restrict__ scal)
{
  constexpr int W1 = 2 * P + 1, NOFF = W1 * W1 * W1, NCC = NC * NC;
  if (scal && scal[S_DONE] != 0.0) return;
  const int64_t nnode = (int64_t)g.n0 * g.n1 * g.n2;
  const int64_t node = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (node >= nnode) return;
  const int i = (int)(node % g.n0), j = (int)((node / g.n0) % g.n1), k = (int)(node / ((int64_t)g.n0 * g.n1));
  const float* sn = S + node * NOFF * NCC;
  float acc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) acc[c] = 0.f;
  for (int off = lane; off < NOFF; off += 64) {
    const int ii = i + off % W1 - P, jj = j + (off / W1) % W1 - P, kk = k + off / (W1 * W1) - P;
    if (ii < 0 || ii >= g.n0 || jj < 0 || jj >= g.n1 || kk < 0 || kk >= g.n2) continue;
    const float* xv = x + ((int64_t)ii + (int64_t)g.n0 * (jj + (int64_t)g.n1 * kk)) * NC;
    const float* sv = sn + off * NCC;
    float xs[NC];
#pragma unroll
    for (int cj = 0; cj < NC; cj++) xs[cj] = xv[cj];
#pragma unroll
    for (int ci = 0; ci < NC; ci++)
#pragma unroll
      for (int cj = 0; cj < NC; cj++) acc[ci] += sv[ci * NC + cj] * xs[cj];
  }
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < NC; c++) y[node * NC + c] = acc[c];
}

// Chebyshev step.  mode 0: r = b, d = c2 D^-1 r, x = d;  mode 1: r -= q,
// d = c1 d + c2 D^-1 r, x += d;  mode 2: r = b - q, d = c2 D^-1 r, x += d.
__global__ __launch_bounds__(256) void k_mg_cheb(int n, int mode, const float* __restrict__ b,
                                                 float* __restrict__ r,
                                                 const float* __restrict__ q,
                                                 float* __restrict__ d, float* __restrict__ x,
                                                 const float* __restrict__ dinv, float c1,
                                                 float c2, const double* __restrict__ scal)
{
  if (scal && scal[S_DONE] != 0.0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (mode == 0) {
      const float ri = b[i];
      const float di = c2 * dinv[i] * ri;
      r[i] = ri;
      d[i] = di;
      x[i] = di;
    } else if (mode == 1) {
      const float ri = r[i] - q[i];
      const float di = c1 * d[i] + c2 * dinv[i] * ri;
      r[i] = ri;
      d[i] = di;
      x[i] += di;
    } else {
      const float ri = b[i] - q[i];
      const float di = c2 * dinv[i] * ri;
      r[i] = ri;
      d[i] = di;
      x[i] += di;
    }
  }
}

struct DirDev {
  const int *clo, *ccnt, *rlo, *rcnt;
  const double *cw, *rw;
};

// bc = P^T (rf - qf) on the coarse grid (masked by the coarse free flags);
// fmap: fine DOF -> vector index (-1 eliminated), nullptr for identity.
// One thread per coarse DOF.
template <int NC>
__global__ __launch_bounds__(256) void k_mg_restrict(const float* __restrict__ rf,
                                                     const float* __restrict__ qf,
                                                     const int* __restrict__ fmap, Grid3 gf,
                                                     float* __restrict__ bc,
                                                     const unsigned char* __restrict__ cfree,
                                                     Grid3 gc, DirDev d0, DirDev d1, DirDev d2,
                                                     const double* __restrict__ scal)
{
  if (scal && scal[S_DONE] != 0.0) return;
  const int64_t nnc = (int64_t)gc.n0 * gc.n1 * gc.n2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnc * NC) return;
  const int64_t node = t / NC;
  const int c = (int)(t - node * NC);
  if (!cfree[t]) {
    bc[t] = 0.f;
    return;
  }
  const int I0 = (int)(node % gc.n0), I1 = (int)((node / gc.n0) % gc.n1);
  const int I2 = (int)(node / ((int64_t)gc.n0 * gc.n1));
  float acc = 0.f;
  const int n2c = d2.ccnt[I2], n1c = d1.ccnt[I1], n0c = d0.ccnt[I0];
  const int l2 = d2.clo[I2], l1 = d1.clo[I1], l0 = d0.clo[I0];
  for (int kg = 0; kg < n2c; kg++) {
    const float w2 = (float)d2.cw[(int64_t)I2 * kMgKC + kg];
    for (int kb = 0; kb < n1c; kb++) {
      const float w12 = w2 * (float)d1.cw[(int64_t)I1 * kMgKC + kb];
      const int64_t frow = (int64_t)gf.n0 * ((l1 + kb) + (int64_t)gf.n1 * (l2 + kg));
      for (int ka = 0; ka < n0c; ka++) {
        const float w = w12 * (float)d0.cw[(int64_t)I0 * kMgKC + ka];
        const int64_t f = (frow + l0 + ka) * NC + c;
        const int64_t idx = fmap ? (int64_t)fmap[f] : f;
        if (idx < 0) continue;
        acc += w * (qf ? rf[idx] - qf[idx] : rf[idx]);
      }
    }
  }
  bc[t] = acc;
}

// xf += P xc (fine free entries only).
template <int NC>
__global__ __launch_bounds__(256) void k_mg_prolong(float* __restrict__ xf,
                                                    const int* __restrict__ fmap, Grid3 gf,
                                                    const float* __restrict__ xc, Grid3 gc,
                                                    DirDev d0, DirDev d1, DirDev d2,
                                                    const double* __restrict__ scal)
{
  if (scal && scal[S_DONE] != 0.0) return;
  const int64_t nnf = (int64_t)gf.n0 * gf.n1 * gf.n2;
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nnf) return;
  const int i0 = (int)(node % gf.n0), i1 = (int)((node / gf.n0) % gf.n1);
  const int i2 = (int)(node / ((int64_t)gf.n0 * gf.n1));
  float acc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) acc[c] = 0.f;
  const int n2r = d2.rcnt[i2], n1r = d1.rcnt[i1], n0r = d0.rcnt[i0];
  for (int kg = 0; kg < n2r; kg++) {
    const int J2 = d2.rlo[i2] + kg;
    const float w2 = (float)d2.rw[(int64_t)i2 * kMgKR + kg];
    for (int kb = 0; kb < n1r; kb++) {
      const int J1 = d1.rlo[i1] + kb;
      const float w12 = w2 * (float)d1.rw[(int64_t)i1 * kMgKR + kb];
      for (int ka = 0; ka < n0r; ka++) {
        const int J0 = d0.rlo[i0] + ka;
        const float w = w12 * (float)d0.rw[(int64_t)i0 * kMgKR + ka];
        const int64_t cn = (int64_t)J0 + (int64_t)gc.n0 * (J1 + (int64_t)gc.n1 * J2);
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] += w * xc[cn * NC + c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int64_t idx = fmap ? (int64_t)fmap[node * NC + c] : node * NC + c;
    if (idx >= 0) xf[idx] += acc[c];
  }
}

// Coarsest level: x[dof_i] = sum_j Ainv(i, j) b[dof_j]; one wave per row.
__global__ __launch_bounds__(256) void k_mg_coarse(int ncf, const int* __restrict__ dofs,
                                                   const float* __restrict__ inv,
                                                   const float* __restrict__ b,
                                                   float* __restrict__ x,
                                                   const double* __restrict__ scal)
{
  if (scal && scal[S_DONE] != 0.0) return;
  const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= ncf) return;
  float s = 0.f;
  for (int j = lane; j < ncf; j += 64) s += inv[(int64_t)i * ncf + j] * b[dofs[j]];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) x[dofs[i]] = s;
}

// lambda_max(D^-1 A) by Lanczos: a few Jacobi-preconditioned CG steps on a
// fixed pseudo-random right-hand side; the CG coefficients give the Lanczos
// tridiagonal matrix (diagonal 1/alpha_j + beta_{j-1}/alpha_{j-1},
// off-diagonal sqrt(beta_j)/alpha_j), whose largest eigenvalue (host) is
// the estimate.  h[0..] holds rz_0, then per step alpha_j, beta_j, rz_{j+1}.
__global__ __launch_bounds__(kThreads) void k_lz_start(int n, float* __restrict__ r,
                                                       float* __restrict__ z,
                                                       float* __restrict__ p,
                                                       const float* __restrict__ dinv,
                                                       double* __restrict__ part)
{
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    uint32_t h = (uint32_t)i * 2654435761u + 12345u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    const float u = (float)(h & 0xFFFFFF) / 16777216.0f - 0.5f;
    const float ri = dinv[i] != 0.f ? u : 0.f;
    const float zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    s += (double)ri * (double)zi;
  }
  const double b = block_sum<double>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ __launch_bounds__(kThreads) void k_lz_sum(const double* __restrict__ part, int nparts,
                                                     double* __restrict__ out)
{
  __shared__ double sh[kThreads / 64];
  const double s = sum_parts(part, nparts, sh);
  if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(kThreads) void k_lz_dot(int n, const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     double* __restrict__ part)
{
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    s += (double)a[i] * (double)b[i];
  const double bs = block_sum<double>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// alpha = rz_j / <p,q>; r -= alpha q; z = D^-1 r; partials of <r,z>
__global__ __launch_bounds__(kThreads) void k_lz_update(int n, float* __restrict__ r,
                                                        float* __restrict__ z,
                                                        const float* __restrict__ q,
                                                        const float* __restrict__ dinv,
                                                        const double* __restrict__ part_pq,
                                                        int nparts, double* __restrict__ h,
                                                        int j, double* __restrict__ part_rz)
{
  __shared__ double sh[kThreads / 64];
  const double pq = sum_parts(part_pq, nparts, sh);
  const double rz = h[3 * j];
  const double alpha = pq > 0.0 ? rz / pq : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) h[3 * j + 1] = alpha;
  const float a = (float)alpha;
  double s = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const float ri = r[i] - a * q[i];
    const float zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    s += (double)ri * (double)zi;
  }
  const double b = block_sum<double>(s, sh);
  if (threadIdx.x == 0) part_rz[blockIdx.x] = b;
}

// beta = rz_{j+1} / rz_j; p = z + beta p
__global__ __launch_bounds__(kThreads) void k_lz_pupdate(int n, float* __restrict__ p,
                                                         const float* __restrict__ z,
                                                         const double* __restrict__ part_rz,
                                                         int nparts, double* __restrict__ h, int j)
{
  __shared__ double sh[kThreads / 64];
  const double rzn = sum_parts(part_rz, nparts, sh);
  const double rz = h[3 * j];
  const double beta = rz > 0.0 ? rzn / rz : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    h[3 * j + 2] = beta;
    h[3 * j + 3] = rzn;
  }
  const float bf = (float)beta;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = z[i] + bf * p[i];
}

template <typename V>
__global__ void k_mg_cast_in(int n, const V* __restrict__ a, float* __restrict__ b)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

template <typename V>
__global__ void k_mg_cast_out(int n, const float* __restrict__ a, V* __restrict__ b)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (V)a[i];
}

// FP16 copy of K / scale for the fine-level smoother (|K_ij| <= max diag for
// an SPD matrix, so with scale = max diag every value lies in [-1, 1]).
__global__ void k_mg_demote16(int64_t n, const double* __restrict__ a, __half* __restrict__ b,
                              double inv_scale)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = __float2half((float)(a[i] * inv_scale));
}

// FP32 inverse diagonal of the FP64 CSR matrix, and per-block maxima of
// the diagonal.
__global__ __launch_bounds__(kThreads) void k_mg_diag(int n, const int64_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ colind,
                                                      const double* __restrict__ val,
                                                      float* __restrict__ dinv,
                                                      double* __restrict__ part)
{
  __shared__ double sh[kThreads / 64];
  double mx = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    int64_t lo = rowptr[i], hi = rowptr[i + 1] - 1;
    double d = 0.0;
    while (lo <= hi) {
      const int64_t mid = (lo + hi) >> 1;
      const int cc = colind[mid];
      if (cc == i) { d = val[mid]; break; }
      if (cc < i) lo = mid + 1; else hi = mid - 1;
    }
    dinv[i] = d > 0.0 ? (float)(1.0 / d) : 0.f;
    mx = fmax(mx, fabs(d));
  }
  for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[wave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kThreads / 64; w++) m = fmax(m, sh[w]);
    part[blockIdx.x] = m;
  }
}

__global__ void k_mg_max(const double* __restrict__ part, int nparts, double* __restrict__ out)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < nparts; i++) m = fmax(m, part[i]);
    *out = m;
  }
}

// ---------------------------------------------------------------------------
// host side

static constexpr int kMgCoarseMax = 400;   // stop coarsening at this many free DOFs
static constexpr int kMgDenseMax = 3000;   // largest coarsest level solved densely
static constexpr int kMgLanczos = 10;      // CG steps of the eigenvalue estimate
static constexpr int kMgParts = 1024;

static inline unsigned grid1(int64_t n) { return (unsigned)((n + 255) / 256); }

// Largest eigenvalue of the Lanczos tridiagonal matrix built from k CG steps
// (h: rz_0, then alpha_j, beta_j, rz_{j+1} per step), by bisection on the
// Sturm sequence count.  Steps after a breakdown (alpha <= 0) are dropped.
static double lanczos_max(const double* h, int k)
{
  std::vector<double> dg, od;
  for (int j = 0; j < k; j++) {
    const double a = h[3 * j + 1];
    if (!(a > 0.0) || !std::isfinite(a)) break;
    double dj = 1.0 / a;
    if (j > 0) dj += h[3 * j - 1] / h[3 * j - 2];
    dg.push_back(dj);
    if (j > 0) od.push_back(std::sqrt(std::max(h[3 * j - 1], 0.0)) / h[3 * j - 2]);
  }
  const int n = (int)dg.size();
  if (n == 0) return 0.0;
  double lo = 0.0, hi = 0.0;
  for (int i = 0; i < n; i++) {
    const double r = (i > 0 ? std::fabs(od[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(od[i]) : 0.0);
    hi = std::max(hi, dg[i] + r);
  }
  auto count_below = [&](double x) { // eigenvalues < x
    int cnt = 0;
    double q = dg[0] - x;
    if (q < 0) cnt++;
    for (int i = 1; i < n; i++) {
      if (q == 0.0) q = 1e-300;
      q = dg[i] - x - od[i - 1] * od[i - 1] / q;
      if (q < 0) cnt++;
    }
    return cnt;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-12 * hi; it++) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= n) hi = mid;
    else lo = mid;
  }
  return hi;
}

static DirDev dirdev(const MgDir& D)
{
  return DirDev{D.clo, D.ccnt, D.rlo, D.rcnt, D.cw, D.rw};
}

static Grid3 grid_of(const MgLevel& L) { return Grid3{L.n[0], L.n[1], L.n[2]}; }

// Structure: levels, transfers, free masks, vectors.  Built once per system.
static int mg_build(Context& c, Multigrid& m)
{
  const PatchRec& P = c.patches[c.patch];
  if (P.p[0] != P.p[1] || P.p[0] != P.p[2])
    return fail(IFEM_GPU_EUNSUPPORTED, "mg: equal degrees per direction required");
  m.p = P.p[0];
  m.nc = c.ncomp;
  const int W1 = 2 * m.p + 1;
  m.noff = W1 * W1 * W1;
  m.lev.clear();
  m.lev.emplace_back();
  {
    MgLevel& L = m.lev[0];
    for (int d = 0; d < 3; d++) {
      L.n[d] = P.n[d];
      L.knots[d] = P.knots[d];
    }
    L.nnode = L.n[0] * L.n[1] * L.n[2];
    L.ndof = L.nnode * m.nc;
    L.nvec = c.neq;
    L.h_free.assign(L.ndof, 0);
    for (int i = 0; i < L.ndof; i++) L.h_free[i] = c.h_eq[i] >= 0;
    L.nfree = c.neq;
  }
  while ((int)m.lev.size() < 16) {
    const MgLevel& F = m.lev.back();
    if (F.nfree <= kMgCoarseMax && m.lev.size() > 1) break;
    MgLevel C;
    bool coarsened = false;
    std::vector<double> Pd[3];
    int nf[3], ncd[3];
    for (int d = 0; d < 3; d++) {
      const std::vector<double>& tf = F.knots[d];
      std::vector<double> tc = coarsen_knots(tf, m.p);
      if (tc.size() < tf.size()) coarsened = true;
      if (!prolongation(tc, tf, m.p, Pd[d], nf[d], ncd[d]))
        return fail(IFEM_GPU_EUNSUPPORTED, "mg: knot vectors are not nested");
      C.knots[d] = tc;
      C.n[d] = ncd[d];
    }
    if (!coarsened) break;
    for (int d = 0; d < 3; d++) {
      int rc = build_dir(Pd[d], nf[d], ncd[d], C.dir[d]);
      if (rc) return rc;
    }
    C.nnode = C.n[0] * C.n[1] * C.n[2];
    C.ndof = C.nnode * m.nc;
    C.nvec = C.ndof;
    // a coarse DOF is kept iff every fine DOF it is made of is free
    std::vector<int> lo[3], hi[3]; // fine index range of each coarse function
    for (int d = 0; d < 3; d++) {
      lo[d].assign(ncd[d], -1);
      hi[d].assign(ncd[d], -1);
      for (int J = 0; J < ncd[d]; J++)
        for (int a = 0; a < nf[d]; a++)
          if (Pd[d][(size_t)a * ncd[d] + J] != 0.0) {
            if (lo[d][J] < 0) lo[d][J] = a;
            hi[d][J] = a;
          }
    }
    C.h_free.assign(C.ndof, 1);
    for (int I2 = 0; I2 < C.n[2]; I2++)
      for (int I1 = 0; I1 < C.n[1]; I1++)
        for (int I0 = 0; I0 < C.n[0]; I0++) {
          const int cn = I0 + C.n[0] * (I1 + C.n[1] * I2);
          for (int a2 = lo[2][I2]; a2 <= hi[2][I2]; a2++)
            for (int a1 = lo[1][I1]; a1 <= hi[1][I1]; a1++)
              for (int a0 = lo[0][I0]; a0 <= hi[0][I0]; a0++) {
                const int fn = a0 + F.n[0] * (a1 + F.n[1] * a2);
                for (int cc = 0; cc < m.nc; cc++)
                  if (!F.h_free[fn * m.nc + cc]) C.h_free[cn * m.nc + cc] = 0;
              }
        }
    C.nfree = 0;
    for (unsigned char f : C.h_free) C.nfree += f;
    m.lev.push_back(std::move(C));
  }
  const MgLevel& Lc = m.lev.back();
  if (m.lev.size() < 2) return fail(IFEM_GPU_EUNSUPPORTED, "mg: the patch cannot be coarsened");
  if (Lc.nfree > kMgDenseMax)
    return fail(IFEM_GPU_EUNSUPPORTED, "mg: coarsest level too large for the dense solve");
  // device buffers
  const int ncc = m.nc * m.nc;
  size_t t1 = 0, t2 = 0;
  for (size_t l = 0; l < m.lev.size(); l++) {
    MgLevel& L = m.lev[l];
    const size_t nv = L.nvec > 0 ? L.nvec : 1;
    float** v[] = {&L.x, &L.b, &L.r, &L.d, &L.q, &L.dinv};
    for (float** a : v) IFEM_HIP(hipMalloc(a, sizeof(float) * nv));
    IFEM_HIP(hipMemset(L.x, 0, sizeof(float) * nv));
    if (l > 0) {
      const size_t ns = (size_t)m.noff * ncc * L.nnode;
      IFEM_HIP(hipMalloc(&L.S64, sizeof(double) * ns));
      IFEM_HIP(hipMalloc(&L.S32, sizeof(float) * ns));
      int rc = upload(L.freem, L.h_free);
      if (rc) return rc;
      const MgLevel& F = m.lev[l - 1];
      t1 = std::max(t1, (size_t)L.n[0] * F.n[1] * F.n[2]);
      t2 = std::max(t2, (size_t)L.n[0] * L.n[1] * F.n[2]);
    }
  }
  m.s0cap = (size_t)m.noff * ncc * m.lev[0].nnode;
  IFEM_HIP(hipMalloc(&m.S0, sizeof(double) * m.s0cap));
  m.t1cap = t1 * m.noff * ncc;
  m.t2cap = t2 * m.noff * ncc;
  IFEM_HIP(hipMalloc(&m.T1, sizeof(double) * m.t1cap));
  IFEM_HIP(hipMalloc(&m.T2, sizeof(double) * m.t2cap));
  std::vector<int> doq(c.neq > 0 ? c.neq : 1, 0);
  for (int i = 0; i < c.ndof; i++)
    if (c.h_eq[i] >= 0) doq[c.h_eq[i]] = i;
  {
    int rc0 = upload(m.dof_of_eq, doq);
    if (rc0) return rc0;
  }
  std::vector<int> cd;
  for (int i = 0; i < Lc.ndof; i++)
    if (Lc.h_free[i]) cd.push_back(i);
  m.ncf = (int)cd.size();
  int rc = upload(m.cdofs, cd);
  if (rc) return rc;
  IFEM_HIP(hipMalloc(&m.cinv, sizeof(float) * (size_t)std::max(1, m.ncf * m.ncf)));
  IFEM_HIP(hipMalloc(&m.part, sizeof(double) * kMgParts));
  IFEM_HIP(hipMalloc(&m.part2, sizeof(double) * kMgParts));
  const size_t nlz = m.lev.size() * (3 * kMgLanczos + 1);
  IFEM_HIP(hipMalloc(&m.lzh, sizeof(double) * nlz));
  IFEM_HIP(hipHostMalloc(&m.h_lzh, sizeof(double) * nlz, hipHostMallocDefault));
  IFEM_HIP(hipHostMalloc(&m.h_coarse, sizeof(double) * (size_t)m.noff * ncc * Lc.nnode,
                         hipHostMallocDefault));
  IFEM_HIP(hipHostMalloc(&m.h_err, sizeof(int), hipHostMallocDefault));
  IFEM_HIP(hipEventCreate(&m.e_copy));
  IFEM_HIP(hipEventCreateWithFlags(&m.e_gal, hipEventDisableTiming));
  IFEM_HIP(hipEventCreateWithFlags(&m.e_side, hipEventDisableTiming));
  IFEM_HIP(hipStreamCreateWithFlags(&m.side, hipStreamNonBlocking)); // default priority
  IFEM_HIP(hipMalloc(&m.part_s, sizeof(double) * kMgParts));
  IFEM_HIP(hipMalloc(&m.part_s2, sizeof(double) * kMgParts));
  IFEM_HIP(hipMalloc(&m.err, sizeof(int)));
  IFEM_HIP(hipEventCreate(&m.e0));
  IFEM_HIP(hipEventCreate(&m.e1));
  m.patch = c.patch;
  m.neq = c.neq;
  m.nnz = c.nnz;
  return IFEM_GPU_OK;
}

// Kernel instantiation for the patch degree and component count.
#define MG_DISPATCH(P_, NC_, CALL)                                             \
  do {                                                                         \
    if ((NC_) == 3) {                                                          \
      if ((P_) == 1) { constexpr int KP = 1, KN = 3; CALL; }                   \
      else if ((P_) == 2) { constexpr int KP = 2, KN = 3; CALL; }              \
      else { constexpr int KP = 3, KN = 3; CALL; }                             \
    } else {                                                                   \
      if ((P_) == 1) { constexpr int KP = 1, KN = 1; CALL; }                   \
      else if ((P_) == 2) { constexpr int KP = 2, KN = 1; CALL; }              \
      else { constexpr int KP = 3, KN = 1; CALL; }                             \
    }                                                                          \
  } while (0)

static void level_spmv(Context& c, Multigrid& m, int l, const float* x, float* y,
                       const double* scal, hipStream_t st = nullptr)
{
  if (!st) st = c.stream;
  if (l == 0) {
    if (m.val16n) launch_spmv16_node3(c, m.val16n, (float)m.kscale, x, y, scal != nullptr, st);
    else launch_spmv16(c, m.val16, (float)m.kscale, x, y, scal != nullptr, st);
    return;
  }
  const MgLevel& L = m.lev[l];
  const unsigned gb = (unsigned)(((int64_t)L.nnode + 3) / 4); // one wave per node
  MG_DISPATCH(m.p, m.nc, (k_mg_spmv<KP, KN><<<gb, 256, 0, st>>>(L.S32, x, y, grid_of(L), scal)));
}

// Lanczos eigenvalue estimate of level l on stream st, partial buffers
// part/part2 (kMgParts each).
static void lanczos_level(Context& c, Multigrid& m, int l, hipStream_t st, double* part,
                          double* part2)
{
  MgLevel& L = m.lev[l];
  const int n = L.nvec;
  double* h = m.lzh + (size_t)l * (3 * kMgLanczos + 1);
  k_lz_start<<<kMgParts, kThreads, 0, st>>>(n, L.r, L.d, L.x, L.dinv, part);
  k_lz_sum<<<1, kThreads, 0, st>>>(part, kMgParts, h);
  for (int j = 0; j < kMgLanczos; j++) {
    level_spmv(c, m, l, L.x, L.q, nullptr, st);
    k_lz_dot<<<kMgParts, kThreads, 0, st>>>(n, L.x, L.q, part);
    k_lz_update<<<kMgParts, kThreads, 0, st>>>(n, L.r, L.d, L.q, L.dinv, part, kMgParts, h, j,
                                               part2);
    k_lz_pupdate<<<kMgParts, kThreads, 0, st>>>(n, L.x, L.d, part2, kMgParts, h, j);
  }
}

// Dense inverse of the coarsest operator (host Cholesky; the level is small).
static int coarse_inverse(Context& c, Multigrid& m)
{
  const MgLevel& L = m.lev.back();
  const int ncc = m.nc * m.nc, nc = m.nc;
  const double* S = m.h_coarse; // node-major, copied by mg_setup
  const int n = m.ncf;
  std::vector<int> pos(L.ndof, -1);
  int k = 0;
  for (int i = 0; i < L.ndof; i++)
    if (L.h_free[i]) pos[i] = k++;
  std::vector<double> A((size_t)n * n, 0.0);
  const int W1 = 2 * m.p + 1, p = m.p;
  for (int node = 0; node < L.nnode; node++) {
    const int i = node % L.n[0], j = (node / L.n[0]) % L.n[1], kk = node / (L.n[0] * L.n[1]);
    for (int off = 0; off < m.noff; off++) {
      const int e0 = off % W1 - p, e1 = (off / W1) % W1 - p, e2 = off / (W1 * W1) - p;
      const int ii = i + e0, jj = j + e1, k2 = kk + e2;
      if (ii < 0 || ii >= L.n[0] || jj < 0 || jj >= L.n[1] || k2 < 0 || k2 >= L.n[2]) continue;
      const int nb = ii + L.n[0] * (jj + L.n[1] * k2);
      for (int ci = 0; ci < nc; ci++)
        for (int cj = 0; cj < nc; cj++) {
          const int r = pos[node * nc + ci], s = pos[nb * nc + cj];
          if (r < 0 || s < 0) continue;
          A[(size_t)r * n + s] = S[((size_t)node * m.noff + off) * ncc + ci * nc + cj];
        }
    }
  }
  for (int r = 0; r < n; r++)
    for (int s = 0; s < r; s++) {
      const double v = 0.5 * (A[(size_t)r * n + s] + A[(size_t)s * n + r]);
      A[(size_t)r * n + s] = A[(size_t)s * n + r] = v;
    }
  // Cholesky A = L L^T (lower triangle in place)
  for (int j = 0; j < n; j++) {
    double dsum = A[(size_t)j * n + j];
    for (int q = 0; q < j; q++) dsum -= A[(size_t)j * n + q] * A[(size_t)j * n + q];
    if (!(dsum > 0.0)) return fail(IFEM_GPU_EDEVICE, "mg: coarsest operator is not positive definite");
    const double ljj = std::sqrt(dsum);
    A[(size_t)j * n + j] = ljj;
    for (int r = j + 1; r < n; r++) {
      double s = A[(size_t)r * n + j];
      for (int q = 0; q < j; q++) s -= A[(size_t)r * n + q] * A[(size_t)j * n + q];
      A[(size_t)r * n + j] = s / ljj;
    }
  }
  // inverse column by column: solve L y = e_k, L^T x = y
  std::vector<float> inv((size_t)n * n);
  std::vector<double> y(n);
  for (int col = 0; col < n; col++) {
    for (int r = 0; r < n; r++) {
      double s = r == col ? 1.0 : 0.0;
      for (int q = 0; q < r; q++) s -= A[(size_t)r * n + q] * y[q];
      y[r] = s / A[(size_t)r * n + r];
    }
    for (int r = n - 1; r >= 0; r--) {
      double s = y[r];
      for (int q = r + 1; q < n; q++) s -= A[(size_t)q * n + r] * y[q];
      y[r] = s / A[(size_t)r * n + r];
    }
    for (int r = 0; r < n; r++) inv[(size_t)r * n + col] = (float)y[r];
  }
  IFEM_HIP(hipMemcpyAsync(m.cinv, inv.data(), sizeof(float) * inv.size(), hipMemcpyHostToDevice,
                          c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream)); // inv is a host temporary
  return IFEM_GPU_OK;
}

int mg_setup(Context& c, int degree)
{
  if (!c.fp64_ready) return fail(IFEM_GPU_EINVAL, "mg: no FP64 matrix assembled");
  if (!c.multi.empty()) return fail(IFEM_GPU_EUNSUPPORTED, "mg: one patch per system only");
  if (c.mpc) return fail(IFEM_GPU_EUNSUPPORTED, "mg: constraints with master DOFs");
  if (degree < 1 || degree > 8) return fail(IFEM_GPU_EINVAL, "mg: smoothing degree must be 1..8");
  if (!c.mg || c.mg->patch != c.patch || c.mg->neq != c.neq || c.mg->nnz != c.nnz) {
    c.mg.reset(new Multigrid());
    int rc = mg_build(c, *c.mg);
    if (rc) {
      c.mg.reset();
      return rc;
    }
  }
  Multigrid& m = *c.mg;
  m.degree = degree;
  IFEM_HIP(hipEventRecord(m.e0, c.stream));
  // fine level: inverse diagonal, then K / max diag in FP16 for the smoother
  MgLevel& F = m.lev[0];
  k_mg_diag<<<kMgParts, kThreads, 0, c.stream>>>(c.neq, c.d_rowptr, c.d_colind, c.d_val, F.dinv,
                                                 m.part);
  k_mg_max<<<1, 64, 0, c.stream>>>(m.part, kMgParts, m.lzh);
  IFEM_HIP(hipMemcpyAsync(m.h_lzh, m.lzh, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  m.kscale = m.h_lzh[0] > 0.0 ? m.h_lzh[0] : 1.0;
  if (c.d_bcol) { // node blocks: one 8-byte load per row and neighbour in the smoother
    if (!m.val16n) IFEM_HIP(hipMalloc(&m.val16n, sizeof(uint2) * (c.nnz > 0 ? c.nnz / 3 : 1)));
    const int rc = demote16_node3(c, 1.0 / m.kscale, m.val16n);
    if (rc) return rc;
  } else {
    if (!m.val16) IFEM_HIP(hipMalloc(&m.val16, sizeof(__half) * (c.nnz > 0 ? c.nnz : 1)));
    k_mg_demote16<<<4096, 256, 0, c.stream>>>(c.nnz, c.d_val, m.val16, 1.0 / m.kscale);
  }
  // fine operator as a stencil, then Galerkin products level by level
  IFEM_HIP(hipMemsetAsync(m.err, 0, sizeof(int), c.stream));
  const int ncc = m.nc * m.nc;
  IFEM_HIP(hipMemsetAsync(m.S0, 0, sizeof(double) * m.s0cap, c.stream));
  if (c.d_bcol && m.nc == 3) {
    const int nnode = c.neq / 3;
    const unsigned gb = (unsigned)(((int64_t)nnode * 64 + 255) / 256);
    if (m.p == 1)
      k_mg_csr_scatter3<1><<<gb, 256, 0, c.stream>>>(nnode, grid_of(F), m.dof_of_eq, c.d_rowptr,
                                                     c.d_bcol, c.d_val, m.S0, m.err);
    else if (m.p == 2)
      k_mg_csr_scatter3<2><<<gb, 256, 0, c.stream>>>(nnode, grid_of(F), m.dof_of_eq, c.d_rowptr,
                                                     c.d_bcol, c.d_val, m.S0, m.err);
    else
      k_mg_csr_scatter3<3><<<gb, 256, 0, c.stream>>>(nnode, grid_of(F), m.dof_of_eq, c.d_rowptr,
                                                     c.d_bcol, c.d_val, m.S0, m.err);
  } else {
    MG_DISPATCH(m.p, m.nc,
                (k_mg_csr_scatter<KP, KN><<<(unsigned)(((int64_t)c.neq * 64 + 255) / 256), 256,
                                            0, c.stream>>>(c.neq, grid_of(F), m.dof_of_eq,
                                                           c.d_rowptr, c.d_colind, c.d_val, m.S0,
                                                           m.err)));
  }
  IFEM_HIP(hipGetLastError());
  for (size_t l = 1; l < m.lev.size(); l++) {
    MgLevel& Fl = m.lev[l - 1];
    MgLevel& C = m.lev[l];
    const double* Sin = l == 1 ? m.S0 : Fl.S64;
    Grid3 g0 = grid_of(Fl);
    Grid3 g1{C.n[0], Fl.n[1], Fl.n[2]};
    Grid3 g2{C.n[0], C.n[1], Fl.n[2]};
    const int64_t n1 = (int64_t)g1.n0 * g1.n1 * g1.n2, n2 = (int64_t)g2.n0 * g2.n1 * g2.n2;
    if ((size_t)n1 * m.noff * ncc > m.t1cap || (size_t)n2 * m.noff * ncc > m.t2cap)
      return fail(IFEM_GPU_EDEVICE, "mg: work buffer too small");
    MG_DISPATCH(m.p, m.nc,
                (k_mg_galerkin<KP, KN><<<grid1(n1 * m.noff), 256, 0, c.stream>>>(
                    Sin, m.T1, g0, 0, C.n[0], C.dir[0].clo, C.dir[0].ccnt, C.dir[0].cw, nullptr,
                    nullptr, nullptr)));
    MG_DISPATCH(m.p, m.nc,
                (k_mg_galerkin<KP, KN><<<grid1(n2 * m.noff), 256, 0, c.stream>>>(
                    m.T1, m.T2, g1, 1, C.n[1], C.dir[1].clo, C.dir[1].ccnt, C.dir[1].cw, nullptr,
                    nullptr, nullptr)));
    MG_DISPATCH(m.p, m.nc,
                (k_mg_galerkin<KP, KN><<<grid1((int64_t)C.nnode * m.noff), 256, 0, c.stream>>>(
                    m.T2, C.S64, g2, 2, C.n[2], C.dir[2].clo, C.dir[2].ccnt, C.dir[2].cw, C.freem,
                    C.S32, C.dinv)));
    IFEM_HIP(hipGetLastError());
  }
  // coarsest operator to the host (pinned, asynchronous) while the GPU runs
  // the eigenvalue estimates
  const MgLevel& Lc = m.lev.back();
  IFEM_HIP(hipMemcpyAsync(m.h_coarse, Lc.S64, sizeof(double) * (size_t)m.noff * ncc * Lc.nnode,
                          hipMemcpyDeviceToHost, c.stream));
  IFEM_HIP(hipMemcpyAsync(m.h_err, m.err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  IFEM_HIP(hipEventRecord(m.e_copy, c.stream));
  // largest eigenvalue of D^-1 A per smoothed level: Lanczos from CG steps.
  // The coarse levels (small, launch-latency bound) run on a side stream
  // beside the fine level (bandwidth bound); they share no data.
  const int nl = (int)m.lev.size() - 1;
  IFEM_HIP(hipEventRecord(m.e_gal, c.stream));
  IFEM_HIP(hipStreamWaitEvent(m.side, m.e_gal, 0));
  for (int l = 1; l < nl; l++) lanczos_level(c, m, l, m.side, m.part_s, m.part_s2);
  IFEM_HIP(hipEventRecord(m.e_side, m.side));
  if (nl > 0) lanczos_level(c, m, 0, c.stream, m.part, m.part2);
  IFEM_HIP(hipStreamWaitEvent(c.stream, m.e_side, 0));
  IFEM_HIP(hipGetLastError());
  IFEM_HIP(hipMemcpyAsync(m.h_lzh, m.lzh, sizeof(double) * (size_t)nl * (3 * kMgLanczos + 1),
                          hipMemcpyDeviceToHost, c.stream));
  // the iteration vectors must start from zero again
  for (int l = 0; l < (int)m.lev.size(); l++)
    k_fill<float><<<grid1(m.lev[l].nvec), 256, 0, c.stream>>>(m.lev[l].nvec, m.lev[l].x, 0.f);
  IFEM_HIP(hipEventSynchronize(m.e_copy));
  if (*m.h_err) return fail(IFEM_GPU_EDEVICE, "mg: matrix pattern differs from the stencil layout");
  int rc = coarse_inverse(c, m);
  if (rc) return rc;
  IFEM_HIP(hipEventRecord(m.e1, c.stream));
  IFEM_HIP(hipEventSynchronize(m.e1));
  for (int l = 0; l < nl; l++) {
    const double lam = lanczos_max(m.h_lzh + (size_t)l * (3 * kMgLanczos + 1), kMgLanczos);
    if (!(lam > 0.0) || !std::isfinite(lam))
      return fail(IFEM_GPU_EDEVICE, "mg: eigenvalue estimate failed");
    m.lev[l].lmax = lam;
  }
  float ms = 0.f;
  IFEM_HIP(hipEventElapsedTime(&ms, m.e0, m.e1));
  m.setup_ms = ms;
  return IFEM_GPU_OK;
}

static void fmap_grid(const Multigrid& m, const Context& c, int l, const int*& fmap)
{
  fmap = l == 0 ? c.d_eq : nullptr;
  (void)m;
}

// One V-cycle on level l: L.x = M_l^-1 L.b.
static void vcycle(Context& c, Multigrid& m, int l, const double* scal)
{
  MgLevel& L = m.lev[l];
  const int n = L.nvec;
  const unsigned gb = std::min<unsigned>(grid1(n), 4096);
  if (l == (int)m.lev.size() - 1) {
    k_mg_coarse<<<(unsigned)((m.ncf + 3) / 4), 256, 0, c.stream>>>(m.ncf, m.cdofs, m.cinv, L.b,
                                                                   L.x, scal);
    return;
  }
  // Chebyshev coefficients on [0.1, 1.1] x lambda_max
  const double hi = 1.05 * L.lmax, lo = 0.1 * L.lmax;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  std::vector<float> c1(m.degree), c2(m.degree);
  double rho = 1.0 / sigma;
  c1[0] = 0.f;
  c2[0] = (float)(1.0 / theta);
  for (int s = 1; s < m.degree; s++) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    c1[s] = (float)(rn * rho);
    c2[s] = (float)(2.0 * rn / delta);
    rho = rn;
  }
  // pre-smoothing from x = 0
  k_mg_cheb<<<gb, 256, 0, c.stream>>>(n, 0, L.b, L.r, nullptr, L.d, L.x, L.dinv, c1[0], c2[0], scal);
  level_spmv(c, m, l, L.d, L.q, scal);
  for (int s = 1; s < m.degree; s++) {
    k_mg_cheb<<<gb, 256, 0, c.stream>>>(n, 1, L.b, L.r, L.q, L.d, L.x, L.dinv, c1[s], c2[s], scal);
    level_spmv(c, m, l, L.d, L.q, scal);
  }
  // restriction of r - q, coarse correction, prolongation
  MgLevel& C = m.lev[l + 1];
  const int* fmap = nullptr;
  fmap_grid(m, c, l, fmap);
  const DirDev d0 = dirdev(C.dir[0]), d1 = dirdev(C.dir[1]), d2 = dirdev(C.dir[2]);
  const unsigned gr = grid1((int64_t)C.nnode * m.nc);
  if (m.nc == 3)
    k_mg_restrict<3><<<gr, 256, 0, c.stream>>>(L.r, L.q, fmap, grid_of(L), C.b, C.freem, grid_of(C),
                                               d0, d1, d2, scal);
  else
    k_mg_restrict<1><<<gr, 256, 0, c.stream>>>(L.r, L.q, fmap, grid_of(L), C.b, C.freem, grid_of(C),
                                               d0, d1, d2, scal);
  vcycle(c, m, l + 1, scal);
  if (m.nc == 3)
    k_mg_prolong<3><<<grid1(L.nnode), 256, 0, c.stream>>>(L.x, fmap, grid_of(L), C.x, grid_of(C), d0,
                                                          d1, d2, scal);
  else
    k_mg_prolong<1><<<grid1(L.nnode), 256, 0, c.stream>>>(L.x, fmap, grid_of(L), C.x, grid_of(C), d0,
                                                          d1, d2, scal);
  // post-smoothing: r = b - A x, then the same Chebyshev steps
  level_spmv(c, m, l, L.x, L.q, scal);
  k_mg_cheb<<<gb, 256, 0, c.stream>>>(n, 2, L.b, L.r, L.q, L.d, L.x, L.dinv, c1[0], c2[0], scal);
  for (int s = 1; s < m.degree; s++) {
    level_spmv(c, m, l, L.d, L.q, scal);
    k_mg_cheb<<<gb, 256, 0, c.stream>>>(n, 1, L.b, L.r, L.q, L.d, L.x, L.dinv, c1[s], c2[s], scal);
  }
}

template <typename V>
static int mg_apply_t(Context& c, const V* r, V* z, const double* scal)
{
  if (!c.mg || c.mg->lev.empty() || c.mg->lev[0].lmax <= 0.0)
    return fail(IFEM_GPU_EINVAL, "mg: no multigrid hierarchy (mg_setup)");
  Multigrid& m = *c.mg;
  MgLevel& F = m.lev[0];
  const int n = c.neq;
  k_mg_cast_in<V><<<grid1(n), 256, 0, c.stream>>>(n, r, F.b);
  vcycle(c, m, 0, scal);
  k_mg_cast_out<V><<<grid1(n), 256, 0, c.stream>>>(n, F.x, z);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

int mg_apply(Context& c, const double* r, double* z, const double* scal)
{
  return mg_apply_t<double>(c, r, z, scal);
}

int mg_apply(Context& c, const float* r, float* z, const double* scal)
{
  return mg_apply_t<float>(c, r, z, scal);
}

} // namespace ifem
