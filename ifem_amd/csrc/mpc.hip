// mpc.hip -- constraint equations with master DOFs (u_s = c0 + sum c_m u_m).
//
// The reference eliminates them element by element while scattering
// (assemSparse with constrained columns, src/LinAlg/SparseMatrix.C:612-688;
// SAM::assembleRHS, src/LinAlg/SAM.C:605-617).  Here the line kernels
// assemble an extended system in which every slave with masters is an extra
// unknown (rows and columns are owned by nodes, as for free DOFs), and one
// kernel condenses it to the free equations:
//   K = T^T K_x T,  b = T^T (b_x - K_x t0),  u_x = T u + t0,
// T the identity on free DOFs and the master coefficients on slave rows, t0
// the constants c0.  This is the same sum as the element-by-element
// elimination, collected per reduced row in a fixed order.  The extended and
// reduced patterns and the condensation tables are built on the host.
#include <algorithm>
#include <cmath>
#include <vector>

#include "devutil.hpp"
#include "internal.hpp"

namespace ifem {

int host_union_pattern(const Context& c, const std::vector<SubPatch>& subs, int neq,
                       std::vector<int64_t>& rowptr, std::vector<int32_t>& colind);
int sub_position_maps(Context& c, std::vector<SubPatch>& subs, const std::vector<int64_t>& rowptr,
                      const std::vector<int32_t>& colind);
int solver_setup(Context& c);

void mpc_free(Context& c)
{
  if (!c.mpc) return;
  MpcSystem& x = *c.mpc;
  hipFree(x.d_rowptr); hipFree(x.d_colind); hipFree(x.d_val); hipFree(x.d_val32); hipFree(x.d_rhs);
  hipFree(x.ext.d_leq); hipFree(x.ext.d_lgval); hipFree(x.ext.d_pmap); hipFree(x.ext.d_pptr);
  hipFree(x.d_cptr); hipFree(x.d_cext); hipFree(x.d_cw);
  hipFree(x.d_sptr); hipFree(x.d_smaster); hipFree(x.d_scoef); hipFree(x.d_sc0);
  hipFree(x.d_slave_of_dof); hipFree(x.d_xdof);
  c.mpc.reset();
}

template <typename T>
static int up(T*& d, const std::vector<T>& h)
{
  IFEM_HIP(hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size())));
  if (!h.empty()) IFEM_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return IFEM_GPU_OK;
}

// Called after the free numbering is bound (c.h_eq: slaves -1; c.h_gval:
// c0 of every constraint).  Constraints without masters stay prescribed
// values; the others become extended unknowns.
int mpc_bind(Context& c, int nceq, const int* mpmceq, const int* mmceq, const double* ttcc)
{
  mpc_free(c);
  const int ndof = c.ndof, neq = c.neq;
  std::unique_ptr<MpcSystem> xp(new MpcSystem());
  MpcSystem& x = *xp;
  x.h_slave_of_dof.assign(ndof, -1);
  x.h_sptr.assign(1, 0);
  std::vector<int> xeq(c.h_eq); // extended numbering per DOF
  std::vector<double> gv(ndof, 0.0);
  // walk the constraints in order: slaves with masters get ne = neq + k
  for (int k = 1; k <= nceq; k++) {
    const int ip = mpmceq[k - 1], jp = mpmceq[k] - 1;
    const int sd = mmceq[ip - 1] - 1;
    if (sd < 0 || sd >= ndof) return fail(IFEM_GPU_EINVAL, "mpc: slave DOF out of range");
    if (jp == ip) { // prescribed value only
      gv[sd] = ttcc[ip - 1];
      continue;
    }
    const int s = x.ns++;
    x.h_slave_of_dof[sd] = s;
    xeq[sd] = neq + s;
    x.h_sc0.push_back(ttcc[ip - 1]);
    for (int q = ip; q < jp; q++) { // 0-based positions of the masters
      const int md = mmceq[q] - 1;
      if (md < 0 || md >= ndof) return fail(IFEM_GPU_EINVAL, "mpc: master DOF out of range");
      if (c.h_eq[md] < 0)
        return fail(IFEM_GPU_EINVAL, "mpc: master DOFs must be free (resolve chains first)");
      x.h_smaster.push_back(c.h_eq[md]);
      x.h_scoef.push_back(ttcc[q]);
    }
    x.h_sptr.push_back((int)x.h_smaster.size());
  }
  x.ne = neq + x.ns;
  x.ext.patch = c.patch;
  // single patch: local DOF = global DOF
  x.ext.h_leq = xeq;
  int rc = up(x.ext.d_leq, x.ext.h_leq);
  bool anyg = false;
  for (double g : gv) anyg = anyg || g != 0.0;
  if (!rc && anyg) rc = up(x.ext.d_lgval, gv);
  if (!rc) rc = up(x.d_sptr, x.h_sptr);
  if (!rc) rc = up(x.d_smaster, x.h_smaster);
  if (!rc) rc = up(x.d_scoef, x.h_scoef);
  if (!rc) rc = up(x.d_sc0, x.h_sc0);
  if (!rc) rc = up(x.d_slave_of_dof, x.h_slave_of_dof);
  if (rc) return rc;
  c.mpc = std::move(xp);
  return IFEM_GPU_OK;
}

// Extended pattern (node couplings over free DOFs and slaves with masters)
// and the reduced pattern: row r holds the images under T of the columns of
// every extended row contributing to r (r itself and the slaves r masters).
int mpc_build_pattern(Context& c)
{
  MpcSystem& x = *c.mpc;
  const int neq = c.neq;
  std::vector<int64_t> xrp;
  std::vector<int32_t> xci;
  // maps of an earlier call on this system are replaced: free them first
  hipFree(x.ext.d_pmap);
  hipFree(x.ext.d_pptr);
  x.ext.d_pmap = nullptr;
  x.ext.d_pptr = nullptr;
  // subs[0] is a shallow copy: the device arrays it points at stay owned by
  // x.ext (mpc_free releases them)
  std::vector<SubPatch> subs(1, x.ext);
  int rc = host_union_pattern(c, subs, x.ne, xrp, xci);
  if (!rc) rc = sub_position_maps(c, subs, xrp, xci);
  x.ext.d_pmap = subs[0].d_pmap;
  x.ext.d_pptr = subs[0].d_pptr;
  if (rc) return rc;
  x.nnz = xrp[x.ne];
  // contributions per reduced row
  std::vector<std::vector<std::pair<int, double>>> contrib(neq > 0 ? neq : 1);
  for (int r = 0; r < neq; r++) contrib[r].push_back({r, 1.0});
  for (int s = 0; s < x.ns; s++)
    for (int q = x.h_sptr[s]; q < x.h_sptr[s + 1]; q++)
      contrib[x.h_smaster[q]].push_back({neq + s, x.h_scoef[q]});
  std::vector<int> cptr(neq + 1, 0), cext;
  std::vector<double> cw;
  std::vector<int64_t> rowptr(neq + 1, 0);
  std::vector<std::vector<int>> rows(neq > 0 ? neq : 1);
  for (int r = 0; r < neq; r++) {
    std::vector<int>& row = rows[r];
    for (const auto& e : contrib[r]) {
      cext.push_back(e.first);
      cw.push_back(e.second);
      for (int64_t q = xrp[e.first]; q < xrp[e.first + 1]; q++) {
        const int col = xci[q];
        if (col < neq) row.push_back(col);
        else
          for (int m = x.h_sptr[col - neq]; m < x.h_sptr[col - neq + 1]; m++)
            row.push_back(x.h_smaster[m]);
      }
    }
    cptr[r + 1] = (int)cext.size();
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    rowptr[r + 1] = rowptr[r] + (int64_t)row.size();
  }
  const int64_t nnz = rowptr[neq];
  std::vector<int32_t> colind(nnz > 0 ? nnz : 1);
  for (int r = 0; r < neq; r++) std::copy(rows[r].begin(), rows[r].end(), colind.begin() + rowptr[r]);
  // device arrays: extended system, reduced system, tables
  hipFree(x.d_rowptr); hipFree(x.d_colind); hipFree(x.d_val); hipFree(x.d_val32); hipFree(x.d_rhs);
  x.d_rowptr = nullptr; x.d_colind = nullptr; x.d_val = nullptr; x.d_val32 = nullptr; x.d_rhs = nullptr;
  hipFree(x.d_cptr); hipFree(x.d_cext); hipFree(x.d_cw);
  x.d_cptr = nullptr; x.d_cext = nullptr; x.d_cw = nullptr;
  rc = up(x.d_rowptr, xrp);
  if (!rc) rc = up(x.d_colind, xci);
  if (!rc) rc = up(x.d_cptr, cptr);
  if (!rc) rc = up(x.d_cext, cext);
  if (!rc) rc = up(x.d_cw, cw);
  if (rc) return rc;
  IFEM_HIP(hipMalloc(&x.d_val, sizeof(double) * (x.nnz > 0 ? x.nnz : 1)));
  IFEM_HIP(hipMalloc(&x.d_rhs, sizeof(double) * (x.ne > 0 ? x.ne : 1)));
  hipFree(c.d_rowptr); hipFree(c.d_colind); hipFree(c.d_val); hipFree(c.d_val32); hipFree(c.d_rhs);
  c.d_rowptr = nullptr; c.d_colind = nullptr; c.d_val = nullptr; c.d_val32 = nullptr; c.d_rhs = nullptr;
  c.fp64_ready = c.fp32_ready = false;
  c.nnz = nnz;
  rc = up(c.d_rowptr, rowptr);
  if (!rc) rc = up(c.d_colind, colind);
  if (rc) return rc;
  IFEM_HIP(hipMalloc(&c.d_val, sizeof(double) * (nnz > 0 ? nnz : 1)));
  IFEM_HIP(hipMalloc(&c.d_rhs, sizeof(double) * (neq > 0 ? neq : 1)));
  IFEM_HIP(hipMemset(c.d_rhs, 0, sizeof(double) * (neq > 0 ? neq : 1)));
  c.pattern_ready = true;
  return solver_setup(c);
}

// One thread per reduced row, sequential and in a fixed order:
// K[r, :] = sum_e w_e T^T-images of the extended row e; b[r] = sum_e w_e
// (b_x[e] - sum over slave columns k of K_x[e, k] c0_k).  val == nullptr:
// right-hand side only; rhs == nullptr: matrix only.
template <typename T>
__global__ __launch_bounds__(256) void k_mpc_condense(
    int neq, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci, T* __restrict__ val,
    double* __restrict__ rhs, const int64_t* __restrict__ xrp, const int32_t* __restrict__ xci,
    const T* __restrict__ xval, const double* __restrict__ xrhs, const int* __restrict__ cptr,
    const int* __restrict__ cext, const double* __restrict__ cw, const int* __restrict__ sptr,
    const int* __restrict__ smaster, const double* __restrict__ scoef,
    const double* __restrict__ sc0)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= neq) return;
  const int64_t b0 = rp[r], b1 = rp[r + 1];
  if (val)
    for (int64_t q = b0; q < b1; q++) val[q] = T(0);
  double racc = 0.0;
  for (int cc = cptr[r]; cc < cptr[r + 1]; cc++) {
    const int e = cext[cc];
    const double w = cw[cc];
    if (rhs) racc += w * xrhs[e];
    for (int64_t t = xrp[e]; t < xrp[e + 1]; t++) {
      const int col = xci[t];
      const double v = w * (double)xval[t];
      if (col < neq) {
        if (!val) continue;
        int64_t lo = b0, hi = b1 - 1;
        while (lo <= hi) {
          const int64_t mid = (lo + hi) >> 1;
          const int cm = ci[mid];
          if (cm == col) { val[mid] += (T)v; break; }
          if (cm < col) lo = mid + 1; else hi = mid - 1;
        }
      } else {
        const int s = col - neq;
        if (rhs) racc -= v * sc0[s];
        if (!val) continue;
        for (int m = sptr[s]; m < sptr[s + 1]; m++) {
          const int mc = smaster[m];
          int64_t lo = b0, hi = b1 - 1;
          while (lo <= hi) {
            const int64_t mid = (lo + hi) >> 1;
            const int cm = ci[mid];
            if (cm == mc) { val[mid] += (T)(v * scoef[m]); break; }
            if (cm < mc) lo = mid + 1; else hi = mid - 1;
          }
        }
      }
    }
  }
  if (rhs) rhs[r] = racc;
}

int mpc_condense(Context& c, bool fp32, bool rhs_only)
{
  MpcSystem& x = *c.mpc;
  const unsigned g = (unsigned)((c.neq + 255) / 256);
  if (c.neq == 0) return IFEM_GPU_OK;
  if (fp32) {
    k_mpc_condense<float><<<g, 256, 0, c.stream>>>(
        c.neq, c.d_rowptr, c.d_colind, c.d_val32, nullptr, x.d_rowptr, x.d_colind, x.d_val32,
        nullptr, x.d_cptr, x.d_cext, x.d_cw, x.d_sptr, x.d_smaster, x.d_scoef, x.d_sc0);
  } else {
    k_mpc_condense<double><<<g, 256, 0, c.stream>>>(
        c.neq, c.d_rowptr, c.d_colind, rhs_only ? nullptr : c.d_val, c.d_rhs, x.d_rowptr,
        x.d_colind, x.d_val, x.d_rhs, x.d_cptr, x.d_cext, x.d_cw, x.d_sptr, x.d_smaster,
        x.d_scoef, x.d_sc0);
  }
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

// Expanded solution per DOF on the device (SAM::expandVector with
// scaleSD = 1, src/LinAlg/SAM.C:783-807).
__global__ void k_mpc_expand(int ndof, const int* __restrict__ eq, const double* __restrict__ x,
                             const int* __restrict__ slave, const int* __restrict__ sptr,
                             const int* __restrict__ smaster, const double* __restrict__ scoef,
                             const double* __restrict__ sc0, const double* __restrict__ gval,
                             double* __restrict__ xdof)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndof) return;
  const int e = eq[i];
  if (e >= 0) {
    xdof[i] = x[e];
    return;
  }
  const int s = slave[i];
  if (s < 0) {
    xdof[i] = gval ? gval[i] : 0.0;
    return;
  }
  double v = sc0[s];
  for (int m = sptr[s]; m < sptr[s + 1]; m++) v += scoef[m] * x[smaster[m]];
  xdof[i] = v;
}

int mpc_expand_device(Context& c)
{
  MpcSystem& x = *c.mpc;
  if (!x.d_xdof) IFEM_HIP(hipMalloc(&x.d_xdof, sizeof(double) * (c.ndof > 0 ? c.ndof : 1)));
  k_mpc_expand<<<(c.ndof + 255) / 256, 256, 0, c.stream>>>(
      c.ndof, c.d_eq, c.w.x, x.d_slave_of_dof, x.d_sptr, x.d_smaster, x.d_scoef, x.d_sc0,
      x.ext.d_lgval, x.d_xdof);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

void mpc_expand_host(const Context& c, const std::vector<double>& xs, double* xdof)
{
  const MpcSystem& x = *c.mpc;
  for (int i = 0; i < c.ndof; i++) {
    if (c.h_eq[i] >= 0) { xdof[i] = xs[c.h_eq[i]]; continue; }
    const int s = x.h_slave_of_dof[i];
    if (s < 0) { xdof[i] = c.h_gval.empty() ? 0.0 : c.h_gval[i]; continue; }
    double v = x.h_sc0[s];
    for (int m = x.h_sptr[s]; m < x.h_sptr[s + 1]; m++) v += x.h_scoef[m] * xs[x.h_smaster[m]];
    xdof[i] = v;
  }
}

} // namespace ifem
