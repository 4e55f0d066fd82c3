// api.hip -- C ABI entry points of libifem_gpu.so: contexts, patch upload,
// SAM binding and data access.  Kernels live in assemble.hip / solver.hip.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ifem {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
  g_last_error = msg;
  return code;
}

// Staging memory of ifem_gpu_spmv_noalloc: x in the first kNoAllocMaxEq
// doubles, y in the second (2 x 4,194,304 doubles = 64 MiB on every GPU the
// library runs on, held until the process ends).  Host code never passes
// this name to a kernel or a copy: it uses the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
constexpr int kNoAllocMaxEq = 4194304;
static __device__ __align__(256) __attribute__((used)) double scratch[2 * (size_t)kNoAllocMaxEq];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), resolved
// once per device; nullptr when the lookup fails (the entry point then
// reports it).
static double* noalloc_scratch(int device)
{
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = (double*)p;
  }
  return g_scratch_ptr[device];
}

static void free_system(Context& c)
{
  solver_free(c);
  c.mg.reset();
  mpc_free(c);
  for (SubPatch& sp : c.multi) {
    hipFree(sp.d_leq);
    hipFree(sp.d_lgval);
    hipFree(sp.d_pmap);
    hipFree(sp.d_pptr);
  }
  c.multi.clear();
  hipFree(c.d_eq);
  hipFree(c.d_gval);
  c.d_gval = nullptr;
  c.h_gval.clear();
  hipFree(c.d_rowptr);
  hipFree(c.d_colind);
  hipFree(c.d_bcol);
  c.d_bcol = nullptr;
  hipFree(c.d_val);
  hipFree(c.d_val32);
  hipFree(c.d_rhs);
  c.d_eq = nullptr;
  c.d_rowptr = nullptr;
  c.d_colind = nullptr;
  c.d_val = nullptr;
  c.d_val32 = nullptr;
  c.d_rhs = nullptr;
  c.pattern_ready = c.fp64_ready = c.fp32_ready = false;
  c.nnz = 0;
  hipFree(c.d_netin);
  hipFree(c.d_rold);
  hipFree(c.d_rold32);
  hipFree(c.d_zn);
  hipFree(c.d_zn32);
  c.d_zn = nullptr;
  c.d_zn32 = nullptr;
  hipFree(c.d_aux);
  c.d_netin = nullptr;
  c.d_rold = nullptr;
  c.d_rold32 = nullptr;
  c.d_aux = nullptr;
  c.learned = false;
}

} // namespace ifem

using namespace ifem;

#define CHECK_CTX(ctx)                                                       \
  do {                                                                       \
    if (!(ctx)) return fail(IFEM_GPU_EINVAL, "null context");                \
    hipError_t e__ = hipSetDevice((ctx)->c.device);                          \
    if (e__ != hipSuccess)                                                   \
      return fail(IFEM_GPU_EDEVICE, std::string("hipSetDevice: ") +          \
                                        hipGetErrorString(e__));             \
  } while (0)

extern "C" {

const char* ifem_gpu_last_error(void) { return g_last_error.c_str(); }

int ifem_gpu_abi_version(void) { return IFEM_GPU_ABI_VERSION; }

int ifem_gpu_ctx_create(int device, void* stream, ifem_gpu_ctx** out)
{
  if (!out) return fail(IFEM_GPU_EINVAL, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(IFEM_GPU_EDEVICE, "no HIP device available");
  if (device < 0 || device >= ndev)
    return fail(IFEM_GPU_EINVAL, "device index out of range");
  IFEM_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  IFEM_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(IFEM_GPU_EUNSUPPORTED,
                std::string("libifem_gpu is built for gfx950, device is ") + prop.gcnArchName);
  ifem_gpu_ctx* ctx = new (std::nothrow) ifem_gpu_ctx();
  if (!ctx) return fail(IFEM_GPU_ENOMEM, "out of host memory");
  ctx->c.device = device;
  if (stream) {
    ctx->c.stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return fail(IFEM_GPU_EDEVICE, "hipStreamCreate failed");
    }
    ctx->c.own_stream = true;
  }
  if (hipEventCreate(&ctx->c.ev0) != hipSuccess || hipEventCreate(&ctx->c.ev1) != hipSuccess ||
      hipEventCreate(&ctx->c.ev_mid) != hipSuccess) {
    delete ctx;
    return fail(IFEM_GPU_EDEVICE, "hipEventCreate failed");
  }
  // set-up of the no-allocation entry point, with the device current
  ctx->c.noalloc_scratch = noalloc_scratch(device);
  *out = ctx;
  return IFEM_GPU_OK;
}

int ifem_gpu_ctx_destroy(ifem_gpu_ctx* ctx)
{
  if (!ctx) return IFEM_GPU_OK;
  hipSetDevice(ctx->c.device);
  hipStreamSynchronize(ctx->c.stream);
  free_system(ctx->c);
  for (auto& p : ctx->c.patches) {
    for (int d = 0; d < 3; d++) hipFree(p.d_knots[d]);
    hipFree(p.d_coefs);
  }
  for (auto& t : ctx->c.basis_cache)
    for (int d = 0; d < 3; d++) {
      hipFree(t.tabN[d]);
      hipFree(t.tabD[d]);
      hipFree(t.span[d]);
    }
  for (int k = 0; k < 3; k++) hipFree(ctx->c.d_modes[k]);
  ctx->c.net.reset();
  ctx->c.batch.reset();
  hipFree(ctx->c.d_asmargs);
  hipFree(ctx->c.d_gpgeo);
  hipFree(ctx->c.d_facew);
  hipEventDestroy(ctx->c.ev0);
  hipEventDestroy(ctx->c.ev1);
  hipEventDestroy(ctx->c.ev_mid);
  if (ctx->c.own_stream) hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return IFEM_GPU_OK;
}

void* ifem_gpu_ctx_stream(ifem_gpu_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

int ifem_gpu_synchronize(ifem_gpu_ctx* ctx)
{
  CHECK_CTX(ctx);
  IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
  return IFEM_GPU_OK;
}

int ifem_gpu_patch_create(ifem_gpu_ctx* ctx, const int order[3], const int ncoef[3],
                          const double* const knots[3], const double* coefs, int rational,
                          int* patch_id)
{
  CHECK_CTX(ctx);
  if (!order || !ncoef || !knots || !coefs || !patch_id)
    return fail(IFEM_GPU_EINVAL, "patch_create: null argument");
  if (rational) return fail(IFEM_GPU_EUNSUPPORTED, "patch_create: rational patches are not supported");
  PatchRec P;
  for (int d = 0; d < 3; d++) {
    const int p = order[d] - 1, n = ncoef[d];
    if (p < 1 || p > kMaxP)
      return fail(IFEM_GPU_EUNSUPPORTED, "patch_create: degree must be 1.." + std::to_string(kMaxP));
    if (n <= p) return fail(IFEM_GPU_EINVAL, "patch_create: ncoef must exceed the degree");
    if (!knots[d]) return fail(IFEM_GPU_EINVAL, "patch_create: null knot vector");
    const double* t = knots[d];
    for (int i = 0; i + 1 < n + p + 1; i++)
      if (t[i + 1] < t[i]) return fail(IFEM_GPU_EINVAL, "patch_create: knots must be non-decreasing");
    for (int i = 0; i < p; i++)
      if (t[i] != t[p] || t[n + i + 1] != t[n])
        return fail(IFEM_GPU_EUNSUPPORTED, "patch_create: knot vectors must be open");
    for (int s = p; s < n; s++)
      if (!(t[s + 1] > t[s]))
        return fail(IFEM_GPU_EUNSUPPORTED, "patch_create: repeated interior knots are not supported");
    P.p[d] = p;
    P.n[d] = n;
    P.nel[d] = n - p;
    P.knots[d].assign(t, t + n + p + 1);
  }
  const size_t nnod = (size_t)P.n[0] * P.n[1] * P.n[2];
  for (int d = 0; d < 3; d++) {
    IFEM_HIP(hipMalloc(&P.d_knots[d], sizeof(double) * P.knots[d].size()));
    IFEM_HIP(hipMemcpy(P.d_knots[d], P.knots[d].data(), sizeof(double) * P.knots[d].size(),
                       hipMemcpyHostToDevice));
  }
  IFEM_HIP(hipMalloc(&P.d_coefs, sizeof(double) * 3 * nnod));
  IFEM_HIP(hipMemcpy(P.d_coefs, coefs, sizeof(double) * 3 * nnod, hipMemcpyHostToDevice));
  ctx->c.patches.push_back(P);
  *patch_id = (int)ctx->c.patches.size() - 1;
  return IFEM_GPU_OK;
}

// Shared by both bind entry points.  gval: per DOF value of prescribed DOFs
// (empty: none), already checked by the caller.
// global_ndof: the DOF count of a multi-patch system (0: the patch's own).
static int bind_system(Context& c, int patch_id, int ncomp, int ndof, const int* meqn, int neq,
                       std::vector<double>&& gval, int global_ndof = 0)
{
  if (patch_id < 0 || patch_id >= (int)c.patches.size())
    return fail(IFEM_GPU_EINVAL, "sam_bind: unknown patch");
  if (ncomp != 1 && ncomp != 3) return fail(IFEM_GPU_EUNSUPPORTED, "sam_bind: ncomp must be 1 or 3");
  const PatchRec& P = c.patches[patch_id];
  const int64_t nnod = (int64_t)P.n[0] * P.n[1] * P.n[2];
  const int64_t want = global_ndof > 0 ? global_ndof : nnod * ncomp;
  if (!meqn || ndof != want) return fail(IFEM_GPU_EINVAL, "sam_bind: ndof != nnod*ncomp");
  std::vector<int> eq(ndof);
  int last = 0;
  for (int i = 0; i < ndof; i++) {
    if (meqn[i] < 0 && gval.empty())
      return fail(IFEM_GPU_EUNSUPPORTED,
                  "sam_bind: constrained (negative) equations need ifem_gpu_sam_bind_constraints");
    if (meqn[i] > 0) {
      if (meqn[i] != last + 1)
        return fail(IFEM_GPU_EUNSUPPORTED,
                    "sam_bind: free DOFs must be numbered 1..neq in increasing DOF order");
      last = meqn[i];
    }
    eq[i] = meqn[i] > 0 ? meqn[i] - 1 : -1;
  }
  if (last != neq) return fail(IFEM_GPU_EINVAL, "sam_bind: neq does not match meqn");
  free_system(c);
  c.patch = patch_id;
  c.ncomp = ncomp;
  c.ndof = ndof;
  c.neq = neq;
  c.h_eq = eq;
  IFEM_HIP(hipMalloc(&c.d_eq, sizeof(int) * ndof));
  IFEM_HIP(hipMemcpy(c.d_eq, eq.data(), sizeof(int) * ndof, hipMemcpyHostToDevice));
  bool any = false;
  for (double g : gval) any = any || g != 0.0;
  if (any) {
    IFEM_HIP(hipMalloc(&c.d_gval, sizeof(double) * ndof));
    IFEM_HIP(hipMemcpy(c.d_gval, gval.data(), sizeof(double) * ndof, hipMemcpyHostToDevice));
    c.h_gval = std::move(gval);
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_sam_bind(ifem_gpu_ctx* ctx, int patch_id, int ncomp, int ndof, const int* meqn,
                      int neq)
{
  CHECK_CTX(ctx);
  return bind_system(ctx->c, patch_id, ncomp, ndof, meqn, neq, {});
}

// Prescribed value per DOF from SAM's constraint arrays (constraints
// without masters only).
static int constraint_values(int ndof, const int* meqn, int nceq, const int* mpmceq,
                             const int* mmceq, const double* ttcc, std::vector<double>& gval,
                             bool* masters = nullptr)
{
  if (masters) *masters = false;
  if (!meqn || ndof < 0 || nceq < 0 || (nceq > 0 && (!mpmceq || !mmceq || !ttcc)))
    return fail(IFEM_GPU_EINVAL, "sam_bind_constraints: null argument");
  // SAM layout (src/LinAlg/SAM.C:260-308): constraint i (1-based) holds the
  // entries mpmceq[i-1] .. mpmceq[i]-1 (1-based) of mmceq/ttcc; the first is
  // the slave DOF and the constant c0, the rest masters and coefficients.
  if (nceq > 0 && mpmceq[0] != 1)
    return fail(IFEM_GPU_EINVAL, "sam_bind_constraints: mpmceq[0] must be 1");
  gval.assign(ndof > 0 ? ndof : 1, 0.0);
  std::vector<char> seen(nceq > 0 ? nceq : 1, 0);
  for (int i = 0; i < ndof; i++) {
    if (meqn[i] >= 0) continue;
    const int ceq = -meqn[i];
    if (ceq > nceq) return fail(IFEM_GPU_EINVAL, "sam_bind_constraints: meqn names an unknown constraint");
    const int ip = mpmceq[ceq - 1], jp = mpmceq[ceq] - 1;
    if (jp < ip) return fail(IFEM_GPU_EINVAL, "sam_bind_constraints: mpmceq is not increasing");
    if (mmceq[ip - 1] != i + 1)
      return fail(IFEM_GPU_EINVAL, "sam_bind_constraints: constraint slave differs from the DOF");
    if (jp > ip) {
      if (!masters)
        return fail(IFEM_GPU_EUNSUPPORTED,
                    "sam_bind_patches: constraints with master DOFs need a single patch");
      *masters = true;
    }
    if (seen[ceq - 1]) return fail(IFEM_GPU_EINVAL, "sam_bind_constraints: constraint used twice");
    seen[ceq - 1] = 1;
    gval[i] = ttcc[ip - 1];
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_sam_bind_constraints(ifem_gpu_ctx* ctx, int patch_id, int ncomp, int ndof,
                                  const int* meqn, int neq, int nceq, const int* mpmceq,
                                  const int* mmceq, const double* ttcc)
{
  CHECK_CTX(ctx);
  std::vector<double> gval;
  bool masters = false;
  int rc = constraint_values(ndof, meqn, nceq, mpmceq, mmceq, ttcc, gval, &masters);
  if (rc) return rc;
  rc = bind_system(ctx->c, patch_id, ncomp, ndof, meqn, neq, std::move(gval));
  if (rc || !masters) return rc;
  // slaves with masters: extended system + condensation
  return mpc_bind(ctx->c, nceq, mpmceq, mmceq, ttcc);
}

int ifem_gpu_sam_bind_patches(ifem_gpu_ctx* ctx, int npatch, const int* patch_ids,
                              const int* const* mlgn, int nnod, int ncomp, const int* meqn,
                              int neq, int nceq, const int* mpmceq, const int* mmceq,
                              const double* ttcc)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (npatch < 1 || !patch_ids || !mlgn || nnod < 1)
    return fail(IFEM_GPU_EINVAL, "sam_bind_patches: null or empty argument");
  const int ndof = nnod * ncomp;
  for (int k = 0; k < npatch; k++) {
    if (patch_ids[k] < 0 || patch_ids[k] >= (int)c.patches.size() || !mlgn[k])
      return fail(IFEM_GPU_EINVAL, "sam_bind_patches: unknown patch");
    const PatchRec& P = c.patches[patch_ids[k]];
    if (P.p[0] != P.p[1] || P.p[0] != P.p[2])
      return fail(IFEM_GPU_EUNSUPPORTED, "sam_bind_patches: equal degrees per direction needed");
    // each patch maps its nodes to distinct global nodes 1..nnod
    const int ln = P.n[0] * P.n[1] * P.n[2];
    std::vector<char> used(nnod, 0);
    for (int i = 0; i < ln; i++) {
      const int g = mlgn[k][i];
      if (g < 1 || g > nnod) return fail(IFEM_GPU_EINVAL, "sam_bind_patches: mlgn out of range");
      if (used[g - 1]) return fail(IFEM_GPU_EUNSUPPORTED, "sam_bind_patches: a patch maps two nodes to one");
      used[g - 1] = 1;
    }
  }
  std::vector<double> gval;
  if (nceq > 0) {
    int rc = constraint_values(ndof, meqn, nceq, mpmceq, mmceq, ttcc, gval);
    if (rc) return rc;
  }
  std::vector<double> gkeep = gval;
  int rc = bind_system(c, patch_ids[0], ncomp, ndof, meqn, neq, std::move(gval), ndof);
  if (rc) return rc;
  const bool anyg = !c.h_gval.empty();
  for (int k = 0; k < npatch; k++) {
    const PatchRec& P = c.patches[patch_ids[k]];
    const int ln = P.n[0] * P.n[1] * P.n[2];
    SubPatch sp;
    sp.patch = patch_ids[k];
    sp.h_leq.resize((size_t)ln * ncomp);
    std::vector<double> lg(anyg ? (size_t)ln * ncomp : 0);
    for (int i = 0; i < ln; i++)
      for (int cc = 0; cc < ncomp; cc++) {
        const int gd = (mlgn[k][i] - 1) * ncomp + cc;
        sp.h_leq[(size_t)i * ncomp + cc] = c.h_eq[gd];
        if (anyg) lg[(size_t)i * ncomp + cc] = gkeep[gd];
      }
    c.multi.push_back(sp);
    SubPatch& s = c.multi.back();
    IFEM_HIP(hipMalloc(&s.d_leq, sizeof(int) * s.h_leq.size()));
    IFEM_HIP(hipMemcpy(s.d_leq, s.h_leq.data(), sizeof(int) * s.h_leq.size(), hipMemcpyHostToDevice));
    if (anyg) {
      IFEM_HIP(hipMalloc(&s.d_lgval, sizeof(double) * lg.size()));
      IFEM_HIP(hipMemcpy(s.d_lgval, lg.data(), sizeof(double) * lg.size(), hipMemcpyHostToDevice));
    }
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_init_assembly(ifem_gpu_ctx* ctx, int64_t* nnz)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (c.patch < 0) return fail(IFEM_GPU_EINVAL, "init_assembly: no SAM bound");
  int rc = build_pattern(c);
  if (rc) return rc;
  if (nnz) *nnz = c.nnz;
  return IFEM_GPU_OK;
}

int ifem_gpu_assemble(ifem_gpu_ctx* ctx, const ifem_gpu_integrand* itg, const int ng[3],
                      int precision)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!itg || !ng) return fail(IFEM_GPU_EINVAL, "assemble: null argument");
  if (!c.pattern_ready) return fail(IFEM_GPU_EINVAL, "assemble: call ifem_gpu_init_assembly first");
  if (precision != IFEM_GPU_FP64 && precision != IFEM_GPU_FP32)
    return fail(IFEM_GPU_EINVAL, "assemble: precision must be 64 or 32");
  const int want = itg->kind == IFEM_GPU_ELASTICITY ? 3 : (itg->kind == IFEM_GPU_POISSON ? 1 : 0);
  if (!want) return fail(IFEM_GPU_EUNSUPPORTED, "assemble: unknown integrand kind");
  if (want != c.ncomp)
    return fail(IFEM_GPU_EINVAL, "assemble: integrand component count differs from the SAM's");
  return assemble_volume(c, *itg, ng, precision);
}

int ifem_gpu_autotune_assembly(ifem_gpu_ctx* ctx, const ifem_gpu_integrand* itg, const int ng[3],
                               int precision, int* half)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!itg || !ng) return fail(IFEM_GPU_EINVAL, "autotune_assembly: null argument");
  if (!c.pattern_ready)
    return fail(IFEM_GPU_EINVAL, "autotune_assembly: call ifem_gpu_init_assembly first");
  if (precision != IFEM_GPU_FP64 && precision != IFEM_GPU_FP32)
    return fail(IFEM_GPU_EINVAL, "autotune_assembly: precision must be 64 or 32");
  const int want = itg->kind == IFEM_GPU_ELASTICITY ? 3 : (itg->kind == IFEM_GPU_POISSON ? 1 : 0);
  if (want != c.ncomp)
    return fail(IFEM_GPU_EINVAL, "autotune_assembly: integrand component count differs from the SAM's");
  if (ng[0] != ng[1] || ng[0] != ng[2])
    return fail(IFEM_GPU_EUNSUPPORTED, "autotune_assembly: equal Gauss counts per direction needed");
  return autotune_assembly(c, *itg, ng, precision, half);
}

int ifem_gpu_set_assembly_variant(ifem_gpu_ctx* ctx, int p, int ng, int ncomp, int precision,
                                  int variant)
{
  CHECK_CTX(ctx);
  if (p < 1 || p > 3 || (ng != p + 1 && ng != p + 2) || (ncomp != 1 && ncomp != 3) ||
      (precision != IFEM_GPU_FP64 && precision != IFEM_GPU_FP32))
    return fail(IFEM_GPU_EINVAL, "set_assembly_variant: (p, ng, ncomp, precision) out of range");
  return set_assembly_variant(ctx->c, p, ng, ncomp, precision == IFEM_GPU_FP32, variant);
}

int ifem_gpu_integrate_norms(ifem_gpu_ctx* ctx, const ifem_gpu_integrand* itg, const int ng[3],
                             const ifem_gpu_function* exact, double norms[6])
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!itg || !ng || !norms) return fail(IFEM_GPU_EINVAL, "integrate_norms: null argument");
  if (!c.pattern_ready) return fail(IFEM_GPU_EINVAL, "integrate_norms: no system");
  const int want = itg->kind == IFEM_GPU_ELASTICITY ? 3 : (itg->kind == IFEM_GPU_POISSON ? 1 : 0);
  if (want != c.ncomp)
    return fail(IFEM_GPU_EINVAL, "integrate_norms: integrand component count differs from the SAM's");
  return integrate_norms(c, *itg, ng, exact, norms);
}

int ifem_gpu_assemble_neumann(ifem_gpu_ctx* ctx, int face, const double* traction, int ng)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!traction) return fail(IFEM_GPU_EINVAL, "assemble_neumann: null traction");
  if (face < 1 || face > 6) return fail(IFEM_GPU_EINVAL, "assemble_neumann: face must be 1..6");
  if (ng < 1 || ng > kMaxNG) return fail(IFEM_GPU_EUNSUPPORTED, "assemble_neumann: ng out of range");
  if (!c.d_rhs) return fail(IFEM_GPU_EINVAL, "assemble_neumann: assemble the volume terms first");
  return assemble_traction(c, 0, face - 1, traction, ng);
}

int ifem_gpu_assemble_neumann_patch(ifem_gpu_ctx* ctx, int index, int face,
                                    const double* traction, int ng)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!traction) return fail(IFEM_GPU_EINVAL, "assemble_neumann: null traction");
  if (face < 1 || face > 6) return fail(IFEM_GPU_EINVAL, "assemble_neumann: face must be 1..6");
  if (ng < 1 || ng > kMaxNG) return fail(IFEM_GPU_EUNSUPPORTED, "assemble_neumann: ng out of range");
  if (!c.d_rhs) return fail(IFEM_GPU_EINVAL, "assemble_neumann: assemble the volume terms first");
  return assemble_traction(c, index, face - 1, traction, ng);
}

void ifem_gpu_solver_defaults(ifem_gpu_solver_params* p)
{
  if (!p) return;
  p->rtol = 1e-6;   // LinSolParams defaults (src/LinAlg/LinSolParams.C:70-76)
  p->atol = 1e-20;
  p->dtol = 1e6;
  p->maxits = 1000;
  p->precision = IFEM_GPU_FP64;
  p->inner_rtol = 1e-4;
  p->check_every = 32;
  p->verbosity = 0;
  p->pc = IFEM_GPU_PC_JACOBI;
  p->mg_smooth = 2;
}

int ifem_gpu_solve(ifem_gpu_ctx* ctx, const ifem_gpu_solver_params* params, ifem_gpu_stats* stats)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!params) return fail(IFEM_GPU_EINVAL, "solve: null params");
  if (!c.fp64_ready) return fail(IFEM_GPU_EINVAL, "solve: no FP64 system assembled");
  if (params->precision == IFEM_GPU_FP32 && !c.fp32_ready)
    return fail(IFEM_GPU_EINVAL, "solve: FP32 solve needs an IFEM_GPU_FP32 assembly");
  return ifem::solve(c, *params, stats);
}

int ifem_gpu_mg_setup(ifem_gpu_ctx* ctx, int smooth, int* nlevels, double* setup_ms)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  int rc = ifem::mg_setup(c, smooth);
  if (rc) return rc;
  if (nlevels) *nlevels = (int)c.mg->lev.size();
  if (setup_ms) *setup_ms = c.mg->setup_ms;
  return IFEM_GPU_OK;
}

int ifem_gpu_mg_level(ifem_gpu_ctx* ctx, int level, int n[3], int* nfree, double* lmax, double* S)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.mg) return fail(IFEM_GPU_EINVAL, "mg_level: no hierarchy (ifem_gpu_mg_setup)");
  const ifem::Multigrid& m = *c.mg;
  if (level < 0 || level >= (int)m.lev.size()) return fail(IFEM_GPU_EINVAL, "mg_level: bad level");
  const ifem::MgLevel& L = m.lev[level];
  if (n)
    for (int d = 0; d < 3; d++) n[d] = L.n[d];
  if (nfree) *nfree = L.nfree;
  if (lmax) *lmax = L.lmax;
  if (S) {
    if (level == 0) return fail(IFEM_GPU_EINVAL, "mg_level: level 0 is the CSR matrix");
    IFEM_HIP(hipStreamSynchronize(c.stream));
    IFEM_HIP(hipMemcpy(S, L.S64, sizeof(double) * (size_t)m.noff * m.nc * m.nc * L.nnode,
                       hipMemcpyDeviceToHost));
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_mg_apply(ifem_gpu_ctx* ctx, const double* r, double* z)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!r || !z) return fail(IFEM_GPU_EINVAL, "mg_apply: null vector");
  if (!c.w.r || !c.w.z) return fail(IFEM_GPU_EINVAL, "mg_apply: no system");
  IFEM_HIP(hipMemcpyAsync(c.w.r, r, sizeof(double) * c.neq, hipMemcpyHostToDevice, c.stream));
  int rc = ifem::mg_apply(c, c.w.r, c.w.z, nullptr);
  if (rc) return rc;
  IFEM_HIP(hipMemcpyAsync(z, c.w.z, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

int ifem_gpu_system_info(ifem_gpu_ctx* ctx, int* neq, int64_t* nnz, int* ncomp)
{
  if (!ctx) return fail(IFEM_GPU_EINVAL, "null context");
  if (neq) *neq = ctx->c.neq;
  if (nnz) *nnz = ctx->c.nnz;
  if (ncomp) *ncomp = ctx->c.ncomp;
  return IFEM_GPU_OK;
}

int ifem_gpu_get_system(ifem_gpu_ctx* ctx, int64_t* rowptr, int32_t* colind, double* values,
                        double* rhs)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.pattern_ready) return fail(IFEM_GPU_EINVAL, "get_system: no pattern");
  IFEM_HIP(hipStreamSynchronize(c.stream));
  if (rowptr)
    IFEM_HIP(hipMemcpy(rowptr, c.d_rowptr, sizeof(int64_t) * (c.neq + 1), hipMemcpyDeviceToHost));
  if (colind)
    IFEM_HIP(hipMemcpy(colind, c.d_colind, sizeof(int32_t) * c.nnz, hipMemcpyDeviceToHost));
  if (values) {
    if (!c.fp64_ready) return fail(IFEM_GPU_EINVAL, "get_system: no FP64 values");
    IFEM_HIP(hipMemcpy(values, c.d_val, sizeof(double) * c.nnz, hipMemcpyDeviceToHost));
  }
  if (rhs) {
    if (!c.d_rhs) return fail(IFEM_GPU_EINVAL, "get_system: no RHS");
    IFEM_HIP(hipMemcpy(rhs, c.d_rhs, sizeof(double) * c.neq, hipMemcpyDeviceToHost));
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_get_values_fp32(ifem_gpu_ctx* ctx, float* values)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.fp32_ready || !values) return fail(IFEM_GPU_EINVAL, "get_values_fp32: no FP32 values");
  IFEM_HIP(hipStreamSynchronize(c.stream));
  IFEM_HIP(hipMemcpy(values, c.d_val32, sizeof(float) * c.nnz, hipMemcpyDeviceToHost));
  return IFEM_GPU_OK;
}

int ifem_gpu_get_solution(ifem_gpu_ctx* ctx, double* x)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.w.x || !x) return fail(IFEM_GPU_EINVAL, "get_solution: no solution");
  IFEM_HIP(hipStreamSynchronize(c.stream));
  IFEM_HIP(hipMemcpy(x, c.w.x, sizeof(double) * c.neq, hipMemcpyDeviceToHost));
  return IFEM_GPU_OK;
}

int ifem_gpu_expand_solution(ifem_gpu_ctx* ctx, double* xdof)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.w.x || !xdof) return fail(IFEM_GPU_EINVAL, "expand_solution: no solution");
  std::vector<double> x(c.neq);
  IFEM_HIP(hipStreamSynchronize(c.stream));
  IFEM_HIP(hipMemcpy(x.data(), c.w.x, sizeof(double) * c.neq, hipMemcpyDeviceToHost));
  if (c.mpc) { // slaves with masters: c0 + sum c_m x_m
    mpc_expand_host(c, x, xdof);
    return IFEM_GPU_OK;
  }
  // eliminated DOFs: 0, or their prescribed value (src/LinAlg/SAM.C:768-810)
  const bool g = !c.h_gval.empty();
  for (int i = 0; i < c.ndof; i++)
    xdof[i] = c.h_eq[i] >= 0 ? x[c.h_eq[i]] : (g ? c.h_gval[i] : 0.0);
  return IFEM_GPU_OK;
}

int ifem_gpu_spmv(ifem_gpu_ctx* ctx, const double* x, double* y)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.fp64_ready || !x || !y) return fail(IFEM_GPU_EINVAL, "spmv: no FP64 system or null vector");
  double *dx = nullptr, *dy = nullptr;
  IFEM_HIP(hipMalloc(&dx, sizeof(double) * c.neq));
  IFEM_HIP(hipMalloc(&dy, sizeof(double) * c.neq));
  int rc = IFEM_GPU_OK;
  if (hipMemcpyAsync(dx, x, sizeof(double) * c.neq, hipMemcpyHostToDevice, c.stream) != hipSuccess)
    rc = fail(IFEM_GPU_EDEVICE, "spmv: copy in failed");
  if (!rc) rc = spmv_device(c, dx, dy);
  if (!rc && hipMemcpyAsync(y, dy, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
    rc = fail(IFEM_GPU_EDEVICE, "spmv: copy out failed");
  if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) rc = fail(IFEM_GPU_EDEVICE, "spmv: sync failed");
  hipFree(dx);
  hipFree(dy);
  return rc;
}

int ifem_gpu_spmv_noalloc(ifem_gpu_ctx* ctx, const double* x, double* y)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.fp64_ready || !x || !y)
    return fail(IFEM_GPU_EINVAL, "spmv_noalloc: no FP64 system or null vector");
  if (c.neq > kNoAllocMaxEq)
    return fail(IFEM_GPU_EINVAL, "spmv_noalloc: the system has " + std::to_string(c.neq) +
                                     " equations, the staging buffer holds at most " +
                                     std::to_string(kNoAllocMaxEq) + " (use ifem_gpu_spmv)");
  if (!c.noalloc_scratch)
    return fail(IFEM_GPU_EDEVICE, "spmv_noalloc: hipGetSymbolAddress of the staging buffer failed");
  double* dx = c.noalloc_scratch;
  double* dy = c.noalloc_scratch + kNoAllocMaxEq;
  if (hipMemcpyAsync(dx, x, sizeof(double) * c.neq, hipMemcpyHostToDevice, c.stream) != hipSuccess)
    return fail(IFEM_GPU_EDEVICE, "spmv_noalloc: copy in failed");
  int rc = spmv_device(c, dx, dy);
  if (!rc && hipMemcpyAsync(y, dy, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
    rc = fail(IFEM_GPU_EDEVICE, "spmv_noalloc: copy out failed");
  if (!rc && hipStreamSynchronize(c.stream) != hipSuccess)
    rc = fail(IFEM_GPU_EDEVICE, "spmv_noalloc: sync failed");
  return rc;
}

int ifem_gpu_spmv_device(ifem_gpu_ctx* ctx, int precision, const void* x, void* y)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!x || !y) return fail(IFEM_GPU_EINVAL, "spmv_device: null vector");
  if (precision == IFEM_GPU_FP64) {
    if (!c.fp64_ready) return fail(IFEM_GPU_EINVAL, "spmv_device: no FP64 system");
    return spmv_device(c, (const double*)x, (double*)y);
  }
  if (precision == IFEM_GPU_FP32) {
    if (!c.fp32_ready) return fail(IFEM_GPU_EINVAL, "spmv_device: no FP32 values");
    return spmv_device32(c, (const float*)x, (float*)y);
  }
  return fail(IFEM_GPU_EINVAL, "spmv_device: precision must be 64 or 32");
}

int ifem_gpu_set_rhs(ifem_gpu_ctx* ctx, const double* b)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (!c.d_rhs || !b) return fail(IFEM_GPU_EINVAL, "set_rhs: no system or null vector");
  IFEM_HIP(hipMemcpyAsync(c.d_rhs, b, sizeof(double) * c.neq, hipMemcpyHostToDevice, c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

int ifem_gpu_device_pointers(ifem_gpu_ctx* ctx, void** rowptr, void** colind, void** values,
                             void** rhs, void** solution)
{
  if (!ctx) return fail(IFEM_GPU_EINVAL, "null context");
  Context& c = ctx->c;
  if (rowptr) *rowptr = c.d_rowptr;
  if (colind) *colind = c.d_colind;
  if (values) *values = c.d_val;
  if (rhs) *rhs = c.d_rhs;
  if (solution) *solution = c.w.x;
  return IFEM_GPU_OK;
}

int ifem_gpu_set_learned_preconditioner(ifem_gpu_ctx* ctx, int enable, int D, int H, int W,
                                        const float* aux, int naux, int features, double scale)
{
  CHECK_CTX(ctx);
  return set_learned(ctx->c, enable, D, H, W, aux, naux, features, scale);
}

int ifem_gpu_last_assembly_parts(ifem_gpu_ctx* ctx, double* geometry_ms, double* lines_ms)
{
  if (!ctx) return fail(IFEM_GPU_EINVAL, "null context");
  if (geometry_ms) *geometry_ms = ctx->c.last_geometry_ms;
  if (lines_ms) *lines_ms = ctx->c.last_lines_ms;
  return IFEM_GPU_OK;
}

int ifem_gpu_last_timings(ifem_gpu_ctx* ctx, double* assemble_ms, double* solve_ms)
{
  if (!ctx) return fail(IFEM_GPU_EINVAL, "null context");
  if (assemble_ms) *assemble_ms = ctx->c.last_assemble_ms;
  if (solve_ms) *solve_ms = ctx->c.last_solve_ms;
  return IFEM_GPU_OK;
}

// ---- batch of independent patches -----------------------------------------

static int check_function(const ifem_gpu_function& f, const char* what)
{
  if (f.type < IFEM_GPU_FN_ZERO || f.type > IFEM_GPU_FN_EXP_FOURIER)
    return fail(IFEM_GPU_EINVAL, std::string("batch_add: unknown ") + what + " type");
  if ((f.type == IFEM_GPU_FN_FOURIER || f.type == IFEM_GPU_FN_EXP_FOURIER) && f.nmodes > 0 && !f.modes)
    return fail(IFEM_GPU_EINVAL, std::string("batch_add: null ") + what + " mode table");
  if (f.nmodes < 0) return fail(IFEM_GPU_EINVAL, std::string("batch_add: negative ") + what + " nmodes");
  return IFEM_GPU_OK;
}

static int upload_modes(const ifem_gpu_function& f, std::vector<double>& host, double** dev)
{
  host.clear();
  *dev = nullptr;
  if ((f.type == IFEM_GPU_FN_FOURIER || f.type == IFEM_GPU_FN_EXP_FOURIER) && f.nmodes > 0) {
    host.assign(f.modes, f.modes + 5 * (size_t)f.nmodes);
    IFEM_HIP(hipMalloc(dev, sizeof(double) * host.size()));
    IFEM_HIP(hipMemcpy(*dev, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice));
  }
  return IFEM_GPU_OK;
}

#define CHECK_BATCH(c, need_ready)                                                         \
  do {                                                                                     \
    if (!(c).batch || (c).batch->pp.empty())                                               \
      return fail(IFEM_GPU_EINVAL, "batch: no patches added");                             \
    if ((need_ready) && !(c).batch->ready)                                                 \
      return fail(IFEM_GPU_EINVAL, "batch: call ifem_gpu_batch_init first");               \
  } while (0)

int ifem_gpu_batch_clear(ifem_gpu_ctx* ctx)
{
  CHECK_CTX(ctx);
  IFEM_HIP(hipStreamSynchronize(ctx->c.stream));
  ctx->c.batch.reset();
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_add(ifem_gpu_ctx* ctx, int patch_id, const int* meqn, int neq,
                       const ifem_gpu_integrand* itg, int ng, int* index)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  if (patch_id < 0 || patch_id >= (int)c.patches.size())
    return fail(IFEM_GPU_EINVAL, "batch_add: unknown patch");
  if (c.batch && c.batch->ready)
    return fail(IFEM_GPU_EINVAL, "batch_add: the batch is initialised; clear it first");
  const PatchRec& P = c.patches[patch_id];
  if (P.p[0] != P.p[1] || P.p[0] != P.p[2] || P.p[0] < 1 || P.p[0] > 3)
    return fail(IFEM_GPU_EUNSUPPORTED, "batch_add: equal degrees 1..3 in all directions are needed");
  if (ng != P.p[0] + 1) return fail(IFEM_GPU_EUNSUPPORTED, "batch_add: the batch path uses ng = p+1");
  if (!itg || itg->kind != IFEM_GPU_POISSON)
    return fail(IFEM_GPU_EUNSUPPORTED, "batch_add: the batch path takes scalar (Poisson) integrands");
  int rc = check_function(itg->kappa, "kappa");
  if (!rc) rc = check_function(itg->source, "source");
  if (rc) return rc;
  const int64_t nnod = (int64_t)P.n[0] * P.n[1] * P.n[2];
  if (!meqn) return fail(IFEM_GPU_EINVAL, "batch_add: null meqn");
  std::vector<int> eq(nnod);
  int last = 0;
  for (int64_t i = 0; i < nnod; i++) {
    if (meqn[i] < 0)
      return fail(IFEM_GPU_EUNSUPPORTED, "batch_add: constrained (negative) equations are not supported");
    if (meqn[i] > 0) {
      if (meqn[i] != last + 1)
        return fail(IFEM_GPU_EUNSUPPORTED, "batch_add: free DOFs must be numbered 1..neq in increasing order");
      last = meqn[i];
    }
    eq[i] = meqn[i] - 1;
  }
  if (last != neq) return fail(IFEM_GPU_EINVAL, "batch_add: neq does not match meqn");
  if (!c.batch) c.batch.reset(new Batch());
  Batch& B = *c.batch;
  int64_t total = 0;
  for (const BatchPatch& q : B.pp) total += q.neq;
  if (total + neq > INT32_MAX) return fail(IFEM_GPU_EUNSUPPORTED, "batch_add: more than 2^31-1 equations");
  // global equation numbers: offset by the equations of the patches before
  for (int& e : eq)
    if (e >= 0) e += (int)total;
  B.pp.emplace_back();
  BatchPatch& bp = B.pp.back();
  bp.patch = patch_id;
  bp.p = P.p[0];
  bp.ng = ng;
  bp.nnod = (int)nnod;
  bp.neq = neq;
  bp.itg = *itg;
  rc = upload_modes(itg->kappa, bp.kappa_modes, &bp.d_kmodes);
  if (!rc) rc = upload_modes(itg->source, bp.source_modes, &bp.d_smodes);
  if (!rc && hipMalloc(&bp.d_eq, sizeof(int) * (nnod > 0 ? nnod : 1)) != hipSuccess)
    rc = fail(IFEM_GPU_EDEVICE, "batch_add: hipMalloc failed");
  if (!rc && hipMemcpy(bp.d_eq, eq.data(), sizeof(int) * nnod, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(IFEM_GPU_EDEVICE, "batch_add: hipMemcpy failed");
  bp.itg.kappa.modes = bp.kappa_modes.empty() ? nullptr : bp.kappa_modes.data();
  bp.itg.source.modes = bp.source_modes.empty() ? nullptr : bp.source_modes.data();
  if (rc) {
    hipFree(bp.d_eq);
    hipFree(bp.d_kmodes);
    hipFree(bp.d_smodes);
    B.pp.pop_back();
    return rc;
  }
  if (index) *index = (int)B.pp.size() - 1;
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_init(ifem_gpu_ctx* ctx, int* neq, int64_t* nnz)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, false);
  if (c.batch->ready) return fail(IFEM_GPU_EINVAL, "batch_init: already initialised");
  int rc = batch_init(c);
  if (rc) return rc;
  if (neq) *neq = c.batch->neq;
  if (nnz) *nnz = c.batch->nnz;
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_assemble(ifem_gpu_ctx* ctx, int precision)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  if (precision != IFEM_GPU_FP64 && precision != IFEM_GPU_FP32)
    return fail(IFEM_GPU_EINVAL, "batch_assemble: precision must be 64 or 32");
  double ms = 0.0;
  int rc = batch_assemble(c, precision, &ms);
  if (rc) return rc;
  c.batch->assemble_ms = ms;
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_autotune(ifem_gpu_ctx* ctx, int precision, int* table)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  if (precision != IFEM_GPU_FP64 && precision != IFEM_GPU_FP32)
    return fail(IFEM_GPU_EINVAL, "batch_autotune: precision must be 64 or 32");
  return batch_autotune(c, precision, table);
}

int ifem_gpu_batch_solve(ifem_gpu_ctx* ctx, const ifem_gpu_solver_params* params,
                         ifem_gpu_stats* stats)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  if (!params) return fail(IFEM_GPU_EINVAL, "batch_solve: null params");
  if (params->precision != IFEM_GPU_FP64 && params->precision != IFEM_GPU_FP32)
    return fail(IFEM_GPU_EINVAL, "batch_solve: precision must be 64 or 32");
  return batch_solve(c, *params, stats);
}

int ifem_gpu_batch_results(ifem_gpu_ctx* ctx, int* iterations, int* converged, double* rel_residual)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  const Batch& B = *c.batch;
  for (size_t q = 0; q < B.pp.size(); q++) {
    if (iterations) iterations[q] = B.iters[q];
    if (converged) converged[q] = B.conv[q];
    if (rel_residual) rel_residual[q] = B.relres[q];
  }
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_patch_info(ifem_gpu_ctx* ctx, int index, int* neq, int64_t* nnz, int* eq0)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  const Batch& B = *c.batch;
  if (index < 0 || index >= (int)B.pp.size()) return fail(IFEM_GPU_EINVAL, "batch: index out of range");
  const BatchPatch& bp = B.pp[index];
  int64_t r[2] = {0, 0};
  IFEM_HIP(hipMemcpy(&r[0], B.d_rowptr + bp.eq0, sizeof(int64_t), hipMemcpyDeviceToHost));
  IFEM_HIP(hipMemcpy(&r[1], B.d_rowptr + bp.eq0 + bp.neq, sizeof(int64_t), hipMemcpyDeviceToHost));
  if (neq) *neq = bp.neq;
  if (nnz) *nnz = r[1] - r[0];
  if (eq0) *eq0 = bp.eq0;
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_get_system(ifem_gpu_ctx* ctx, int index, int64_t* rowptr, int32_t* colind,
                              double* values, float* values32, double* rhs)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  const Batch& B = *c.batch;
  if (index < 0 || index >= (int)B.pp.size()) return fail(IFEM_GPU_EINVAL, "batch: index out of range");
  const BatchPatch& bp = B.pp[index];
  IFEM_HIP(hipStreamSynchronize(c.stream));
  std::vector<int64_t> rp(bp.neq + 1);
  IFEM_HIP(hipMemcpy(rp.data(), B.d_rowptr + bp.eq0, sizeof(int64_t) * (bp.neq + 1), hipMemcpyDeviceToHost));
  const int64_t z0 = rp[0], nz = rp[bp.neq] - rp[0];
  if (rowptr)
    for (int i = 0; i <= bp.neq; i++) rowptr[i] = rp[i] - z0;
  if (colind && nz > 0) {
    IFEM_HIP(hipMemcpy(colind, B.d_colind + z0, sizeof(int32_t) * nz, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < nz; k++) colind[k] -= bp.eq0;
  }
  if (values && nz > 0) {
    if (!B.fp64_ready) return fail(IFEM_GPU_EINVAL, "batch_get_system: no FP64 values");
    IFEM_HIP(hipMemcpy(values, B.d_val + z0, sizeof(double) * nz, hipMemcpyDeviceToHost));
  }
  if (values32 && nz > 0) {
    if (!B.fp32_ready) return fail(IFEM_GPU_EINVAL, "batch_get_system: no FP32 values");
    IFEM_HIP(hipMemcpy(values32, B.d_val32 + z0, sizeof(float) * nz, hipMemcpyDeviceToHost));
  }
  if (rhs && bp.neq > 0)
    IFEM_HIP(hipMemcpy(rhs, B.d_rhs + bp.eq0, sizeof(double) * bp.neq, hipMemcpyDeviceToHost));
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_get_solution(ifem_gpu_ctx* ctx, int index, double* x)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  const Batch& B = *c.batch;
  if (index < 0 || index >= (int)B.pp.size()) return fail(IFEM_GPU_EINVAL, "batch: index out of range");
  const BatchPatch& bp = B.pp[index];
  IFEM_HIP(hipStreamSynchronize(c.stream));
  if (x && bp.neq > 0)
    IFEM_HIP(hipMemcpy(x, B.x + bp.eq0, sizeof(double) * bp.neq, hipMemcpyDeviceToHost));
  return IFEM_GPU_OK;
}

int ifem_gpu_batch_timings(ifem_gpu_ctx* ctx, double* symbolic_ms, double* assemble_ms,
                           double* solve_ms)
{
  CHECK_CTX(ctx);
  Context& c = ctx->c;
  CHECK_BATCH(c, true);
  if (symbolic_ms) *symbolic_ms = c.batch->symbolic_ms;
  if (assemble_ms) *assemble_ms = c.batch->assemble_ms;
  if (solve_ms) *solve_ms = c.batch->solve_ms;
  return IFEM_GPU_OK;
}

} // extern "C"
