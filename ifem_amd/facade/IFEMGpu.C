// IFEMGpu.C -- reference-interface classes on the libifem_gpu C ABI.
#include "IFEMGpu.h"
#include "ifem_hdf5.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>

namespace ifem_gpu {

static bool report(int rc, const char* where)
{
  if (rc == IFEM_GPU_OK) return true;
  std::cerr << " *** " << where << ": " << ifem_gpu_last_error() << " (code " << rc << ")"
            << std::endl;
  return false;
}

// ---------------------------------------------------------------------------
std::string SettingMap::getStringValue(const std::string& key) const
{
  auto it = values.find(key);
  return it == values.end() ? std::string() : it->second;
}

int SettingMap::getIntValue(const std::string& key) const
{
  auto it = values.find(key);
  return it == values.end() ? 0 : std::atoi(it->second.c_str());
}

double SettingMap::getDoubleValue(const std::string& key) const
{
  auto it = values.find(key);
  return it == values.end() ? 0.0 : std::atof(it->second.c_str());
}

LinSolParams::LinSolParams(LinAlg::LinearSystemType ls) : linSys(ls)
{
  addValue("type", "gmres");
  addValue("pc", "default");
  addValue("rtol", "1e-6");
  addValue("atol", "1e-20");
  addValue("dtol", "1e6");
  addValue("maxits", "1000");
  addValue("verbosity", "1");
  addValue("precision", "fp64");
  addValue("inner_rtol", "1e-4");
}

// Minimal reader of <tag>value</tag> children (no attributes, no nesting
// beyond one level), enough for the <linearsolver> keys.
bool LinSolParams::read(const std::string& xml)
{
  static const char* keys[] = {"type", "pc", "rtol", "atol", "dtol", "maxits", "verbosity",
                               "precision", "inner_rtol", "multigrid_no_smooth"};
  for (const char* k : keys) {
    const std::string open = std::string("<") + k + ">", close = std::string("</") + k + ">";
    size_t a = xml.find(open);
    if (a == std::string::npos) continue;
    size_t b = xml.find(close, a);
    if (b == std::string::npos) {
      std::cerr << " *** LinSolParams::read: unterminated <" << k << ">" << std::endl;
      return false;
    }
    std::string v = xml.substr(a + open.size(), b - a - open.size());
    size_t s = v.find_first_not_of(" \t\r\n"), e = v.find_last_not_of(" \t\r\n");
    addValue(k, s == std::string::npos ? std::string() : v.substr(s, e - s + 1));
  }
  return true;
}

// ---------------------------------------------------------------------------
static void set_function(ifem_gpu_function& f, const std::vector<double>& modes)
{
  if (f.type == IFEM_GPU_FN_FOURIER || f.type == IFEM_GPU_FN_EXP_FOURIER) {
    f.nmodes = (int)(modes.size() / 5);
    f.modes = modes.data();
  } else {
    f.nmodes = 0;
    f.modes = nullptr;
  }
}

bool PoissonIntegrand::gpuDescription(ifem_gpu_integrand& d) const
{
  std::memset(&d, 0, sizeof d);
  d.kind = IFEM_GPU_POISSON;
  d.kappa = kappa;
  d.source = source;
  set_function(d.kappa, kappaModes);
  set_function(d.source, sourceModes);
  return true;
}

bool ElasticityIntegrand::gpuDescription(ifem_gpu_integrand& d) const
{
  std::memset(&d, 0, sizeof d);
  d.kind = IFEM_GPU_ELASTICITY;
  d.E = E;
  d.nu = nu;
  for (int i = 0; i < 3; i++) d.body[i] = body[i];
  return true;
}

void ElasticityIntegrand::setTraction(int lIndex, double tx, double ty, double tz)
{
  traction[lIndex] = {tx, ty, tz};
}

bool ElasticityIntegrand::gpuTraction(int lIndex, double t[3]) const
{
  auto it = traction.find(lIndex);
  if (it == traction.end()) return false;
  for (int i = 0; i < 3; i++) t[i] = it->second[i];
  return true;
}

// ---------------------------------------------------------------------------
SystemMatrix* SystemMatrix::create(void* adm, LinAlg::MatrixType mtype, const LinSolParams& spar)
{
  (void)adm;
  if (mtype != LinAlg::GPU) {
    std::cerr << " *** SystemMatrix::create: only LinAlg::GPU is available here" << std::endl;
    return nullptr;
  }
  return new GPUMatrix(spar);
}

GPUMatrix::GPUMatrix(const LinSolParams& spar) : par(spar)
{
  report(ifem_gpu_ctx_create(0, nullptr, &ctx), "GPUMatrix");
}

GPUMatrix::~GPUMatrix()
{
  if (ctx) ifem_gpu_ctx_destroy(ctx);
}

void GPUMatrix::initAssembly(const SAM& sam, char)
{
  ok = false;
  if (!ctx || !sam.patch) {
    std::cerr << " *** GPUMatrix::initAssembly: no context or no patch in the SAM" << std::endl;
    return;
  }
  const ASMs3D& P = *sam.patch;
  const double* kp[3] = {P.knots[0].data(), P.knots[1].data(), P.knots[2].data()};
  int pid = -1;
  if (!report(ifem_gpu_patch_create(ctx, P.order, P.n, kp, P.coefs.data(), 0, &pid),
              "GPUMatrix::initAssembly"))
    return;
  const int brc = sam.nceq > 0
      ? ifem_gpu_sam_bind_constraints(ctx, pid, sam.ncomp, sam.ndof, sam.meqn.data(), sam.neq,
                                      sam.nceq, sam.mpmceq.data(), sam.mmceq.data(),
                                      sam.ttcc.data())
      : ifem_gpu_sam_bind(ctx, pid, sam.ncomp, sam.ndof, sam.meqn.data(), sam.neq);
  if (!report(brc, "GPUMatrix::initAssembly")) return;
  if (!report(ifem_gpu_init_assembly(ctx, &nnz), "GPUMatrix::initAssembly")) return;
  neq = sam.neq;
  ok = true;
}

bool GPUMatrix::multiply(const SystemVector& B, SystemVector& C) const
{
  if (!ok || B.dim() < neq) return false;
  C.redim(neq);
  return report(ifem_gpu_spmv(ctx, B.getRef(), C.getPtr()), "GPUMatrix::multiply");
}

bool GPUMatrix::solve(SystemVector& b, Real* rc)
{
  if (!ok || b.dim() != neq) return false;
  const std::string type = par.getStringValue("type");
  if (type != "cg") {
    std::cerr << " *** GPUMatrix::solve: solver type '" << type
              << "' is not available on the GPU (use <type>cg</type>)" << std::endl;
    return false;
  }
  const std::string pc = par.getStringValue("pc");
  // PETSc's multigrid choices map to the spline multigrid of the library
  const bool mg = pc == "gamg" || pc == "ml" || pc == "mg" || pc == "multigrid";
  if (pc != "jacobi" && pc != "default" && !mg) {
    std::cerr << " *** GPUMatrix::solve: preconditioner '" << pc << "' is not available"
              << std::endl;
    return false;
  }
  ifem_gpu_solver_params p;
  ifem_gpu_solver_defaults(&p);
  if (mg) {
    p.pc = IFEM_GPU_PC_MG;
    if (par.hasValue("multigrid_no_smooth")) p.mg_smooth = par.getIntValue("multigrid_no_smooth");
  }
  p.rtol = par.getDoubleValue("rtol");
  p.atol = par.getDoubleValue("atol");
  p.dtol = par.getDoubleValue("dtol");
  p.maxits = par.getIntValue("maxits");
  p.verbosity = par.getIntValue("verbosity");
  p.precision = par.getStringValue("precision") == "fp32" ? IFEM_GPU_FP32 : IFEM_GPU_FP64;
  if (par.hasValue("inner_rtol")) p.inner_rtol = par.getDoubleValue("inner_rtol");
  if (!report(ifem_gpu_set_rhs(ctx, b.getRef()), "GPUMatrix::solve")) return false;
  const int st = ifem_gpu_solve(ctx, &p, &stats);
  if (st != IFEM_GPU_OK && st != IFEM_GPU_ENOCONV) return report(st, "GPUMatrix::solve");
  if (!report(ifem_gpu_get_solution(ctx, b.getPtr()), "GPUMatrix::solve")) return false;
  if (rc) *rc = stats.rel_residual;
  if (p.verbosity > 1)
    std::cout << "\tGPU CG: " << stats.iterations << " iterations, relative residual "
              << stats.rel_residual << std::endl;
  return report(st, "GPUMatrix::solve");
}

// ---------------------------------------------------------------------------
bool SAM::expandSolution(const SystemVector& solVec, std::vector<Real>& dofVec,
                         Real scaleSD) const
{
  if ((int)solVec.dim() < neq) return false;
  dofVec.assign(ndof, 0.0);
  const Real* x = solVec.getRef();
  for (int i = 0; i < ndof; i++)
    if (meqn[i] > 0) dofVec[i] = x[meqn[i] - 1];
    else if (meqn[i] < 0) dofVec[i] = scaleSD * ttcc[mpmceq[-meqn[i] - 1] - 1]; // c0
    else dofVec[i] = 0.0; // homogeneous Dirichlet value
  return true;
}

Real SAM::normL2(const std::vector<Real>& x) const
{
  if (x.empty()) return 0.0;
  double s = 0.0;
  for (double v : x) s += v * v;
  return std::sqrt(s / x.size());
}

bool SAMpatch::init(const ASMs3D& pch, int nc)
{
  patch = &pch;
  ncomp = nc;
  nnod = pch.getNoNodes();
  ndof = nnod * nc;
  msc.assign(ndof, 1);
  const int n1 = pch.n[0], n2 = pch.n[1], n3 = pch.n[2];
  auto on_face = [&](int f, int i, int j, int k) {
    switch (f) {
    case 1: return i == 0;
    case 2: return i == n1 - 1;
    case 3: return j == 0;
    case 4: return j == n2 - 1;
    case 5: return k == 0;
    default: return k == n3 - 1;
    }
  };
  // prescribed value per DOF (NaN: none); constraint equations are kept in
  // slave-DOF order, as the reference's MPC set orders them
  std::vector<double> value(ndof, std::nan(""));
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        int m = 0;
        for (const auto& d : pch.dirichlet)
          if (on_face(d.first, i, j, k)) m |= d.second;
        const int node = i + n1 * (j + n2 * k);
        for (int c = 0; c < nc; c++)
          if ((m >> c) & 1) msc[node * nc + c] = 0;
        for (const auto& pr : pch.prescribed)
          if (on_face(pr.face, i, j, k))
            for (int c = 0; c < nc; c++)
              if ((pr.cmask >> c) & 1 && msc[node * nc + c] == 1) value[node * nc + c] = pr.value;
      }
  nceq = 0;
  mpmceq.assign(1, 1);
  mmceq.clear();
  ttcc.clear();
  meqn.assign(ndof, 0);
  for (int i = 0; i < ndof; i++)
    if (!std::isnan(value[i])) {
      msc[i] = 0;
      meqn[i] = -(++nceq);
      mmceq.push_back(i + 1);
      ttcc.push_back(value[i]);
      mpmceq.push_back(mpmceq.back() + 1);
    }
  // SAM::initSystemEquations: free DOFs numbered in DOF order
  neq = 0;
  for (int i = 0; i < ndof; i++)
    if (msc[i] == 1) meqn[i] = ++neq;
  return true;
}

// ---------------------------------------------------------------------------
bool AlgEqSystem::init(LinAlg::MatrixType mtype, const LinSolParams& spar)
{
  A.reset(SystemMatrix::create(nullptr, mtype, spar));
  if (!A) return false;
  A->initAssembly(sam);
  b.redim(sam.getNoEquations());
  GPUMatrix* G = dynamic_cast<GPUMatrix*>(A.get());
  return G && G->ready();
}

void AlgEqSystem::initialize(char initLHS)
{
  if (initLHS && A) A->init();
  b.init();
}

bool AlgEqSystem::finalize(bool)
{
  GPUMatrix* G = dynamic_cast<GPUMatrix*>(A.get());
  if (!G || !G->ready()) return false;
  if (!A->endAssembly()) return false;
  return report(ifem_gpu_get_system(G->context(), nullptr, nullptr, nullptr, b.getPtr()),
                "AlgEqSystem::finalize");
}

// ---------------------------------------------------------------------------
ASMs3D::ASMs3D(const int ord[3], const int ncoef[3], const std::vector<double> kn[3],
               const std::vector<double>& cf)
  : coefs(cf)
{
  for (int d = 0; d < 3; d++) {
    order[d] = ord[d];
    n[d] = ncoef[d];
    knots[d] = kn[d];
  }
}

bool ASMs3D::write(std::ostream& os, int basis) const
{
  if (basis != 1) return false; // one basis per patch here
  const double* kn[3] = {knots[0].data(), knots[1].data(), knots[2].data()};
  size_t len = 0;
  if (ifem_g2_format(order, n, kn, coefs.data(), nullptr, 0, &len)) return false;
  std::string text(len, '\0');
  if (ifem_g2_format(order, n, kn, coefs.data(), &text[0], len, &len)) return false;
  os << text;
  return os.good();
}

std::string Integrand::getField1Name(size_t idx, const char* prefix) const
{
  std::string name = "primary solution " + std::to_string(1 + idx);
  return prefix ? prefix + std::string(" ") + name : name;
}

std::string PoissonIntegrand::getField1Name(size_t, const char* prefix) const
{
  return prefix ? prefix + std::string(" u") : std::string("u");
}

std::string ElasticityIntegrand::getField1Name(size_t idx, const char* prefix) const
{
  std::string name = idx < 3 ? std::string("u_") + char('x' + idx) : std::string("u");
  return prefix ? prefix + std::string(" ") + name : name;
}

std::unique_ptr<ASMs3D> ASMs3D::read(std::istream& is)
{
  int cls = 0, v1 = 0, v2 = 0, v3 = 0, dim = 0, rational = 0;
  if (!(is >> cls >> v1 >> v2 >> v3 >> dim >> rational)) return nullptr;
  if (cls != 700 || dim != 3) {
    std::cerr << " *** ASMs3D::read: not a 3D spline volume (class " << cls << ")" << std::endl;
    return nullptr;
  }
  if (rational) {
    std::cerr << " *** ASMs3D::read: rational volumes are not supported on the GPU" << std::endl;
    return nullptr;
  }
  int ord[3], ncf[3];
  std::vector<double> kn[3];
  for (int d = 0; d < 3; d++) {
    if (!(is >> ncf[d] >> ord[d]) || ncf[d] < ord[d] || ord[d] < 1) return nullptr;
    kn[d].resize(ncf[d] + ord[d]);
    for (double& t : kn[d])
      if (!(is >> t)) return nullptr;
  }
  std::vector<double> cf((size_t)ncf[0] * ncf[1] * ncf[2] * 3);
  for (double& x : cf)
    if (!(is >> x)) return nullptr;
  return std::unique_ptr<ASMs3D>(new ASMs3D(ord, ncf, kn, cf));
}

// Control point (i, j, k) of an n[0] x n[1] x n[2] net, i fastest.
static inline double* cpt(std::vector<double>& c, const int nn[3], int i, int j, int k)
{
  return &c[3 * ((size_t)i + (size_t)nn[0] * (j + (size_t)nn[1] * k))];
}

bool ASMs3D::uniformRefine(int dir, int nInsert)
{
  if (dir < 0 || dir > 2 || nInsert < 0) return false;
  const std::vector<double> t0 = knots[dir];
  for (size_t s = 0; s + 1 < t0.size(); s++) {
    const double a = t0[s], b = t0[s + 1];
    if (!(b > a)) continue;
    for (int m = 1; m <= nInsert; m++) {
      // Boehm insertion of u into direction dir
      const double u = a + (b - a) * m / (nInsert + 1);
      const int p = order[dir] - 1;
      const std::vector<double>& t = knots[dir];
      int k = (int)(std::upper_bound(t.begin(), t.end(), u) - t.begin()) - 1;
      int nn[3] = {n[0], n[1], n[2]}, nw[3] = {n[0], n[1], n[2]};
      nw[dir]++;
      std::vector<double> nc((size_t)nw[0] * nw[1] * nw[2] * 3);
      for (int kk = 0; kk < nw[2]; kk++)
        for (int jj = 0; jj < nw[1]; jj++)
          for (int ii = 0; ii < nw[0]; ii++) {
            int idx[3] = {ii, jj, kk};
            const int i = idx[dir];
            double* q = cpt(nc, nw, ii, jj, kk);
            auto src = [&](int ia) {
              int id2[3] = {ii, jj, kk};
              id2[dir] = ia;
              return cpt(coefs, nn, id2[0], id2[1], id2[2]);
            };
            if (i <= k - p) {
              const double* s0 = src(i);
              for (int c = 0; c < 3; c++) q[c] = s0[c];
            } else if (i <= k) {
              const double al = (u - t[i]) / (t[i + p] - t[i]);
              const double *s1 = src(i), *s0 = src(i - 1);
              for (int c = 0; c < 3; c++) q[c] = al * s1[c] + (1.0 - al) * s0[c];
            } else {
              const double* s0 = src(i - 1);
              for (int c = 0; c < 3; c++) q[c] = s0[c];
            }
          }
      knots[dir].insert(knots[dir].begin() + k + 1, u);
      coefs.swap(nc);
      n[dir]++;
    }
  }
  return true;
}

bool ASMs3D::raiseOrder(int ru, int rv, int rw)
{
  const int r[3] = {ru, rv, rw};
  for (int d = 0; d < 3; d++)
    for (int step = 0; step < r[d]; step++) {
      const int p = order[d] - 1;
      if (n[d] != p + 1) {
        std::cerr << " *** ASMs3D::raiseOrder: interior knots in direction " << d + 1
                  << " (raise the order before refining)" << std::endl;
        return false;
      }
      int nn[3] = {n[0], n[1], n[2]}, nw[3] = {n[0], n[1], n[2]};
      nw[d] = p + 2;
      std::vector<double> nc((size_t)nw[0] * nw[1] * nw[2] * 3, 0.0);
      for (int kk = 0; kk < nw[2]; kk++)
        for (int jj = 0; jj < nw[1]; jj++)
          for (int ii = 0; ii < nw[0]; ii++) {
            int idx[3] = {ii, jj, kk};
            const int i = idx[d];
            const double a = (double)i / (p + 1);
            double* q = cpt(nc, nw, ii, jj, kk);
            int id2[3] = {ii, jj, kk};
            if (i > 0) {
              id2[d] = i - 1;
              const double* s0 = cpt(coefs, nn, id2[0], id2[1], id2[2]);
              for (int c = 0; c < 3; c++) q[c] += a * s0[c];
            }
            if (i <= p) {
              id2[d] = i;
              const double* s1 = cpt(coefs, nn, id2[0], id2[1], id2[2]);
              for (int c = 0; c < 3; c++) q[c] += (1.0 - a) * s1[c];
            }
          }
      const double t0 = knots[d].front(), t1 = knots[d].back();
      knots[d].assign(p + 2, t0);
      knots[d].insert(knots[d].end(), p + 2, t1);
      coefs.swap(nc);
      n[d] = p + 2;
      order[d] = p + 2;
    }
  return true;
}

std::unique_ptr<ASMs3D> ASMs3D::unitCube(int p, int nel)
{
  std::vector<double> t;
  for (int i = 0; i <= p; i++) t.push_back(0.0);
  for (int i = 1; i < nel; i++) t.push_back((double)i / nel);
  for (int i = 0; i <= p; i++) t.push_back(1.0);
  const int nc = nel + p;
  std::vector<double> g(nc);
  for (int i = 0; i < nc; i++) {
    double s = 0.0;
    for (int j = 1; j <= p; j++) s += t[i + j];
    g[i] = s / p; // Greville abscissae: identity geometry
  }
  std::vector<double> cf;
  cf.reserve((size_t)nc * nc * nc * 3);
  for (int k = 0; k < nc; k++)
    for (int j = 0; j < nc; j++)
      for (int i = 0; i < nc; i++) {
        cf.push_back(g[i]);
        cf.push_back(g[j]);
        cf.push_back(g[k]);
      }
  const int ord[3] = {p + 1, p + 1, p + 1}, ncf[3] = {nc, nc, nc};
  const std::vector<double> kn[3] = {t, t, t};
  return std::unique_ptr<ASMs3D>(new ASMs3D(ord, ncf, kn, cf));
}

// The GPU matrix behind a global integral: only an AlgEqSystem with an
// initialised GPU matrix can receive the GPU's integrals (any other
// GlobalIntegral would need element matrices on the host, which the GPU path
// never forms).
static GPUMatrix* gpu_matrix(GlobalIntegral& glInt, const char* where)
{
  AlgEqSystem* eq = dynamic_cast<AlgEqSystem*>(&glInt);
  GPUMatrix* G = eq ? dynamic_cast<GPUMatrix*>(eq->getMatrix()) : nullptr;
  if (!G || !G->ready()) {
    std::cerr << " *** " << where << ": the global integral has no initialised GPU matrix"
              << std::endl;
    return nullptr;
  }
  return G;
}

int ASMs3D::getNoGaussPt(int p, bool neumann) const
{
  if (nGauss <= -p) return 0;
  if (nGauss > 0 && nGauss < 11) return nGauss;
  if (neumann) return p;
  return p + nGauss % 10;
}

bool ASMs3D::integrate(Integrand& integrand, GlobalIntegral& glInt, const TimeDomain&)
{
  GPUMatrix* G = gpu_matrix(glInt, "ASMs3D::integrate");
  if (!G) return false;
  ifem_gpu_integrand d;
  if (!integrand.gpuDescription(d)) {
    std::cerr << " *** ASMs3D::integrate: integrand kind not available on the GPU" << std::endl;
    return false;
  }
  // per direction from the spline order, as BasisFunctionCache::
  // setupQuadrature does (src/ASM/ASMs3D.C:4024-4036); the GPU refuses
  // counts it does not support
  const int ngv[3] = {getNoGaussPt(order[0]), getNoGaussPt(order[1]), getNoGaussPt(order[2])};
  if (!report(ifem_gpu_assemble(G->context(), &d, ngv, IFEM_GPU_FP64), "ASMs3D::integrate"))
    return false;
  // the FP32 inner solver needs the FP32 element integrals as well
  if (G->wantsFP32())
    return report(ifem_gpu_assemble(G->context(), &d, ngv, IFEM_GPU_FP32), "ASMs3D::integrate");
  return true;
}

bool ASMs3D::integrate(Integrand& integrand, int lIndex, GlobalIntegral& glInt,
                       const TimeDomain&)
{
  GPUMatrix* G = gpu_matrix(glInt, "ASMs3D::integrate(face)");
  if (!G) return false;
  double t[3];
  if (!integrand.gpuTraction(lIndex % 10, t)) return true; // no load on this face
  // face rule of src/ASM/ASMs3D.C:2593 in the two tangent directions
  const int face = lIndex % 10;
  const int d0 = face <= 2 ? 1 : 0, d1 = face <= 4 ? 2 : 1;
  const int ng = getNoGaussPt(order[d0], true);
  if (getNoGaussPt(order[d1], true) != ng) {
    std::cerr << " *** ASMs3D::integrate(face): unequal Gauss counts on the face" << std::endl;
    return false;
  }
  return report(ifem_gpu_assemble_neumann(G->context(), face, t, ng),
                "ASMs3D::integrate(face)");
}

} // namespace ifem_gpu
