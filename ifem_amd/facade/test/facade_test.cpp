// facade_test.cpp -- SIM-style driver through the reference-interface facade:
// model -> SAMpatch -> AlgEqSystem(GPU) -> ASMs3D::integrate (volume, face)
// -> SystemMatrix::solve -> SAM::expandSolution.  Prints one JSON line per
// case for tests/test_facade.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "HDF5Writer.h"
#include "IFEMGpu.h"

using namespace ifem_gpu;

static int poisson(int nel)
{
  auto pch = ASMs3D::unitCube(2, nel);
  for (int f = 1; f <= 6; f++) pch->constrainFace(f, 1);
  SAMpatch sam;
  sam.init(*pch, 1);
  LinSolParams par;
  par.read("<linearsolver><type>cg</type><pc>jacobi</pc><rtol>1e-10</rtol>"
           "<maxits>5000</maxits></linearsolver>");
  AlgEqSystem eq(sam);
  if (!eq.init(LinAlg::GPU, par)) return 1;
  PoissonIntegrand itg;
  itg.source.type = IFEM_GPU_FN_SINPROD;
  itg.source.scale = 3.0 * M_PI * M_PI;
  TimeDomain time;
  eq.initialize(true);
  if (!pch->integrate(itg, eq, time)) return 2;
  if (!eq.finalize()) return 3;
  StdVector& b = *static_cast<StdVector*>(eq.getVector());
  double bsum = 0.0;
  for (double v : b.v) bsum += v;
  Real rc = 0.0;
  if (!eq.getMatrix()->solve(b, &rc)) return 4;
  std::vector<Real> sol;
  sam.expandSolution(b, sol);
  double err = 0.0;
  for (int k = 0; k < pch->n[2]; k++)
    for (int j = 0; j < pch->n[1]; j++)
      for (int i = 0; i < pch->n[0]; i++) {
        const int node = i + pch->n[0] * (j + pch->n[1] * k);
        const double* X = &pch->coefs[3 * node];
        double ue = std::sin(M_PI * X[0]) * std::sin(M_PI * X[1]) * std::sin(M_PI * X[2]);
        err = std::fmax(err, std::fabs(sol[node] - ue));
      }
  // result export (HDF5Writer::writeSIM + writeTimeInfo) when asked for
  int h5 = -1;
  if (const char* out = std::getenv("IFEM_FACADE_HDF5")) {
    HDF5Writer w(out);
    w.openFile(0);
    h5 = w.writeSIM(0, "Poisson", {pch.get()}, {sol}, itg) && w.writeTimeInfo(0, 0.0) ? 1 : 0;
    w.closeFile(0);
  }
  const GPUMatrix* G = static_cast<const GPUMatrix*>(eq.getMatrix());
  std::printf("{\"case\": \"poisson\", \"hdf5\": %d, \"nel\": %d, \"neq\": %d, \"nnz\": %lld, \"bsum\": %.17g, "
              "\"iterations\": %d, \"rel_residual\": %.6g, \"max_nodal_err\": %.6g, "
              "\"normL2\": %.17g}\n",
              h5, nel, sam.getNoEquations(), (long long)G->nonZeros(), bsum, G->lastStats().iterations,
              rc, err, sam.normL2(sol));
  return 0;
}

// GoTools .g2 trilinear unit cube, raised to p = 2 and refined to 16^3
// elements: must equal the generated 16^3 p = 2 cube.
static int g2_cube()
{
  std::istringstream is("700 1 0 0\n3 0\n2 2\n0 0 1 1\n2 2\n0 0 1 1\n2 2\n0 0 1 1\n"
                        "0 0 0\n1 0 0\n0 1 0\n1 1 0\n0 0 1\n1 0 1\n0 1 1\n1 1 1\n");
  auto pch = ASMs3D::read(is);
  if (!pch) return 1;
  if (!pch->raiseOrder(1, 1, 1)) return 2;
  for (int d = 0; d < 3; d++)
    if (!pch->uniformRefine(d, 15)) return 3;
  auto ref = ASMs3D::unitCube(2, 16);
  double dk = 0.0, dc = 0.0;
  for (int d = 0; d < 3; d++) {
    if (pch->n[d] != ref->n[d] || pch->order[d] != ref->order[d]) return 4;
    for (size_t i = 0; i < ref->knots[d].size(); i++)
      dk = std::fmax(dk, std::fabs(pch->knots[d][i] - ref->knots[d][i]));
  }
  for (size_t i = 0; i < ref->coefs.size(); i++)
    dc = std::fmax(dc, std::fabs(pch->coefs[i] - ref->coefs[i]));
  std::printf("{\"case\": \"g2\", \"n\": %d, \"knot_diff\": %.6g, \"coef_diff\": %.6g}\n",
              pch->n[0], dk, dc);
  return 0;
}

static int elasticity(int nel, const char* precision)
{
  auto pch = ASMs3D::unitCube(2, nel);
  pch->constrainFace(1, 7); // clamped at x = 0
  SAMpatch sam;
  sam.init(*pch, 3);
  LinSolParams par;
  par.addValue("type", "cg");
  par.addValue("pc", "jacobi");
  par.addValue("rtol", "1e-8");
  par.addValue("maxits", "50000");
  par.addValue("precision", precision);
  AlgEqSystem eq(sam);
  if (!eq.init(LinAlg::GPU, par)) return 11;
  ElasticityIntegrand itg(1.0, 0.3);
  itg.setTraction(2, 0.0, 0.0, -1.0); // face 2: x = 1
  TimeDomain time;
  eq.initialize(true);
  if (!pch->integrate(itg, eq, time)) return 12;
  for (int f = 1; f <= 6; f++)
    if (!pch->integrate(itg, f, eq, time)) return 13;
  if (!eq.finalize()) return 14;
  StdVector& b = *static_cast<StdVector*>(eq.getVector());
  double fz = 0.0;
  for (size_t i = 2; i < b.v.size(); i += 3) fz += b.v[i];
  Real rc = 0.0;
  if (!eq.getMatrix()->solve(b, &rc)) return 15;
  std::vector<Real> sol;
  sam.expandSolution(b, sol);
  double tip = 0.0;
  for (size_t i = 2; i < sol.size(); i += 3) tip = std::fmin(tip, sol[i]);
  const GPUMatrix* G = static_cast<const GPUMatrix*>(eq.getMatrix());
  std::printf("{\"case\": \"elasticity\", \"precision\": \"%s\", \"nel\": %d, \"neq\": %d, "
              "\"force_z\": %.17g, \"iterations\": %d, \"rel_residual\": %.6g, "
              "\"min_uz\": %.17g, \"normL2\": %.17g}\n",
              precision, nel, sam.getNoEquations(), fz, G->lastStats().iterations, rc, tip,
              sam.normL2(sol));
  return 0;
}

// Uniaxial strain by prescribed displacements: u_x = 0 on x = 0 and d on
// x = 1, symmetry planes y = 0 (u_y = 0) and z = 0 (u_z = 0), traction free
// elsewhere; the exact u = (d x, -nu d y, -nu d z) is in the spline space.
static int uniaxial(int nel)
{
  const double d = 0.01, nu = 0.3;
  auto pch = ASMs3D::unitCube(2, nel);
  pch->constrainFace(3, 2);
  pch->constrainFace(5, 4);
  pch->prescribeFace(1, 1, 0.0);
  pch->prescribeFace(2, 1, d);
  SAMpatch sam;
  sam.init(*pch, 3);
  LinSolParams par;
  par.addValue("type", "cg");
  par.addValue("pc", "jacobi");
  par.addValue("rtol", "1e-12");
  par.addValue("maxits", "50000");
  AlgEqSystem eq(sam);
  if (!eq.init(LinAlg::GPU, par)) return 21;
  ElasticityIntegrand itg(1.0, nu);
  TimeDomain time;
  eq.initialize(true);
  if (!pch->integrate(itg, eq, time)) return 22;
  if (!eq.finalize()) return 23;
  StdVector& b = *static_cast<StdVector*>(eq.getVector());
  Real rc = 0.0;
  if (!eq.getMatrix()->solve(b, &rc)) return 24;
  std::vector<Real> sol;
  sam.expandSolution(b, sol);
  double err = 0.0;
  for (int node = 0; node < pch->getNoNodes(); node++) {
    const double* X = &pch->coefs[3 * node];
    const double ue[3] = {d * X[0], -nu * d * X[1], -nu * d * X[2]};
    for (int c = 0; c < 3; c++) err = std::fmax(err, std::fabs(sol[3 * node + c] - ue[c]));
  }
  std::printf("{\"case\": \"uniaxial\", \"nel\": %d, \"neq\": %d, \"nceq\": %d, "
              "\"max_err\": %.6g, \"rel_residual\": %.6g}\n",
              nel, sam.getNoEquations(), sam.getNoConstraints(), err, rc);
  return 0;
}

// ASMs3D::getNoGaussPt for (order, nGauss, neumann) triples given as
// arguments; prints the counts as one JSON line (no GPU call).
static int gauss_rule(int argc, char** argv)
{
  auto pch = ASMs3D::unitCube(1, 1);
  std::printf("{\"case\": \"gauss_rule\", \"ng\": [");
  for (int a = 2; a + 2 < argc; a += 3) {
    pch->setGauss(std::atoi(argv[a + 1]));
    std::printf("%s%d", a > 2 ? ", " : "",
                pch->getNoGaussPt(std::atoi(argv[a]), std::atoi(argv[a + 2]) != 0));
  }
  std::printf("]}\n");
  return 0;
}

int main(int argc, char** argv)
{
  if (argc > 1 && std::string(argv[1]) == "gauss") return gauss_rule(argc, argv);
  int rc = g2_cube();
  if (!rc) rc = poisson(16);
  if (!rc) rc = elasticity(8, "fp64");
  if (!rc) rc = elasticity(8, "fp32");
  if (!rc) rc = uniaxial(6);
  // unknown solver type is refused (bool false), as the reference does for
  // unavailable back-ends
  if (!rc) {
    auto pch = ASMs3D::unitCube(1, 2);
    SAMpatch sam;
    sam.init(*pch, 1);
    LinSolParams par; // default type gmres
    AlgEqSystem eq(sam);
    eq.init(LinAlg::GPU, par);
    PoissonIntegrand itg;
    TimeDomain time;
    pch->integrate(itg, eq, time);
    StdVector& b = *static_cast<StdVector*>(eq.getVector());
    std::printf("{\"case\": \"gmres_refused\", \"ok\": %s}\n",
                eq.getMatrix()->solve(b) ? "false" : "true");
  }
  return rc;
}
