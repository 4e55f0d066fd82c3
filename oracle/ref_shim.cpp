// ref_shim.cpp -- C entry points around the reference's own, unmodified
// sources, compiled where they lie under the reference tree by
// oracle/Makefile into oracle/_ref/libifemref.so.
//
// TEST INFRASTRUCTURE ONLY: the tests use it to pin the oracle's restatement
// of the Gauss tables, utl::Jacobian and SAM equation numbering / DOF
// couplings against the reference code itself.  Nothing in the product loads
// it and it does not travel as a dependency of any product path.
//
// Wrapped reference functions:
//   GaussQuadrature::getCoord/getWeight   src/Utility/GaussQuadrature.C
//   utl::Jacobian                         src/Utility/CoordinateMapping.C:24-63
//   SAM::initSystemEquations              src/LinAlg/SAM.C:225-336
//   SAM::getDofCouplings                  src/LinAlg/SAM.C:377-427
//   SAM::getElmEqns                       src/LinAlg/SAM.C:655-696
//   SAM::expandVector                     src/LinAlg/SAM.C:783-807
#include "CoordinateMapping.h"
#include "GaussQuadrature.h"
#include "SAM.h"
#include "matrix.h"

#include <cstdint>
#include <cstring>
#include <set>
#include <vector>

namespace {

// A SAM filled directly from element connectivity and status codes, the way
// the reference's own unit tests build hand-made SAMs
// (src/LinAlg/Test/TestDiagMatrix.C:27-46).
class ShimSAM : public SAM
{
public:
  // With nceq > 0, the constraint equations in SAM's own layout (the slave
  // DOFs carry status 0, as SAMpatch::initConstraintEqs leaves them).
  ShimSAM(int nnod_, int ncomp, int nel_, const int* mpmnpc_, const int* mmnpc_,
          const int* msc_, int nceq_ = 0, const int* mpmceq_ = nullptr,
          const int* mmceq_ = nullptr, const double* ttcc_ = nullptr)
  {
    nnod = nnod_;
    nel = nel_;
    ndof = nnod_ * ncomp;
    nmmnpc = mpmnpc_[nel_] - 1;
    mpmnpc = new int[nel_ + 1];
    std::memcpy(mpmnpc, mpmnpc_, sizeof(int) * (nel_ + 1));
    mmnpc = new int[nmmnpc];
    std::memcpy(mmnpc, mmnpc_, sizeof(int) * nmmnpc);
    madof = new int[nnod_ + 1];
    for (int i = 0; i <= nnod_; i++) madof[i] = 1 + i * ncomp;
    msc = new int[ndof];
    std::memcpy(msc, msc_, sizeof(int) * ndof);
    if (nceq_ > 0) {
      nceq = nceq_;
      nmmceq = mpmceq_[nceq_] - 1;
      mpmceq = new int[nceq_ + 1];
      std::memcpy(mpmceq, mpmceq_, sizeof(int) * (nceq_ + 1));
      mmceq = new int[nmmceq];
      std::memcpy(mmceq, mmceq_, sizeof(int) * nmmceq);
      ttcc = new Real[nmmceq];
      std::memcpy(ttcc, ttcc_, sizeof(Real) * nmmceq);
    }
    ok = this->initSystemEquations();
  }
  bool ok = false;
  const int* eqns() const { return meqn; }
  // SAM::expandVector with scaleSD = 1 (what SAM::expandSolution uses by
  // default)
  bool expand(const double* x, Vector& dofs) const { return expandVector(x, dofs, 1.0); }
};

} // namespace

extern "C" {

int ref_gauss(int n, double* xi, double* w)
{
  const double* x = GaussQuadrature::getCoord(n);
  const double* g = GaussQuadrature::getWeight(n);
  if (!x || !g) return -1;
  std::memcpy(xi, x, sizeof(double) * n);
  std::memcpy(w, g, sizeof(double) * n);
  return 0;
}

// X: 3 x nen row-major, dNdu: nen x 3 row-major; outputs Jinv 3x3 row-major
// and dNdX nen x 3 row-major.  Returns detJ as utl::Jacobian does.
double ref_jacobian(int nen, const double* X, const double* dNdu, double* Jinv,
                    double* dNdX)
{
  utl::matrix<double> J, dx, Xm(3, nen), Dm(nen, 3);
  for (int i = 0; i < 3; i++)
    for (int b = 0; b < nen; b++) Xm(i + 1, b + 1) = X[i * nen + b];
  for (int b = 0; b < nen; b++)
    for (int k = 0; k < 3; k++) Dm(b + 1, k + 1) = dNdu[b * 3 + k];
  double det = utl::Jacobian(J, dx, Xm, Dm, true);
  if (J.rows() == 3 && J.cols() == 3)
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) Jinv[i * 3 + k] = J(i + 1, k + 1);
  if (det != 0.0)
    for (int b = 0; b < nen; b++)
      for (int m = 0; m < 3; m++) dNdX[b * 3 + m] = dx(b + 1, m + 1);
  return det;
}

// Equation numbers and CSR couplings from the reference SAM.  mpmnpc/mmnpc
// are 1-based as in SAM.  Returns nnz, or -1 on failure / -2 if cap is too
// small.  meqn gets ndof entries, rowptr neq+1 entries.
int64_t ref_sam_couplings(int nnod, int ncomp, int nel, const int* mpmnpc,
                          const int* mmnpc, const int* msc, int* meqn,
                          int* neq_out, int64_t* rowptr, int32_t* colind,
                          int64_t cap)
{
  ShimSAM sam(nnod, ncomp, nel, mpmnpc, mmnpc, msc);
  if (!sam.ok) return -1;
  const int ndof = nnod * ncomp;
  std::memcpy(meqn, sam.eqns(), sizeof(int) * ndof);
  const int neq = sam.getNoEquations();
  *neq_out = neq;
  std::vector<std::set<int>> dofc;
  sam.getDofCouplings(dofc);
  int64_t nnz = 0;
  rowptr[0] = 0;
  for (int r = 0; r < neq; r++) {
    if (nnz + (int64_t)dofc[r].size() > cap) return -2;
    for (int c : dofc[r]) colind[nnz++] = c - 1;
    rowptr[r + 1] = nnz;
  }
  return nnz;
}

// Element equation numbers of element iel (1-based), SAM::getElmEqns.
int ref_sam_elmeqns(int nnod, int ncomp, int nel, const int* mpmnpc,
                    const int* mmnpc, const int* msc, int iel, int* meen)
{
  ShimSAM sam(nnod, ncomp, nel, mpmnpc, mmnpc, msc);
  if (!sam.ok) return -1;
  std::vector<int> v;
  if (!sam.getElmEqns(v, iel)) return -1;
  for (size_t i = 0; i < v.size(); i++) meen[i] = v[i];
  return (int)v.size();
}

// Couplings, equation numbers and solution expansion with constraint
// equations (masters included): SAM::initSystemEquations,
// SAM::getDofCouplings (src/LinAlg/SAM.C:377-427) and SAM::expandVector
// (src/LinAlg/SAM.C:783-807, scaleSD = 1) on xeq (neq entries) -> xdof.
int64_t ref_sam_couplings_mpc(int nnod, int ncomp, int nel, const int* mpmnpc,
                              const int* mmnpc, const int* msc, int nceq, const int* mpmceq,
                              const int* mmceq, const double* ttcc, int* meqn, int* neq_out,
                              int64_t* rowptr, int32_t* colind, int64_t cap, const double* xeq,
                              double* xdof)
{
  ShimSAM sam(nnod, ncomp, nel, mpmnpc, mmnpc, msc, nceq, mpmceq, mmceq, ttcc);
  if (!sam.ok) return -1;
  const int ndof = nnod * ncomp;
  std::memcpy(meqn, sam.eqns(), sizeof(int) * ndof);
  const int neq = sam.getNoEquations();
  *neq_out = neq;
  std::vector<std::set<int>> dofc;
  sam.getDofCouplings(dofc);
  int64_t nnz = 0;
  rowptr[0] = 0;
  for (int r = 0; r < neq; r++) {
    if (nnz + (int64_t)dofc[r].size() > cap) return -2;
    for (int c : dofc[r]) colind[nnz++] = c - 1;
    rowptr[r + 1] = nnz;
  }
  if (xeq && xdof) {
    Vector dofs;
    if (!sam.expand(xeq, dofs)) return -3;
    for (int i = 0; i < ndof; i++) xdof[i] = dofs[i];
  }
  return nnz;
}

} // extern "C"
