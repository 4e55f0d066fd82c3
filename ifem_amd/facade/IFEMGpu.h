// IFEMGpu.h -- C++ classes with the reference's interface names for the
// isogeometric assembly + linear solve path, implemented on the C ABI of
// libifem_gpu.so (include/ifem_gpu.h).  A SIM-style driver written against
// the reference classes calls these with the same method names and meaning:
//
//   ASMs3D            ASMbase::integrate / ASMs3D::integrate (volume, face)
//                       src/ASM/ASMbase.h:555-563, src/ASM/ASMs3D.h:467-475
//   Integrand         Integrand::evalInt physics; GPU-recognised kinds
//                       PoissonIntegrand, ElasticityIntegrand
//                       src/ASM/Integrand.h:206,344
//   SAM / SAMpatch    equation numbering, expandSolution, normL2
//                       src/LinAlg/SAM.h, src/ASM/SAMpatch.C:21-161
//   GlobalIntegral    system-level integrated quantity (initialize,
//                       finalize, assemble, threadSafe) src/ASM/GlobalIntegral.h
//   AlgEqSystem       GlobalIntegral owning the system (init, initialize,
//                       finalize, getMatrix, getVector) src/ASM/AlgEqSystem.C
//   SystemMatrix      create(adm, MatrixType, LinSolParams), initAssembly,
//                       init, endAssembly, multiply, solve
//                       src/LinAlg/SystemMatrix.h:219-388
//   LinSolParams      key/value solver settings src/LinAlg/LinSolParams.h:31-57
//
// Errors follow the reference: bool returns plus a message on std::cerr; no
// exceptions cross these calls.  Integrands the GPU does not recognise make
// integrate() return false -- there is no CPU fallback.
#pragma once

#include <istream>
#include <map>
#include <ostream>
#include <memory>
#include <string>
#include <vector>

#include "ifem_gpu.h"

namespace ifem_gpu {

typedef double Real;

namespace LinAlg {
//! Matrix types of src/LinAlg/LinAlgenums.h:21-31 plus the GPU type.
enum MatrixType { DENSE = 0, SPR = 1, SPARSE = 2, SAMG = 3, PETSC = 4, ISTL = 5,
                  UMFPACK = 6, DIAG = 7, GPU = 8 };
enum LinearSystemType { GENERAL_MATRIX = 0, SYMMETRIC = 1, SPD = 2 };
} // namespace LinAlg

//! Time level (src/ASM/TimeDomain.h); stationary problems use the defaults.
struct TimeDomain {
  double t = 0.0, dt = 0.0;
  int it = 0;
  bool first = true;
};

//! Key/value settings store (src/LinAlg/LinSolParams.h:31-57).
class SettingMap {
public:
  void addValue(const std::string& key, const std::string& value) { values[key] = value; }
  std::string getStringValue(const std::string& key) const;
  int getIntValue(const std::string& key) const;
  double getDoubleValue(const std::string& key) const;
  bool hasValue(const std::string& key) const { return values.count(key) > 0; }

protected:
  std::map<std::string, std::string> values;
};

//! Linear solver parameters with the reference defaults
//! (src/LinAlg/LinSolParams.C:68-77).  Keys honoured on the GPU: type (cg),
//! pc (jacobi | default), rtol, atol, dtol, maxits, and the additions
//! precision (fp64 | fp32), inner_rtol.
class LinSolParams : public SettingMap {
public:
  explicit LinSolParams(LinAlg::LinearSystemType ls = LinAlg::GENERAL_MATRIX);
  //! Reads the children of a <linearsolver> element given as XML text
  //! (the tags of LinSolParams::read, src/LinAlg/LinSolParams.C:158-195).
  bool read(const std::string& xml);
  LinAlg::LinearSystemType getLinearSystemType() const { return linSys; }

private:
  LinAlg::LinearSystemType linSys;
};

//! Parameters of one integration point (src/ASM/FiniteElement.h subset).
struct FiniteElement {
  double detJxW = 0.0;
  std::vector<double> N;
};

class LocalIntegral {
public:
  virtual ~LocalIntegral() {}
};

//! Physics callback base (src/ASM/Integrand.h).  The GPU executes the
//! integrand kinds that describe themselves through gpuDescription().
class Integrand {
public:
  virtual ~Integrand() {}
  //! Number of solution components per node.
  virtual int getNoFields() const = 0;
  //! Fills the device description; false if the GPU does not know the kind.
  virtual bool gpuDescription(ifem_gpu_integrand& d) const = 0;
  //! Neumann traction of face lIndex (false if none).
  virtual bool gpuTraction(int lIndex, double t[3]) const { (void)lIndex; (void)t; return false; }
  //! Name of primary solution field idx ("primary solution <idx+1>",
  //! IntegrandBase::getField1Name, src/ASM/IntegrandBase.C:398-403); the
  //! result export names the whole primary solution with idx = 11.
  virtual std::string getField1Name(size_t idx, const char* prefix = nullptr) const;
};

//! -div(kappa grad u) = f (Weak::Laplacian/LaplacianCoeff + Weak::Source,
//! Apps/Utils/EqualOrderOperators.C:158-211,243-255).
class PoissonIntegrand : public Integrand {
public:
  PoissonIntegrand() { kappa = {}; source = {}; }
  int getNoFields() const override { return 1; }
  bool gpuDescription(ifem_gpu_integrand& d) const override;
  //! "u" (the scalar solution).
  std::string getField1Name(size_t idx, const char* prefix = nullptr) const override;
  ifem_gpu_function kappa;   //!< conductivity (ZERO: 1)
  ifem_gpu_function source;  //!< source f
  std::vector<double> kappaModes, sourceModes; //!< storage for mode tables
};

//! Isotropic linear elasticity (textbook B^T C B) with surface tractions.
class ElasticityIntegrand : public Integrand {
public:
  ElasticityIntegrand(double E_ = 1.0, double nu_ = 0.3) : E(E_), nu(nu_) {}
  int getNoFields() const override { return 3; }
  bool gpuDescription(ifem_gpu_integrand& d) const override;
  bool gpuTraction(int lIndex, double t[3]) const override;
  //! "u_x", "u_y", "u_z" for idx 0..2, "u" for the displacement vector.
  std::string getField1Name(size_t idx, const char* prefix = nullptr) const override;
  void setTraction(int lIndex, double tx, double ty, double tz);
  double E, nu;
  double body[3] = {0.0, 0.0, 0.0};
  std::map<int, std::vector<double>> traction;
};

class SAM;
class SystemMatrix;
class SystemVector;
class GPUMatrix;

//! Abstract right-hand side / solution vector (src/LinAlg/SystemMatrix.h).
class SystemVector {
public:
  virtual ~SystemVector() {}
  virtual size_t dim() const = 0;
  virtual void redim(size_t n) = 0;
  virtual Real* getPtr() = 0;
  virtual const Real* getRef() const = 0;
  virtual void init(Real value = Real(0)) = 0;
};

//! Contiguous host vector with 1-based operator() (StdVector).
class StdVector : public SystemVector {
public:
  explicit StdVector(size_t n = 0) : v(n, 0.0) {}
  size_t dim() const override { return v.size(); }
  void redim(size_t n) override { v.assign(n, 0.0); }
  Real* getPtr() override { return v.data(); }
  const Real* getRef() const override { return v.data(); }
  void init(Real value = Real(0)) override { std::fill(v.begin(), v.end(), value); }
  Real& operator()(size_t i) { return v[i - 1]; }
  const Real& operator()(size_t i) const { return v[i - 1]; }
  std::vector<Real> v;
};

//! Abstract system matrix (src/LinAlg/SystemMatrix.h:219-388).
class SystemMatrix {
public:
  virtual ~SystemMatrix() {}
  //! Factory of SystemMatrix::create (src/LinAlg/SystemMatrix.C:105-174);
  //! only LinAlg::GPU is provided by this library.
  static SystemMatrix* create(void* adm, LinAlg::MatrixType mtype, const LinSolParams& spar);
  virtual LinAlg::MatrixType getType() const = 0;
  virtual size_t dim(int idim = 1) const = 0;
  virtual void initAssembly(const SAM& sam, char preAssembly = 0) = 0;
  virtual void init() = 0;
  virtual bool endAssembly() { return true; }
  virtual bool multiply(const SystemVector& B, SystemVector& C) const = 0;
  //! Solves A x = b; b is overwritten by x (reference semantics).
  virtual bool solve(SystemVector& b, Real* rc = nullptr) = 0;
};

//! The GPU system matrix: CSR in HBM, assembled and solved by libifem_gpu.
class GPUMatrix : public SystemMatrix {
public:
  explicit GPUMatrix(const LinSolParams& spar);
  ~GPUMatrix() override;
  LinAlg::MatrixType getType() const override { return LinAlg::GPU; }
  size_t dim(int idim = 1) const override { (void)idim; return neq; }
  void initAssembly(const SAM& sam, char preAssembly = 0) override;
  void init() override {}
  bool multiply(const SystemVector& B, SystemVector& C) const override;
  bool solve(SystemVector& b, Real* rc = nullptr) override;

  ifem_gpu_ctx* context() const { return ctx; }
  bool ready() const { return ok; }
  int64_t nonZeros() const { return nnz; }
  //! True when the solver parameters ask for the FP32 inner iteration.
  bool wantsFP32() const { return par.getStringValue("precision") == "fp32"; }
  //! Statistics of the last solve.
  const ifem_gpu_stats& lastStats() const { return stats; }

private:
  const LinSolParams& par;
  ifem_gpu_ctx* ctx = nullptr;
  bool ok = false;
  size_t neq = 0;
  int64_t nnz = 0;
  ifem_gpu_stats stats{};
};

//! Assembly manager: equation numbering of the patch
//! (src/LinAlg/SAM.h, src/LinAlg/SAM.C:225-336, 768-880).
class SAM {
public:
  virtual ~SAM() {}
  int getNoNodes() const { return nnod; }
  int getNoDOFs() const { return ndof; }
  int getNoEquations() const { return neq; }
  const std::vector<int>& getMEQN() const { return meqn; }
  int getNoConstraints() const { return nceq; }
  //! Equation vector -> DOF vector, fixed DOFs = 0, prescribed DOFs
  //! scaleSD * c0 (SAM::expandSolution, src/LinAlg/SAM.C:768-810).
  bool expandSolution(const SystemVector& solVec, std::vector<Real>& dofVec,
                      Real scaleSD = Real(1)) const;
  //! RMS norm ||x||_2 / sqrt(n) (SAM::normL2, src/LinAlg/SAM.C:860-880).
  Real normL2(const std::vector<Real>& x) const;

protected:
  int nnod = 0, ndof = 0, neq = 0, ncomp = 1, nceq = 0;
  std::vector<int> msc, meqn;
  //! constraint equations (1-based, src/LinAlg/SAM.h:330-350): entries
  //! mpmceq[k-1] .. mpmceq[k]-1 of mmceq / ttcc belong to constraint k
  std::vector<int> mpmceq, mmceq;
  std::vector<Real> ttcc;
  friend class GPUMatrix;
  const class ASMs3D* patch = nullptr;
};

class ASMs3D;

//! SAM of one structured patch (SAMpatch::init / initNodeDofs with
//! homogeneous Dirichlet codes, src/ASM/SAMpatch.C:21-161; prescribed values
//! as constraint equations without masters, SAMpatch::initConstraintEqs,
//! src/ASM/SAMpatch.C:259-357).
class SAMpatch : public SAM {
public:
  bool init(const ASMs3D& pch, int ncomp);
};

//! System-level integrated quantity (src/ASM/GlobalIntegral.h:28-50): the
//! target ASMbase::integrate assembles into.
class GlobalIntegral {
public:
  virtual ~GlobalIntegral() {}
  //! Initializes the integrated quantity to zero.
  virtual void initialize(char) {}
  //! Finalizes the integrated quantity after element assembly.
  virtual bool finalize(bool) { return true; }
  //! Adds a LocalIntegral into the global object.
  virtual bool assemble(const LocalIntegral*, int) { return true; }
  //! True if all elements can be assembled in parallel.
  virtual bool threadSafe() const { return false; }
};

//! GlobalIntegral owning one system matrix and one right-hand side
//! (src/ASM/AlgEqSystem.C:33-351).  The GPU integrates and scatters in one
//! pass, so no element matrices reach assemble(): integrate() fills the
//! matrix of an AlgEqSystem directly.
class AlgEqSystem : public GlobalIntegral {
public:
  explicit AlgEqSystem(const SAM& s) : sam(s) {}
  bool init(LinAlg::MatrixType mtype, const LinSolParams& spar);
  void initialize(char initLHS = 1) override;
  bool finalize(bool newLHS = true) override;
  SystemMatrix* getMatrix() { return A.get(); }
  SystemVector* getVector() { return &b; }
  const SAM& getSAM() const { return sam; }
  bool threadSafe() const override { return true; }

private:
  const SAM& sam;
  std::unique_ptr<SystemMatrix> A;
  StdVector b;
  friend class ASMs3D;
};

//! Structured trivariate spline patch (src/ASM/ASMs3D.h).
class ASMs3D {
public:
  //! Patch from knot vectors (GoTools order = p+1) and control points.
  ASMs3D(const int order[3], const int ncoef[3], const std::vector<double> knots[3],
         const std::vector<double>& coefs);
  //! Unit-cube patch of degree p with nel elements per direction
  //! (DefaultGeometry3D + refine/raiseorder, src/SIM/ModelGenerator.C:203-273).
  static std::unique_ptr<ASMs3D> unitCube(int p, int nel);
  //! Reads one non-rational SplineVolume (GoTools .g2, class 700) as
  //! ASMs3D::read does (src/ASM/ASMs3D.C:117-160); nullptr on failure.
  static std::unique_ptr<ASMs3D> read(std::istream& is);
  //! Writes the patch as GoTools text (ASMs3D::write, src/ASM/ASMs3D.C:163-173):
  //! "700 1 0 0", the spline volume, numbers with 15 significant digits.
  bool write(std::ostream& os, int basis = 1) const;
  //! Inserts nInsert equally spaced knots in every knot span of direction
  //! dir = 0..2 (ASM3D::uniformRefine, src/SIM/SIM3D.C:104-128), by knot
  //! insertion, which leaves the geometry unchanged.
  bool uniformRefine(int dir, int nInsert);
  //! Raises the degree by (ru, rv, rw) (ASM3D::raiseOrder,
  //! src/SIM/SIM3D.C:152-170); only for directions without interior knots
  //! (raise before refining, as IFEM inputs do).
  bool raiseOrder(int ru, int rv, int rw);

  //! Marks components cmask (bit c) of face 1..6 as homogeneous Dirichlet
  //! (ASMbase::fix/constrainFace with fixHomogeneousDirichlet).
  void constrainFace(int face, int cmask) { dirichlet[face] |= cmask; }
  //! Prescribes the value on components cmask of face 1..6 (constrainFace
  //! with a constant inhomogeneous Dirichlet function: one constraint
  //! equation u = value per DOF, ASMs3D::constrainFace with code > 0,
  //! src/ASM/ASMs3D.C:1025-1118).  A DOF already fixed homogeneously stays
  //! fixed (ASMbase::addMPC, src/ASM/ASMbase.C:527-545); a DOF on several
  //! prescribed faces takes the last value.
  void prescribeFace(int face, int cmask, double value)
  { prescribed.push_back({face, cmask, value}); }
  //! Integration rule code of the patch (ASMbase::setGauss; SIMbase passes
  //! SIMoptions::nGauss[0], default 4): 1..10 a fixed count, else see
  //! getNoGaussPt.
  void setGauss(int ng) { nGauss = ng; }
  //! Gauss points in a direction of spline order p (= degree + 1), as
  //! ASMbase::getNoGaussPt (src/ASM/ASMbase.C:1788-1798): 0 for
  //! nGauss <= -p, nGauss for 1..10, p on faces (neumann), else
  //! p + nGauss%10 (C remainder).
  int getNoGaussPt(int p, bool neumann = false) const;

  //! Volume integration (ASMs3D::integrate, src/ASM/ASMs3D.C:2071-2335).
  bool integrate(Integrand& integrand, GlobalIntegral& glInt, const TimeDomain& time);
  //! Face integration (src/ASM/ASMs3D.C:2547-2812), lIndex = face 1..6.
  bool integrate(Integrand& integrand, int lIndex, GlobalIntegral& glInt, const TimeDomain& time);

  int getNoNodes() const { return n[0] * n[1] * n[2]; }
  int order[3];
  int n[3];
  std::vector<double> knots[3];
  std::vector<double> coefs;
  std::map<int, int> dirichlet;
  struct Prescribed { int face, cmask; double value; };
  std::vector<Prescribed> prescribed;
  int nGauss = 0; //!< ASMbase's initial value (src/ASM/ASMbase.C:65): order points

private:
  bool uploaded(GPUMatrix* M) const;
};

} // namespace ifem_gpu
