/*
 * ifem_oracle.c -- CPU restatement of IFEM's 3D isogeometric assembly and
 * Jacobi-PCG hot path (plain C99 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product library never links or calls it.
 *
 * Each function names the reference file:line it restates.  Spline
 * evaluation in the reference is delegated to GoTools (unvendored); the basis
 * code here restates the published Cox-de Boor algorithm (Piegl & Tiller,
 * The NURBS Book, A2.1-A2.3) for non-rational splines.
 */
#include "ifem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EPSZ 1.0e-12 /* epsZ of the reference build (CMakeLists.txt:538) */
#define ORC_MAXP 4
#define ORC_MAXNEN ((ORC_MAXP + 1) * (ORC_MAXP + 1) * (ORC_MAXP + 1))
#define ORC_MAXNG 10

/* ---- a1: Gauss-Legendre rule ------------------------------------------- */
/* The reference stores tables G1..G10 (src/Utility/GaussQuadrature.C:17-60);
 * here the same rule is computed by Newton iteration on the Legendre
 * polynomial in long double and rounded to double; the tests pin it to the
 * reference's table values. */
int orc_gauss(int n, double* xi, double* w)
{
  if (n < 1 || n > ORC_MAXNG) return -1;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; i++) {
    long double x = cosl(pi * (i + 0.75L) / (n + 0.5L));
    long double dp = 0.0L;
    for (int it = 0; it < 100; it++) {
      long double p0 = 1.0L, p1 = x;
      for (int k = 1; k < n; k++) {
        long double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0L);
      long double dx = p1 / dp;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    if (2 * i + 1 == n) x = 0.0L; /* exact middle node of odd rules */
    {
      long double p0 = 1.0L, p1 = x;
      for (int k = 1; k < n; k++) {
        long double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0L);
    }
    /* roots come out descending; store ascending like the reference tables */
    int j = n - 1 - i;
    xi[j] = (double)x;
    w[j] = (double)(2.0L / ((1.0L - x * x) * dp * dp));
  }
  return 0;
}

/* ---- a2: ASMbase::getNoGaussPt (src/ASM/ASMbase.C:1788-1798) ------------ */
/* order = spline order p+1, as ASMs3D passes it (getOrder,
 * src/ASM/ASMs3D.C:1898-1906; calls at 4033 and 2593).  C's % keeps the sign
 * of nGauss, as in the reference. */
int orc_num_gauss(int order, int nGauss, int neumann)
{
  if (nGauss <= -order) return 0;
  if (nGauss > 0 && nGauss < 11) return nGauss;
  if (neumann) return order;
  return order + nGauss % 10;
}

/* ---- a3: SplineUtils::getGaussParameters (src/Utility/SplineUtils.C:217) */
double orc_gauss_param(double u0, double u1, double xi)
{
  return 0.5 * ((u1 - u0) * xi + u1 + u0);
}

/* ---- a4: knot span + Cox-de Boor values and first derivatives ---------- */
/* GoTools SplineVolume::computeBasis (called at src/ASM/ASMs3D.C:4080,4087)
 * restated from Piegl & Tiller A2.1 (span), A2.2 (values), A2.3 (k=1). */
int orc_find_span(int p, int n, const double* t, double u)
{
  if (u >= t[n]) {
    int s = n - 1;
    while (s > p && t[s] >= t[s + 1]) s--;
    return s;
  }
  if (u <= t[p]) return p;
  int lo = p, hi = n;
  int mid = (lo + hi) / 2;
  while (u < t[mid] || u >= t[mid + 1]) {
    if (u < t[mid]) hi = mid; else lo = mid;
    mid = (lo + hi) / 2;
  }
  return mid;
}

void orc_basis_funs(int p, const double* t, int s, double u, double* N, double* dN)
{
  double left[ORC_MAXP + 2], right[ORC_MAXP + 2];
  double ndu[ORC_MAXP + 1][ORC_MAXP + 1];
  ndu[0][0] = 1.0;
  for (int j = 1; j <= p; j++) {
    left[j] = u - t[s + 1 - j];
    right[j] = t[s + j] - u;
    double saved = 0.0;
    for (int r = 0; r < j; r++) {
      ndu[j][r] = right[r + 1] + left[j - r];
      double temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
  for (int r = 0; r <= p; r++) N[r] = ndu[r][p];
  for (int r = 0; r <= p; r++) {
    double d = 0.0;
    if (p >= 1) {
      if (r >= 1) d += ndu[r - 1][p - 1] / ndu[p][r - 1];
      if (r <= p - 1) d -= ndu[r][p - 1] / ndu[p][r];
    }
    dN[r] = p * d;
  }
}

/* ---- a5: elements and MNPC (src/ASM/ASMs3D.C:3899-3924, 2140-2142) ----- */
int orc_num_elements(const orc_patch* P, int nel[3])
{
  for (int d = 0; d < 3; d++) {
    int p = P->p[d], n = P->n[d];
    const double* t = P->knots[d];
    if (p < 1 || p > ORC_MAXP || n <= p) return -1;
    for (int s = p; s < n; s++)
      if (!(t[s + 1] > t[s])) return -2; /* repeated interior knot */
    nel[d] = n - p;
  }
  return 0;
}

void orc_element_nodes(const orc_patch* P, int e1, int e2, int e3, int* mnpc)
{
  const int n1 = P->n[0], n2 = P->n[1];
  int a = 0;
  for (int a3 = 0; a3 <= P->p[2]; a3++)
    for (int a2 = 0; a2 <= P->p[1]; a2++)
      for (int a1 = 0; a1 <= P->p[0]; a1++, a++)
        mnpc[a] = (e1 + a1) + n1 * ((e2 + a2) + n2 * (e3 + a3));
}

/* ---- a8: utl::Jacobian (src/Utility/CoordinateMapping.C:24-63) with
 *          matrix::det / matrix::inverse (src/LinAlg/matrix.h:791-862) ---- */
static double det3(const double J[9])
{
#define T(i, j) J[((i) - 1) * 3 + ((j) - 1)]
  return T(1, 1) * (T(2, 2) * T(3, 3) - T(3, 2) * T(2, 3))
       - T(1, 2) * (T(2, 1) * T(3, 3) - T(3, 1) * T(2, 3))
       + T(1, 3) * (T(2, 1) * T(3, 2) - T(3, 1) * T(2, 2));
}

static double inverse3(double J[9])
{
  double Det = det3(J);
  if (Det <= ORC_EPSZ && Det >= -ORC_EPSZ) return 0.0;
  double B[9];
#define B_(i, j) B[((i) - 1) * 3 + ((j) - 1)]
  B_(1, 1) = (T(2, 2) * T(3, 3) - T(3, 2) * T(2, 3)) / Det;
  B_(2, 1) = -(T(2, 1) * T(3, 3) - T(3, 1) * T(2, 3)) / Det;
  B_(3, 1) = (T(2, 1) * T(3, 2) - T(3, 1) * T(2, 2)) / Det;
  B_(1, 2) = -(T(1, 2) * T(3, 3) - T(3, 2) * T(1, 3)) / Det;
  B_(2, 2) = (T(1, 1) * T(3, 3) - T(3, 1) * T(1, 3)) / Det;
  B_(3, 2) = -(T(1, 1) * T(3, 2) - T(3, 1) * T(1, 2)) / Det;
  B_(1, 3) = (T(1, 2) * T(2, 3) - T(2, 2) * T(1, 3)) / Det;
  B_(2, 3) = -(T(1, 1) * T(2, 3) - T(2, 1) * T(1, 3)) / Det;
  B_(3, 3) = (T(1, 1) * T(2, 2) - T(2, 1) * T(1, 2)) / Det;
#undef B_
  memcpy(J, B, sizeof B);
  return Det;
#undef T
}

double orc_jacobian(int nen, const double* X, const double* dNdu,
                    double* Jinv, double* dNdX)
{
  double J[9];
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) {
      double s = 0.0;
      for (int b = 0; b < nen; b++) s += X[i * nen + b] * dNdu[b * 3 + k];
      J[i * 3 + k] = s; /* J = X * dNdu */
    }
  double detJ = inverse3(J);
  memcpy(Jinv, J, sizeof J);
  if (detJ == 0.0) return 0.0;
  for (int b = 0; b < nen; b++)
    for (int m = 0; m < 3; m++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += dNdu[b * 3 + k] * J[k * 3 + m];
      dNdX[b * 3 + m] = s; /* dNdX = dNdu * J^-1 */
    }
  return detJ;
}

/* ---- field evaluation (the configs' closed-form f, kappa) ---------------- */
double orc_eval_function(const orc_function* f, const double X[3])
{
  const double pi = 3.14159265358979323846;
  switch (f->type) {
  case ORC_FN_CONST:
    return f->scale;
  case ORC_FN_SINPROD:
    return f->scale * sin(pi * X[0]) * sin(pi * X[1]) * sin(pi * X[2]);
  case ORC_FN_FOURIER:
  case ORC_FN_EXP_FOURIER: {
    double s = 0.0;
    for (int m = 0; m < f->nmodes; m++) {
      const double* md = f->modes + 5 * m;
      s += md[4] * cos(md[0] * X[0] + md[1] * X[1] + md[2] * X[2] + md[3]);
    }
    double g = f->mean + f->scale * s;
    return f->type == ORC_FN_EXP_FOURIER ? exp(g) : g;
  }
  default:
    return 0.0;
  }
}

/* ---- a9/a8/a6/a7: one element (src/ASM/ASMs3D.C:2140-2312) -------------- */
/* Volume loop of ASMs3D::integrate for one element: Gauss parameters,
 * basis (cache), Jacobian, detJxW *= dV*wi*wj*wk, then Integrand::evalInt:
 * Poisson   EK += kappa*detJxW*dNdX*dNdX^T, EV += f*detJxW*N
 *           (Apps/Utils/EqualOrderOperators.C:158-174,243-255);
 * elasticity EK(a i, b j) += detJxW*(lam g_ai g_bj + mu g_aj g_bi
 *           + mu delta_ij g_a.g_b), EV(a i) += detJxW*body_i*N_a
 *           (textbook isotropic Hooke; no reference integrand exists). */
int orc_element_matrices(const orc_patch* P, const orc_integrand* itg,
                         const int ng[3], int e1, int e2, int e3,
                         double* EK, double* EV)
{
  const int p1 = P->p[0], p2 = P->p[1], p3 = P->p[2];
  const int nen = (p1 + 1) * (p2 + 1) * (p3 + 1);
  const int ncomp = itg->kind == ORC_ELASTICITY ? 3 : 1;
  const int nedof = ncomp * nen;
  int mnpc[ORC_MAXNEN];
  double Xnod[3 * ORC_MAXNEN];
  double xg[3][ORC_MAXNG], wg[3][ORC_MAXNG];
  const int e[3] = {e1, e2, e3};
  for (int d = 0; d < 3; d++)
    if (orc_gauss(ng[d], xg[d], wg[d])) return -1;

  orc_element_nodes(P, e1, e2, e3, mnpc);
  for (int a = 0; a < nen; a++)
    for (int i = 0; i < 3; i++) Xnod[i * nen + a] = P->coefs[3 * mnpc[a] + i];

  /* dV = 0.125 * parametric volume (ASMs3D.C:2145, 1531-1558) */
  double span[3];
  for (int d = 0; d < 3; d++) {
    const int s = e[d] + P->p[d];
    span[d] = P->knots[d][s + 1] - P->knots[d][s];
  }
  const double dV = 0.125 * (span[0] * span[1] * span[2]);

  double lam = 0.0, mu = 0.0;
  if (ncomp == 3) {
    lam = itg->E * itg->nu / ((1.0 + itg->nu) * (1.0 - 2.0 * itg->nu));
    mu = itg->E / (2.0 * (1.0 + itg->nu));
  }

  memset(EK, 0, sizeof(double) * nedof * nedof);
  memset(EV, 0, sizeof(double) * nedof);

  double N1[ORC_MAXNG][ORC_MAXP + 1], dN1[ORC_MAXNG][ORC_MAXP + 1];
  double N2[ORC_MAXNG][ORC_MAXP + 1], dN2[ORC_MAXNG][ORC_MAXP + 1];
  double N3[ORC_MAXNG][ORC_MAXP + 1], dN3[ORC_MAXNG][ORC_MAXP + 1];
  for (int d = 0; d < 3; d++) {
    const int s = e[d] + P->p[d];
    const double* t = P->knots[d];
    for (int g = 0; g < ng[d]; g++) {
      double u = orc_gauss_param(t[s], t[s + 1], xg[d][g]);
      double* Nd = d == 0 ? N1[g] : (d == 1 ? N2[g] : N3[g]);
      double* dNd = d == 0 ? dN1[g] : (d == 1 ? dN2[g] : dN3[g]);
      orc_basis_funs(P->p[d], t, s, u, Nd, dNd);
    }
  }

  double N[ORC_MAXNEN], dNdu[3 * ORC_MAXNEN], dNdX[3 * ORC_MAXNEN], Jinv[9];
  for (int k = 0; k < ng[2]; k++)
    for (int j = 0; j < ng[1]; j++)
      for (int i = 0; i < ng[0]; i++) {
        int a = 0;
        for (int a3 = 0; a3 <= p3; a3++)
          for (int a2 = 0; a2 <= p2; a2++)
            for (int a1 = 0; a1 <= p1; a1++, a++) {
              N[a] = N1[i][a1] * N2[j][a2] * N3[k][a3];
              dNdu[3 * a + 0] = dN1[i][a1] * N2[j][a2] * N3[k][a3];
              dNdu[3 * a + 1] = N1[i][a1] * dN2[j][a2] * N3[k][a3];
              dNdu[3 * a + 2] = N1[i][a1] * N2[j][a2] * dN3[k][a3];
            }
        double detJxW = orc_jacobian(nen, Xnod, dNdu, Jinv, dNdX);
        if (detJxW == 0.0) continue; /* skip singular points (ASMs3D.C:2286) */
        double X[3];
        for (int d = 0; d < 3; d++) {
          double s = 0.0;
          for (int b = 0; b < nen; b++) s += Xnod[d * nen + b] * N[b];
          X[d] = s; /* X = Xnod * N (ASMs3D.C:2303) */
        }
        detJxW *= dV * wg[0][i] * wg[1][j] * wg[2][k];

        if (ncomp == 1) {
          double kap = itg->kappa.type == ORC_FN_ZERO ? 1.0 : orc_eval_function(&itg->kappa, X);
          double alpha = kap * detJxW;
          for (int a2 = 0; a2 < nen; a2++)
            for (int b = 0; b < nen; b++) {
              double s = dNdX[3 * a2] * dNdX[3 * b] + dNdX[3 * a2 + 1] * dNdX[3 * b + 1]
                       + dNdX[3 * a2 + 2] * dNdX[3 * b + 2];
              EK[a2 * nedof + b] += alpha * s;
            }
          double fw = orc_eval_function(&itg->source, X) * detJxW;
          for (int a2 = 0; a2 < nen; a2++) EV[a2] += fw * N[a2];
        } else {
          for (int a2 = 0; a2 < nen; a2++) {
            const double* ga = dNdX + 3 * a2;
            for (int b = 0; b < nen; b++) {
              const double* gb = dNdX + 3 * b;
              double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
              for (int ii = 0; ii < 3; ii++)
                for (int jj = 0; jj < 3; jj++) {
                  double v = lam * ga[ii] * gb[jj] + mu * ga[jj] * gb[ii];
                  if (ii == jj) v += mu * dot;
                  EK[(3 * a2 + ii) * nedof + 3 * b + jj] += detJxW * v;
                }
            }
          }
          for (int a2 = 0; a2 < nen; a2++)
            for (int ii = 0; ii < 3; ii++)
              EV[3 * a2 + ii] += detJxW * itg->body[ii] * N[a2];
        }
      }
  return 0;
}

/* ---- a17: SAM::initSystemEquations (src/LinAlg/SAM.C:225-336) ----------- */
int orc_sam_meqn(int ndof, const int* msc, int* meqn)
{
  int neq = 0;
  for (int i = 0; i < ndof; i++) meqn[i] = msc[i] == 1 ? ++neq : 0;
  return neq;
}

/* Status codes for homogeneous Dirichlet faces: ASMbase::fix with
 * fixHomogeneousDirichlet (src/ASM/ASMbase.C:886-887) gives msc = 0,
 * SAMpatch::initNodeDofs (src/ASM/SAMpatch.C:84-161). */
void orc_face_msc(const orc_patch* P, int ncomp, const int fixmask[6], int* msc)
{
  const int n1 = P->n[0], n2 = P->n[1], n3 = P->n[2];
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        int node = i + n1 * (j + n2 * k);
        int m = 0;
        if (i == 0) m |= fixmask[0];
        if (i == n1 - 1) m |= fixmask[1];
        if (j == 0) m |= fixmask[2];
        if (j == n2 - 1) m |= fixmask[3];
        if (k == 0) m |= fixmask[4];
        if (k == n3 - 1) m |= fixmask[5];
        for (int c = 0; c < ncomp; c++) msc[node * ncomp + c] = (m >> c) & 1 ? 0 : 1;
      }
}

/* ---- a16: SAM::getDofCouplings + SparseMatrix::optimiseCols -------------
 * (src/LinAlg/SAM.C:377-427; src/LinAlg/SparseMatrix.C:984-1023).  The
 * coupling set of equation i is every free equation of every element that
 * holds equation i; it is built row by row from the elements of the row's
 * node, and stored sorted (the CSC of a symmetric pattern equals its CSR). */
static int cmp_int(const void* a, const void* b)
{
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

static int row_couplings(const orc_patch* P, int ncomp, const int* meqn,
                         int node, int* buf)
{
  const int n1 = P->n[0], n2 = P->n[1];
  int nel[3];
  orc_num_elements(P, nel);
  const int i = node % n1, j = (node / n1) % n2, k = node / (n1 * n2);
  int cnt = 0;
  int mnpc[ORC_MAXNEN];
  const int nen = (P->p[0] + 1) * (P->p[1] + 1) * (P->p[2] + 1);
  for (int e3 = k - P->p[2]; e3 <= k; e3++) {
    if (e3 < 0 || e3 >= nel[2]) continue;
    for (int e2 = j - P->p[1]; e2 <= j; e2++) {
      if (e2 < 0 || e2 >= nel[1]) continue;
      for (int e1 = i - P->p[0]; e1 <= i; e1++) {
        if (e1 < 0 || e1 >= nel[0]) continue;
        orc_element_nodes(P, e1, e2, e3, mnpc);
        for (int a = 0; a < nen; a++)
          for (int c = 0; c < ncomp; c++) {
            int eq = meqn[mnpc[a] * ncomp + c];
            if (eq > 0) buf[cnt++] = eq - 1;
          }
      }
    }
  }
  qsort(buf, cnt, sizeof(int), cmp_int);
  int u = 0;
  for (int q = 0; q < cnt; q++)
    if (u == 0 || buf[q] != buf[u - 1]) buf[u++] = buf[q];
  return u;
}

int64_t orc_pattern(const orc_patch* P, int ncomp, const int* meqn, int neq,
                    int64_t* rowptr, int32_t* colind)
{
  int nel[3];
  if (orc_num_elements(P, nel)) return -1;
  const int nnod = P->n[0] * P->n[1] * P->n[2];
  const int nen = (P->p[0] + 1) * (P->p[1] + 1) * (P->p[2] + 1);
  const int bufsz = nen * nen * ncomp;
  int* rowlen = (int*)calloc(neq > 0 ? neq : 1, sizeof(int));
  int bad = 0;
#pragma omp parallel
  {
    int* buf = (int*)malloc(sizeof(int) * bufsz);
#pragma omp for schedule(static)
    for (int node = 0; node < nnod; node++) {
      int len = -1;
      for (int c = 0; c < ncomp; c++) {
        int eq = meqn[node * ncomp + c];
        if (eq <= 0) continue;
        if (len < 0) len = row_couplings(P, ncomp, meqn, node, buf);
        if (eq > neq) bad = 1; else rowlen[eq - 1] = len;
      }
    }
    free(buf);
  }
  if (bad) { free(rowlen); return -1; }
  rowptr[0] = 0;
  for (int r = 0; r < neq; r++) rowptr[r + 1] = rowptr[r] + rowlen[r];
  if (colind) {
#pragma omp parallel
    {
      int* buf = (int*)malloc(sizeof(int) * bufsz);
#pragma omp for schedule(static)
      for (int node = 0; node < nnod; node++) {
        int len = -1;
        for (int c = 0; c < ncomp; c++) {
          int eq = meqn[node * ncomp + c];
          if (eq <= 0) continue;
          if (len < 0) len = row_couplings(P, ncomp, meqn, node, buf);
          memcpy(colind + rowptr[eq - 1], buf, sizeof(int) * len);
        }
      }
      free(buf);
    }
  }
  int64_t nnz = rowptr[neq];
  free(rowlen);
  return nnz;
}

/* ---- a14: assemSparse, free-free part (src/LinAlg/SparseMatrix.C:614-637)
 * with SparseMatrix::operator() lookup (std::find in the column,
 * src/LinAlg/SparseMatrix.C:283-289); a15: SAM::assembleRHS
 * (src/LinAlg/SAM.C:605-617).  Homogeneous Dirichlet DOFs have meqn 0 and
 * are skipped (src/LinAlg/SAM.C:246-248). */
static inline double* sm_entry(const int64_t* rowptr, const int32_t* colind,
                               double* A, int r, int c)
{
  for (int64_t q = rowptr[r]; q < rowptr[r + 1]; q++)
    if (colind[q] == c) return A + q;
  return NULL;
}

/* With prescribed values (constraint equations without master DOFs,
 * meqn < 0, value c0 = ttcc of the constraint's first entry): the constrained
 * columns move to the right-hand side, b_i -= c0_j * EK(i, j) for free rows i
 * (src/LinAlg/SparseMatrix.C:639-661).  gval[j] holds c0 of element DOF j
 * (0 for fixed or free DOFs); NULL means none. */
static int assem_element(int nedof, const double* EK, const double* EV,
                         const int* meen, const int64_t* rowptr,
                         const int32_t* colind, double* A, double* b, const double* gval)
{
  if (gval)
    for (int j = 0; j < nedof; j++) {
      if (meen[j] >= 1 || gval[j] == 0.0) continue;
      for (int i = 0; i < nedof; i++)
        if (meen[i] > 0) b[meen[i] - 1] -= gval[j] * EK[i * nedof + j];
    }
  for (int j = 0; j < nedof; j++) {
    int jeq = meen[j];
    if (jeq < 1) continue;
    double* d = sm_entry(rowptr, colind, A, jeq - 1, jeq - 1);
    if (!d) return -1;
    *d += EK[j * nedof + j];
    for (int i = 0; i < j; i++) {
      int ieq = meen[i];
      if (ieq < 1) continue;
      double* u = sm_entry(rowptr, colind, A, ieq - 1, jeq - 1);
      double* l = sm_entry(rowptr, colind, A, jeq - 1, ieq - 1);
      if (!u || !l) return -1;
      *u += EK[i * nedof + j];
      *l += EK[j * nedof + i];
    }
  }
  for (int i = 0; i < nedof; i++)
    if (meen[i] > 0) b[meen[i] - 1] += EV[i];
  return 0;
}

/* ---- a11/a12: ASMs3D::integrate element loop with ThreadGroups stripes ---
 * (src/ASM/ASMs3D.C:2110-2130; src/Utility/ThreadGroups.C:253-351,444-479):
 * the strip direction leaves the fewest elements over when cut into
 * 2*threads stripes; threads are reduced until stripes are >= p wide; thread
 * t owns stripe 2t (group 0) and 2t+1 (group 1), so stripes of one group share
 * no node; OpenMP runs the stripes of a group in parallel, groups in turn. */
static int strip_direction(long n1, long n2, long n3, long parts)
{
  long r1 = n1 - (n1 / parts) * parts, r2 = n2 - (n2 / parts) * parts, r3 = n3 - (n3 / parts) * parts;
  if (r1 * n2 * n3 < n1 * r2 * n3 && r1 * n2 * n3 < n1 * n2 * r3) return 0;
  if (n1 * r2 * n3 < r1 * n2 * n3 && n1 * r2 * n3 < n1 * n2 * r3) return 1;
  if (n1 * n2 * r3 < r1 * n2 * n3 && n1 * n2 * r3 < n1 * r2 * n3) return 2;
  if (r1 * n2 * n3 > n1 * r2 * n3) return n2 > n3 ? 1 : 2;
  if (n1 * r2 * n3 > n1 * n2 * r3) return n1 > n3 ? 0 : 2;
  if (n1 * n2 * r3 > r1 * n2 * n3) return n1 > n2 ? 0 : 1;
  if (n1 >= n2 && n1 >= n3) return 0;
  if (n2 >= n1 && n2 >= n3) return 1;
  return 2;
}

int orc_assemble_range(const orc_patch* P, const orc_integrand* itg,
                       const int ng[3], int ncomp, const int* meqn, int neq,
                       const int64_t* rowptr, const int32_t* colind,
                       double* A, double* b, int nthreads,
                       int e3_begin, int e3_end)
{
  return orc_assemble_range_g(P, itg, ng, ncomp, meqn, neq, rowptr, colind, A, b, nthreads,
                              e3_begin, e3_end, NULL);
}

int orc_assemble_range_g(const orc_patch* P, const orc_integrand* itg,
                         const int ng[3], int ncomp, const int* meqn, int neq,
                         const int64_t* rowptr, const int32_t* colind,
                         double* A, double* b, int nthreads,
                         int e3_begin, int e3_end, const double* gdof)
{
  (void)neq;
  int nel[3];
  if (orc_num_elements(P, nel)) return -1;
  if (e3_begin < 0) e3_begin = 0;
  if (e3_end > nel[2]) e3_end = nel[2];
  const int nz = e3_end - e3_begin;
  if (nz <= 0) return 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const int nen = (P->p[0] + 1) * (P->p[1] + 1) * (P->p[2] + 1);
  const int nedof = ncomp * nen;
  const int ext[3] = {nel[0], nel[1], nz};
  int threads = nthreads;
  int parts = threads > 1 ? 2 * threads : 1;
  const int dir = strip_direction(ext[0], ext[1], ext[2], parts);
  const int els = ext[dir];
  int stripsize = 0;
  while (threads > 1 && (stripsize = els / parts) < P->p[dir]) {
    threads--;
    parts -= 2;
  }
  /* stripe bounds along dir: [lo, hi) for group g of thread t */
  int* lo = (int*)malloc(sizeof(int) * 2 * (threads > 0 ? threads : 1));
  int* hi = (int*)malloc(sizeof(int) * 2 * (threads > 0 ? threads : 1));
  if (threads <= 1) {
    threads = 1;
    lo[0] = 0; hi[0] = els;
    lo[1] = 0; hi[1] = 0;
  } else {
    int remainder = els - stripsize * parts;
    int* sz = (int*)malloc(sizeof(int) * 2 * threads);
    for (int t = 0; t < 2 * threads; t++) sz[t] = stripsize;
    for (int i = 1; i <= remainder; i++) sz[(i % 2) * threads + threads - (i + 1) / 2]++;
    int offs = 0;
    for (int t = 0; t < threads; t++)
      for (int g = 0; g < 2; g++) {
        lo[g * threads + t] = offs;
        offs += sz[g * threads + t];
        hi[g * threads + t] = offs;
      }
    free(sz);
  }
  int err = 0;
  for (int g = 0; g < (threads > 1 ? 2 : 1); g++) {
#pragma omp parallel num_threads(threads)
    {
      double* EK = (double*)malloc(sizeof(double) * nedof * nedof);
      double* EV = (double*)malloc(sizeof(double) * nedof);
      int* meen = (int*)malloc(sizeof(int) * nedof);
      double* gval = gdof ? (double*)malloc(sizeof(double) * nedof) : NULL;
      int mnpc[ORC_MAXNEN];
#pragma omp for schedule(static)
      for (int t = 0; t < threads; t++) {
        int bl[3] = {0, 0, 0}, bh[3] = {ext[0], ext[1], ext[2]};
        bl[dir] = lo[g * threads + t];
        bh[dir] = hi[g * threads + t];
        for (int i3 = bl[2]; i3 < bh[2]; i3++)
          for (int e2 = bl[1]; e2 < bh[1]; e2++)
            for (int e1 = bl[0]; e1 < bh[0]; e1++) {
              const int e3 = e3_begin + i3;
              if (orc_element_matrices(P, itg, ng, e1, e2, e3, EK, EV)) { err = 1; continue; }
              orc_element_nodes(P, e1, e2, e3, mnpc);
              for (int a = 0; a < nen; a++)
                for (int c = 0; c < ncomp; c++) {
                  meen[a * ncomp + c] = meqn[mnpc[a] * ncomp + c];
                  if (gval) gval[a * ncomp + c] = gdof[mnpc[a] * ncomp + c];
                }
              if (assem_element(nedof, EK, EV, meen, rowptr, colind, A, b, gval)) err = 1;
            }
      }
      free(EK);
      free(EV);
      free(meen);
      free(gval);
    }
  }
  free(lo);
  free(hi);
  return err ? -1 : 0;
}

int orc_assemble(const orc_patch* P, const orc_integrand* itg, const int ng[3],
                 int ncomp, const int* meqn, int neq, const int64_t* rowptr,
                 const int32_t* colind, double* A, double* b, int nthreads)
{
  memset(A, 0, sizeof(double) * rowptr[neq]);
  memset(b, 0, sizeof(double) * neq);
  return orc_assemble_range(P, itg, ng, ncomp, meqn, neq, rowptr, colind, A, b,
                            nthreads, 0, 1 << 30);
}

/* ---- a21: ASMs3D::integrate over a face (src/ASM/ASMs3D.C:2547-2812) with
 * the face utl::Jacobian (src/Utility/CoordinateMapping.C:90-133): surface
 * dilation dS = |J_t1 x J_t2|, detJxW = dS*dA*w_i*w_j with dA = 0.25*area;
 * traction integrand EV(a i) += detJxW*t_i*N_a. */
int orc_assemble_traction(const orc_patch* P, int face, const double t[3],
                          int ng, int ncomp, const int* meqn, double* b)
{
  int nel[3];
  if (orc_num_elements(P, nel) || face < 0 || face > 5) return -1;
  const int dn = face / 2, side = face % 2;
  const int faceDir = side ? dn + 1 : -(dn + 1);
  const int t1 = 1 + abs(faceDir) % 3, t2 = 1 + t1 % 3; /* 1-based */
  const int tt1 = (t1 > t2 ? t2 : t1) - 1, tt2 = (t1 > t2 ? t1 : t2) - 1;
  const int nen = (P->p[0] + 1) * (P->p[1] + 1) * (P->p[2] + 1);
  double xg[ORC_MAXNG], wg[ORC_MAXNG];
  if (orc_gauss(ng, xg, wg)) return -1;
  int mnpc[ORC_MAXNEN];
  double Xnod[3 * ORC_MAXNEN], N[ORC_MAXNEN], dNdu[3 * ORC_MAXNEN];
  int lo[3] = {0, 0, 0}, hi[3] = {nel[0], nel[1], nel[2]};
  lo[dn] = side ? nel[dn] - 1 : 0;
  hi[dn] = lo[dn] + 1;
  for (int e3 = lo[2]; e3 < hi[2]; e3++)
    for (int e2 = lo[1]; e2 < hi[1]; e2++)
      for (int e1 = lo[0]; e1 < hi[0]; e1++) {
        const int e[3] = {e1, e2, e3};
        orc_element_nodes(P, e1, e2, e3, mnpc);
        for (int a = 0; a < nen; a++)
          for (int i = 0; i < 3; i++) Xnod[i * nen + a] = P->coefs[3 * mnpc[a] + i];
        double span[3];
        for (int d = 0; d < 3; d++) {
          const int s = e[d] + P->p[d];
          span[d] = P->knots[d][s + 1] - P->knots[d][s];
        }
        const double dA = 0.25 * (span[tt1] * span[tt2]);
        double EV[3 * ORC_MAXNEN];
        memset(EV, 0, sizeof EV);
        for (int jg = 0; jg < ng; jg++)
          for (int ig = 0; ig < ng; ig++) {
            double Nd[3][ORC_MAXP + 1], dNd[3][ORC_MAXP + 1];
            int gidx[3];
            gidx[tt1] = ig;
            gidx[tt2] = jg;
            for (int d = 0; d < 3; d++) {
              const double* kt = P->knots[d];
              const int s = e[d] + P->p[d];
              double u = d == dn ? (side ? kt[P->n[d]] : kt[P->p[d]])
                                 : orc_gauss_param(kt[s], kt[s + 1], xg[gidx[d]]);
              orc_basis_funs(P->p[d], kt, s, u, Nd[d], dNd[d]);
            }
            int a = 0;
            for (int a3 = 0; a3 <= P->p[2]; a3++)
              for (int a2 = 0; a2 <= P->p[1]; a2++)
                for (int a1 = 0; a1 <= P->p[0]; a1++, a++) {
                  N[a] = Nd[0][a1] * Nd[1][a2] * Nd[2][a3];
                  dNdu[3 * a + 0] = dNd[0][a1] * Nd[1][a2] * Nd[2][a3];
                  dNdu[3 * a + 1] = Nd[0][a1] * dNd[1][a2] * Nd[2][a3];
                  dNdu[3 * a + 2] = Nd[0][a1] * Nd[1][a2] * dNd[2][a3];
                }
            double J[9];
            for (int i = 0; i < 3; i++)
              for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int bb = 0; bb < nen; bb++) s += Xnod[i * nen + bb] * dNdu[bb * 3 + k];
                J[i * 3 + k] = s;
              }
            const int c1 = t1 - 1, c2 = t2 - 1;
            double v1[3] = {J[0 * 3 + c1], J[1 * 3 + c1], J[2 * 3 + c1]};
            double v2[3] = {J[0 * 3 + c2], J[1 * 3 + c2], J[2 * 3 + c2]};
            double nrm[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                             v1[0] * v2[1] - v1[1] * v2[0]};
            double dS = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            if (inverse3(J) == 0.0) dS = 0.0;
            if (dS == 0.0) continue;
            double detJxW = dS * (dA * wg[ig] * wg[jg]);
            for (int a2 = 0; a2 < nen; a2++)
              for (int c = 0; c < ncomp; c++) EV[a2 * ncomp + c] += detJxW * t[c] * N[a2];
          }
        for (int a2 = 0; a2 < nen; a2++)
          for (int c = 0; c < ncomp; c++) {
            int eq = meqn[mnpc[a2] * ncomp + c];
            if (eq > 0) b[eq - 1] += EV[a2 * ncomp + c];
          }
      }
  return 0;
}

/* ---- a19: SpMV (SparseMatrix::multiply, src/LinAlg/SparseMatrix.C:578-605,
 * row-oriented branch) ---------------------------------------------------- */
void orc_spmv(int n, const int64_t* rowptr, const int32_t* colind,
              const double* A, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int64_t q = rowptr[i]; q < rowptr[i + 1]; q++) s += A[q] * x[colind[q]];
    y[i] = s;
  }
}

static double dot(int n, const double* a, const double* b)
{
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static void jacobi_diag(int n, const int64_t* rowptr, const int32_t* colind,
                        const double* A, double* dinv)
{
  for (int i = 0; i < n; i++) {
    double d = 0.0;
    for (int64_t q = rowptr[i]; q < rowptr[i + 1]; q++)
      if (colind[q] == i) d = A[q];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
}

/* ---- a18: CG + Jacobi (PETSc KSPCG/PCJACOBI, dune CGSolver/SeqJac, selected
 * at src/LinAlg/PETScMatrix.C:980-990, src/LinAlg/ISTLSolParams.C:228-229),
 * convergence contract of LinSolParams (rtol, atol, maxits;
 * src/LinAlg/LinSolParams.C:70-76), x = 0 start (src/LinAlg/ISTLMatrix.C:190).
 * The recursive residual drives the loop; the true residual b - A x decides. */
int orc_pcg_jacobi(int n, const int64_t* rowptr, const int32_t* colind,
                   const double* A, const double* b, double* x,
                   double rtol, double atol, int maxits, double* resid_out)
{
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double* dinv = (double*)malloc(sizeof(double) * n);
  jacobi_diag(n, rowptr, colind, A, dinv);
  const double bnorm = sqrt(dot(n, b, b));
  const double tol = fmax(rtol * bnorm, atol);
  memset(x, 0, sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  int it = 0, done = 0;
  double rel = bnorm > 0 ? 1.0 : 0.0;
  while (!done) {
    for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    double rz = dot(n, r, z);
    double rnorm = sqrt(dot(n, r, r));
    while (rnorm > tol && it < maxits) {
      orc_spmv(n, rowptr, colind, A, p, q);
      double alpha = rz / dot(n, p, q);
      for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
      it++;
      rnorm = sqrt(dot(n, r, r));
      if (rnorm <= tol) break;
      for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
      double rzn = dot(n, r, z);
      double beta = rzn / rz;
      rz = rzn;
      for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    /* true residual */
    orc_spmv(n, rowptr, colind, A, x, q);
    for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
    double tn = sqrt(dot(n, r, r));
    rel = bnorm > 0 ? tn / bnorm : tn;
    if (tn <= tol || it >= maxits) done = 1;
  }
  if (resid_out) *resid_out = rel;
  free(r); free(z); free(p); free(q); free(dinv);
  return rel <= fmax(rtol, atol / (bnorm > 0 ? bnorm : 1.0)) ? it : -it;
}

void orc_pcg_jacobi_fixed(int n, const int64_t* rowptr, const int32_t* colind,
                          const double* A, const double* b, double* x, int iters)
{
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double* dinv = (double*)malloc(sizeof(double) * n);
  jacobi_diag(n, rowptr, colind, A, dinv);
  memset(x, 0, sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
  double rz = dot(n, r, z);
  for (int it = 0; it < iters; it++) {
    orc_spmv(n, rowptr, colind, A, p, q);
    double alpha = rz / dot(n, p, q);
    for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    double rzn = dot(n, r, z);
    double beta = rzn / rz;
    rz = rzn;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  free(r); free(z); free(p); free(q); free(dinv);
}

/* ---- a20: SAM::expandSolution/expandVector (src/LinAlg/SAM.C:768-810) --- */
void orc_expand_solution(int ndof, const int* meqn, const double* xeq, double* xdof)
{
  for (int i = 0; i < ndof; i++) xdof[i] = meqn[i] > 0 ? xeq[meqn[i] - 1] : 0.0;
}

/* With prescribed values: constrained DOFs take c0 (SAM::expandSolution,
 * src/LinAlg/SAM.C:768-810, for constraints without master DOFs). */
void orc_expand_solution_g(int ndof, const int* meqn, const double* xeq, const double* gdof,
                           double* xdof)
{
  for (int i = 0; i < ndof; i++) xdof[i] = meqn[i] > 0 ? xeq[meqn[i] - 1] : gdof[i];
}

/* Gradient of a field (analytic derivative of orc_eval_function). */
void orc_eval_function_grad(const orc_function* f, const double X[3], double g[3])
{
  const double pi = 3.14159265358979323846;
  g[0] = g[1] = g[2] = 0.0;
  switch (f->type) {
  case ORC_FN_SINPROD: {
    const double s0 = sin(pi * X[0]), s1 = sin(pi * X[1]), s2 = sin(pi * X[2]);
    const double c0 = cos(pi * X[0]), c1 = cos(pi * X[1]), c2 = cos(pi * X[2]);
    g[0] = f->scale * pi * c0 * s1 * s2;
    g[1] = f->scale * pi * s0 * c1 * s2;
    g[2] = f->scale * pi * s0 * s1 * c2;
    return;
  }
  case ORC_FN_FOURIER:
  case ORC_FN_EXP_FOURIER: {
    double s = 0.0;
    for (int m = 0; m < f->nmodes; m++) {
      const double* md = f->modes + 5 * m;
      const double arg = md[0] * X[0] + md[1] * X[1] + md[2] * X[2] + md[3];
      s += md[4] * cos(arg);
      for (int d = 0; d < 3; d++) g[d] -= f->scale * md[4] * md[d] * sin(arg);
    }
    if (f->type == ORC_FN_EXP_FOURIER) {
      const double e = exp(f->mean + f->scale * s);
      for (int d = 0; d < 3; d++) g[d] *= e;
    }
    return;
  }
  default:
    return;
  }
}

/* ---- norms: NormBase/ElmNorm/GlbNorm (src/ASM/IntegrandBase.h:378-530,
 * src/ASM/ElmNorm.h:28, src/ASM/GlbNorm.C:33-79 with the SQRT final
 * operation).  Element sums over the Gauss points of the volume loop
 * (same quadrature, Jacobian and singular-point rule as the assembly),
 * summed over the elements. ---------------------------------------------- */
int orc_norms(const orc_patch* P, const orc_integrand* itg, const int ng[3],
              const double* xdof, const orc_function* exact, double* out)
{
  const int p1 = P->p[0], p2 = P->p[1], p3 = P->p[2];
  const int nen = (p1 + 1) * (p2 + 1) * (p3 + 1);
  const int ncomp = itg->kind == ORC_ELASTICITY ? 3 : 1;
  int nel[3];
  orc_num_elements(P, nel);
  double lam = 0.0, mu = 0.0;
  if (ncomp == 3) {
    lam = itg->E * itg->nu / ((1.0 + itg->nu) * (1.0 - 2.0 * itg->nu));
    mu = itg->E / (2.0 * (1.0 + itg->nu));
  }
  double xg[3][ORC_MAXNG], wg[3][ORC_MAXNG];
  for (int d = 0; d < 3; d++)
    if (orc_gauss(ng[d], xg[d], wg[d])) return -1;
  double sum[6] = {0, 0, 0, 0, 0, 0};
  int mnpc[ORC_MAXNEN];
  double Xnod[3 * ORC_MAXNEN], N[ORC_MAXNEN], dNdu[3 * ORC_MAXNEN], dNdX[3 * ORC_MAXNEN], Jinv[9];
  for (int e3 = 0; e3 < nel[2]; e3++)
    for (int e2 = 0; e2 < nel[1]; e2++)
      for (int e1 = 0; e1 < nel[0]; e1++) {
        const int e[3] = {e1, e2, e3};
        orc_element_nodes(P, e1, e2, e3, mnpc);
        for (int a = 0; a < nen; a++)
          for (int i = 0; i < 3; i++) Xnod[i * nen + a] = P->coefs[3 * mnpc[a] + i];
        double span[3];
        for (int d = 0; d < 3; d++) {
          const int s = e[d] + P->p[d];
          span[d] = P->knots[d][s + 1] - P->knots[d][s];
        }
        const double dV = 0.125 * (span[0] * span[1] * span[2]);
        double N1[ORC_MAXNG][ORC_MAXP + 1], dN1[ORC_MAXNG][ORC_MAXP + 1];
        double N2[ORC_MAXNG][ORC_MAXP + 1], dN2[ORC_MAXNG][ORC_MAXP + 1];
        double N3[ORC_MAXNG][ORC_MAXP + 1], dN3[ORC_MAXNG][ORC_MAXP + 1];
        for (int d = 0; d < 3; d++) {
          const int s = e[d] + P->p[d];
          const double* t = P->knots[d];
          for (int g = 0; g < ng[d]; g++) {
            const double u = orc_gauss_param(t[s], t[s + 1], xg[d][g]);
            double* Nd = d == 0 ? N1[g] : (d == 1 ? N2[g] : N3[g]);
            double* dNd = d == 0 ? dN1[g] : (d == 1 ? dN2[g] : dN3[g]);
            orc_basis_funs(P->p[d], t, s, u, Nd, dNd);
          }
        }
        double esum[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < ng[2]; k++)
          for (int j = 0; j < ng[1]; j++)
            for (int i = 0; i < ng[0]; i++) {
              int a = 0;
              for (int a3 = 0; a3 <= p3; a3++)
                for (int a2 = 0; a2 <= p2; a2++)
                  for (int a1 = 0; a1 <= p1; a1++, a++) {
                    N[a] = N1[i][a1] * N2[j][a2] * N3[k][a3];
                    dNdu[3 * a + 0] = dN1[i][a1] * N2[j][a2] * N3[k][a3];
                    dNdu[3 * a + 1] = N1[i][a1] * dN2[j][a2] * N3[k][a3];
                    dNdu[3 * a + 2] = N1[i][a1] * N2[j][a2] * dN3[k][a3];
                  }
              double w = orc_jacobian(nen, Xnod, dNdu, Jinv, dNdX);
              if (w == 0.0) continue;
              w *= dV * wg[0][i] * wg[1][j] * wg[2][k];
              double X[3], uh[3] = {0, 0, 0}, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
              for (int d = 0; d < 3; d++) {
                double s = 0.0;
                for (int b = 0; b < nen; b++) s += Xnod[d * nen + b] * N[b];
                X[d] = s;
              }
              for (int b = 0; b < nen; b++)
                for (int c = 0; c < ncomp; c++) {
                  const double xb = xdof[ncomp * mnpc[b] + c];
                  uh[c] += N[b] * xb;
                  for (int d = 0; d < 3; d++) G[c][d] += dNdX[3 * b + d] * xb;
                }
              if (ncomp == 1) {
                const double kap =
                    itg->kappa.type == ORC_FN_ZERO ? 1.0 : orc_eval_function(&itg->kappa, X);
                esum[0] += w * uh[0] * uh[0];
                esum[1] += w * kap * (G[0][0] * G[0][0] + G[0][1] * G[0][1] + G[0][2] * G[0][2]);
                if (exact) {
                  double ge[3];
                  const double ue = orc_eval_function(exact, X);
                  orc_eval_function_grad(exact, X, ge);
                  const double d0 = ge[0] - G[0][0], d1 = ge[1] - G[0][1], d2 = ge[2] - G[0][2];
                  esum[2] += w * ue * ue;
                  esum[3] += w * kap * (ge[0] * ge[0] + ge[1] * ge[1] + ge[2] * ge[2]);
                  esum[4] += w * (ue - uh[0]) * (ue - uh[0]);
                  esum[5] += w * kap * (d0 * d0 + d1 * d1 + d2 * d2);
                }
              } else {
                double eps[3][3], tr = 0.0, ee = 0.0;
                for (int r = 0; r < 3; r++)
                  for (int c = 0; c < 3; c++) eps[r][c] = 0.5 * (G[r][c] + G[c][r]);
                for (int r = 0; r < 3; r++) tr += eps[r][r];
                for (int r = 0; r < 3; r++)
                  for (int c = 0; c < 3; c++) ee += eps[r][c] * eps[r][c];
                esum[0] += w * (uh[0] * uh[0] + uh[1] * uh[1] + uh[2] * uh[2]);
                esum[1] += w * (lam * tr * tr + 2.0 * mu * ee);
              }
            }
        for (int m = 0; m < 6; m++) sum[m] += esum[m];
      }
  for (int m = 0; m < 6; m++) out[m] = sum[m] < 0.0 ? -sqrt(-sum[m]) : sqrt(sum[m]);
  return 0;
}

/* ---- constraint equations with master DOFs ---------------------------------
 * SAM's layout (src/LinAlg/SAM.h:330-350): meqn < 0 marks the slave of
 * constraint k = -meqn; its entries are the 1-based positions
 * mpmceq[k-1] .. mpmceq[k]-1 of mmceq/ttcc, the first the slave itself with
 * the constant c0, the others its masters (free DOFs) and coefficients. */

/* Element equation numbers meen from the DOF numbers (SAM::getElmEqns,
 * src/LinAlg/SAM.C:655-696: meqn of each element DOF). */
static void elm_eqns(const orc_patch* P, const orc_sam* s, int ncomp, int e1, int e2, int e3,
                     int* meen, int* nedof)
{
  int mnpc[ORC_MAXNEN];
  orc_element_nodes(P, e1, e2, e3, mnpc);
  const int nen = (P->p[0] + 1) * (P->p[1] + 1) * (P->p[2] + 1);
  for (int a = 0; a < nen; a++)
    for (int c = 0; c < ncomp; c++) meen[a * ncomp + c] = s->meqn[mnpc[a] * ncomp + c];
  *nedof = nen * ncomp;
}

typedef struct { int* v; int n, cap; } ivec;
static void iv_push(ivec* a, int x)
{
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 64;
    a->v = (int*)realloc(a->v, sizeof(int) * a->cap);
  }
  a->v[a->n++] = x;
}

/* SAM::getDofCouplings with constraint equations (src/LinAlg/SAM.C:377-427),
 * 0-based columns, sorted and unique per row. */
int64_t orc_pattern_sam(const orc_patch* P, int ncomp, const orc_sam* s, int neq,
                        int64_t* rowptr, int32_t* colind)
{
  int nel[3];
  if (orc_num_elements(P, nel)) return -1;
  ivec* rows = (ivec*)calloc(neq > 0 ? neq : 1, sizeof(ivec));
  int meen[3 * ORC_MAXNEN], nedof = 0;
  for (int e3 = 0; e3 < nel[2]; e3++)
    for (int e2 = 0; e2 < nel[1]; e2++)
      for (int e1 = 0; e1 < nel[0]; e1++) {
        elm_eqns(P, s, ncomp, e1, e2, e3, meen, &nedof);
        for (int j = 0; j < nedof; j++) {
          const int jeq = meen[j];
          if (jeq > 0) {
            iv_push(&rows[jeq - 1], jeq);
            for (int i = 0; i < j; i++)
              if (meen[i] > 0) {
                iv_push(&rows[meen[i] - 1], jeq);
                iv_push(&rows[jeq - 1], meen[i]);
              }
          } else if (jeq < 0) {
            const int jc = -jeq;
            for (int jp = s->mpmceq[jc - 1]; jp < s->mpmceq[jc] - 1; jp++) {
              if (s->mmceq[jp] <= 0) continue;
              const int mj = s->meqn[s->mmceq[jp] - 1];
              for (int i = 0; i < nedof; i++) {
                const int ieq = meen[i];
                if (ieq > 0) {
                  iv_push(&rows[ieq - 1], mj);
                  iv_push(&rows[mj - 1], ieq);
                } else if (ieq < 0) {
                  const int ic = -ieq;
                  for (int ip = s->mpmceq[ic - 1]; ip < s->mpmceq[ic] - 1; ip++)
                    if (s->mmceq[ip] > 0) iv_push(&rows[s->meqn[s->mmceq[ip] - 1] - 1], mj);
                }
              }
            }
          }
        }
      }
  int64_t nnz = 0;
  rowptr[0] = 0;
  for (int r = 0; r < neq; r++) {
    ivec* a = &rows[r];
    qsort(a->v, a->n, sizeof(int), cmp_int);
    int u = 0;
    for (int q = 0; q < a->n; q++)
      if (u == 0 || a->v[q] != a->v[u - 1]) a->v[u++] = a->v[q];
    if (colind)
      for (int q = 0; q < u; q++) colind[nnz + q] = a->v[q] - 1;
    nnz += u;
    rowptr[r + 1] = nnz;
    free(a->v);
  }
  free(rows);
  return nnz;
}

static int add_entry(const int64_t* rowptr, const int32_t* colind, double* A, int r, int c,
                     double v)
{
  double* d = sm_entry(rowptr, colind, A, r, c);
  if (!d) return -1;
  *d += v;
  return 0;
}

/* SparseMatrix assemSparse with constraint equations
 * (src/LinAlg/SparseMatrix.C:612-688) and SAM::assembleRHS for the element
 * vector (src/LinAlg/SAM.C:605-617), one element. */
static int assem_element_sam(int nedof, const double* EK, const double* EV, const int* meen,
                             const orc_sam* s, const int64_t* rowptr, const int32_t* colind,
                             double* A, double* b)
{
  int err = 0;
  /* free-free block */
  for (int j = 0; j < nedof; j++) {
    const int jeq = meen[j];
    if (jeq < 1) continue;
    err |= add_entry(rowptr, colind, A, jeq - 1, jeq - 1, EK[j * nedof + j]);
    for (int i = 0; i < j; i++) {
      const int ieq = meen[i];
      if (ieq < 1) continue;
      err |= add_entry(rowptr, colind, A, ieq - 1, jeq - 1, EK[i * nedof + j]);
      err |= add_entry(rowptr, colind, A, jeq - 1, ieq - 1, EK[j * nedof + i]);
    }
  }
  /* constrained (prescribed and dependent) columns */
  for (int j = 0; j < nedof; j++) {
    const int jc = -meen[j];
    if (jc < 1) continue;
    const int jp0 = s->mpmceq[jc - 1];
    const double c0 = s->ttcc[jp0 - 1];
    for (int i = 0; i < nedof; i++) {
      const int ieq = meen[i];
      if (ieq > 0) b[ieq - 1] -= c0 * EK[i * nedof + j];
      else if (ieq < 0) {
        const int ic = -ieq;
        for (int ip = s->mpmceq[ic - 1]; ip < s->mpmceq[ic] - 1; ip++)
          if (s->mmceq[ip] > 0)
            b[s->meqn[s->mmceq[ip] - 1] - 1] -= c0 * s->ttcc[ip] * EK[i * nedof + j];
      }
    }
    for (int jp = s->mpmceq[jc - 1]; jp < s->mpmceq[jc] - 1; jp++) {
      if (s->mmceq[jp] <= 0) continue;
      const int jeq = s->meqn[s->mmceq[jp] - 1];
      for (int i = 0; i < nedof; i++) {
        const int ieq = meen[i];
        if (ieq > 0) {
          err |= add_entry(rowptr, colind, A, ieq - 1, jeq - 1, s->ttcc[jp] * EK[i * nedof + j]);
          err |= add_entry(rowptr, colind, A, jeq - 1, ieq - 1, s->ttcc[jp] * EK[j * nedof + i]);
        } else if (ieq < 0) {
          const int ic = -ieq;
          for (int ip = s->mpmceq[ic - 1]; ip < s->mpmceq[ic] - 1; ip++)
            if (s->mmceq[ip] > 0)
              err |= add_entry(rowptr, colind, A, s->meqn[s->mmceq[ip] - 1] - 1, jeq - 1,
                               s->ttcc[ip] * s->ttcc[jp] * EK[i * nedof + j]);
        }
      }
    }
  }
  /* element vector */
  for (int i = 0; i < nedof; i++) {
    const int ieq = meen[i];
    if (ieq > 0) b[ieq - 1] += EV[i];
    else if (ieq < 0) {
      const int ic = -ieq;
      for (int ip = s->mpmceq[ic - 1]; ip < s->mpmceq[ic] - 1; ip++)
        if (s->mmceq[ip] > 0) b[s->meqn[s->mmceq[ip] - 1] - 1] += s->ttcc[ip] * EV[i];
    }
  }
  return err;
}

/* Element loop (serial) adding into A, b (not zeroed: several patches can
 * be added with their own local meqn). */
int orc_assemble_sam(const orc_patch* P, const orc_integrand* itg, const int ng[3], int ncomp,
                     const orc_sam* s, int neq, const int64_t* rowptr, const int32_t* colind,
                     double* A, double* b)
{
  (void)neq;
  int nel[3];
  if (orc_num_elements(P, nel)) return -1;
  const int nen = (P->p[0] + 1) * (P->p[1] + 1) * (P->p[2] + 1);
  const int nedof = nen * ncomp;
  double* EK = (double*)malloc(sizeof(double) * nedof * nedof);
  double* EV = (double*)malloc(sizeof(double) * nedof);
  int meen[3 * ORC_MAXNEN], nd = 0, err = 0;
  for (int e3 = 0; e3 < nel[2] && !err; e3++)
    for (int e2 = 0; e2 < nel[1] && !err; e2++)
      for (int e1 = 0; e1 < nel[0] && !err; e1++) {
        if (orc_element_matrices(P, itg, ng, e1, e2, e3, EK, EV)) { err = 1; break; }
        elm_eqns(P, s, ncomp, e1, e2, e3, meen, &nd);
        if (assem_element_sam(nedof, EK, EV, meen, s, rowptr, colind, A, b)) err = 1;
      }
  free(EK);
  free(EV);
  return err;
}

/* SAM::expandVector with scaleSD = 1 (src/LinAlg/SAM.C:783-807). */
void orc_expand_sam(const orc_sam* s, const double* xeq, double* xdof)
{
  for (int i = 0; i < s->ndof; i++) {
    xdof[i] = 0.0;
    const int ieq = s->meqn[i];
    if (ieq > 0) xdof[i] += xeq[ieq - 1];
    else if (ieq < 0) {
      const int ic = -ieq;
      int ip = s->mpmceq[ic - 1];
      xdof[i] += s->ttcc[ip - 1];
      for (; ip < s->mpmceq[ic] - 1; ip++)
        if (s->mmceq[ip] > 0) {
          const int me = s->meqn[s->mmceq[ip] - 1];
          if (me > 0) xdof[i] += s->ttcc[ip] * xeq[me - 1];
        }
    }
  }
}
