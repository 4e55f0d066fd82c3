// assemble.hip -- symbolic pattern, numeric volume assembly and Neumann
// traction kernels for one tensor-product B-spline patch on gfx950.
//
// Numeric assembly is row-owned ("owner computes"): one workgroup owns the
// line of control points (i = 0..n1-1, j, k), walks the element layers
// e1 = 0..nel1-1 and, for each layer, integrates the (p+1)^2 elements of
// the layer that touch the line.  Contributions go into an LDS window that
// holds p+1 complete rows in dense neighbour layout (2p+1)^3 x ncomp^2; after
// layer e1 the row of node e1 can get no more contributions and is written
// to the CSR arrays once, compacted over eliminated / absent neighbours.
// Every K entry is written exactly once, coalesced, without atomics, and
// the summation order is fixed, so results are bitwise reproducible.
//
// This replaces the reference's element loop with colour groups and the
// per-entry std::find scatter (src/ASM/ASMs3D.C:2071-2335,
// src/LinAlg/SparseMatrix.C:614-688); the per-GP arithmetic follows
// utl::Jacobian (src/Utility/CoordinateMapping.C:24-63) and the integrands
// of Integrand::evalInt.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "devutil.hpp"
#include "internal.hpp"

namespace ifem {

// ---------------------------------------------------------------------------
// Gauss-Legendre rule by Newton iteration on the Legendre polynomial in long
// double (same nodes/weights as the reference tables G1..G10,
// src/Utility/GaussQuadrature.C:17-60, to the last bit or one ulp).
static void gauss_rule(int n, double* xi, double* w)
{
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
    if (2 * i + 1 == n) x = 0.0L;
    long double p0 = 1.0L, p1 = x;
    for (int k = 1; k < n; k++) {
      long double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1.0L);
    xi[n - 1 - i] = (double)x;
    w[n - 1 - i] = (double)(2.0L / ((1.0L - x * x) * dp * dp));
  }
}

// ---------------------------------------------------------------------------
// 1D B-spline values and first derivatives on span s (Cox-de Boor,
// Piegl & Tiller A2.2/A2.3), the per-direction factor of GoTools'
// SplineVolume::computeBasisGrid (src/ASM/ASMs3D.C:4114-4143).
template <int P>
__device__ inline void bspline_1d(const double* __restrict__ t, int s, double u, double* N,
                                  double* dN)
{
  double left[P + 1], right[P + 1], ndu[P + 1][P + 1];
  ndu[0][0] = 1.0;
#pragma unroll
  for (int j = 1; j <= P; j++) {
    left[j] = u - t[s + 1 - j];
    right[j] = t[s + j] - u;
    double saved = 0.0;
#pragma unroll
    for (int r = 0; r < j; r++) {
      ndu[j][r] = right[r + 1] + left[j - r];
      double temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
#pragma unroll
  for (int r = 0; r <= P; r++) {
    N[r] = ndu[r][P];
    double d = 0.0;
    if (r >= 1) d += ndu[r - 1][P - 1] / ndu[P][r - 1];
    if (r <= P - 1) d -= ndu[r][P - 1] / ndu[P][r];
    dN[r] = P * d;
  }
}

struct GaussArgs {
  double xi[kMaxNG];
  double w[kMaxNG];
};

// Tables per direction: for element e and Gauss point g, N and dN/du of the
// p+1 nonzero functions, plus the span length.
template <int P>
__global__ void k_basis_tables(const double* __restrict__ t, int nel, int ng, GaussArgs ga,
                               double* __restrict__ tabN, double* __restrict__ tabD,
                               double* __restrict__ span)
{
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nel * ng) return;
  const int e = id / ng, g = id % ng;
  const int s = e + P;
  const double u0 = t[s], u1 = t[s + 1];
  const double u = 0.5 * ((u1 - u0) * ga.xi[g] + u1 + u0); // SplineUtils.C:217
  double N[P + 1], dN[P + 1];
  bspline_1d<P>(t, s, u, N, dN);
#pragma unroll
  for (int a = 0; a <= P; a++) {
    tabN[id * (P + 1) + a] = N[a];
    tabD[id * (P + 1) + a] = dN[a];
  }
  if (g == 0) span[e] = u1 - u0;
}

// ---------------------------------------------------------------------------
// cos(x) for the Fourier fields: Cody-Waite argument reduction modulo pi
// (W. J. Cody and W. Waite, Software Manual for the Elementary Functions,
// 1980: pi split into three doubles, each product subtracted by a fused
// multiply-add; the first difference is exact), cos x = (-1)^n cos r with
// |r| <= pi/2, and the Taylor series of cos to degree 20 (truncation below
// 2e-17 at |r| = pi/2).  Absolute error below 1.3 ulp(1) and within 2 ulp of
// the C library's cos where |cos x| > 1/2, for |x| <= 1e9 (2e7 samples per
// range, also next to multiples of pi/2 and pi); larger |x| goes to the
// library routine, kept out of line so that its registers do not count
// against the callers.  The 128-mode log-normal conductivity spends most of
// the geometry pass in this function.
__device__ __noinline__ double cos_library(double x)
{
  return cos(x);
}

__device__ __noinline__ double sin_library(double x)
{
  return sin(x);
}

// Reduction modulo pi: x = n pi + r, |r| <= pi/2; returns n, r^2 in s.
__device__ inline double reduce_pi(double x, double& s)
{
  const double n = rint(x * 0x1.45f306dc9c883p-2); // x / pi
  double r = fma(-n, 0x1.921fb54442d18p+1, x);     // pi, three parts
  r = fma(-n, 0x1.1a62633145c07p-53, r);
  r = fma(-n, -0x1.f1976b7ed8fbcp-109, r);
  s = r * r;
  return n;
}

// Taylor coefficients of cos in r^2, highest first: 1/20!, -1/18!, ..., 1.
__device__ constexpr double kCosTaylor[11] = {
    1.0 / 2432902008176640000.0, -1.0 / 6402373705728000.0, 1.0 / 20922789888000.0,
    -1.0 / 87178291200.0,        1.0 / 479001600.0,         -1.0 / 3628800.0,
    1.0 / 40320.0,               -1.0 / 720.0,              1.0 / 24.0,
    -0.5,                        1.0};

// |x| <= 2^30 (no check)
__device__ inline double cos_reduced(double x)
{
  double s;
  const double n = reduce_pi(x, s);
  double c = kCosTaylor[0];
#pragma unroll
  for (int k = 1; k < 11; k++) c = fma(c, s, kCosTaylor[k]);
  return ((int)n & 1) ? -c : c;
}

__device__ inline double cos_cw(double x)
{
  return fabs(x) <= 0x1p30 ? cos_reduced(x) : cos_library(x);
}

// Mode tables are read through the constant address space: where the table
// pointer is the same for the whole wave the compiler loads the modes with
// scalar loads (no vector registers, no per-lane memory instructions).
typedef const __attribute__((address_space(4))) double* ConstModes;

// sum_m a_m cos(k_m.x + phi_m), four modes at a time, summed in mode order.
// FAST: every argument is known to be within 2^30 (no library fallback, so
// no call in the loop); the four cosines run the operations of cos_reduced
// with their polynomial steps interleaved (four independent chains).
template <bool FAST>
__device__ inline double fourier_sum(const DevFunction& f, double x, double y, double z)
{
  double acc = 0.0;
  ConstModes md = (ConstModes)f.modes;
  int m = 0;
  for (; m + 4 <= f.nmodes; m += 4, md += 20) {
    double c[4];
    if (FAST) {
      double n[4], s[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double arg = md[5 * u] * x + md[5 * u + 1] * y + md[5 * u + 2] * z + md[5 * u + 3];
        n[u] = reduce_pi(arg, s[u]);
        c[u] = kCosTaylor[0];
      }
#pragma unroll
      for (int k = 1; k < 11; k++)
#pragma unroll
        for (int u = 0; u < 4; u++) c[u] = fma(c[u], s[u], kCosTaylor[k]);
#pragma unroll
      for (int u = 0; u < 4; u++) c[u] = ((int)n[u] & 1) ? -c[u] : c[u];
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double arg = md[5 * u] * x + md[5 * u + 1] * y + md[5 * u + 2] * z + md[5 * u + 3];
        c[u] = cos_cw(arg);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc += md[5 * u + 4] * c[u];
  }
  for (; m < f.nmodes; m++, md += 5) {
    const double arg = md[0] * x + md[1] * y + md[2] * z + md[3];
    acc += md[4] * (FAST ? cos_reduced(arg) : cos_cw(arg));
  }
  return acc;
}

// Scalar fields evaluated at Gauss points (same definitions as the oracle).
__device__ inline double eval_field(const DevFunction& f, double x, double y, double z)
{
  switch (f.type) {
  case IFEM_GPU_FN_CONST:
    return f.scale;
  case IFEM_GPU_FN_SINPROD:
    return f.scale * sin_library(M_PI * x) * sin_library(M_PI * y) * sin_library(M_PI * z);
  case IFEM_GPU_FN_FOURIER:
  case IFEM_GPU_FN_EXP_FOURIER: {
    // the bound (with a margin for rounding) decides once per point
    const double xm = fmax(fabs(x), fmax(fabs(y), fabs(z)));
    const double s = f.kbound * xm + f.pbound <= 0x1p29 ? fourier_sum<true>(f, x, y, z)
                                                        : fourier_sum<false>(f, x, y, z);
    const double g = f.mean + f.scale * s;
    return f.type == IFEM_GPU_FN_EXP_FOURIER ? exp(g) : g;
  }
  default:
    return 0.0;
  }
}

struct AsmArgs {
  int n1, n2, n3;
  int nel1, nel2, nel3;
  const double* coefs;                 // nnod*3
  const double *tabN1, *tabD1, *span1; // direction tables
  const double *tabN2, *tabD2, *span2;
  const double *tabN3, *tabD3, *span3;
  double gw[kMaxNG];                   // Gauss weights
  const int* eq;                       // per DOF: 0-based eq or -1
  const int64_t* rowptr;
  double* val64;
  float* val32;
  double* rhs;                         // FP64 RHS (nullptr: not written)
  DevFunction kappa, source;
  double body[3];
  double lam, mu;
  int variant;                         // batch kernel: 2*(p-1) + half chunk
  int ng;                              // Gauss points per direction
  int ncomp;                           // 1 scalar, 3 vector (elasticity)
  // geometry record per (element, Gauss point), [element][gp][GeoRec::N]
  // with element e1 + nel1*(e2 + nel2*e3) and gp q1 + ng*(q2 + ng*q3),
  // written by k_gp_geometry, read by the line kernels
  double* gp_geo;
  double kappa_c, source_c;            // values of constant fields
  // per DOF prescribed value of eliminated DOFs (FP64 pass only; nullptr:
  // homogeneous); the eliminated columns move to the RHS, b_i -= K_ij g_j
  const double* gval;
  // several patches in one system: rows are shared across patches, so
  // values are added at the column found by binary search in the sorted row
  const int32_t* colind;
  int accum;
  // accumulate mode with precomputed positions (SubPatch::d_pmap / d_pptr)
  const int64_t* pmap;
  const int64_t* pptr;
};

// Geometry record of one Gauss point.  Vector problems: J^-1 (row-major, 9)
// and w*detJ.  Scalar problems: M = w*detJ*kappa(X) * J^-1 J^-T (symmetric:
// m00 m01 m02 m11 m12 m22) and w*detJ*f(X), padded to 8.  Vector records are
// stored per element as 5 rows of ng^3 pairs (values 2k, 2k+1 of point q of
// element e in pair (e*5 + k)*ng^3 + q), so that the threads of consecutive
// Gauss points write and read consecutive 16-byte pairs; scalar records (64
// bytes) one after the other (point id at 4*id pairs).
template <int NC>
struct GeoRec {
  static constexpr int N = NC == 1 ? 8 : 10;
  static constexpr int NP = N / 2;
  static constexpr bool ROWS = NC == 3;
};

// The line kernels read their AsmArgs from a device buffer through the
// constant address space: the fields are uniform, read-only loads that the
// compiler may re-issue instead of holding them all in registers.
typedef const __attribute__((address_space(4))) AsmArgs ConstArgs;

__device__ inline ConstArgs* const_args(const AsmArgs* p)
{
  return (ConstArgs*)p;
}

// Chunk of layer elements processed together so that CH*NG^3 ~ one block;
// HALF = 1 selects half of that (an autotuning variant: less LDS, more
// phases); HALF = 2 keeps the full chunk and splits its Gauss points into
// two slabs of q3 rows (half the geometry LDS: three workgroups per CU).
template <int P, int NG, int HALF = 0>
struct Chunk {
  static constexpr int P1 = P + 1;
  static constexpr int NB = P1 * P1;
  static constexpr int GP3 = NG * NG * NG;
  static constexpr int full = (256 / GP3) < 1 ? 1 : ((256 / GP3) > NB ? NB : (256 / GP3));
  static constexpr int v = HALF == 1 ? ((full + 1) / 2) : full;
  static constexpr int NSL = HALF == 2 ? 2 : 1;                    // Gauss-point slabs
  static constexpr int SLQ = NG * NG * ((NG + NSL - 1) / NSL);     // points per slab (max)
};

// Geometry pre-pass: one thread per (element, Gauss point) computes, once
// per element, what the line kernels would otherwise recompute in each of the
// (p+1)^2 lines touching the element: J = X dN/du with the reference's
// cofactor det/inverse (utl::Jacobian, src/Utility/CoordinateMapping.C:24-63,
// src/LinAlg/matrix.h:791-862, |detJ| <= 1e-12 drops the point),
// detJxW = detJ*dV*w_i*w_j*w_k (src/ASM/ASMs3D.C:2145,2285), and for scalar
// problems the material kappa(X) and source f(X) at X = sum_b N_b x_b
// (src/ASM/ASMs3D.C:2303), folded into the record GeoRec.
__device__ inline bool field_varies(const DevFunction& f)
{
  return f.type == IFEM_GPU_FN_SINPROD ||
         ((f.type == IFEM_GPU_FN_FOURIER || f.type == IFEM_GPU_FN_EXP_FOURIER) && f.nmodes > 0);
}

__device__ inline DevFunction load_function(const __attribute__((address_space(4))) DevFunction& f)
{
  DevFunction d;
  d.type = f.type;
  d.nmodes = f.nmodes;
  d.scale = f.scale;
  d.mean = f.mean;
  d.modes = f.modes;
  d.kbound = f.kbound;
  d.pbound = f.pbound;
  return d;
}

template <int P, bool VEC = false>
__device__ inline void gp_geometry_at(ConstArgs& a, int64_t id, bool want_rhs)
{
  constexpr int P1 = P + 1;
  const int ng = a.ng;
  const int gp3 = ng * ng * ng;
  const int64_t e = id / gp3;
  const int q = (int)(id - e * gp3);
  const int q1 = q % ng, q2 = (q / ng) % ng, q3 = q / (ng * ng);
  const int e1 = (int)(e % a.nel1), e2 = (int)((e / a.nel1) % a.nel2),
            e3 = (int)(e / ((int64_t)a.nel1 * a.nel2));
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double X[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int b3 = 0; b3 < P1; b3++)
#pragma unroll 1
    for (int b2 = 0; b2 < P1; b2++)
#pragma unroll
      for (int b1 = 0; b1 < P1; b1++) {
        const double n1 = a.tabN1[(e1 * ng + q1) * P1 + b1], d1 = a.tabD1[(e1 * ng + q1) * P1 + b1];
        const double n2 = a.tabN2[(e2 * ng + q2) * P1 + b2], d2 = a.tabD2[(e2 * ng + q2) * P1 + b2];
        const double n3 = a.tabN3[(e3 * ng + q3) * P1 + b3], d3 = a.tabD3[(e3 * ng + q3) * P1 + b3];
        const double g0 = d1 * n2 * n3, g1 = n1 * d2 * n3, g2 = n1 * n2 * d3;
        const double nv = n1 * n2 * n3;
        const int64_t node =
            (int64_t)(e1 + b1) + (int64_t)a.n1 * ((e2 + b2) + (int64_t)a.n2 * (e3 + b3));
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const double xi = a.coefs[3 * node + i];
          J[3 * i + 0] += xi * g0;
          J[3 * i + 1] += xi * g1;
          J[3 * i + 2] += xi * g2;
          X[i] += xi * nv;
        }
      }
  const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[6] * J[5]) +
                     J[2] * (J[3] * J[7] - J[6] * J[4]);
  double Ji[9];
  double wdet;
  if (det <= 1e-12 && det >= -1e-12) {
#pragma unroll
    for (int m = 0; m < 9; m++) Ji[m] = 0.0;
    wdet = 0.0;
  } else {
    Ji[0] = (J[4] * J[8] - J[7] * J[5]) / det;
    Ji[3] = -(J[3] * J[8] - J[6] * J[5]) / det;
    Ji[6] = (J[3] * J[7] - J[6] * J[4]) / det;
    Ji[1] = -(J[1] * J[8] - J[7] * J[2]) / det;
    Ji[4] = (J[0] * J[8] - J[6] * J[2]) / det;
    Ji[7] = -(J[0] * J[7] - J[6] * J[1]) / det;
    Ji[2] = (J[1] * J[5] - J[4] * J[2]) / det;
    Ji[5] = -(J[0] * J[5] - J[3] * J[2]) / det;
    Ji[8] = (J[0] * J[4] - J[3] * J[1]) / det;
    const double dV = 0.125 * (a.span1[e1] * a.span2[e2] * a.span3[e3]);
    wdet = det * (dV * a.gw[q1] * a.gw[q2] * a.gw[q3]);
  }
  if (VEC || a.ncomp == 3) {
    double2* out = reinterpret_cast<double2*>(a.gp_geo) + e * gp3 * GeoRec<3>::NP + q;
#pragma unroll
    for (int k = 0; k < 4; k++) out[k * gp3] = make_double2(Ji[2 * k], Ji[2 * k + 1]);
    out[4 * gp3] = make_double2(Ji[8], wdet);
    return;
  }
  const DevFunction kf = load_function(a.kappa), sf = load_function(a.source);
  const double kap = field_varies(kf) ? eval_field(kf, X[0], X[1], X[2]) : a.kappa_c;
  const double src =
      want_rhs ? (field_varies(sf) ? eval_field(sf, X[0], X[1], X[2]) : a.source_c) : 0.0;
  const double sk = wdet * kap;
  double2* out = reinterpret_cast<double2*>(a.gp_geo) + id * GeoRec<1>::NP;
  out[0] = make_double2(sk * (Ji[0] * Ji[0] + Ji[1] * Ji[1] + Ji[2] * Ji[2]),
                        sk * (Ji[0] * Ji[3] + Ji[1] * Ji[4] + Ji[2] * Ji[5]));
  out[1] = make_double2(sk * (Ji[0] * Ji[6] + Ji[1] * Ji[7] + Ji[2] * Ji[8]),
                        sk * (Ji[3] * Ji[3] + Ji[4] * Ji[4] + Ji[5] * Ji[5]));
  out[2] = make_double2(sk * (Ji[3] * Ji[6] + Ji[4] * Ji[7] + Ji[5] * Ji[8]),
                        sk * (Ji[6] * Ji[6] + Ji[7] * Ji[7] + Ji[8] * Ji[8]));
  out[3] = make_double2(wdet * src, 0.0);
}

// Scalar problems: built for four waves per SIMD (at most 128 registers):
// the Fourier-mode loop is latency-bound, and more waves hide more of it
// (measured: 4 waves beat 2, 3, 5, 6 and 8 on the 256-patch diffusion batch;
// beyond 4 the geometry part spills).  Vector problems (no fields): their
// own instance without the field code, built for five waves per SIMD (64^3
// elasticity: 0.34 ms against 0.48 ms with the shared instance).
template <int P>
__global__ __launch_bounds__(256, 4) __attribute__((flatten)) void k_gp_geometry(
    const AsmArgs* __restrict__ args, int64_t total, int want_rhs)
{
  ConstArgs& a = *const_args(args);
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x)
    gp_geometry_at<P>(a, id, want_rhs != 0);
}

template <int P>
__global__ __launch_bounds__(256, 5) void k_gp_geometry_vec(const AsmArgs* __restrict__ args,
                                                            int64_t total)
{
  ConstArgs& a = *const_args(args);
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x)
    gp_geometry_at<P, true>(a, id, false);
}

// The value of the wave's first active lane (64-bit).
__device__ __forceinline__ int64_t readfirstlane64(int64_t v)
{
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Batch, one launch per degree P: the patches of that degree are
// plist[0..np), patch plist[i] holding points [gq0[i], gq0[i+1]) of the
// launch's total.  The point function is inlined (flatten: no call, so no
// register saves to scratch per point; the library fallbacks stay calls).
template <int P>
__global__ __launch_bounds__(256, 4) __attribute__((flatten)) void k_gp_geometry_batch(
    const AsmArgs* __restrict__ args, const int* __restrict__ plist,
    const int64_t* __restrict__ gq0, int np, int64_t total, int want_rhs)
{
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    // The wave's first point (its lowest index) picks a patch by a
    // wave-uniform search; a wave whose points all lie in that patch (all
    // but the waves that straddle a patch boundary) runs with the patch index
    // as a wave-uniform value, so that the patch's arguments and mode tables
    // come through scalar loads.
    const int64_t idw = readfirstlane64(id);
    int lw = 0, hw = np - 1;
    while (lw < hw) {
      const int mid = (lw + hw + 1) >> 1;
      if (gq0[mid] <= idw) lw = mid;
      else hw = mid - 1;
    }
    const int64_t wend = lw + 1 < np ? gq0[lw + 1] : total;
    if (__ballot(id >= wend) == 0) {
      const int pl = __builtin_amdgcn_readfirstlane(plist[lw]);
      gp_geometry_at<P>(const_args(args)[pl], id - readfirstlane64(gq0[lw]), want_rhs != 0);
    } else {
      int lo = 0, hi = np - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gq0[mid] <= id) lo = mid;
        else hi = mid - 1;
      }
      gp_geometry_at<P>(const_args(args)[plist[lo]], id - gq0[lo], want_rhs != 0);
    }
  }
}

// Value of a constant field (kappa of type ZERO means 1).
static double const_value(const ifem_gpu_function& f, double zero_value)
{
  switch (f.type) {
  case IFEM_GPU_FN_CONST: return f.scale;
  case IFEM_GPU_FN_FOURIER: return f.mean;
  case IFEM_GPU_FN_EXP_FOURIER: return std::exp(f.mean);
  case IFEM_GPU_FN_ZERO: return zero_value;
  default: return 0.0;
  }
}

static int grid_for(int64_t total)
{
  const int64_t g = (total + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

constexpr int kBlock = 256;

// Block barrier after LDS writes that other waves read right after it: the
// wave first waits for its own LDS operations to complete (s_waitcnt
// lgkmcnt(0)).  The compiler emits no such wait before the barrier of a loop
// whose body ends in LDS stores (the element rounds of the window update),
// and without it a round's reads could see a window entry before the
// previous round's store to it had landed (observed as lost contributions on
// entries shared by consecutive bundle elements, in workgroups sharing a CU).
__device__ __forceinline__ void lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __syncthreads();
}

// LDS of one line workgroup.  T: arithmetic type of the element integrals
// (double: FP64 path, float: FP32 path).  NC: components (1 Poisson,
// 3 elasticity).  SPLIT: number of blocks per line sharing the row
// components (NC/SPLIT rows each).  HALF: half-size element chunks.
// Equation number of neighbour DOF t (node offset t / NC in the (2p+1)^3
// window around node (i, jc, kc), component t % NC); -2 outside the patch.
template <int P, int NC>
__device__ __forceinline__ int window_eq(ConstArgs& a, int i, int jc, int kc, int t)
{
  constexpr int W1 = 2 * P + 1;
  const int nbr = t / NC, cc = t % NC;
  const int ii = i + nbr % W1 - P, jj = jc + (nbr / W1) % W1 - P, kk = kc + nbr / (W1 * W1) - P;
  if (t >= W1 * W1 * W1 * NC || ii < 0 || ii >= a.n1 || jj < 0 || jj >= a.n2 || kk < 0 || kk >= a.n3)
    return -2;
  return a.eq[((int64_t)ii + (int64_t)a.n1 * (jj + (int64_t)a.n2 * kk)) * NC + cc];
}

template <int P, int NG, int NC, typename T, int SPLIT, bool HAS_RHS, int HALF, bool ORDL = false>
struct LineLds {
  static constexpr int P1 = P + 1;
  static constexpr int NEN = P1 * P1 * P1;
  static constexpr int GP3 = NG * NG * NG;
  static constexpr int NB = P1 * P1;
  static constexpr int W1 = 2 * P + 1;
  static constexpr int WN = W1 * W1 * W1;
  static constexpr int NCR = NC / SPLIT;
  static constexpr bool SACC = (NC == 3 && SPLIT == 1); // accumulate S_ij, apply lam/mu later
  static constexpr int NS = NC == 1 ? 1 : (SACC ? 9 : 3);
  static constexpr int CH = Chunk<P, NG, HALF>::v;
  static constexpr int NSL = Chunk<P, NG, HALF>::NSL;
  static constexpr int SLQ = Chunk<P, NG, HALF>::SLQ;
  static constexpr int GEO = (NC == 3 ? 9 : 0) + 3 * P1;
  static constexpr int NRHS = HAS_RHS ? P1 * NC : 1;
  // scalar problems with more than 27 Gauss points per element (the ordered,
  // not pipelined path): phase B by sum factorisation, see line_impl
  static constexpr bool SF =
      NC == 1 && GP3 > 27 && HALF != 1 && CH * NEN <= 256 && P >= 2 && P1 <= NG;

  // every member 16-byte aligned: lets the compiler use 128-bit LDS loads
  alignas(16) T win[P1][NCR][WN][NC];
  alignas(16) T geo[CH * SLQ][GEO];
  // ORDL (ordered window updates): no stage; RHS stage of w*detJ*N_a only
  alignas(16) T stage[ORDL ? 1 : CH][NEN][P1][NS];
  alignas(16) T rstage[HAS_RHS ? CH * SLQ : 1][ORDL ? (HAS_RHS ? P1 : 1) : NRHS];
  alignas(16) T rwin[P1][NC];
  alignas(16) T tN1[NG][P1];
  alignas(16) T tD1[NG][P1];
  alignas(16) T tN2[P1][NG][P1];
  alignas(16) T tD2[P1][NG][P1];
  alignas(16) T tN3[P1][NG][P1];
  alignas(16) T tD3[P1][NG][P1];
  alignas(16) int be2[NB];
  alignas(16) int be3[NB];
  static constexpr int NT = (WN * NC + 63) / 64;
  // neighbour equation numbers kept in LDS for phase D (not for p = 3,
  // whose LDS budget goes to the element chunk, nor for the slab variant,
  // whose budget is three workgroups per CU)
  static constexpr bool PFD = P < 3 && NSL == 1;
  // two buffers: row e1's (being filled) and row e1-1's (written out at the
  // top of layer e1)
  alignas(16) int eqw[2][PFD ? NT * 64 : 1];
  alignas(16) int64_t pbase[2][NCR];
  int prow[2][NCR];
  int nbundle;
};

// The LDS arrays come in as separate __restrict__ pointers so that, after
// inlining, the compiler knows they do not alias (fields of one struct with
// run-time indices would otherwise be assumed to overlap).
template <int P, int NG, int NC, typename T, int SPLIT, bool HAS_RHS, int HALF, bool PIPE, bool ORD,
          bool SEMI = false>
__device__ __forceinline__ void line_impl(
    ConstArgs& a, const int line, const int split,
    T (*__restrict__ win)[LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::NCR]
                         [LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::WN][NC],
    T (*__restrict__ geo)[LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::GEO],
    T (*__restrict__ stage)[LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::NEN][P + 1]
                           [LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::NS],
    T* __restrict__ rstage,
    T (*__restrict__ rwin)[NC],
    T (*__restrict__ tN1)[P + 1], T (*__restrict__ tD1)[P + 1],
    T (*__restrict__ tN2)[NG][P + 1], T (*__restrict__ tD2)[NG][P + 1],
    T (*__restrict__ tN3)[NG][P + 1], T (*__restrict__ tD3)[NG][P + 1],
    int* __restrict__ be2, int* __restrict__ be3,
    int (*__restrict__ eqw)[LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::PFD
                            ? LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::NT * 64 : 1],
    int64_t (*__restrict__ pbase)[LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::NCR],
    int (*__restrict__ prow)[LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>::NCR],
    int* __restrict__ nbundlep)
{
  using LL = LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>;
  constexpr int P1 = LL::P1;
  constexpr int NEN = LL::NEN;
  constexpr int GP3 = LL::GP3;
  constexpr int W1 = LL::W1;
  constexpr int WN = LL::WN;
  constexpr int NCR = LL::NCR;
  constexpr bool SACC = LL::SACC;
  constexpr int NS = LL::NS;
  constexpr int CH = LL::CH;
  constexpr int NSL = LL::NSL; // Gauss-point slabs per chunk (geometry rows in LDS per slab)
  constexpr int SLQ = LL::SLQ;
  static_assert(NSL == 1 || (ORD && PIPE), "Gauss-point slabs need the ordered, pipelined path");
  // ORD: the whole bundle in one chunk, column sums added to the window
  // element by element (no stage); the RHS stage then holds w*detJ*N_a and
  // the body force is applied after the sum
  static_assert(!ORD || CH * LL::NEN <= kBlock, "ordered updates need one item per thread");
  constexpr int RS = ORD ? (HAS_RHS ? P1 : 1) : LL::NRHS; // RHS stage row length
  const int rc = split * NCR; // first row component of the block

  const int tid = threadIdx.x;
  const int jc = line % a.n2, kc = line / a.n2;
  // split: row components [split*NCR, split*NCR+NCR)

  // bundle of elements touching the line, fixed order (e3 outer, e2 inner)
  if (tid == 0) {
    int nb = 0;
    for (int m3 = 0; m3 < P1; m3++)
      for (int m2 = 0; m2 < P1; m2++) {
        const int e2 = jc - P + m2, e3 = kc - P + m3;
        if (e2 >= 0 && e2 < a.nel2 && e3 >= 0 && e3 < a.nel3) {
          be2[nb] = e2;
          be3[nb] = e3;
          nb++;
        }
      }
    *nbundlep = nb;
  }
  // direction 2/3 tables for elements jc-P..jc and kc-P..kc
  for (int t = tid; t < P1 * NG * P1; t += kBlock) {
    const int m = t / (NG * P1), r = t % (NG * P1);
    const int g = r / P1, q = r % P1;
    const int e2 = jc - P + m, e3 = kc - P + m;
    T n2 = 0, d2 = 0, n3 = 0, d3 = 0;
    if (e2 >= 0 && e2 < a.nel2) {
      n2 = (T)a.tabN2[(e2 * NG + g) * P1 + q];
      d2 = (T)a.tabD2[(e2 * NG + g) * P1 + q];
    }
    if (e3 >= 0 && e3 < a.nel3) {
      n3 = (T)a.tabN3[(e3 * NG + g) * P1 + q];
      d3 = (T)a.tabD3[(e3 * NG + g) * P1 + q];
    }
    tN2[m][g][q] = n2;
    tD2[m][g][q] = d2;
    tN3[m][g][q] = n3;
    tD3[m][g][q] = d3;
  }
  for (int t = tid; t < P1 * NCR * WN * NC; t += kBlock) (&win[0][0][0][0])[t] = T(0);
  for (int t = tid; t < P1 * NC; t += kBlock) (&rwin[0][0])[t] = T(0);
  __syncthreads();
  const int nb = *nbundlep;

  const T lam = (T)a.lam, mu = (T)a.mu;

  constexpr int NT = LL::NT; // phase D: 64-column steps per row
  // ---- phase D: write row i of the window out (row e1-1 at the top of layer
  // e1, while phase A runs; the last layer's rows at its end).  pf: the
  // equation numbers of the row are in buffer pb (from the layer top);
  // zero_rhs: clear the row's right-hand side entries after reading them.
  auto row_out = [&](int i, bool pf, int pb, bool zero_rhs) {
    const int slot = i % P1;
    const int64_t node = (int64_t)i + (int64_t)a.n1 * (jc + (int64_t)a.n2 * kc);
    const int wave = tid >> 6, lane = tid & 63;
    if (wave < NCR) {
      const int ic = split * NCR + wave;
      const int req = pf ? prow[pb][wave] : a.eq[node * NC + ic];
      const bool pfe = LL::PFD && pf;
      if (req >= 0) {
        const int64_t base = pf ? pbase[pb][wave] : a.rowptr[req];
        const double* __restrict__ gv = a.gval;
        double corr = 0.0;
        int64_t pos = 0;
#pragma unroll 1
        for (int k = 0; k < NT; k++) {
          const int t = k * 64 + lane;
          const int ev = pfe ? eqw[pb][t] : window_eq<P, NC>(a, i, jc, kc, t);
          const bool valid = ev >= 0;
          const unsigned long long m = __ballot(valid);
          if (valid && a.accum && a.pmap) {
            const T v = win[slot][wave][t / NC][t % NC];
            const int64_t at = a.pmap[a.pptr[node * NC + ic] + pos + __popcll(m & ((1ull << lane) - 1ull))];
            if (a.val32) a.val32[at] += (float)v;
            else a.val64[at] += (double)v;
          } else if (valid && a.accum) {
            const T v = win[slot][wave][t / NC][t % NC];
            int64_t lo = base, hi = a.rowptr[req + 1] - 1;
            while (lo <= hi) {
              const int64_t mid = (lo + hi) >> 1;
              const int cm = a.colind[mid];
              if (cm == ev) {
                if (a.val32) a.val32[mid] += (float)v;
                else a.val64[mid] += (double)v;
                break;
              }
              if (cm < ev) lo = mid + 1;
              else hi = mid - 1;
            }
          } else if (valid) {
            const int64_t at = base + pos + __popcll(m & ((1ull << lane) - 1ull));
            const T v = win[slot][wave][t / NC][t % NC];
            if (a.val32) a.val32[at] = (float)v;
            else a.val64[at] = (double)v;
          } else if (gv && ev == -1) { // eliminated column inside the patch
            const int nbr = t / NC;
            const int ii = i + nbr % W1 - P, jj = jc + (nbr / W1) % W1 - P, kk = kc + nbr / (W1 * W1) - P;
            const int64_t gdof = ((int64_t)ii + (int64_t)a.n1 * (jj + (int64_t)a.n2 * kk)) * NC + t % NC;
            corr += (double)win[slot][wave][t / NC][t % NC] * gv[gdof];
          }
          pos += __popcll(m);
        }
        if (gv) {
#pragma unroll
          for (int off = 32; off >= 1; off >>= 1) corr += __shfl_xor(corr, off, 64);
        }
        if ((HAS_RHS || gv) && a.rhs && lane == 0) {
          if (a.accum) a.rhs[req] += (double)rwin[slot][ic] - corr;
          else a.rhs[req] = (double)rwin[slot][ic] - corr;
        }
      }
      if (zero_rhs && lane == 0) rwin[slot][ic] = T(0);
    }
    if (zero_rhs && tid == 0)
      for (int c = 0; c < NC; c++)
        if (c < split * NCR || c >= split * NCR + NCR) rwin[slot][c] = T(0);
  };
  for (int e1 = 0; e1 < a.nel1; e1++) {
    // ---- direction-1 tables of this layer ----------------------------------
    // with the equation numbers of row e1 and of its neighbour columns for
    // phase D (their loads overlap those of the tables)
    for (int t = tid; t < NG * P1; t += kBlock) {
      (&tN1[0][0])[t] = (T)a.tabN1[e1 * NG * P1 + t];
      (&tD1[0][0])[t] = (T)a.tabD1[e1 * NG * P1 + t];
    }
    const int buf = e1 & 1;
    // equation numbers of row e1's neighbour columns for its phase D (at the
    // top of the next layer): for elasticity without a body force loaded now
    // and stored after phase B (after phase A in the unordered variants; EQE),
    // so that their latency is not exposed at the barrier below; elsewhere the
    // two registers would cost a wave per SIMD, and they are stored here
    constexpr bool EQE = LL::PFD && NC == 3 && !HAS_RHS;
    constexpr int NEQV = EQE ? (NT * 64 + kBlock - 1) / kBlock : 1;
    int eqv[NEQV];
#pragma unroll
    for (int k = 0; k < NEQV; k++) {
      const int t = tid + k * kBlock;
      eqv[k] = (EQE && t < NT * 64) ? window_eq<P, NC>(a, e1, jc, kc, t) : -2;
    }
    if (LL::PFD && !EQE)
      for (int t = tid; t < NT * 64; t += kBlock) eqw[buf][t] = window_eq<P, NC>(a, e1, jc, kc, t);
    if (tid < NCR) {
      const int64_t node = (int64_t)e1 + (int64_t)a.n1 * (jc + (int64_t)a.n2 * kc);
      prow[buf][tid] = a.eq[node * NC + split * NCR + tid];
    }
    __syncthreads();
    // row start of row e1, read while phase A runs
    int64_t pf_base = 0;
    if (tid < NCR && prow[buf][tid] >= 0) pf_base = a.rowptr[prow[buf][tid]];
    // row e1-1, complete since the last layer, goes out now: its loads and
    // stores overlap phase A (its window slot is cleared after phase A, and
    // its right-hand side entries by the threads that read them)
    if (e1 > 0) row_out(e1 - 1, true, buf ^ 1, true);

    for (int c0 = 0; c0 < nb; c0 += CH) {
      const int nch = (nb - c0) < CH ? (nb - c0) : CH;

      // ---- phase A: geometry per (element, Gauss point) of the points
      // [qa, qa + ngs) (all of them unless the chunk runs in slabs) --------
      auto phase_a = [&](const int qa, const int ngs) {
      for (int t = tid; t < nch * ngs; t += kBlock) {
        const int le = t / ngs, q = qa + t % ngs;
        const int e2 = be2[c0 + le], e3 = be3[c0 + le];
        const int m2 = e2 - jc + P, m3 = e3 - kc + P; // table slots
        const int q1 = q % NG, q2 = (q / NG) % NG, q3 = q / (NG * NG);
        // geometry record of this (element, Gauss point), computed once per
        // element by k_gp_geometry: J^-1 and w*detJ (vector), or
        // M = w*detJ*kappa*J^-1 J^-T and w*detJ*f (scalar)
        const int64_t ge = (int64_t)e1 + (int64_t)a.nel1 * (e2 + (int64_t)a.nel2 * e3);
        constexpr int PS = GeoRec<NC>::ROWS ? GP3 : 1; // pair stride of a record
        const double2* __restrict__ rec =
            reinterpret_cast<const double2*>(a.gp_geo) + ge * (GP3 * GeoRec<NC>::NP) +
            q * (GeoRec<NC>::ROWS ? 1 : GeoRec<NC>::NP);
        T R[GeoRec<NC>::N];
#pragma unroll
        for (int k = 0; k < GeoRec<NC>::NP; k++) {
          const double2 v = rec[k * PS];
          R[2 * k] = (T)v.x;
          R[2 * k + 1] = (T)v.y;
        }
        if (NC == 3) {
#pragma unroll
          for (int m = 0; m < 9; m++) geo[t][m] = R[m];
        }
        const int a2r = jc - e2, a3r = kc - e3; // row node's local (a2, a3)
        const T n2a = tN2[m2][q2][a2r], d2a = tD2[m2][q2][a2r];
        const T n3a = tN3[m3][q3][a3r], d3a = tD3[m3][q3][a3r];
        if constexpr (LL::SF) {
          // W_kl = M_kl * A2_k(q2) * A3_k(q3) for the line's row functions in
          // directions 2 and 3 (A_k: derivative in direction k, else value)
          const T M[9] = {R[0], R[1], R[2], R[1], R[3], R[4], R[2], R[4], R[5]};
          const T a2[3] = {n2a, d2a, n2a}, a3[3] = {n3a, n3a, d3a};
          // stored [element][kl][q1][q2][q3]: the (q2, q3) slab of each q1
          // is contiguous, and its T2 later overwrites it in place
          T* __restrict__ wf = &geo[0][0] + (int64_t)le * 9 * GP3 + (q1 * NG + q2) * NG + q3;
#pragma unroll
          for (int k = 0; k < 3; k++)
#pragma unroll
            for (int l = 0; l < 3; l++) wf[(3 * k + l) * GP3] = M[3 * k + l] * a2[k] * a3[k];
        }
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++) {
          const T n1 = tN1[q1][a1], d1 = tD1[q1][a1];
          const T r0 = d1 * n2a * n3a, r1 = n1 * d2a * n3a, r2 = n1 * n2a * d3a;
          if (NC == 1 && LL::SF) {
            // written once below the a1 loop
          } else if (NC == 1) {
            // h_a = M grad_ref N_a: S_ab = grad_ref N_b . h_a
            geo[t][3 * a1 + 0] = R[0] * r0 + R[1] * r1 + R[2] * r2;
            geo[t][3 * a1 + 1] = R[1] * r0 + R[3] * r1 + R[4] * r2;
            geo[t][3 * a1 + 2] = R[2] * r0 + R[4] * r1 + R[5] * r2;
          } else {
            // w g_a = w J^-T grad_ref N_a  (dNdX = dNdu * J^-1)
            const T w = R[9];
            geo[t][9 + 3 * a1 + 0] = w * (r0 * R[0] + r1 * R[3] + r2 * R[6]);
            geo[t][9 + 3 * a1 + 1] = w * (r0 * R[1] + r1 * R[4] + r2 * R[7]);
            geo[t][9 + 3 * a1 + 2] = w * (r0 * R[2] + r1 * R[5] + r2 * R[8]);
          }
          if (HAS_RHS) {
            const T Na = n1 * n2a * n3a;
            if (NC == 1) {
              rstage[t * RS + a1] = R[6] * Na;
            } else if (ORD) {
              rstage[t * RS + a1] = R[9] * Na;
            } else {
#pragma unroll
              for (int c = 0; c < NC; c++) rstage[t * RS + a1 * NC + c] = R[9] * (T)a.body[c] * Na;
            }
          }
        }
      }
      };
      // ---- RHS: fixed-order reduction over (element, GP) of a slab ---------
      auto rhs_reduce = [&](const int ngs) {
        const int v = tid >> 4, l = tid & 15;
        T s = T(0);
        const int vs = (ORD && NC == 3) ? v / NC : v;
        if (v < P1 * NC)
          for (int pr = l; pr < nch * ngs; pr += 16) s += rstage[pr * RS + vs];
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 16);
        if (v < P1 * NC && l == 0) {
          const int a1 = v / NC, c = v % NC;
          rwin[(e1 + a1) % P1][c] += (ORD && NC == 3) ? s * (T)a.body[c] : s;
        }
      };
      phase_a(0, NSL == 1 ? GP3 : SLQ);
      if (c0 == 0 && tid < NCR) pbase[buf][tid] = pf_base;
      __syncthreads();
      // (ordered updates: after phase B, further from the loads)
      if (EQE && !ORD && c0 == 0) {
#pragma unroll
        for (int k = 0; k < NEQV; k++) {
          const int t = tid + k * kBlock;
          if (t < NT * 64) eqw[buf][t] = eqv[k];
        }
      }
      if (c0 == 0 && e1 > 0) { // slot of row e1-1, read out before the barrier
        const int zs = (e1 - 1) % P1;
        for (int t = tid; t < NCR * WN * NC; t += kBlock) (&win[zs][0][0][0])[t] = T(0);
      }

      if (HAS_RHS) rhs_reduce(NSL == 1 ? GP3 : SLQ);

      // ---- phase B: element-row x column-node products ----------------------
      if constexpr (ORD) {
      // One (element, column node b) item per thread for the chunk (the
      // whole bundle when it fits), Gauss loop as in the pipelined or the
      // row-prefetching variant below; the sums stay in
      // registers and go into the row window one element after the other
      // (element order fixed, so every window entry is summed in that order).
      const bool has = tid < nch * NEN;
      const int le = has ? tid / NEN : 0, b = has ? tid % NEN : 0;
      const int b1 = b % P1, b2 = (b / P1) % P1, b3 = b / (P1 * P1);
      const int e2 = be2[c0 + le], e3 = be3[c0 + le];
      T acc[P1][NS];
#pragma unroll
      for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
        for (int s = 0; s < NS; s++) acc[a1][s] = 0;
        if constexpr (PIPE) {
        const int m2 = e2 - jc + P, m3 = e3 - kc + P;
        T vn1[NG], vd1[NG], vn2[NG], vd2[NG], vn3[NG], vd3[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) {
          vn1[q] = tN1[q][b1]; vd1[q] = tD1[q][b1];
          vn2[q] = tN2[m2][q][b2]; vd2[q] = tD2[m2][q][b2];
          vn3[q] = tN3[m3][q][b3]; vd3[q] = tD3[m3][q][b3];
        }
        // Gauss points in NSL slabs: the geometry rows of slab sl > 0 replace
        // those of slab sl-1 in LDS (every thread takes part in the barriers)
        // (a compile-time slab index, so that each slab's Gauss loop unrolls)
        auto slab_b = [&](auto slc) {
        constexpr int qa = decltype(slc)::value * SLQ;
        constexpr int ngs = (GP3 - qa) < SLQ ? (GP3 - qa) : SLQ;
        if (has) {
        // one workgroup per CU and SIMD pair: the next point's geometry row
        // is prefetched into registers; with slabs (three workgroups per CU)
        // the rows are read from LDS where they are used, the other
        // workgroups hiding the latency
        constexpr bool PREF = NSL == 1;
        const T* __restrict__ Gbase = geo[le * ngs];
        T Gc[PREF ? LL::GEO : 1];
        if constexpr (PREF) {
#pragma unroll
          for (int m = 0; m < LL::GEO; m++) Gc[m] = Gbase[m];
        }
#pragma unroll
        for (int qq = 0; qq < ngs; qq++) {
          const int q = qa + qq;
          T Gn[PREF ? LL::GEO : 1];
          if (PREF && qq + 1 < ngs) {
#pragma unroll
            for (int m = 0; m < LL::GEO; m++) Gn[m] = Gbase[(qq + 1) * LL::GEO + m];
          }
          const int q1 = q % NG, q2 = (q / NG) % NG, q3 = q / (NG * NG);
          const T n23 = vn2[q2] * vn3[q3], d23 = vd2[q2] * vn3[q3], n2d3 = vn2[q2] * vd3[q3];
          const T r0 = vd1[q1] * n23, r1 = vn1[q1] * d23, r2 = vn1[q1] * n2d3;
          const T* G = PREF ? Gc : Gbase + qq * LL::GEO;
          if (NC == 1) {
#pragma unroll
            for (int a1 = 0; a1 < P1; a1++)
              acc[a1][0] += r0 * G[3 * a1] + r1 * G[3 * a1 + 1] + r2 * G[3 * a1 + 2];
          } else {
            const T gb0 = r0 * G[0] + r1 * G[3] + r2 * G[6];
            const T gb1 = r0 * G[1] + r1 * G[4] + r2 * G[7];
            const T gb2 = r0 * G[2] + r1 * G[5] + r2 * G[8];
#pragma unroll
            for (int a1 = 0; a1 < P1; a1++) {
              const T wa0 = G[9 + 3 * a1], wa1 = G[10 + 3 * a1], wa2 = G[11 + 3 * a1];
              if constexpr (SACC) {
                acc[a1][0] += wa0 * gb0; acc[a1][1] += wa0 * gb1; acc[a1][2] += wa0 * gb2;
                acc[a1][3] += wa1 * gb0; acc[a1][4] += wa1 * gb1; acc[a1][5] += wa1 * gb2;
                acc[a1][6] += wa2 * gb0; acc[a1][7] += wa2 * gb1; acc[a1][8] += wa2 * gb2;
              } else {
                const T wa[3] = {wa0, wa1, wa2}, gb[3] = {gb0, gb1, gb2};
                const T dot = wa0 * gb0 + wa1 * gb1 + wa2 * gb2;
#pragma unroll
                for (int jj = 0; jj < 3; jj++)
                  acc[a1][jj] += lam * wa[rc] * gb[jj] + mu * wa[jj] * gb[rc] +
                                 (jj == rc ? mu * dot : T(0));
              }
            }
          }
          if (PREF && qq + 1 < ngs) {
#pragma unroll
            for (int m = 0; m < LL::GEO; m++) Gc[m] = Gn[m];
          }
        }
        }
        };
        slab_b(std::integral_constant<int, 0>{});
        if constexpr (NSL > 1) {
          static_assert(NSL == 2, "two slabs");
          constexpr int qa1 = SLQ, ngs1 = GP3 - SLQ;
          __syncthreads();
          phase_a(qa1, ngs1);
          __syncthreads();
          if (HAS_RHS) rhs_reduce(ngs1);
          slab_b(std::integral_constant<int, 1>{});
        }
        } else if constexpr (LL::SF) {
        // Sum factorisation (scalar problems): for the line's row functions
        // a = (a1, a2, a3) with a2, a3 fixed by the line,
        //   S_ab = sum_kl sum_q W_kl(q) A1_k(a1,q1) B1_l(b1,q1) B2_l(b2,q2) B3_l(b3,q3)
        // contracted over q3 (T1), then q2 (T2, in the geometry rows' LDS),
        // then q1 with the direction-1 functions, each step in fixed order:
        // 9*NG*NG*P1 + 9*NG*P1*P1 + 3*NG*(P1+3)*NEN operations per element
        // and line instead of 3*GP3*P1*NEN.
        // one thread per (element, kl, q1): its (q2, q3) slab of W into
        // registers, T1(q2, b3) over q3, T2(b2, b3) over q2, written back
        // over the slab (no other thread touches it, so no barrier between
        // the read and the write)
        T* __restrict__ gf = &geo[0][0]; // W, then T2, [element][kl][q1][...] with stride GP3
        for (int t = tid; t < nch * 9 * NG; t += kBlock) {
          const int q1 = t % NG, kl = (t / NG) % 9, lq = t / (9 * NG);
          const int l = kl % 3;
          const int mm2 = be2[c0 + lq] - jc + P, mm3 = be3[c0 + lq] - kc + P;
          T* __restrict__ w = gf + (int64_t)(lq * 9 + kl) * GP3 + q1 * NG * NG;
          T wv[NG][NG];
#pragma unroll
          for (int q2 = 0; q2 < NG; q2++)
#pragma unroll
            for (int q3 = 0; q3 < NG; q3++) wv[q2][q3] = w[q2 * NG + q3];
          T t1[NG][P1];
#pragma unroll
          for (int q2 = 0; q2 < NG; q2++)
#pragma unroll
            for (int bb3 = 0; bb3 < P1; bb3++) {
              T sum = 0;
#pragma unroll
              for (int q3 = 0; q3 < NG; q3++)
                sum += wv[q2][q3] * (l == 2 ? tD3[mm3][q3][bb3] : tN3[mm3][q3][bb3]);
              t1[q2][bb3] = sum;
            }
#pragma unroll
          for (int bb2 = 0; bb2 < P1; bb2++)
#pragma unroll
            for (int bb3 = 0; bb3 < P1; bb3++) {
              T sum = 0;
#pragma unroll
              for (int q2 = 0; q2 < NG; q2++)
                sum += t1[q2][bb3] * (l == 1 ? tD2[mm2][q2][bb2] : tN2[mm2][q2][bb2]);
              w[bb2 * P1 + bb3] = sum; // T2(q1, b2, b3)
            }
        }
        __syncthreads();
        if (has) {
          // per (k, q1): U = T2_k0 B1'(b1) + (T2_k1 + T2_k2) B1(b1) (direction 1
          // carries the derivative for l = 0 only), then U A1_k(a1) per row
          const T* __restrict__ v = gf + (int64_t)le * 9 * GP3 + b2 * P1 + b3;
          T bn[NG], bd[NG];
#pragma unroll
          for (int q1 = 0; q1 < NG; q1++) {
            bn[q1] = tN1[q1][b1];
            bd[q1] = tD1[q1][b1];
          }
#pragma unroll
          for (int k = 0; k < 3; k++) {
#pragma unroll
            for (int q1 = 0; q1 < NG; q1++) {
              const T* __restrict__ vk = v + 3 * k * GP3 + q1 * NG * NG;
              const T u = vk[0] * bd[q1] + (vk[GP3] + vk[2 * GP3]) * bn[q1];
#pragma unroll
              for (int a1 = 0; a1 < P1; a1++)
                acc[a1][0] += u * (k == 0 ? tD1[q1][a1] : tN1[q1][a1]);
            }
          }
        }
        } else {
      if (has) {
        const int m2 = e2 - jc + P, m3 = e3 - kc + P;
        T vn1[NG], vd1[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) {
          vn1[q] = tN1[q][b1];
          vd1[q] = tD1[q][b1];
        }
        const T* __restrict__ Gbase = geo[le * GP3];
        T Gc[LL::GEO];
#pragma unroll
        for (int m = 0; m < LL::GEO; m++) Gc[m] = Gbase[m];
        // fully unrolled up to 4x4 rows: the loads of later points can then be
        // issued well ahead of their use; the two copies differ only in the
        // unroll directive
        if constexpr (NG * NG <= 16) {
#pragma unroll
          for (int g23 = 0; g23 < NG * NG; g23++) {
            const int q2 = g23 % NG, q3 = g23 / NG;
            const T n3 = tN3[m3][q3][b3], d3 = tD3[m3][q3][b3];
            const T n2 = tN2[m2][q2][b2], d2 = tD2[m2][q2][b2];
            const T n23 = n2 * n3, d23 = d2 * n3, n2d3 = n2 * d3;
#pragma unroll
            for (int q1 = 0; q1 < NG; q1++) {
              const int qn = g23 * NG + q1 + 1 < GP3 ? g23 * NG + q1 + 1 : GP3 - 1;
              T Gn[LL::GEO];
#pragma unroll
              for (int m = 0; m < LL::GEO; m++) Gn[m] = Gbase[qn * LL::GEO + m];
              const T r0 = vd1[q1] * n23, r1 = vn1[q1] * d23, r2 = vn1[q1] * n2d3;
              const T* G = Gc;
              if (NC == 1) {
#pragma unroll
                for (int a1 = 0; a1 < P1; a1++)
                  acc[a1][0] += r0 * G[3 * a1] + r1 * G[3 * a1 + 1] + r2 * G[3 * a1 + 2];
              } else {
                const T gb0 = r0 * G[0] + r1 * G[3] + r2 * G[6];
                const T gb1 = r0 * G[1] + r1 * G[4] + r2 * G[7];
                const T gb2 = r0 * G[2] + r1 * G[5] + r2 * G[8];
#pragma unroll
                for (int a1 = 0; a1 < P1; a1++) {
                  const T wa0 = G[9 + 3 * a1], wa1 = G[10 + 3 * a1], wa2 = G[11 + 3 * a1];
                  if constexpr (SACC) {
                    acc[a1][0] += wa0 * gb0; acc[a1][1] += wa0 * gb1; acc[a1][2] += wa0 * gb2;
                    acc[a1][3] += wa1 * gb0; acc[a1][4] += wa1 * gb1; acc[a1][5] += wa1 * gb2;
                    acc[a1][6] += wa2 * gb0; acc[a1][7] += wa2 * gb1; acc[a1][8] += wa2 * gb2;
                  } else {
                    const T wa[3] = {wa0, wa1, wa2}, gb[3] = {gb0, gb1, gb2};
                    const T dot = wa0 * gb0 + wa1 * gb1 + wa2 * gb2;
#pragma unroll
                    for (int jj = 0; jj < 3; jj++)
                      acc[a1][jj] += lam * wa[rc] * gb[jj] + mu * wa[jj] * gb[rc] +
                                     (jj == rc ? mu * dot : T(0));
                  }
                }
              }
#pragma unroll
              for (int m = 0; m < LL::GEO; m++) Gc[m] = Gn[m];
            }
          }
        } else {
          for (int g23 = 0; g23 < NG * NG; g23++) {
            const int q2 = g23 % NG, q3 = g23 / NG;
            const T n3 = tN3[m3][q3][b3], d3 = tD3[m3][q3][b3];
            const T n2 = tN2[m2][q2][b2], d2 = tD2[m2][q2][b2];
            const T n23 = n2 * n3, d23 = d2 * n3, n2d3 = n2 * d3;
#pragma unroll
            for (int q1 = 0; q1 < NG; q1++) {
              const int qn = g23 * NG + q1 + 1 < GP3 ? g23 * NG + q1 + 1 : GP3 - 1;
              T Gn[LL::GEO];
#pragma unroll
              for (int m = 0; m < LL::GEO; m++) Gn[m] = Gbase[qn * LL::GEO + m];
              const T r0 = vd1[q1] * n23, r1 = vn1[q1] * d23, r2 = vn1[q1] * n2d3;
              const T* G = Gc;
              if (NC == 1) {
#pragma unroll
                for (int a1 = 0; a1 < P1; a1++)
                  acc[a1][0] += r0 * G[3 * a1] + r1 * G[3 * a1 + 1] + r2 * G[3 * a1 + 2];
              } else {
                const T gb0 = r0 * G[0] + r1 * G[3] + r2 * G[6];
                const T gb1 = r0 * G[1] + r1 * G[4] + r2 * G[7];
                const T gb2 = r0 * G[2] + r1 * G[5] + r2 * G[8];
#pragma unroll
                for (int a1 = 0; a1 < P1; a1++) {
                  const T wa0 = G[9 + 3 * a1], wa1 = G[10 + 3 * a1], wa2 = G[11 + 3 * a1];
                  if constexpr (SACC) {
                    acc[a1][0] += wa0 * gb0; acc[a1][1] += wa0 * gb1; acc[a1][2] += wa0 * gb2;
                    acc[a1][3] += wa1 * gb0; acc[a1][4] += wa1 * gb1; acc[a1][5] += wa1 * gb2;
                    acc[a1][6] += wa2 * gb0; acc[a1][7] += wa2 * gb1; acc[a1][8] += wa2 * gb2;
                  } else {
                    const T wa[3] = {wa0, wa1, wa2}, gb[3] = {gb0, gb1, gb2};
                    const T dot = wa0 * gb0 + wa1 * gb1 + wa2 * gb2;
#pragma unroll
                    for (int jj = 0; jj < 3; jj++)
                      acc[a1][jj] += lam * wa[rc] * gb[jj] + mu * wa[jj] * gb[rc] +
                                     (jj == rc ? mu * dot : T(0));
                  }
                }
              }
#pragma unroll
              for (int m = 0; m < LL::GEO; m++) Gc[m] = Gn[m];
            }
          }
        }
      }
        }
      // rounds over the bundle elements: in round r the items of element r
      // add their column sums (distinct window entries within a round)
      // p <= 2 (STAGED): the geometry rows are no longer needed once the sums
      // are in registers; their LDS takes the sums of GS elements at a time,
      // which are then reduced per window entry in fixed element order (as
      // phase C below) and added to the window, group after group.  p = 3:
      // element rounds (measured faster there).
      if (EQE && c0 == 0) {
#pragma unroll
        for (int k = 0; k < NEQV; k++) {
          const int t = tid + k * kBlock;
          if (t < NT * 64) eqw[buf][t] = eqv[k];
        }
      }
      constexpr bool STAGED = P <= 2;
      if constexpr (STAGED) {
      constexpr int GS0 = (CH * SLQ * LL::GEO) / (NEN * P1 * NS);
      constexpr int GS = GS0 < 1 ? 1 : (GS0 > CH ? CH : GS0);
      T (*stg)[NEN][P1][NS] = reinterpret_cast<T (*)[NEN][P1][NS]>(&geo[0][0]);
      for (int g0 = 0; g0 < nch; g0 += GS) {
        const int g1 = g0 + GS < nch ? g0 + GS : nch;
        lds_barrier(); // geometry reads (or the previous group) done
        if (has && le >= g0 && le < g1) {
#pragma unroll
          for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
            for (int s = 0; s < NS; s++) stg[le - g0][b][a1][s] = acc[a1][s];
        }
        __syncthreads();
        for (int t = tid; t < P1 * P1 * W1 * W1; t += kBlock) {
          const int a1 = t % P1, r = t / P1;
          const int cb1 = r % P1, r2 = r / P1;
          const int cjr = r2 % W1, ckr = r2 / W1;
          const int cjj = jc - P + cjr, ckk = kc - P + ckr;
          T sm[NS];
#pragma unroll
          for (int m = 0; m < NS; m++) sm[m] = 0;
          bool any = false;
          for (int l = g0; l < g1; l++) {
            const int cb2 = cjj - be2[c0 + l], cb3 = ckk - be3[c0 + l];
            if (cb2 < 0 || cb2 > P || cb3 < 0 || cb3 > P) continue;
            any = true;
            const int cb = cb1 + P1 * (cb2 + P1 * cb3);
#pragma unroll
            for (int m = 0; m < NS; m++) sm[m] += stg[l - g0][cb][a1][m];
          }
          if (!any) continue;
          const int slot = (e1 + a1) % P1;
          const int di = cb1 - a1, dj = cjj - jc, dk = ckk - kc;
          const int nbr = ((dk + P) * W1 + (dj + P)) * W1 + (di + P);
          if (NC == 1) {
            win[slot][0][nbr][0] += sm[0];
          } else if (SACC) {
            const T tr = sm[0] + sm[4] + sm[8];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
              for (int j = 0; j < 3; j++)
                win[slot][i][nbr][j] += lam * sm[3 * i + j] + mu * sm[3 * j + i] + (i == j ? mu * tr : T(0));
          } else {
#pragma unroll
            for (int j = 0; j < 3; j++) win[slot][0][nbr][j] += sm[j];
          }
        }
      }
      } else {
      // Each thread's window entries are distinct, so all of them are read
      // first and written after (one LDS round trip per round, not one per
      // entry).
      const int cj = e2 + b2, ck = e3 + b3;
      constexpr int NV = NC == 1 ? 1 : (SACC ? 9 : 3); // window values per row function
      T* wp[P1];
#pragma unroll
      for (int a1 = 0; a1 < P1; a1++) {
        const int slot = (e1 + a1) % P1;
        const int di = b1 - a1, dj = cj - jc, dk = ck - kc;
        const int nbr = ((dk + P) * W1 + (dj + P)) * W1 + (di + P);
        wp[a1] = &win[slot][0][nbr][0];
      }
      T add[P1][NV];
#pragma unroll
      for (int a1 = 0; a1 < P1; a1++) {
        const T* sv = acc[a1];
        if (NC == 1) {
          add[a1][0] = sv[0];
        } else if (SACC) {
          const T tr = sv[0] + sv[4] + sv[8];
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
              add[a1][3 * i + j] = lam * sv[3 * i + j] + mu * sv[3 * j + i] + (i == j ? mu * tr : T(0));
        } else {
#pragma unroll
          for (int j = 0; j < 3; j++) add[a1][j] = sv[j];
        }
      }
      // offset of value v of a row function inside win[slot]: (i, j) of the
      // 3x3 block -> [i][nbr][j]
      constexpr int ISTR = LL::WN * NC;
      for (int r = 0; r < nch; r++) {
        lds_barrier();
        if (has && le == r) {
          T cur[P1][NV];
#pragma unroll
          for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
            for (int v = 0; v < NV; v++)
              cur[a1][v] = wp[a1][(NV == 9 ? v / 3 : 0) * ISTR + (NV == 9 ? v % 3 : v)];
#pragma unroll
          for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
            for (int v = 0; v < NV; v++)
              wp[a1][(NV == 9 ? v / 3 : 0) * ISTR + (NV == 9 ? v % 3 : v)] = cur[a1][v] + add[a1][v];
        }
      }
      }
      } else if constexpr (SEMI) {
      // Larger Gauss rules: loop over the (q2, q3) rows with the q1 row
      // unrolled; the next point's geometry row is loaded into registers while
      // the current one is used (across rows too), the column node's
      // direction-1 basis values are held in registers.
      for (int t = tid; t < nch * NEN; t += kBlock) {
        const int le = t / NEN, b = t % NEN;
        const int b1 = b % P1, b2 = (b / P1) % P1, b3 = b / (P1 * P1);
        const int e2 = be2[c0 + le], e3 = be3[c0 + le];
        const int m2 = e2 - jc + P, m3 = e3 - kc + P;
        T vn1[NG], vd1[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) {
          vn1[q] = tN1[q][b1];
          vd1[q] = tD1[q][b1];
        }
        T acc[P1][NS];
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
          for (int s = 0; s < NS; s++) acc[a1][s] = 0;
        const T* __restrict__ Gbase = geo[le * GP3];
        T Gc[LL::GEO];
#pragma unroll
        for (int m = 0; m < LL::GEO; m++) Gc[m] = Gbase[m];
        for (int g23 = 0; g23 < NG * NG; g23++) {
          const int q2 = g23 % NG, q3 = g23 / NG;
          const T n3 = tN3[m3][q3][b3], d3 = tD3[m3][q3][b3];
          const T n2 = tN2[m2][q2][b2], d2 = tD2[m2][q2][b2];
          const T n23 = n2 * n3, d23 = d2 * n3, n2d3 = n2 * d3;
#pragma unroll
          for (int q1 = 0; q1 < NG; q1++) {
            // next row (the last point reloads itself: in bounds, unused)
            const int qn = g23 * NG + q1 + 1 < GP3 ? g23 * NG + q1 + 1 : GP3 - 1;
            T Gn[LL::GEO];
#pragma unroll
            for (int m = 0; m < LL::GEO; m++) Gn[m] = Gbase[qn * LL::GEO + m];
            const T r0 = vd1[q1] * n23, r1 = vn1[q1] * d23, r2 = vn1[q1] * n2d3;
            const T* G = Gc;
            if (NC == 1) {
#pragma unroll
              for (int a1 = 0; a1 < P1; a1++)
                acc[a1][0] += r0 * G[3 * a1] + r1 * G[3 * a1 + 1] + r2 * G[3 * a1 + 2];
            } else {
              const T gb0 = r0 * G[0] + r1 * G[3] + r2 * G[6];
              const T gb1 = r0 * G[1] + r1 * G[4] + r2 * G[7];
              const T gb2 = r0 * G[2] + r1 * G[5] + r2 * G[8];
#pragma unroll
              for (int a1 = 0; a1 < P1; a1++) {
                const T wa0 = G[9 + 3 * a1], wa1 = G[10 + 3 * a1], wa2 = G[11 + 3 * a1];
                if (SACC) {
                  acc[a1][0] += wa0 * gb0; acc[a1][1] += wa0 * gb1; acc[a1][2] += wa0 * gb2;
                  acc[a1][3] += wa1 * gb0; acc[a1][4] += wa1 * gb1; acc[a1][5] += wa1 * gb2;
                  acc[a1][6] += wa2 * gb0; acc[a1][7] += wa2 * gb1; acc[a1][8] += wa2 * gb2;
                } else {
                  const T wa[3] = {wa0, wa1, wa2}, gb[3] = {gb0, gb1, gb2};
                  const int ic = split * NCR;
                  const T dot = wa0 * gb0 + wa1 * gb1 + wa2 * gb2;
#pragma unroll
                  for (int jj = 0; jj < 3; jj++)
                    acc[a1][jj] += lam * wa[ic] * gb[jj] + mu * wa[jj] * gb[ic] +
                                   (jj == ic ? mu * dot : T(0));
                }
              }
            }
#pragma unroll
            for (int m = 0; m < LL::GEO; m++) Gc[m] = Gn[m];
          }
        }
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
          for (int s = 0; s < NS; s++) stage[le][b][a1][s] = acc[a1][s];
      }
      } else if constexpr (!PIPE) {
      // one thread per (element, column node b), Gauss points in fixed order
      for (int t = tid; t < nch * NEN; t += kBlock) {
        const int le = t / NEN, b = t % NEN;
        const int b1 = b % P1, b2 = (b / P1) % P1, b3 = b / (P1 * P1);
        const int e2 = be2[c0 + le], e3 = be3[c0 + le];
        const int m2 = e2 - jc + P, m3 = e3 - kc + P;
        T acc[P1][NS];
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
          for (int s = 0; s < NS; s++) acc[a1][s] = 0;
        for (int q3 = 0; q3 < NG; q3++) {
          const T n3 = tN3[m3][q3][b3], d3 = tD3[m3][q3][b3];
          for (int q2 = 0; q2 < NG; q2++) {
            const T n2 = tN2[m2][q2][b2], d2 = tD2[m2][q2][b2];
            const T n23 = n2 * n3, d23 = d2 * n3, n2d3 = n2 * d3;
#pragma unroll
            for (int q1 = 0; q1 < NG; q1++) {
              const int pr = le * GP3 + (q3 * NG + q2) * NG + q1;
              const T n1 = tN1[q1][b1], d1 = tD1[q1][b1];
              const T r0 = d1 * n23, r1 = n1 * d23, r2 = n1 * n2d3;
              const T* G = geo[pr];
              if (NC == 1) {
#pragma unroll
                for (int a1 = 0; a1 < P1; a1++)
                  acc[a1][0] += r0 * G[3 * a1] + r1 * G[3 * a1 + 1] + r2 * G[3 * a1 + 2];
              } else {
                const T gb0 = r0 * G[0] + r1 * G[3] + r2 * G[6];
                const T gb1 = r0 * G[1] + r1 * G[4] + r2 * G[7];
                const T gb2 = r0 * G[2] + r1 * G[5] + r2 * G[8];
#pragma unroll
                for (int a1 = 0; a1 < P1; a1++) {
                  const T wa0 = G[9 + 3 * a1], wa1 = G[10 + 3 * a1], wa2 = G[11 + 3 * a1];
                  if (SACC) {
                    acc[a1][0] += wa0 * gb0; acc[a1][1] += wa0 * gb1; acc[a1][2] += wa0 * gb2;
                    acc[a1][3] += wa1 * gb0; acc[a1][4] += wa1 * gb1; acc[a1][5] += wa1 * gb2;
                    acc[a1][6] += wa2 * gb0; acc[a1][7] += wa2 * gb1; acc[a1][8] += wa2 * gb2;
                  } else {
                    const T wa[3] = {wa0, wa1, wa2}, gb[3] = {gb0, gb1, gb2};
                    const int ic = split * NCR;
                    const T dot = wa0 * gb0 + wa1 * gb1 + wa2 * gb2;
#pragma unroll
                    for (int jj = 0; jj < 3; jj++)
                      acc[a1][jj] += lam * wa[ic] * gb[jj] + mu * wa[jj] * gb[ic] +
                                     (jj == ic ? mu * dot : T(0));
                  }
                }
              }
            }
          }
        }
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
          for (int s = 0; s < NS; s++) stage[le][b][a1][s] = acc[a1][s];
      }
      } else {
      // One thread per (element, column node b); the Gauss points are fully
      // unrolled with the next point's geometry row loaded into registers
      // while the current one is used, and the column node's 1D basis
      // values held in registers.
      for (int t = tid; t < nch * NEN; t += kBlock) {
        const int le = t / NEN, b = t % NEN;
        const int b1 = b % P1, b2 = (b / P1) % P1, b3 = b / (P1 * P1);
        const int e2 = be2[c0 + le], e3 = be3[c0 + le];
        const int m2 = e2 - jc + P, m3 = e3 - kc + P;
        T vn1[NG], vd1[NG], vn2[NG], vd2[NG], vn3[NG], vd3[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) {
          vn1[q] = tN1[q][b1]; vd1[q] = tD1[q][b1];
          vn2[q] = tN2[m2][q][b2]; vd2[q] = tD2[m2][q][b2];
          vn3[q] = tN3[m3][q][b3]; vd3[q] = tD3[m3][q][b3];
        }
        T acc[P1][NS];
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
          for (int s = 0; s < NS; s++) acc[a1][s] = 0;
        const T* __restrict__ Gbase = geo[le * GP3];
        T Gc[LL::GEO];
#pragma unroll
        for (int m = 0; m < LL::GEO; m++) Gc[m] = Gbase[m];
#pragma unroll
        for (int q = 0; q < GP3; q++) {
          T Gn[LL::GEO];
          if (q + 1 < GP3) {
#pragma unroll
            for (int m = 0; m < LL::GEO; m++) Gn[m] = Gbase[(q + 1) * LL::GEO + m];
          }
          const int q1 = q % NG, q2 = (q / NG) % NG, q3 = q / (NG * NG);
          const T n23 = vn2[q2] * vn3[q3], d23 = vd2[q2] * vn3[q3], n2d3 = vn2[q2] * vd3[q3];
          const T r0 = vd1[q1] * n23, r1 = vn1[q1] * d23, r2 = vn1[q1] * n2d3;
          const T* G = Gc;
          if (NC == 1) {
#pragma unroll
            for (int a1 = 0; a1 < P1; a1++)
              acc[a1][0] += r0 * G[3 * a1] + r1 * G[3 * a1 + 1] + r2 * G[3 * a1 + 2];
          } else {
            const T gb0 = r0 * G[0] + r1 * G[3] + r2 * G[6];
            const T gb1 = r0 * G[1] + r1 * G[4] + r2 * G[7];
            const T gb2 = r0 * G[2] + r1 * G[5] + r2 * G[8];
#pragma unroll
            for (int a1 = 0; a1 < P1; a1++) {
              const T wa0 = G[9 + 3 * a1], wa1 = G[10 + 3 * a1], wa2 = G[11 + 3 * a1];
              if (SACC) {
                acc[a1][0] += wa0 * gb0; acc[a1][1] += wa0 * gb1; acc[a1][2] += wa0 * gb2;
                acc[a1][3] += wa1 * gb0; acc[a1][4] += wa1 * gb1; acc[a1][5] += wa1 * gb2;
                acc[a1][6] += wa2 * gb0; acc[a1][7] += wa2 * gb1; acc[a1][8] += wa2 * gb2;
              } else {
                const T wa[3] = {wa0, wa1, wa2}, gb[3] = {gb0, gb1, gb2};
                const int ic = split * NCR;
                const T dot = wa0 * gb0 + wa1 * gb1 + wa2 * gb2;
#pragma unroll
                for (int jj = 0; jj < 3; jj++)
                  acc[a1][jj] += lam * wa[ic] * gb[jj] + mu * wa[jj] * gb[ic] +
                                 (jj == ic ? mu * dot : T(0));
              }
            }
          }
          if (q + 1 < GP3) {
#pragma unroll
            for (int m = 0; m < LL::GEO; m++) Gc[m] = Gn[m];
          }
        }
#pragma unroll
        for (int a1 = 0; a1 < P1; a1++)
#pragma unroll
          for (int s = 0; s < NS; s++) stage[le][b][a1][s] = acc[a1][s];
      }
      }
      __syncthreads();

      // ---- phase C: ordered reduction of the chunk into the row window ------
      if (!ORD)
      for (int t = tid; t < P1 * P1 * W1 * W1; t += kBlock) {
        const int a1 = t % P1, r = t / P1;
        const int b1 = r % P1, r2 = r / P1;
        const int cjr = r2 % W1, ckr = r2 / W1;
        const int cj = jc - P + cjr, ck = kc - P + ckr;
        T s[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) s[m] = 0;
        bool any = false;
        for (int le = 0; le < nch; le++) {
          const int b2 = cj - be2[c0 + le], b3 = ck - be3[c0 + le];
          if (b2 < 0 || b2 > P || b3 < 0 || b3 > P) continue;
          any = true;
          const int b = b1 + P1 * (b2 + P1 * b3);
#pragma unroll
          for (int m = 0; m < NS; m++) s[m] += stage[le][b][a1][m];
        }
        if (!any) continue;
        const int slot = (e1 + a1) % P1;
        const int di = b1 - a1, dj = cj - jc, dk = ck - kc;
        const int nbr = ((dk + P) * W1 + (dj + P)) * W1 + (di + P);
        if (NC == 1) {
          win[slot][0][nbr][0] += s[0];
        } else if (SACC) {
          const T tr = s[0] + s[4] + s[8];
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
              win[slot][i][nbr][j] += lam * s[3 * i + j] + mu * s[3 * j + i] + (i == j ? mu * tr : T(0));
        } else {
#pragma unroll
          for (int j = 0; j < 3; j++) win[slot][0][nbr][j] += s[j];
        }
      }
      __syncthreads();
    }

    // ---- phase D (last layer): rows e1 .. e1+p are complete ------------------
    if (e1 == a.nel1 - 1) {
      for (int i = e1; i <= e1 + P; i++) {
        const int slot = i % P1;
        row_out(i, i == e1, buf, false);
        __syncthreads();
        for (int t = tid; t < NCR * WN * NC; t += kBlock) (&win[slot][0][0][0])[t] = T(0);
        if (tid < NC) rwin[slot][tid] = T(0);
        __syncthreads();
      }
    }
  }
}

// Single patch: the LDS arrays are separate __shared__ variables.  The
// Gauss-point slab variant (HALF = 2) is built for three workgroups per CU.
template <int P, int NG, int NC, typename T, int SPLIT, bool HAS_RHS, int HALF>
__global__ __launch_bounds__(kBlock, HALF == 2 ? 3 : 1) void k_assemble_lines(
    const AsmArgs* __restrict__ args)
{
  ConstArgs& a = *const_args(args);
  using LL = LineLds<P, NG, NC, T, SPLIT, HAS_RHS, HALF>;
  constexpr int P1 = P + 1;
  // register-pipelined phase B where the unrolled Gauss loop stays small;
  // ordered window updates (no stage) where the whole bundle is one chunk
  constexpr bool PIPE = LL::GP3 <= 27;
  constexpr bool ORD = (HALF == 0 || HALF == 2) && LL::CH * LL::NEN <= kBlock;
  __shared__ T win[P1][LL::NCR][LL::WN][NC];
  __shared__ T geo[LL::CH * LL::SLQ][LL::GEO];
  __shared__ T stage[ORD ? 1 : LL::CH][LL::NEN][P1][LL::NS];
  __shared__ T rstage[HAS_RHS ? LL::CH * LL::SLQ * (ORD ? P1 : LL::NRHS) : 1];
  __shared__ T rwin[P1][NC];
  __shared__ T tN1[NG][P1], tD1[NG][P1];
  __shared__ T tN2[P1][NG][P1], tD2[P1][NG][P1];
  __shared__ T tN3[P1][NG][P1], tD3[P1][NG][P1];
  __shared__ int be2[LL::NB], be3[LL::NB];
  __shared__ int eqw[2][LL::PFD ? LL::NT * 64 : 1];
  __shared__ int64_t pbase[2][LL::NCR];
  __shared__ int prow[2][LL::NCR];
  __shared__ int nbundle;
  line_impl<P, NG, NC, T, SPLIT, HAS_RHS, HALF, PIPE, ORD, !PIPE>(a, blockIdx.x, blockIdx.y, win, geo, stage,
                                                      rstage, rwin, tN1, tD1, tN2, tD2, tN3, tD3,
                                                      be2, be3, eqw, pbase, prow, &nbundle);
}

// Persistent batch assembly over many independent patches, one launch per
// degree p = 1..3 (ng = p+1) present in the batch, each with the single-patch
// kernel's own phase-B and window-update variants and its own register and
// LDS budget.  Work items are (patch, line) pairs sorted by decreasing cost
// within the degree; blocks take the next item from the degree's queue
// counter, so load balances over patches of uneven size.  Each line's rows
// are written by exactly one block, so the result does not depend on which
// block takes which item.  HALF: the element-chunk variant chosen for the
// degree (AsmArgs::variant = 2*(p-1) + HALF for every patch of it).
template <int P, typename T, bool HAS_RHS, int HALF>
__global__ __launch_bounds__(kBlock, 3) void k_assemble_batch(const AsmArgs* __restrict__ args,
                                                           const int2* __restrict__ items,
                                                           int nitems, int* __restrict__ queue)
{
  constexpr int NG = P + 1;
  using LL = LineLds<P, NG, 1, T, 1, HAS_RHS, HALF>;
  constexpr int P1 = P + 1;
  constexpr bool PIPE = LL::GP3 <= 27;
  constexpr bool ORD = (HALF == 0 || HALF == 2) && LL::CH * LL::NEN <= kBlock;
  __shared__ T win[P1][LL::NCR][LL::WN][1];
  __shared__ T geo[LL::CH * LL::SLQ][LL::GEO];
  __shared__ T stage[ORD ? 1 : LL::CH][LL::NEN][P1][LL::NS];
  __shared__ T rstage[HAS_RHS ? LL::CH * LL::SLQ * (ORD ? P1 : LL::NRHS) : 1];
  __shared__ T rwin[P1][1];
  __shared__ T tN1[NG][P1], tD1[NG][P1];
  __shared__ T tN2[P1][NG][P1], tD2[P1][NG][P1];
  __shared__ T tN3[P1][NG][P1], tD3[P1][NG][P1];
  __shared__ int be2[LL::NB], be3[LL::NB];
  __shared__ int eqw[2][LL::PFD ? LL::NT * 64 : 1];
  __shared__ int64_t pbase[2][LL::NCR];
  __shared__ int prow[2][LL::NCR];
  __shared__ int nbundle;
  __shared__ int next;
  while (true) {
    if (threadIdx.x == 0) next = atomicAdd(queue, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(next);
    if (it >= nitems) break;
    const int2 w = items[it];
    ConstArgs& a = const_args(args)[w.x];
    line_impl<P, NG, 1, T, 1, HAS_RHS, HALF, PIPE, ORD, !PIPE>(a, w.y, 0, win, geo, stage, rstage,
                                                               rwin, tN1, tD1, tN2, tD2, tN3, tD3,
                                                               be2, be3, eqw, pbase, prow, &nbundle);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Symbolic pattern: row lengths, scan, columns.  Two nodes couple when their
// index boxes are within p in every direction (open knots, simple interior
// knots); eliminated DOFs are dropped.  Column order = DOF order =
// equation order (the SAM numbering is monotone), i.e. sorted.
__global__ void k_row_lengths(int n1, int n2, int n3, int p, int nc, const int* __restrict__ eq,
                              int64_t* __restrict__ rowlen, int ndof)
{
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndof) return;
  const int r = eq[d];
  if (r < 0) return;
  const int node = d / nc;
  const int i = node % n1, j = (node / n1) % n2, k = node / (n1 * n2);
  int64_t cnt = 0;
  for (int kk = max(0, k - p); kk <= min(n3 - 1, k + p); kk++)
    for (int jj = max(0, j - p); jj <= min(n2 - 1, j + p); jj++)
      for (int ii = max(0, i - p); ii <= min(n1 - 1, i + p); ii++) {
        const int64_t nn = (int64_t)ii + (int64_t)n1 * (jj + (int64_t)n2 * kk);
        for (int c = 0; c < nc; c++) cnt += eq[nn * nc + c] >= 0;
      }
  rowlen[r] = cnt;
}

// One wave per row DOF; lanes write consecutive columns (ballot compaction).
__global__ void k_fill_columns(int n1, int n2, int n3, int p, int nc, const int* __restrict__ eq,
                               const int64_t* __restrict__ rowptr, int32_t* __restrict__ colind,
                               int ndof)
{
  const int d = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (d >= ndof) return;
  const int r = eq[d];
  if (r < 0) return;
  const int node = d / nc;
  const int i = node % n1, j = (node / n1) % n2, k = node / (n1 * n2);
  const int W1 = 2 * p + 1;
  const int total = W1 * W1 * W1 * nc;
  int64_t pos = rowptr[r];
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    bool valid = false;
    int col = 0;
    if (t < total) {
      const int nbr = t / nc, cc = t % nc;
      const int ii = i + nbr % W1 - p, jj = j + (nbr / W1) % W1 - p, kk = k + nbr / (W1 * W1) - p;
      if (ii >= 0 && ii < n1 && jj >= 0 && jj < n2 && kk >= 0 && kk < n3) {
        const int64_t nn = (int64_t)ii + (int64_t)n1 * (jj + (int64_t)n2 * kk);
        col = eq[nn * nc + cc];
        valid = col >= 0;
      }
    }
    const unsigned long long m = __ballot(valid);
    if (valid) colind[pos + __popcll(m & ((1ull << lane) - 1ull))] = col;
    pos += __popcll(m);
  }
}

// Exclusive scan of int64 (block-local scan, scan of block sums, add).
constexpr int kScanBlock = 1024;

__global__ void k_scan_local(const int64_t* __restrict__ in, int64_t* __restrict__ out,
                             int64_t* __restrict__ bsum, int n)
{
  __shared__ int64_t s[kScanBlock];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  s[threadIdx.x] = i < n ? in[i] : 0;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int64_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  if (i < n) out[i + 1] = s[threadIdx.x]; // inclusive into out[1..]
  if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_sums(int64_t* __restrict__ bsum, int nb)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int64_t run = 0;
    for (int b = 0; b < nb; b++) {
      const int64_t v = bsum[b];
      bsum[b] = run;
      run += v;
    }
  }
}

__global__ void k_scan_add(int64_t* __restrict__ out, const int64_t* __restrict__ bsum, int n)
{
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  if (i < n) out[i + 1] += bsum[blockIdx.x];
  if (i == 0) out[0] = 0;
}

// Several patches: the global pattern is the union of the patches' node
// couplings (nodes within +-p in every direction of one patch), built on the
// host (SAM::getDofCouplings over all patches, src/LinAlg/SAM.C:377-427).
int host_union_pattern(const Context& c, const std::vector<SubPatch>& subs, int neq,
                       std::vector<int64_t>& rowptr, std::vector<int32_t>& colind)
{
  const int nc = c.ncomp;
  std::vector<std::vector<int>> cols(neq > 0 ? neq : 1);
  for (const SubPatch& sp : subs) {
    const PatchRec& P = c.patches[sp.patch];
    const int p = P.p[0], n1 = P.n[0], n2 = P.n[1], n3 = P.n[2];
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          const int node = i + n1 * (j + n2 * k);
          for (int ci = 0; ci < nc; ci++) {
            const int r = sp.h_leq[node * nc + ci];
            if (r < 0) continue;
            std::vector<int>& row = cols[r];
            for (int kk = std::max(0, k - p); kk <= std::min(n3 - 1, k + p); kk++)
              for (int jj = std::max(0, j - p); jj <= std::min(n2 - 1, j + p); jj++)
                for (int ii = std::max(0, i - p); ii <= std::min(n1 - 1, i + p); ii++) {
                  const int nb = ii + n1 * (jj + n2 * kk);
                  for (int cj = 0; cj < nc; cj++) {
                    const int e = sp.h_leq[nb * nc + cj];
                    if (e >= 0) row.push_back(e);
                  }
                }
          }
        }
  }
  rowptr.assign(neq + 1, 0);
  for (int r = 0; r < neq; r++) {
    std::vector<int>& row = cols[r];
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    rowptr[r + 1] = rowptr[r] + (int64_t)row.size();
  }
  const int64_t nnz = rowptr[neq];
  colind.assign(nnz > 0 ? nnz : 1, 0);
  for (int r = 0; r < neq; r++)
    std::copy(cols[r].begin(), cols[r].end(), colind.begin() + rowptr[r]);
  return IFEM_GPU_OK;
}

// Positions of every (local row, valid window column) of each patch in the
// shared CSR, in the order the line kernels visit them (window offsets
// (dk, dj, di) ascending, components minor), found once by binary search.
int sub_position_maps(Context& c, std::vector<SubPatch>& subs, const std::vector<int64_t>& rowptr,
                      const std::vector<int32_t>& colind)
{
  const int nc = c.ncomp;
  for (SubPatch& sp : subs) {
    const PatchRec& P = c.patches[sp.patch];
    const int p = P.p[0], n1 = P.n[0], n2 = P.n[1], n3 = P.n[2];
    const int64_t nd = (int64_t)n1 * n2 * n3 * nc;
    std::vector<int64_t> pptr(nd + 1, 0), pmap;
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          const int node = i + n1 * (j + n2 * k);
          for (int ci = 0; ci < nc; ci++) {
            const int64_t d = (int64_t)node * nc + ci;
            const int r = sp.h_leq[d];
            if (r >= 0) {
              const int32_t* rb = colind.data() + rowptr[r];
              const int32_t* re = colind.data() + rowptr[r + 1];
              for (int kk = k - p; kk <= k + p; kk++)
                for (int jj = j - p; jj <= j + p; jj++)
                  for (int ii = i - p; ii <= i + p; ii++) {
                    if (ii < 0 || ii >= n1 || jj < 0 || jj >= n2 || kk < 0 || kk >= n3) continue;
                    const int nb = ii + n1 * (jj + n2 * kk);
                    for (int cj = 0; cj < nc; cj++) {
                      const int e = sp.h_leq[(int64_t)nb * nc + cj];
                      if (e < 0) continue;
                      const int32_t* f = std::lower_bound(rb, re, e);
                      if (f == re || *f != e)
                        return fail(IFEM_GPU_EINVAL, "pattern: column missing from the union row");
                      pmap.push_back(rowptr[r] + (f - rb));
                    }
                  }
            }
            pptr[d + 1] = (int64_t)pmap.size();
          }
        }
    hipFree(sp.d_pmap);
    hipFree(sp.d_pptr);
    sp.d_pmap = sp.d_pptr = nullptr;
    IFEM_HIP(hipMalloc(&sp.d_pptr, sizeof(int64_t) * pptr.size()));
    IFEM_HIP(hipMemcpy(sp.d_pptr, pptr.data(), sizeof(int64_t) * pptr.size(), hipMemcpyHostToDevice));
    IFEM_HIP(hipMalloc(&sp.d_pmap, sizeof(int64_t) * (pmap.empty() ? 1 : pmap.size())));
    if (!pmap.empty())
      IFEM_HIP(hipMemcpy(sp.d_pmap, pmap.data(), sizeof(int64_t) * pmap.size(), hipMemcpyHostToDevice));
  }
  return IFEM_GPU_OK;
}

int device_union_pattern(Context& c, std::vector<SubPatch>& subs, int neq, int64_t** rowptr_out,
                         int32_t** colind_out, int64_t* nnz_out);

// Several patches sharing nodes: the union pattern and the per-patch
// position maps are built on the device (pattern.hip).
static int build_pattern_multi(Context& c)
{
  const int neq = c.neq;
  int64_t* rowptr = nullptr;
  int32_t* colind = nullptr;
  int64_t nnz = 0;
  int rc0 = device_union_pattern(c, c.multi, neq, &rowptr, &colind, &nnz);
  if (rc0) return rc0;
  hipFree(c.d_rowptr);
  hipFree(c.d_colind);
  hipFree(c.d_val);
  hipFree(c.d_val32);
  hipFree(c.d_rhs);
  c.d_rowptr = rowptr;
  c.d_colind = colind;
  c.d_val = nullptr;
  c.d_val32 = nullptr;
  c.d_rhs = nullptr;
  c.fp64_ready = c.fp32_ready = false;
  c.nnz = nnz;
  IFEM_HIP(hipMalloc(&c.d_val, sizeof(double) * (nnz > 0 ? nnz : 1)));
  IFEM_HIP(hipMalloc(&c.d_rhs, sizeof(double) * (neq > 0 ? neq : 1)));
  IFEM_HIP(hipMemsetAsync(c.d_rhs, 0, sizeof(double) * (neq > 0 ? neq : 1), c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  c.pattern_ready = true;
  return solver_setup(c);
}

int build_pattern(Context& c)
{
  if (c.mpc) return mpc_build_pattern(c);
  if (!c.multi.empty()) return build_pattern_multi(c);
  const PatchRec& P = c.patches[c.patch];
  if (P.p[0] != P.p[1] || P.p[0] != P.p[2])
    return fail(IFEM_GPU_EUNSUPPORTED, "init_assembly: the GPU path needs equal degrees in all directions");
  const int p = P.p[0];
  hipFree(c.d_rowptr);
  hipFree(c.d_colind);
  hipFree(c.d_val);
  hipFree(c.d_val32);
  hipFree(c.d_rhs);
  c.d_rowptr = nullptr;
  c.d_colind = nullptr;
  c.d_val = nullptr;
  c.d_val32 = nullptr;
  c.d_rhs = nullptr;
  c.fp64_ready = c.fp32_ready = false;
  const int neq = c.neq;
  int64_t* rowlen = nullptr;
  int64_t* bsum = nullptr;
  const int nb = (neq + kScanBlock - 1) / kScanBlock;
  IFEM_HIP(hipMalloc(&c.d_rowptr, sizeof(int64_t) * (neq + 1)));
  IFEM_HIP(hipMalloc(&rowlen, sizeof(int64_t) * (neq > 0 ? neq : 1)));
  IFEM_HIP(hipMalloc(&bsum, sizeof(int64_t) * (nb > 0 ? nb : 1)));
  IFEM_HIP(hipMemsetAsync(c.d_rowptr, 0, sizeof(int64_t) * (neq + 1), c.stream));
  if (neq > 0) {
    k_row_lengths<<<(c.ndof + 255) / 256, 256, 0, c.stream>>>(P.n[0], P.n[1], P.n[2], p, c.ncomp,
                                                              c.d_eq, rowlen, c.ndof);
    k_scan_local<<<nb, kScanBlock, 0, c.stream>>>(rowlen, c.d_rowptr, bsum, neq);
    k_scan_sums<<<1, 64, 0, c.stream>>>(bsum, nb);
    k_scan_add<<<nb, kScanBlock, 0, c.stream>>>(c.d_rowptr, bsum, neq);
  }
  IFEM_HIP(hipGetLastError());
  int64_t nnz = 0;
  IFEM_HIP(hipMemcpyAsync(&nnz, c.d_rowptr + neq, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  hipFree(rowlen);
  hipFree(bsum);
  c.nnz = nnz;
  IFEM_HIP(hipMalloc(&c.d_colind, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
  if (neq > 0) {
    const int waves = 4;
    k_fill_columns<<<(c.ndof + waves - 1) / waves, 64 * waves, 0, c.stream>>>(
        P.n[0], P.n[1], P.n[2], p, c.ncomp, c.d_eq, c.d_rowptr, c.d_colind, c.ndof);
  }
  IFEM_HIP(hipGetLastError());
  IFEM_HIP(hipMalloc(&c.d_val, sizeof(double) * (nnz > 0 ? nnz : 1)));
  IFEM_HIP(hipMalloc(&c.d_rhs, sizeof(double) * (neq > 0 ? neq : 1)));
  IFEM_HIP(hipMemsetAsync(c.d_rhs, 0, sizeof(double) * (neq > 0 ? neq : 1), c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  c.pattern_ready = true;
  return solver_setup(c);
}

// ---------------------------------------------------------------------------
// Host dispatch of the numeric assembly.
// Basis tables of patch Pt for ng Gauss points per direction (owned by the
// context's cache; built by k_basis_tables on first use).
using Tables = Context::BasisTables;

template <int P>
static int make_tables(Context& c, const PatchRec& Pt, int ng, Tables& tb, GaussArgs& ga)
{
  for (const Tables& t : c.basis_cache)
    if (t.knots0 == Pt.d_knots[0] && t.ng == ng) {
      tb = t;
      return IFEM_GPU_OK;
    }
  Tables t;
  t.knots0 = Pt.d_knots[0];
  t.ng = ng;
  for (int d = 0; d < 3; d++) {
    const size_t cnt = (size_t)Pt.nel[d] * ng * (P + 1);
    if (hipMalloc(&t.tabN[d], sizeof(double) * cnt) != hipSuccess ||
        hipMalloc(&t.tabD[d], sizeof(double) * cnt) != hipSuccess ||
        hipMalloc(&t.span[d], sizeof(double) * Pt.nel[d]) != hipSuccess) {
      for (int e = 0; e <= d; e++) {
        hipFree(t.tabN[e]);
        hipFree(t.tabD[e]);
        hipFree(t.span[e]);
      }
      return fail(IFEM_GPU_EDEVICE, "assemble: out of device memory for the basis tables");
    }
  }
  for (int d = 0; d < 3; d++) {
    const int tot = Pt.nel[d] * ng;
    k_basis_tables<P><<<(tot + 127) / 128, 128, 0, c.stream>>>(Pt.d_knots[d], Pt.nel[d], ng, ga,
                                                              t.tabN[d], t.tabD[d], t.span[d]);
  }
  c.basis_cache.push_back(t);
  tb = t;
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

// slot: 0 conductivity, 1 source, 2 exact solution (one device buffer each,
// reused by later calls on the same stream)
static int upload_function(Context& c, const ifem_gpu_function& f, DevFunction& df, int slot)
{
  df.type = f.type;
  df.nmodes = f.nmodes;
  df.scale = f.scale;
  df.mean = f.mean;
  df.modes = nullptr;
  if (f.type < 0 || f.type > IFEM_GPU_FN_EXP_FOURIER)
    return fail(IFEM_GPU_EINVAL, "assemble: unknown field type");
  if ((f.type == IFEM_GPU_FN_FOURIER || f.type == IFEM_GPU_FN_EXP_FOURIER) && f.nmodes > 0) {
    if (!f.modes) return fail(IFEM_GPU_EINVAL, "assemble: null mode table");
    const size_t need = (size_t)5 * f.nmodes;
    if (c.modes_cap[slot] < need) {
      hipFree(c.d_modes[slot]);
      c.d_modes[slot] = nullptr;
      c.modes_cap[slot] = 0;
      IFEM_HIP(hipMalloc(&c.d_modes[slot], sizeof(double) * need));
      c.modes_cap[slot] = need;
    }
    double* d = c.d_modes[slot];
    IFEM_HIP(hipMemcpyAsync(d, f.modes, sizeof(double) * need, hipMemcpyHostToDevice, c.stream));
    df.modes = d;
    mode_bounds(f.modes, f.nmodes, df.kbound, df.pbound);
  } else {
    df.nmodes = 0;
  }
  return IFEM_GPU_OK;
}

// Element-chunk variant of the line kernel for (p, ng, ncomp, precision):
// the tuned (or set) choice if ifem_gpu_autotune_assembly or
// ifem_gpu_set_assembly_variant ran, else the built-in default below.
int& asm_half_slot(Context& c, int p, int ng, int nc, bool fp32)
{
  return c.asm_half[p][ng > p + 1 ? 1 : 0][nc == 3 ? 1 : 0][fp32 ? 1 : 0];
}

static int asm_half(Context& c, int p, int ng, int nc, bool fp32)
{
  const int h = asm_half_slot(c, p, ng, nc, fp32);
  if (h >= 0) return h;
  // FP32 elasticity at p = 2, 3x3x3 points: Gauss-point slabs (measured on
  // 64^3: 3.9 ms against 5.2 ms for full chunks; in FP64 the slabs lose,
  // 7.3 against 6.4 ms, and in scalar problems too).  Full chunks take the
  // ordered-update path (no stage) for p <= 2 with ng <= 3; elsewhere FP64
  // elasticity needs half chunks to keep two workgroups per CU.
  if (fp32 && nc == 3 && p == 2 && ng == 3) return 2;
  const bool ordered = p <= 2 && ng <= 3;
  return (!fp32 && nc == 3 && !ordered && p < 3) ? 1 : 0;
}

template <int P, int NG, int NC, typename T, int HALF>
static void launch_lines_h(Context& c, const AsmArgs& args, bool has_rhs)
{
  // One block per line holds all three row components where the window fits
  // the LDS (p = 3 with full chunks of 4x4x4 points: 144 KB in FP64, one
  // workgroup per CU), measured 1.7-1.9x faster than three blocks that each
  // recompute the geometry rows and the column gradients; half chunks and
  // 5x5x5 rules keep three blocks (their window and chunk exceed 160 KB).
  constexpr int SPLIT = (NC == 3 && P >= 3 && (HALF || NG > 4)) ? 3 : 1;
  dim3 grid(args.n2 * args.n3, SPLIT);
  const AsmArgs* d = static_cast<const AsmArgs*>(c.d_asmargs);
  if (has_rhs)
    k_assemble_lines<P, NG, NC, T, SPLIT, true, HALF><<<grid, kBlock, 0, c.stream>>>(d);
  else
    k_assemble_lines<P, NG, NC, T, SPLIT, false, HALF><<<grid, kBlock, 0, c.stream>>>(d);
}

// The Gauss-point slab variant exists where the pipelined, ordered path
// runs (p <= 2, at most 3x3x3 points).
static inline bool slab_variant_ok(int p, int ng) { return p <= 2 && ng * ng * ng <= 27; }

template <int P, int NG, int NC, typename T>
static int launch_lines(Context& c, const AsmArgs& args, bool has_rhs)
{
  const int h = asm_half(c, P, NG, NC, sizeof(T) == 4);
  if constexpr (P <= 2 && NG * NG * NG <= 27) {
    if (h == 2) {
      launch_lines_h<P, NG, NC, T, 2>(c, args, has_rhs);
      IFEM_HIP(hipGetLastError());
      return IFEM_GPU_OK;
    }
  }
  if (h == 1)
    launch_lines_h<P, NG, NC, T, 1>(c, args, has_rhs);
  else
    launch_lines_h<P, NG, NC, T, 0>(c, args, has_rhs);
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

template <int P, int NG>
static int dispatch_nc(Context& c, const AsmArgs& args, int nc, bool fp32, bool has_rhs)
{
  if (nc == 1) return fp32 ? launch_lines<P, NG, 1, float>(c, args, has_rhs)
                           : launch_lines<P, NG, 1, double>(c, args, has_rhs);
  return fp32 ? launch_lines<P, NG, 3, float>(c, args, has_rhs)
              : launch_lines<P, NG, 3, double>(c, args, has_rhs);
}

template <int P>
static int dispatch_ng(Context& c, const AsmArgs& args, int ng, int nc, bool fp32, bool has_rhs)
{
  if (ng == P + 1) return dispatch_nc<P, P + 1>(c, args, nc, fp32, has_rhs);
  if (ng == P + 2) return dispatch_nc<P, P + 2>(c, args, nc, fp32, has_rhs);
  return fail(IFEM_GPU_EUNSUPPORTED, "assemble: the GPU path supports ng = p+1 or p+2");
}

// One patch: Pt with eq (local DOF -> equation or -1) and gval (local DOF
// -> prescribed value, or nullptr).  accum: the system holds several
// patches; values and right-hand side are added (zeroed by the caller).
template <int P>
static int assemble_p(Context& c, const SysView& sv, const PatchRec& Pt, const int* eq,
                      const double* gval, int accum, const int64_t* pmap, const int64_t* pptr,
                      const ifem_gpu_integrand& itg, int ng, int precision)
{
  GaussArgs ga{};
  gauss_rule(ng, ga.xi, ga.w);
  Tables tb;
  int rc = make_tables<P>(c, Pt, ng, tb, ga);
  if (rc) return rc;
  AsmArgs a{};
  a.n1 = Pt.n[0];
  a.n2 = Pt.n[1];
  a.n3 = Pt.n[2];
  a.nel1 = Pt.nel[0];
  a.nel2 = Pt.nel[1];
  a.nel3 = Pt.nel[2];
  a.coefs = Pt.d_coefs;
  a.tabN1 = tb.tabN[0]; a.tabD1 = tb.tabD[0]; a.span1 = tb.span[0];
  a.tabN2 = tb.tabN[1]; a.tabD2 = tb.tabD[1]; a.span2 = tb.span[1];
  a.tabN3 = tb.tabN[2]; a.tabD3 = tb.tabD[2]; a.span3 = tb.span[2];
  for (int g = 0; g < kMaxNG; g++) a.gw[g] = ga.w[g];
  a.eq = eq;
  a.rowptr = sv.rowptr;
  a.colind = sv.colind;
  a.accum = accum;
  a.pmap = accum ? pmap : nullptr;
  a.pptr = accum ? pptr : nullptr;
  const bool fp32 = precision == IFEM_GPU_FP32;
  if (fp32) {
    a.val32 = sv.val32;
    a.val64 = nullptr;
    a.rhs = nullptr;
  } else {
    a.val64 = sv.val;
    a.val32 = nullptr;
    a.rhs = sv.rhs;
    a.gval = gval;
  }
  rc = upload_function(c, itg.kappa, a.kappa, 0);
  if (!rc) rc = upload_function(c, itg.source, a.source, 1);
  if (rc) return rc;
  for (int i = 0; i < 3; i++) a.body[i] = itg.body[i];
  if (itg.kind == IFEM_GPU_ELASTICITY) {
    a.lam = itg.E * itg.nu / ((1.0 + itg.nu) * (1.0 - 2.0 * itg.nu));
    a.mu = itg.E / (2.0 * (1.0 + itg.nu));
  }
  bool has_rhs = false;
  if (!fp32) {
    if (c.ncomp == 1) has_rhs = itg.source.type != IFEM_GPU_FN_ZERO;
    else has_rhs = itg.body[0] != 0.0 || itg.body[1] != 0.0 || itg.body[2] != 0.0;
    if (!accum)
      IFEM_HIP(hipMemsetAsync(sv.rhs, 0, sizeof(double) * (sv.neq > 0 ? sv.neq : 1), c.stream));
  }
  // geometry records of all (element, Gauss point) pairs, then the lines
  a.ng = ng;
  a.ncomp = c.ncomp;
  a.kappa_c = const_value(itg.kappa, 1.0);
  a.source_c = const_value(itg.source, 0.0);
  const int64_t ngp = (int64_t)Pt.nel[0] * Pt.nel[1] * Pt.nel[2] * ng * ng * ng;
  const size_t need = (size_t)(ngp > 0 ? ngp : 1) * (c.ncomp == 3 ? GeoRec<3>::N : GeoRec<1>::N);
  if (c.gpgeo_cap < need) {
    hipFree(c.d_gpgeo);
    c.d_gpgeo = nullptr;
    c.gpgeo_cap = 0;
    IFEM_HIP(hipMalloc(&c.d_gpgeo, sizeof(double) * need));
    c.gpgeo_cap = need;
  }
  a.gp_geo = c.d_gpgeo;
  if (!c.d_asmargs) IFEM_HIP(hipMalloc(&c.d_asmargs, sizeof(AsmArgs)));
  IFEM_HIP(hipMemcpyAsync(c.d_asmargs, &a, sizeof(AsmArgs), hipMemcpyHostToDevice, c.stream));
  IFEM_HIP(hipEventRecord(c.ev0, c.stream));
  if (ngp > 0 && c.ncomp == 3)
    k_gp_geometry_vec<P><<<grid_for(ngp), 256, 0, c.stream>>>(
        static_cast<const AsmArgs*>(c.d_asmargs), ngp);
  else if (ngp > 0)
    k_gp_geometry<P><<<grid_for(ngp), 256, 0, c.stream>>>(static_cast<const AsmArgs*>(c.d_asmargs),
                                                         ngp, has_rhs ? 1 : 0);
  IFEM_HIP(hipEventRecord(c.ev_mid, c.stream));
  rc = dispatch_ng<P>(c, a, ng, c.ncomp, fp32, has_rhs);
  if (rc) return rc;
  IFEM_HIP(hipEventRecord(c.ev1, c.stream));
  IFEM_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f, ms_geo = 0.f;
  IFEM_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  IFEM_HIP(hipEventElapsedTime(&ms_geo, c.ev0, c.ev_mid));
  c.last_assemble_ms = ms;
  c.last_geometry_ms = ms_geo;
  c.last_lines_ms = ms - ms_geo;
  return IFEM_GPU_OK;
}

int set_assembly_variant(Context& c, int p, int ng, int nc, bool fp32, int variant)
{
  if (variant < -1 || variant > 2 || (variant == 2 && !slab_variant_ok(p, ng)))
    return fail(IFEM_GPU_EUNSUPPORTED, "set_assembly_variant: no such variant for this degree and rule");
  asm_half_slot(c, p, ng, nc, fp32) = variant;
  return IFEM_GPU_OK;
}

// Times the element-chunk variants of the line kernel on this system (full
// chunks, half chunks, and for p <= 2 with at most 3x3x3 points full chunks
// in two Gauss-point slabs) and keeps the fastest for (p, ng, ncomp,
// precision); the system is left assembled with the chosen variant.
int autotune_assembly(Context& c, const ifem_gpu_integrand& itg, const int ng[3], int precision,
                      int* half_out)
{
  const PatchRec& Pt = c.patches[c.patch];
  const bool fp32 = precision == IFEM_GPU_FP32;
  if (ng[0] < 1 || ng[0] > kMaxNG) return fail(IFEM_GPU_EUNSUPPORTED, "autotune: ng out of range");
  int& slot = asm_half_slot(c, Pt.p[0], ng[0], c.ncomp, fp32);
  double best = 1e300;
  int best_h = 0;
  const int hmax = slab_variant_ok(Pt.p[0], ng[0]) ? 2 : 1;
  for (int h = 0; h <= hmax; h++) {
    slot = h;
    double t = 1e300;
    for (int rep = 0; rep < 3; rep++) {
      int rc = assemble_volume(c, itg, ng, precision);
      if (rc) return rc;
      t = std::fmin(t, c.last_assemble_ms);
    }
    if (t < best) {
      best = t;
      best_h = h;
    }
  }
  slot = best_h;
  if (half_out) *half_out = best_h;
  return assemble_volume(c, itg, ng, precision);
}

static int assemble_one(Context& c, const SysView& sv, const PatchRec& Pt, const int* eq,
                        const double* gval, int accum, const ifem_gpu_integrand& itg, int ng,
                        int precision, const int64_t* pmap = nullptr,
                        const int64_t* pptr = nullptr)
{
  switch (Pt.p[0]) {
  case 1: return assemble_p<1>(c, sv, Pt, eq, gval, accum, pmap, pptr, itg, ng, precision);
  case 2: return assemble_p<2>(c, sv, Pt, eq, gval, accum, pmap, pptr, itg, ng, precision);
  case 3: return assemble_p<3>(c, sv, Pt, eq, gval, accum, pmap, pptr, itg, ng, precision);
  default: return fail(IFEM_GPU_EUNSUPPORTED, "assemble: degree out of range");
  }
}

// Patches added one after the other into the zeroed system sv.
static int assemble_sum(Context& c, SysView& sv, const std::vector<SubPatch>& subs,
                        const ifem_gpu_integrand& itg, int ng, int precision)
{
  const bool fp32 = precision == IFEM_GPU_FP32;
  const size_t nz = (size_t)(sv.nnz > 0 ? sv.nnz : 1);
  if (fp32) {
    IFEM_HIP(hipMemsetAsync(sv.val32, 0, sizeof(float) * nz, c.stream));
  } else {
    IFEM_HIP(hipMemsetAsync(sv.val, 0, sizeof(double) * nz, c.stream));
    IFEM_HIP(hipMemsetAsync(sv.rhs, 0, sizeof(double) * (sv.neq > 0 ? sv.neq : 1), c.stream));
  }
  double tot = 0.0, geo = 0.0;
  for (const SubPatch& sp : subs) {
    int rc = assemble_one(c, sv, c.patches[sp.patch], sp.d_leq, sp.d_lgval, 1, itg, ng, precision,
                          sp.d_pmap, sp.d_pptr);
    if (rc) return rc;
    tot += c.last_assemble_ms;
    geo += c.last_geometry_ms;
  }
  c.last_assemble_ms = tot;
  c.last_geometry_ms = geo;
  c.last_lines_ms = tot - geo;
  return IFEM_GPU_OK;
}

int assemble_volume(Context& c, const ifem_gpu_integrand& itg, const int ng[3], int precision)
{
  if (ng[0] != ng[1] || ng[0] != ng[2])
    return fail(IFEM_GPU_EUNSUPPORTED, "assemble: the GPU path needs equal Gauss counts per direction");
  if (ng[0] < 1 || ng[0] > kMaxNG) return fail(IFEM_GPU_EUNSUPPORTED, "assemble: ng out of range");
  const bool fp32 = precision == IFEM_GPU_FP32;
  if (fp32 && !c.d_val32)
    IFEM_HIP(hipMalloc(&c.d_val32, sizeof(float) * (c.nnz > 0 ? c.nnz : 1)));
  SysView sv{c.d_rowptr, c.d_colind, c.d_val, c.d_val32, c.d_rhs, c.nnz, c.neq};
  if (c.mpc) {
    // constraint equations with masters: the extended system (slaves as
    // unknowns) is assembled, then condensed to the free equations
    MpcSystem& x = *c.mpc;
    if (fp32 && !x.d_val32)
      IFEM_HIP(hipMalloc(&x.d_val32, sizeof(float) * (x.nnz > 0 ? x.nnz : 1)));
    SysView xv{x.d_rowptr, x.d_colind, x.d_val, x.d_val32, x.d_rhs, x.nnz, x.ne};
    // a shallow copy: x.ext owns the device arrays (SubPatch frees nothing)
    std::vector<SubPatch> subs(1, x.ext);
    int rc = assemble_sum(c, xv, subs, itg, ng[0], precision);
    if (rc) return rc;
    rc = mpc_condense(c, fp32, false);
    if (rc) return rc;
  } else if (c.multi.empty()) {
    int rc = assemble_one(c, sv, c.patches[c.patch], c.d_eq, c.d_gval, 0, itg, ng[0], precision);
    if (rc) return rc;
  } else {
    int rc = assemble_sum(c, sv, c.multi, itg, ng[0], precision);
    if (rc) return rc;
  }
  if (fp32) c.fp32_ready = true;
  else c.fp64_ready = true;
  return IFEM_GPU_OK;
}

// ---------------------------------------------------------------------------
// Neumann traction on one face: one thread per (face node, component); sums
// over the face elements holding the node and their ng x ng Gauss points in
// fixed order (src/ASM/ASMs3D.C:2547-2812; face utl::Jacobian,
// src/Utility/CoordinateMapping.C:90-133: dS = |J_t1 x J_t2|, zero if J is
// singular; detJxW = dS * dA * w_i * w_j with dA = 0.25 * area).
// Pass 1: one thread per (face element, Gauss point): w = dS * dA * w_i * w_j
// (0 where J is singular), face element (ea, eb) along the tangent
// directions, w[((eb*nel_ta + ea)*ng + gb)*ng + ga].
template <int P>
__global__ void k_traction_w(int n1, int n2, int n3, int nel1, int nel2, int nel3,
                             const double* __restrict__ coefs, const double* k1,
                             const double* k2, const double* k3, int face, int ng, GaussArgs ga,
                             double* __restrict__ wout)
{
  constexpr int P1 = P + 1;
  const int dn = face / 2, side = face % 2;
  const int n[3] = {n1, n2, n3}, nel[3] = {nel1, nel2, nel3};
  const double* kn[3] = {k1, k2, k3};
  const int ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2; // tangent directions, ascending
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int ntot = nel[ta] * nel[tb] * ng * ng;
  if (id >= ntot) return;
  const int gA = id % ng, gb = (id / ng) % ng, fe = id / (ng * ng);
  const int ea = fe % nel[ta], eb = fe / nel[ta];
  // faceDir = +-(dn+1): t1 = 1 + |faceDir| % 3, t2 = 1 + t1 % 3 (1-based)
  const int T1 = (dn + 1) % 3, T2 = (T1 + 1) % 3;
  int e[3];
  e[dn] = side ? nel[dn] - 1 : 0;
  e[ta] = ea;
  e[tb] = eb;
  double spn[3];
  for (int d = 0; d < 3; d++) spn[d] = kn[d][e[d] + P + 1] - kn[d][e[d] + P];
  const double dA = 0.25 * (spn[ta] * spn[tb]);
  double Nd[3][P1], dNd[3][P1];
  for (int d = 0; d < 3; d++) {
    const int s = e[d] + P;
    const double* kt = kn[d];
    double u;
    if (d == dn) u = side ? kt[n[d]] : kt[P];
    else {
      const double xi = ga.xi[d == ta ? gA : gb];
      u = 0.5 * ((kt[s + 1] - kt[s]) * xi + kt[s + 1] + kt[s]);
    }
    bspline_1d<P>(kt, s, u, Nd[d], dNd[d]);
  }
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b3 = 0; b3 < P1; b3++)
    for (int b2 = 0; b2 < P1; b2++)
      for (int b1 = 0; b1 < P1; b1++) {
        const int64_t nb = (int64_t)(e[0] + b1) + (int64_t)n1 * ((e[1] + b2) + (int64_t)n2 * (e[2] + b3));
        const double g0 = dNd[0][b1] * Nd[1][b2] * Nd[2][b3];
        const double g1 = Nd[0][b1] * dNd[1][b2] * Nd[2][b3];
        const double g2 = Nd[0][b1] * Nd[1][b2] * dNd[2][b3];
        for (int i = 0; i < 3; i++) {
          const double x = coefs[3 * nb + i];
          J[3 * i + 0] += x * g0;
          J[3 * i + 1] += x * g1;
          J[3 * i + 2] += x * g2;
        }
      }
  const double v1[3] = {J[T1], J[3 + T1], J[6 + T1]};
  const double v2[3] = {J[T2], J[3 + T2], J[6 + T2]};
  const double nx = v1[1] * v2[2] - v1[2] * v2[1];
  const double ny = v1[2] * v2[0] - v1[0] * v2[2];
  const double nz = v1[0] * v2[1] - v1[1] * v2[0];
  double dS = sqrt(nx * nx + ny * ny + nz * nz);
  const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[6] * J[5]) +
                     J[2] * (J[3] * J[7] - J[6] * J[4]);
  if (det <= 1e-12 && det >= -1e-12) dS = 0.0;
  wout[id] = dS == 0.0 ? 0.0 : dS * (dA * ga.w[gA] * ga.w[gb]);
}

// Pass 2: one thread per face node sums w * N_a over the face elements and
// Gauss points holding it (fixed order), then adds t_c times the sum to the
// node's free components.
template <int P>
__global__ void k_traction_sum(int n1, int n2, int n3, int nel1, int nel2, int nel3,
                               const double* k1, const double* k2, const double* k3, int face,
                               int ng, GaussArgs ga, const double* __restrict__ wf, double t0,
                               double t1, double t2, int nc, const int* __restrict__ eq,
                               double* __restrict__ rhs)
{
  constexpr int P1 = P + 1;
  const int dn = face / 2, side = face % 2;
  const int n[3] = {n1, n2, n3}, nel[3] = {nel1, nel2, nel3};
  const double* kn[3] = {k1, k2, k3};
  const int ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n[ta] * n[tb]) return;
  const int ia = id % n[ta], ib = id / n[ta];
  int idx[3];
  idx[dn] = side ? n[dn] - 1 : 0;
  idx[ta] = ia;
  idx[tb] = ib;
  const int64_t node = (int64_t)idx[0] + (int64_t)n1 * (idx[1] + (int64_t)n2 * idx[2]);
  // basis value of the node in the normal direction on the face
  const int en = side ? nel[dn] - 1 : 0;
  double Nn[P1], dNn[P1];
  {
    const double* kt = kn[dn];
    bspline_1d<P>(kt, en + P, side ? kt[n[dn]] : kt[P], Nn, dNn);
  }
  const double nface = Nn[idx[dn] - en];
  double sum = 0.0;
  for (int eb = max(0, ib - P); eb <= min(nel[tb] - 1, ib); eb++)
    for (int ea = max(0, ia - P); ea <= min(nel[ta] - 1, ia); ea++) {
      const double* w = wf + (int64_t)(eb * nel[ta] + ea) * ng * ng;
      for (int gb = 0; gb < ng; gb++) {
        double Nb[P1], dNb[P1];
        const double* kb = kn[tb];
        const int sb = eb + P;
        bspline_1d<P>(kb, sb, 0.5 * ((kb[sb + 1] - kb[sb]) * ga.xi[gb] + kb[sb + 1] + kb[sb]), Nb,
                      dNb);
        for (int gA = 0; gA < ng; gA++) {
          const double wg = w[gb * ng + gA];
          if (wg == 0.0) continue;
          double Na[P1], dNa[P1];
          const double* ka = kn[ta];
          const int sa = ea + P;
          bspline_1d<P>(ka, sa, 0.5 * ((ka[sa + 1] - ka[sa]) * ga.xi[gA] + ka[sa + 1] + ka[sa]),
                        Na, dNa);
          sum += wg * (Na[ia - ea] * Nb[ib - eb] * nface);
        }
      }
    }
  const double tv[3] = {t0, t1, t2};
  for (int c = 0; c < nc; c++) {
    const int r = eq[node * nc + c];
    if (r >= 0) rhs[r] += tv[c] * sum;
  }
}

int assemble_traction(Context& c, int index, int face0, const double* t, int ng)
{
  if (index < 0 || index >= (c.multi.empty() ? 1 : (int)c.multi.size()))
    return fail(IFEM_GPU_EINVAL, "traction: patch index out of range");
  const PatchRec& Pt = c.patches[c.multi.empty() ? c.patch : c.multi[index].patch];
  const int* eqmap = c.multi.empty() ? c.d_eq : c.multi[index].d_leq;
  double* rhs = c.d_rhs;
  if (c.mpc) { // into the extended right-hand side, condensed below
    if (!c.fp64_ready)
      return fail(IFEM_GPU_EINVAL,
                  "assemble_neumann: constraints with masters need the FP64 system first");
    eqmap = c.mpc->ext.d_leq;
    rhs = c.mpc->d_rhs;
  }
  GaussArgs ga{};
  gauss_rule(ng, ga.xi, ga.w);
  const int dn = face0 / 2;
  const int ta = dn == 0 ? 1 : 0, tb = dn == 2 ? 1 : 2;
  const int nfn = Pt.n[ta] * Pt.n[tb];
  const int nw = Pt.nel[ta] * Pt.nel[tb] * ng * ng;
  const double t1 = c.ncomp > 1 ? t[1] : 0.0, t2 = c.ncomp > 2 ? t[2] : 0.0;
  if (c.facew_cap < (size_t)nw) {
    hipFree(c.d_facew);
    c.d_facew = nullptr;
    c.facew_cap = 0;
    IFEM_HIP(hipMalloc(&c.d_facew, sizeof(double) * (nw > 0 ? nw : 1)));
    c.facew_cap = nw;
  }
#define LAUNCH_TR(PP)                                                                          \
  k_traction_w<PP><<<(nw + 255) / 256, 256, 0, c.stream>>>(                                     \
      Pt.n[0], Pt.n[1], Pt.n[2], Pt.nel[0], Pt.nel[1], Pt.nel[2], Pt.d_coefs, Pt.d_knots[0],    \
      Pt.d_knots[1], Pt.d_knots[2], face0, ng, ga, c.d_facew);                                  \
  k_traction_sum<PP><<<(nfn + 127) / 128, 128, 0, c.stream>>>(                                  \
      Pt.n[0], Pt.n[1], Pt.n[2], Pt.nel[0], Pt.nel[1], Pt.nel[2], Pt.d_knots[0], Pt.d_knots[1], \
      Pt.d_knots[2], face0, ng, ga, c.d_facew, t[0], t1, t2, c.ncomp, eqmap, rhs)
  switch (Pt.p[0]) {
  case 1: LAUNCH_TR(1); break;
  case 2: LAUNCH_TR(2); break;
  case 3: LAUNCH_TR(3); break;
  default: return fail(IFEM_GPU_EUNSUPPORTED, "traction: degree out of range");
  }
#undef LAUNCH_TR
  IFEM_HIP(hipGetLastError());
  if (c.mpc) {
    int rc = mpc_condense(c, false, true);
    if (rc) return rc;
  }
  IFEM_HIP(hipStreamSynchronize(c.stream));
  return IFEM_GPU_OK;
}

// ---------------------------------------------------------------------------
// Batch of independent scalar patches (block-diagonal global system).  Each
// patch is integrated as SIMbase::assembleSystem does per patch
// (src/SIM/SIMbase.C:1229-1400), but every patch has its own equations and
// all patches go through one persistent launch.

static int batch_tables(Context& c, const PatchRec& Pt, BatchPatch& bp)
{
  GaussArgs ga{};
  gauss_rule(bp.ng, ga.xi, ga.w);
  for (int d = 0; d < 3; d++) {
    const size_t cnt = (size_t)Pt.nel[d] * bp.ng * (bp.p + 1);
    IFEM_HIP(hipMalloc(&bp.tabN[d], sizeof(double) * cnt));
    IFEM_HIP(hipMalloc(&bp.tabD[d], sizeof(double) * cnt));
    IFEM_HIP(hipMalloc(&bp.span[d], sizeof(double) * Pt.nel[d]));
    const int tot = Pt.nel[d] * bp.ng;
    const dim3 g((tot + 127) / 128), b(128);
    switch (bp.p) {
    case 1: k_basis_tables<1><<<g, b, 0, c.stream>>>(Pt.d_knots[d], Pt.nel[d], bp.ng, ga, bp.tabN[d], bp.tabD[d], bp.span[d]); break;
    case 2: k_basis_tables<2><<<g, b, 0, c.stream>>>(Pt.d_knots[d], Pt.nel[d], bp.ng, ga, bp.tabN[d], bp.tabD[d], bp.span[d]); break;
    default: k_basis_tables<3><<<g, b, 0, c.stream>>>(Pt.d_knots[d], Pt.nel[d], bp.ng, ga, bp.tabN[d], bp.tabD[d], bp.span[d]); break;
    }
  }
  IFEM_HIP(hipGetLastError());
  return IFEM_GPU_OK;
}

static DevFunction batch_function(const ifem_gpu_function& f, const double* d_modes)
{
  DevFunction df;
  df.type = f.type;
  df.nmodes = d_modes ? f.nmodes : 0;
  df.scale = f.scale;
  df.mean = f.mean;
  df.modes = d_modes;
  if (d_modes && f.modes) mode_bounds(f.modes, f.nmodes, df.kbound, df.pbound);
  else df.kbound = df.pbound = 1e300; // unknown: the checked loop
  return df;
}

// Per-patch assembly arguments for both precisions (the chunk variant of
// each degree comes from the batch tuning table).
int batch_upload_args(Context& c)
{
  Batch& B = *c.batch;
  const int np = (int)B.pp.size();
  std::vector<AsmArgs> a64(np), a32(np);
  for (int q = 0; q < np; q++) {
    const BatchPatch& bp = B.pp[q];
    const PatchRec& Pt = c.patches[bp.patch];
    AsmArgs a{};
    a.n1 = Pt.n[0];
    a.n2 = Pt.n[1];
    a.n3 = Pt.n[2];
    a.nel1 = Pt.nel[0];
    a.nel2 = Pt.nel[1];
    a.nel3 = Pt.nel[2];
    a.coefs = Pt.d_coefs;
    a.tabN1 = bp.tabN[0]; a.tabD1 = bp.tabD[0]; a.span1 = bp.span[0];
    a.tabN2 = bp.tabN[1]; a.tabD2 = bp.tabD[1]; a.span2 = bp.span[1];
    a.tabN3 = bp.tabN[2]; a.tabD3 = bp.tabD[2]; a.span3 = bp.span[2];
    GaussArgs ga{};
    gauss_rule(bp.ng, ga.xi, ga.w);
    for (int g = 0; g < kMaxNG; g++) a.gw[g] = ga.w[g];
    a.eq = bp.d_eq;
    a.rowptr = B.d_rowptr;
    a.kappa = batch_function(bp.itg.kappa, bp.d_kmodes);
    a.source = batch_function(bp.itg.source, bp.d_smodes);
    a.variant = 2 * (bp.p - 1) + (B.tune[bp.p].half ? 1 : 0);
    a.ng = bp.ng;
    a.ncomp = 1;
    const int64_t g0 = B.h_gp0.empty() ? 0 : B.h_gp0[q];
    a.gp_geo = B.d_geo + g0 * GeoRec<1>::N;
    a.kappa_c = const_value(bp.itg.kappa, 1.0);
    a.source_c = const_value(bp.itg.source, 0.0);
    a.val64 = B.d_val;
    a.val32 = nullptr;
    a.rhs = B.d_rhs;
    a64[q] = a;
    a.val64 = nullptr;
    a.val32 = B.d_val32;
    a.rhs = nullptr;
    a32[q] = a;
  }
  if (!B.d_args64) IFEM_HIP(hipMalloc(&B.d_args64, sizeof(AsmArgs) * (np > 0 ? np : 1)));
  if (!B.d_args32) IFEM_HIP(hipMalloc(&B.d_args32, sizeof(AsmArgs) * (np > 0 ? np : 1)));
  IFEM_HIP(hipMemcpy(B.d_args64, a64.data(), sizeof(AsmArgs) * np, hipMemcpyHostToDevice));
  IFEM_HIP(hipMemcpy(B.d_args32, a32.data(), sizeof(AsmArgs) * np, hipMemcpyHostToDevice));
  return IFEM_GPU_OK;
}

// Symbolic phase of the batch: per-patch basis tables, global CSR pattern
// (patch q's rows and columns are its equations shifted by eq0), and the
// (patch, line) work list sorted by decreasing estimated cost.
int batch_symbolic(Context& c)
{
  Batch& B = *c.batch;
  const int np = (int)B.pp.size();
  int eq0 = 0;
  for (BatchPatch& bp : B.pp) {
    bp.eq0 = eq0;
    eq0 += bp.neq;
  }
  B.neq = eq0;
  const int neq = B.neq;
  IFEM_HIP(hipEventRecord(c.ev0, c.stream));
  for (BatchPatch& bp : B.pp) {
    int rc = batch_tables(c, c.patches[bp.patch], bp);
    if (rc) return rc;
  }
  int64_t* rowlen = nullptr;
  int64_t* bsum = nullptr;
  const int nb = (neq + kScanBlock - 1) / kScanBlock;
  IFEM_HIP(hipMalloc(&B.d_rowptr, sizeof(int64_t) * (neq + 1)));
  IFEM_HIP(hipMalloc(&rowlen, sizeof(int64_t) * (neq > 0 ? neq : 1)));
  IFEM_HIP(hipMalloc(&bsum, sizeof(int64_t) * (nb > 0 ? nb : 1)));
  IFEM_HIP(hipMemsetAsync(B.d_rowptr, 0, sizeof(int64_t) * (neq + 1), c.stream));
  for (BatchPatch& bp : B.pp) {
    const PatchRec& Pt = c.patches[bp.patch];
    if (bp.neq > 0)
      k_row_lengths<<<(bp.nnod + 255) / 256, 256, 0, c.stream>>>(Pt.n[0], Pt.n[1], Pt.n[2], bp.p, 1,
                                                                   bp.d_eq, rowlen, bp.nnod);
  }
  if (neq > 0) {
    k_scan_local<<<nb, kScanBlock, 0, c.stream>>>(rowlen, B.d_rowptr, bsum, neq);
    k_scan_sums<<<1, 64, 0, c.stream>>>(bsum, nb);
    k_scan_add<<<nb, kScanBlock, 0, c.stream>>>(B.d_rowptr, bsum, neq);
  }
  IFEM_HIP(hipGetLastError());
  int64_t nnz = 0;
  IFEM_HIP(hipMemcpyAsync(&nnz, B.d_rowptr + neq, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  IFEM_HIP(hipStreamSynchronize(c.stream));
  hipFree(rowlen);
  hipFree(bsum);
  B.nnz = nnz;
  IFEM_HIP(hipMalloc(&B.d_colind, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
  for (BatchPatch& bp : B.pp) {
    const PatchRec& Pt = c.patches[bp.patch];
    if (bp.neq > 0)
      k_fill_columns<<<(bp.nnod + 3) / 4, 256, 0, c.stream>>>(Pt.n[0], Pt.n[1], Pt.n[2], bp.p, 1,
                                                              bp.d_eq, B.d_rowptr, B.d_colind, bp.nnod);
  }
  IFEM_HIP(hipGetLastError());
  IFEM_HIP(hipMalloc(&B.d_val, sizeof(double) * (nnz > 0 ? nnz : 1)));
  IFEM_HIP(hipMalloc(&B.d_val32, sizeof(float) * (nnz > 0 ? nnz : 1)));
  IFEM_HIP(hipMalloc(&B.d_rhs, sizeof(double) * (neq > 0 ? neq : 1)));
  IFEM_HIP(hipMemsetAsync(B.d_rhs, 0, sizeof(double) * (neq > 0 ? neq : 1), c.stream));

  // geometry records of every patch's (element, Gauss point) pairs
  B.h_gp0.assign(np, 0);
  int64_t ngp = 0;
  for (int q = 0; q < np; q++) {
    const BatchPatch& bp = B.pp[q];
    const PatchRec& Pt = c.patches[bp.patch];
    B.h_gp0[q] = ngp;
    ngp += (int64_t)Pt.nel[0] * Pt.nel[1] * Pt.nel[2] * bp.ng * bp.ng * bp.ng;
  }
  B.ngp = ngp;
  IFEM_HIP(hipMalloc(&B.d_geo, sizeof(double) * GeoRec<1>::N * (ngp > 0 ? ngp : 1)));
  // per degree: its patches and the first point of each among that degree's
  // points (one geometry launch per degree)
  std::vector<int> glist;
  std::vector<int64_t> gq0;
  for (int p = 1; p <= 3; p++) {
    B.gdeg_off[p] = (int)glist.size();
    int64_t t = 0;
    for (int q = 0; q < np; q++) {
      const BatchPatch& bp = B.pp[q];
      if (bp.p != p) continue;
      const PatchRec& Pt = c.patches[bp.patch];
      glist.push_back(q);
      gq0.push_back(t);
      t += (int64_t)Pt.nel[0] * Pt.nel[1] * Pt.nel[2] * bp.ng * bp.ng * bp.ng;
    }
    B.gdeg_n[p] = (int)glist.size() - B.gdeg_off[p];
    B.gdeg_total[p] = t;
  }
  IFEM_HIP(hipMalloc(&B.d_glist, sizeof(int) * (np > 0 ? np : 1)));
  IFEM_HIP(hipMalloc(&B.d_gq0, sizeof(int64_t) * (np > 0 ? np : 1)));
  IFEM_HIP(hipMemcpy(B.d_glist, glist.data(), sizeof(int) * glist.size(), hipMemcpyHostToDevice));
  IFEM_HIP(hipMemcpy(B.d_gq0, gq0.data(), sizeof(int64_t) * gq0.size(), hipMemcpyHostToDevice));

  // work list: one item per node line (j, k) of each patch
  struct Item { double cost; int patch, line; };
  std::vector<Item> items;
  for (int q = 0; q < np; q++) {
    const BatchPatch& bp = B.pp[q];
    if (bp.neq == 0) continue;
    const PatchRec& Pt = c.patches[bp.patch];
    const int P1 = bp.p + 1;
    const double per_elem = (double)P1 * P1 * P1 * bp.ng * bp.ng * bp.ng;
    for (int k = 0; k < Pt.n[2]; k++)
      for (int j = 0; j < Pt.n[1]; j++) {
        const int c2 = std::min(j, Pt.nel[1] - 1) - std::max(0, j - bp.p) + 1;
        const int c3 = std::min(k, Pt.nel[2] - 1) - std::max(0, k - bp.p) + 1;
        items.push_back({per_elem * Pt.nel[0] * c2 * c3, q, j + Pt.n[1] * k});
      }
  }
  // grouped by degree (one launch each, highest first), by decreasing cost
  // within a degree
  std::stable_sort(items.begin(), items.end(), [&](const Item& x, const Item& y) {
    const int px = B.pp[x.patch].p, py = B.pp[y.patch].p;
    return px != py ? px > py : x.cost > y.cost;
  });
  std::vector<int2> hi(items.size());
  for (size_t i = 0; i < items.size(); i++) hi[i] = make_int2(items[i].patch, items[i].line);
  B.nitems = (int)hi.size();
  for (int p = 0; p < 4; p++) B.item0[p] = B.item1[p] = 0;
  for (size_t i = 0; i < items.size(); i++) {
    const int p = B.pp[items[i].patch].p;
    if (B.item1[p] == 0) B.item0[p] = (int)i;
    B.item1[p] = (int)i + 1;
  }
  IFEM_HIP(hipMalloc(&B.d_items, sizeof(int2) * (hi.empty() ? 1 : hi.size())));
  if (!hi.empty())
    IFEM_HIP(hipMemcpy(B.d_items, hi.data(), sizeof(int2) * hi.size(), hipMemcpyHostToDevice));
  IFEM_HIP(hipMalloc(&B.d_queue, sizeof(int) * 4));
  hipDeviceProp_t prop;
  IFEM_HIP(hipGetDeviceProperties(&prop, c.device));
  B.ncu = prop.multiProcessorCount;
  int rc = batch_upload_args(c);
  if (rc) return rc;
  IFEM_HIP(hipEventRecord(c.ev1, c.stream));
  IFEM_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  IFEM_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  B.symbolic_ms = ms;
  return IFEM_GPU_OK;
}

// Numeric assembly of every patch of the batch in one persistent launch:
// FP64 writes K and b, FP32 writes the FP32 copy of K.  Returns the launch's
// device time in *ms.
// One persistent launch over the items of degree P, chunk variant HALF: as
// many blocks as stay resident.
template <int P, typename T, bool HAS_RHS, int HALF>
static int launch_batch_degree(Context& c, Batch& B, const void* args)
{
  const int n = B.item1[P] - B.item0[P];
  if (n <= 0) return IFEM_GPU_OK;
  int per_cu = 0;
  IFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&k_assemble_batch<P, T, HAS_RHS, HALF>), kBlock, 0));
  const int grid = std::max(1, std::min(n, std::max(1, per_cu) * B.ncu));
  k_assemble_batch<P, T, HAS_RHS, HALF><<<grid, kBlock, 0, c.stream>>>(
      static_cast<const AsmArgs*>(args), B.d_items + B.item0[P], n, B.d_queue + P);
  return IFEM_GPU_OK;
}

template <typename T, bool HAS_RHS>
static int launch_batch(Context& c, Batch& B, const void* args)
{
  int rc = IFEM_GPU_OK;
  if (!rc) rc = B.tune[3].half ? launch_batch_degree<3, T, HAS_RHS, 1>(c, B, args)
                               : launch_batch_degree<3, T, HAS_RHS, 0>(c, B, args);
  if (!rc) rc = B.tune[2].half ? launch_batch_degree<2, T, HAS_RHS, 1>(c, B, args)
                               : launch_batch_degree<2, T, HAS_RHS, 0>(c, B, args);
  if (!rc) rc = B.tune[1].half ? launch_batch_degree<1, T, HAS_RHS, 1>(c, B, args)
                               : launch_batch_degree<1, T, HAS_RHS, 0>(c, B, args);
  return rc;
}

int batch_assemble(Context& c, int precision, double* ms_out)
{
  Batch& B = *c.batch;
  const bool fp32 = precision == IFEM_GPU_FP32;
  k_fill<int><<<1, 64, 0, c.stream>>>(4, B.d_queue, 0);
  IFEM_HIP(hipEventRecord(c.ev0, c.stream));
  // geometry pass: always for FP64 (the full assembly); an FP32 assembly
  // reuses the records of an earlier assembly of the same batch (the
  // material part of the records is the same in both precisions)
  if (B.ngp > 0 && !(fp32 && B.geo_valid)) {
    const AsmArgs* ga = static_cast<const AsmArgs*>(fp32 ? B.d_args32 : B.d_args64);
    for (int p = 3; p >= 1; p--) {
      const int64_t tot = B.gdeg_total[p];
      if (tot <= 0) continue;
      const int* pl = B.d_glist + B.gdeg_off[p];
      const int64_t* q0 = B.d_gq0 + B.gdeg_off[p];
      if (p == 3)
        k_gp_geometry_batch<3><<<grid_for(tot), 256, 0, c.stream>>>(ga, pl, q0, B.gdeg_n[p], tot, fp32 ? 0 : 1);
      else if (p == 2)
        k_gp_geometry_batch<2><<<grid_for(tot), 256, 0, c.stream>>>(ga, pl, q0, B.gdeg_n[p], tot, fp32 ? 0 : 1);
      else
        k_gp_geometry_batch<1><<<grid_for(tot), 256, 0, c.stream>>>(ga, pl, q0, B.gdeg_n[p], tot, fp32 ? 0 : 1);
    }
  }
  if (B.nitems > 0) {
    const int rc = fp32 ? launch_batch<float, false>(c, B, B.d_args32)
                        : launch_batch<double, true>(c, B, B.d_args64);
    if (rc) return rc;
  }
  IFEM_HIP(hipGetLastError());
  IFEM_HIP(hipEventRecord(c.ev1, c.stream));
  IFEM_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  IFEM_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  if (ms_out) *ms_out = ms;
  if (fp32) B.fp32_ready = true;
  else B.fp64_ready = true;
  B.geo_valid = true;
  return IFEM_GPU_OK;
}

// ---------------------------------------------------------------------------
// Norms of the current solution, NormBase / ElmNorm / GlbNorm with the SQRT
// final operation (src/ASM/IntegrandBase.h:378-530, src/ASM/GlbNorm.C:33-79):
// one thread per (element, Gauss point) evaluates u_h = sum_b N_b u_b and
// grad u_h = J^-T sum_b u_b grad_ref N_b with the volume loop's quadrature,
// Jacobian and singular-point rule, and accumulates
//   [0] |u_h|^2, [1] energy density of u_h (kappa |grad u_h|^2, or
//   eps:C:eps), and with an exact scalar solution u: [2] u^2, [3] kappa
//   |grad u|^2, [4] (u - u_h)^2, [5] kappa |grad(u - u_h)|^2
// times detJxW, reduced per block in fixed order.
__device__ inline void eval_field_grad(const DevFunction& f, double x, double y, double z,
                                       double* g)
{
  g[0] = g[1] = g[2] = 0.0;
  switch (f.type) {
  case IFEM_GPU_FN_SINPROD: {
    const double s0 = sin(M_PI * x), s1 = sin(M_PI * y), s2 = sin(M_PI * z);
    const double c0 = cos(M_PI * x), c1 = cos(M_PI * y), c2 = cos(M_PI * z);
    g[0] = f.scale * M_PI * c0 * s1 * s2;
    g[1] = f.scale * M_PI * s0 * c1 * s2;
    g[2] = f.scale * M_PI * s0 * s1 * c2;
    return;
  }
  case IFEM_GPU_FN_FOURIER:
  case IFEM_GPU_FN_EXP_FOURIER: {
    double s = 0.0;
    for (int m = 0; m < f.nmodes; m++) {
      const double* md = f.modes + 5 * m;
      const double arg = md[0] * x + md[1] * y + md[2] * z + md[3];
      s += md[4] * cos(arg);
      for (int d = 0; d < 3; d++) g[d] -= f.scale * md[4] * md[d] * sin(arg);
    }
    if (f.type == IFEM_GPU_FN_EXP_FOURIER) {
      const double e = exp(f.mean + f.scale * s);
      for (int d = 0; d < 3; d++) g[d] *= e;
    }
    return;
  }
  default:
    return;
  }
}

struct NormArgs {
  AsmArgs a;
  const double* x;     // solution, equation order
  const double* xdof;  // or the expanded solution per DOF (nullptr: use x)
  DevFunction exact;
  int has_exact;
  double* part;     // 6 x gridDim.x partial sums
};

template <int P, int NC>
__global__ __launch_bounds__(256) void k_norms(NormArgs na, int64_t total)
{
  constexpr int P1 = P + 1;
  const AsmArgs& a = na.a;
  const int ng = a.ng, gp3 = ng * ng * ng;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = id / gp3;
    const int q = (int)(id - e * gp3);
    const int q1 = q % ng, q2 = (q / ng) % ng, q3 = q / (ng * ng);
    const int e1 = (int)(e % a.nel1), e2 = (int)((e / a.nel1) % a.nel2),
              e3 = (int)(e / ((int64_t)a.nel1 * a.nel2));
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double X[3] = {0, 0, 0}, uh[NC], gr[NC][3];
    for (int c = 0; c < NC; c++) {
      uh[c] = 0.0;
      gr[c][0] = gr[c][1] = gr[c][2] = 0.0;
    }
    for (int b3 = 0; b3 < P1; b3++)
      for (int b2 = 0; b2 < P1; b2++)
#pragma unroll
        for (int b1 = 0; b1 < P1; b1++) {
          const double n1 = a.tabN1[(e1 * ng + q1) * P1 + b1], d1 = a.tabD1[(e1 * ng + q1) * P1 + b1];
          const double n2 = a.tabN2[(e2 * ng + q2) * P1 + b2], d2 = a.tabD2[(e2 * ng + q2) * P1 + b2];
          const double n3 = a.tabN3[(e3 * ng + q3) * P1 + b3], d3 = a.tabD3[(e3 * ng + q3) * P1 + b3];
          const double g0 = d1 * n2 * n3, g1 = n1 * d2 * n3, g2 = n1 * n2 * d3, nv = n1 * n2 * n3;
          const int64_t node =
              (int64_t)(e1 + b1) + (int64_t)a.n1 * ((e2 + b2) + (int64_t)a.n2 * (e3 + b3));
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const double xi = a.coefs[3 * node + i];
            J[3 * i + 0] += xi * g0;
            J[3 * i + 1] += xi * g1;
            J[3 * i + 2] += xi * g2;
            X[i] += xi * nv;
          }
#pragma unroll
          for (int c = 0; c < NC; c++) {
            const int eqn = a.eq[node * NC + c];
            const double ub = na.xdof ? na.xdof[node * NC + c]
                                      : (eqn >= 0 ? na.x[eqn] : (a.gval ? a.gval[node * NC + c] : 0.0));
            uh[c] += nv * ub;
            gr[c][0] += g0 * ub;
            gr[c][1] += g1 * ub;
            gr[c][2] += g2 * ub;
          }
        }
    const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[6] * J[5]) +
                       J[2] * (J[3] * J[7] - J[6] * J[4]);
    if (det <= 1e-12 && det >= -1e-12) continue;
    double Ji[9];
    Ji[0] = (J[4] * J[8] - J[7] * J[5]) / det;
    Ji[3] = -(J[3] * J[8] - J[6] * J[5]) / det;
    Ji[6] = (J[3] * J[7] - J[6] * J[4]) / det;
    Ji[1] = -(J[1] * J[8] - J[7] * J[2]) / det;
    Ji[4] = (J[0] * J[8] - J[6] * J[2]) / det;
    Ji[7] = -(J[0] * J[7] - J[6] * J[1]) / det;
    Ji[2] = (J[1] * J[5] - J[4] * J[2]) / det;
    Ji[5] = -(J[0] * J[5] - J[3] * J[2]) / det;
    Ji[8] = (J[0] * J[4] - J[3] * J[1]) / det;
    const double dV = 0.125 * (a.span1[e1] * a.span2[e2] * a.span3[e3]);
    const double w = det * (dV * a.gw[q1] * a.gw[q2] * a.gw[q3]);
    // grad u_h: G_cj = sum_i ghat_ci Ji[i][j]
    double G[NC][3];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        G[c][j] = gr[c][0] * Ji[j] + gr[c][1] * Ji[3 + j] + gr[c][2] * Ji[6 + j];
    if (NC == 1) {
      const double kap = field_varies(a.kappa) ? eval_field(a.kappa, X[0], X[1], X[2]) : a.kappa_c;
      acc[0] += w * uh[0] * uh[0];
      acc[1] += w * kap * (G[0][0] * G[0][0] + G[0][1] * G[0][1] + G[0][2] * G[0][2]);
      if (na.has_exact) {
        double ge[3];
        const double ue = eval_field(na.exact, X[0], X[1], X[2]);
        eval_field_grad(na.exact, X[0], X[1], X[2], ge);
        const double d0 = ge[0] - G[0][0], d1 = ge[1] - G[0][1], d2 = ge[2] - G[0][2];
        acc[2] += w * ue * ue;
        acc[3] += w * kap * (ge[0] * ge[0] + ge[1] * ge[1] + ge[2] * ge[2]);
        acc[4] += w * (ue - uh[0]) * (ue - uh[0]);
        acc[5] += w * kap * (d0 * d0 + d1 * d1 + d2 * d2);
      }
    } else {
      double tr = 0.0, ee = 0.0;
#pragma unroll
      for (int r = 0; r < 3; r++) tr += G[r][r];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
          const double eps = 0.5 * (G[r][cc] + G[cc][r]);
          ee += eps * eps;
        }
      double u2 = 0.0;
#pragma unroll
      for (int c = 0; c < NC; c++) u2 += uh[c] * uh[c];
      acc[0] += w * u2;
      acc[1] += w * (a.lam * tr * tr + 2.0 * a.mu * ee);
    }
  }
  __shared__ double sh[4];
  for (int m = 0; m < 6; m++) {
    double v = acc[m];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) na.part[m * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
  }
}

// Squared norm integrals of one patch added to sq[6].
static int norms_patch(Context& c, const PatchRec& Pt, const int* eqmap, const double* gval,
                       const double* xdof, const ifem_gpu_integrand& itg, const int ng[3],
                       const ifem_gpu_function* exact, double* sq);

int integrate_norms(Context& c, const ifem_gpu_integrand& itg, const int ng[3],
                    const ifem_gpu_function* exact, double* out)
{
  double sq[6] = {0, 0, 0, 0, 0, 0};
  if (c.mpc) { // slaves with masters: integrate the expanded solution
    int rc = mpc_expand_device(c);
    if (!rc) rc = norms_patch(c, c.patches[c.patch], c.d_eq, c.d_gval, c.mpc->d_xdof, itg, ng,
                              exact, sq);
    if (rc) return rc;
  } else if (c.multi.empty()) {
    int rc = norms_patch(c, c.patches[c.patch], c.d_eq, c.d_gval, nullptr, itg, ng, exact, sq);
    if (rc) return rc;
  } else {
    for (const SubPatch& sp : c.multi) {
      int rc = norms_patch(c, c.patches[sp.patch], sp.d_leq, sp.d_lgval, nullptr, itg, ng, exact,
                           sq);
      if (rc) return rc;
    }
  }
  for (int m = 0; m < 6; m++) out[m] = sq[m] < 0.0 ? -std::sqrt(-sq[m]) : std::sqrt(sq[m]);
  return IFEM_GPU_OK;
}

static int norms_patch(Context& c, const PatchRec& Pt, const int* eqmap, const double* gval,
                       const double* xdof, const ifem_gpu_integrand& itg, const int ng[3],
                       const ifem_gpu_function* exact, double* sq)
{
  const int p = Pt.p[0];
  if (ng[0] != ng[1] || ng[0] != ng[2] || ng[0] < 1 || ng[0] > kMaxNG)
    return fail(IFEM_GPU_EUNSUPPORTED, "norms: equal Gauss counts 1..5 per direction needed");
  if (exact && c.ncomp != 1)
    return fail(IFEM_GPU_EUNSUPPORTED, "norms: exact solutions are scalar (Poisson) only");
  const int g = ng[0];
  GaussArgs ga{};
  gauss_rule(g, ga.xi, ga.w);
  Tables tb;
  int rc = 0;
  switch (p) {
  case 1: rc = make_tables<1>(c, Pt, g, tb, ga); break;
  case 2: rc = make_tables<2>(c, Pt, g, tb, ga); break;
  case 3: rc = make_tables<3>(c, Pt, g, tb, ga); break;
  default: return fail(IFEM_GPU_EUNSUPPORTED, "norms: degree out of range");
  }
  if (rc) return rc;
  NormArgs na{};
  AsmArgs& a = na.a;
  a.n1 = Pt.n[0]; a.n2 = Pt.n[1]; a.n3 = Pt.n[2];
  a.nel1 = Pt.nel[0]; a.nel2 = Pt.nel[1]; a.nel3 = Pt.nel[2];
  a.coefs = Pt.d_coefs;
  a.tabN1 = tb.tabN[0]; a.tabD1 = tb.tabD[0]; a.span1 = tb.span[0];
  a.tabN2 = tb.tabN[1]; a.tabD2 = tb.tabD[1]; a.span2 = tb.span[1];
  a.tabN3 = tb.tabN[2]; a.tabD3 = tb.tabD[2]; a.span3 = tb.span[2];
  for (int k = 0; k < kMaxNG; k++) a.gw[k] = ga.w[k];
  a.eq = eqmap;
  a.gval = gval;
  a.ng = g;
  a.ncomp = c.ncomp;
  rc = upload_function(c, itg.kappa, a.kappa, 0);
  if (rc) return rc;
  a.kappa_c = const_value(itg.kappa, 1.0);
  if (itg.kind == IFEM_GPU_ELASTICITY) {
    a.lam = itg.E * itg.nu / ((1.0 + itg.nu) * (1.0 - 2.0 * itg.nu));
    a.mu = itg.E / (2.0 * (1.0 + itg.nu));
  }
  if (exact) {
    rc = upload_function(c, *exact, na.exact, 2);
    if (rc) return rc;
    na.has_exact = 1;
  }
  na.x = c.w.x;
  na.xdof = xdof;
  const int64_t total = (int64_t)Pt.nel[0] * Pt.nel[1] * Pt.nel[2] * g * g * g;
  const int grid = 1024;
  double* part = nullptr;
  IFEM_HIP(hipMalloc(&part, sizeof(double) * 6 * grid));
  na.part = part;
  if (c.ncomp == 3) {
    switch (p) {
    case 1: k_norms<1, 3><<<grid, 256, 0, c.stream>>>(na, total); break;
    case 2: k_norms<2, 3><<<grid, 256, 0, c.stream>>>(na, total); break;
    default: k_norms<3, 3><<<grid, 256, 0, c.stream>>>(na, total); break;
    }
  } else {
    switch (p) {
    case 1: k_norms<1, 1><<<grid, 256, 0, c.stream>>>(na, total); break;
    case 2: k_norms<2, 1><<<grid, 256, 0, c.stream>>>(na, total); break;
    default: k_norms<3, 1><<<grid, 256, 0, c.stream>>>(na, total); break;
    }
  }
  const hipError_t le = hipGetLastError();
  std::vector<double> h(6 * grid);
  const hipError_t ce = hipMemcpyAsync(h.data(), part, sizeof(double) * 6 * grid,
                                       hipMemcpyDeviceToHost, c.stream);
  const hipError_t se = hipStreamSynchronize(c.stream);
  hipFree(part);
  if (le != hipSuccess || ce != hipSuccess || se != hipSuccess)
    return fail(IFEM_GPU_EDEVICE, "norms: device error");
  for (int m = 0; m < 6; m++) {
    double sum = 0.0;
    for (int b = 0; b < grid; b++) sum += h[m * grid + b];
    sq[m] += sum;
  }
  return IFEM_GPU_OK;
}

} // namespace ifem
