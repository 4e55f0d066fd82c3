"""Model generation for single tensor-product spline patches.

Host-side input preparation, the counterpart of the reference's default
geometry generator and SAM set-up:

* ``cube_patch``: a trivariate B-spline block on the unit cube with open
  uniform knot vectors and control points at the Greville abscissae, i.e. the
  identity geometry that ``DefaultGeometry3D`` followed by ``refine`` and
  ``raiseorder`` produces (src/SIM/ModelGenerator.C:203-273), plus an optional
  seeded jitter of the interior control points.
* ``face_status`` / ``equation_numbers``: status codes for homogeneous
  Dirichlet faces and SAM's equation numbering
  (src/ASM/ASMbase.C:884-908, src/LinAlg/SAM.C:225-336);
  ``prescribed_equations``: SAM arrays with prescribed (inhomogeneous
  Dirichlet) values as constraint equations without masters.
* ``Function`` / ``Integrand``: the recognised integrand kinds of the GPU path
  (Poisson with coefficient and source fields, isotropic elasticity).

Random numbers come from splitmix64 so the CPU and GPU sides see identical
inputs for a given seed.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` consecutive splitmix64 outputs for ``seed`` (uint64 array)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in [0, 1) with 53 random bits each."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal(seed: int, count: int) -> np.ndarray:
    """Standard normal deviates by Box-Muller on splitmix64 uniforms."""
    m = (count + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:count]


def open_uniform_knots(p: int, nel: int) -> np.ndarray:
    """Open knot vector on [0,1] with nel equal spans and simple interior knots."""
    inner = [i / nel for i in range(1, nel)]
    return np.array([0.0] * (p + 1) + inner + [1.0] * (p + 1), dtype=np.float64)


def greville(p: int, t: np.ndarray) -> np.ndarray:
    n = len(t) - p - 1
    return np.array([t[i + 1:i + p + 1].sum() / p for i in range(n)], dtype=np.float64)


@dataclasses.dataclass
class Patch:
    """Non-rational trivariate B-spline patch (GoTools SplineVolume analogue).

    ``p`` are degrees (GoTools order minus one), ``n`` control points per
    direction, ``coefs`` the control points with node (i,j,k) at row
    ``(k*n2 + j)*n1 + i`` (src/ASM/ASMs3D.C:1600-1615 layout).
    """
    p: List[int]
    n: List[int]
    knots: List[np.ndarray]
    coefs: np.ndarray

    @property
    def nel(self) -> List[int]:
        return [self.n[d] - self.p[d] for d in range(3)]

    @property
    def nnod(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def nen(self) -> int:
        return (self.p[0] + 1) * (self.p[1] + 1) * (self.p[2] + 1)

    @property
    def order(self) -> List[int]:
        return [q + 1 for q in self.p]


def cube_patch(p, nel, jitter: float = 0.0, seed: int = 0) -> Patch:
    """Unit-cube patch of degree p with nel elements per direction.

    ``p`` and ``nel`` are ints or 3-sequences.  With ``jitter`` > 0 every
    control point that lies on no face of the cube is moved by
    ``(2u-1)*jitter*h_d`` per coordinate (h_d = 1/nel_d), u drawn from
    splitmix64(seed) in node order (i fastest), x then y then z per node.
    """
    ps = [p] * 3 if isinstance(p, int) else list(p)
    ne = [nel] * 3 if isinstance(nel, int) else list(nel)
    knots = [open_uniform_knots(ps[d], ne[d]) for d in range(3)]
    g = [greville(ps[d], knots[d]) for d in range(3)]
    n = [len(x) for x in g]
    X, Y, Z = np.meshgrid(g[0], g[1], g[2], indexing="ij")
    coefs = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                      Z.transpose(2, 1, 0).ravel()], axis=1).astype(np.float64)
    if jitter > 0.0:
        i = np.arange(n[0]); j = np.arange(n[1]); k = np.arange(n[2])
        K, J, I = np.meshgrid(k, j, i, indexing="ij")
        interior = ((I > 0) & (I < n[0] - 1) & (J > 0) & (J < n[1] - 1)
                    & (K > 0) & (K < n[2] - 1)).ravel()
        idx = np.nonzero(interior)[0]
        u = uniform(seed, 3 * len(idx)).reshape(-1, 3)
        h = np.array([1.0 / ne[0], 1.0 / ne[1], 1.0 / ne[2]])
        coefs[idx] += (2.0 * u - 1.0) * jitter * h
    return Patch(ps, n, knots, np.ascontiguousarray(coefs))


# IFEM face numbering 1..6 is 0..5 here: u=0, u=1, v=0, v=1, w=0, w=1.
FACE_U0, FACE_U1, FACE_V0, FACE_V1, FACE_W0, FACE_W1 = range(6)


def face_status(patch: Patch, ncomp: int, fixmask: Sequence[int]) -> np.ndarray:
    """SAM status codes (1 free, 0 fixed) for homogeneous Dirichlet faces.

    ``fixmask[f]`` is a bit mask of the components fixed on face f.
    """
    n1, n2, n3 = patch.n
    K, J, I = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    I = I.ravel(); J = J.ravel(); K = K.ravel()
    m = np.zeros(I.shape, dtype=np.int64)
    m |= np.where(I == 0, fixmask[0], 0)
    m |= np.where(I == n1 - 1, fixmask[1], 0)
    m |= np.where(J == 0, fixmask[2], 0)
    m |= np.where(J == n2 - 1, fixmask[3], 0)
    m |= np.where(K == 0, fixmask[4], 0)
    m |= np.where(K == n3 - 1, fixmask[5], 0)
    msc = np.ones((len(I), ncomp), dtype=np.int32)
    for c in range(ncomp):
        msc[:, c] = np.where((m >> c) & 1, 0, 1)
    return msc.ravel()


def gauss_points(order: int, nGauss: int, neumann: bool = False) -> int:
    """Gauss points per direction, ASMbase::getNoGaussPt
    (src/ASM/ASMbase.C:1788-1798).  ``order`` is the spline order p+1, as
    ASMs3D passes it (getOrder, src/ASM/ASMs3D.C:1898-1906, used at
    src/ASM/ASMs3D.C:4033 for volumes and 2593 for faces).  1..10 is a fixed
    count; nGauss <= -order gives 0; otherwise faces take ``order`` and
    volumes ``order + nGauss % 10`` with C's remainder (sign of nGauss)."""
    nGauss = int(nGauss)
    if nGauss <= -order:
        return 0
    if 0 < nGauss < 11:
        return nGauss
    if neumann:
        return order
    rem = abs(nGauss) % 10
    return order + (-rem if nGauss < 0 else rem)


def equation_numbers(msc: np.ndarray):
    """SAM::initSystemEquations for status codes 0/1: returns (meqn, neq)."""
    free = msc == 1
    meqn = np.zeros(len(msc), dtype=np.int32)
    meqn[free] = np.arange(1, int(free.sum()) + 1, dtype=np.int32)
    return meqn, int(free.sum())


@dataclasses.dataclass
class SamArrays:
    """SAM's equation arrays with constraint equations (1-based, SAM layout):
    meqn per DOF (> 0 equation, 0 fixed, -k slave of constraint k); the
    entries of constraint k are mpmceq[k-1] .. mpmceq[k]-1 of mmceq (DOF
    numbers, slave first) and ttcc (constant c0 first, then coefficients)."""
    meqn: np.ndarray
    neq: int
    mpmceq: np.ndarray
    mmceq: np.ndarray
    ttcc: np.ndarray

    @property
    def nceq(self) -> int:
        return len(self.mpmceq) - 1

    def prescribed_values(self) -> np.ndarray:
        """Per DOF c0 of its constraint (0 for free and fixed DOFs)."""
        g = np.zeros(len(self.meqn))
        k = -self.meqn[self.meqn < 0]
        g[self.meqn < 0] = self.ttcc[self.mpmceq[k - 1] - 1]
        return g


def prescribed_equations(msc: np.ndarray, values: np.ndarray) -> SamArrays:
    """SAM arrays for status codes 0/1 plus prescribed values: every free DOF
    with a finite ``values`` entry becomes the slave of a constraint equation
    without masters, u_i = c0 (ASMbase::addMPC + SAMpatch::initConstraintEqs,
    src/ASM/ASMbase.C:527-560, src/ASM/SAMpatch.C:259-357: constraints in
    slave-DOF order, status of the slaves reset to 0, then
    SAM::initSystemEquations numbers the free DOFs, src/LinAlg/SAM.C:225-336).
    A value on a DOF that is already fixed (msc 0) is ignored, as addMPC
    ignores constraints on FIXED DOFs."""
    msc = np.asarray(msc, dtype=np.int32).copy()
    values = np.asarray(values, dtype=np.float64)
    pres = np.nonzero(np.isfinite(values) & (msc == 1))[0]
    msc[pres] = 0
    meqn, neq = equation_numbers(msc)
    meqn[pres] = -np.arange(1, len(pres) + 1, dtype=np.int32)
    mpmceq = np.arange(1, len(pres) + 2, dtype=np.int32)
    return SamArrays(meqn, neq, mpmceq, (pres + 1).astype(np.int32),
                     values[pres].astype(np.float64))


@dataclasses.dataclass
class MultiPatch:
    """Patches side by side along x that share their common faces: patch q
    covers [x_q, x_{q+1}] x [0,1]^2; global nodes form one (gx, ny, nz) grid
    (i fastest) in which the shared faces appear once (SAM's MLGN)."""
    patches: List[Patch]
    mlgn: List[np.ndarray]     # per patch: local node -> global node (1-based)
    dims: List[int]            # global node grid (gx, ny, nz)

    @property
    def nnod(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]


def x_split_patches(p: int, nels: Sequence[int], nyz: int, jitter: float = 0.0,
                    seed: int = 0) -> MultiPatch:
    """len(nels) degree-p patches along x with nels[q] x nyz x nyz elements,
    widths proportional to nels[q]; interior control points of every patch
    jittered as in cube_patch (seed + q), patch faces left exact so the shared
    faces coincide."""
    tot = float(sum(nels))
    x0 = 0.0
    patches, mlgn = [], []
    off = 0
    ny = nyz + p
    n_tot = sum(ne + p for ne in nels) - (len(nels) - 1)
    for q, ne in enumerate(nels):
        P = cube_patch(p, [ne, nyz, nyz], jitter, seed + q)
        w = ne / tot
        P.coefs[:, 0] = x0 + w * P.coefs[:, 0]
        x0 += w
        n1 = P.n[0]
        k, j, i = np.meshgrid(np.arange(ny), np.arange(ny), np.arange(n1), indexing="ij")
        mlgn.append((off + i.ravel() + n_tot * (j.ravel() + ny * k.ravel()) + 1).astype(np.int32))
        off += n1 - 1
        patches.append(P)
    return MultiPatch(patches, mlgn, [n_tot, ny, ny])


def constraint_equations(msc: np.ndarray, constraints) -> SamArrays:
    """SAM arrays for status codes 0/1 plus linear constraint equations
    u_s = c0 + sum_m c_m u_m, given as (slave DOF, c0, [(master DOF, c_m)...])
    with 0-based DOFs (SAMpatch::initConstraintEqs, src/ASM/SAMpatch.C:259-357:
    constraints in slave-DOF order; a constraint on a fixed DOF is ignored; a
    fixed master is dropped, leaving its constant; the slaves' status is reset
    to 0 before SAM::initSystemEquations numbers the free DOFs)."""
    msc = np.asarray(msc, dtype=np.int32).copy()
    fixed0 = msc == 0
    cons = sorted([c for c in constraints if not fixed0[c[0]]], key=lambda c: c[0])
    slaves = [c[0] for c in cons]
    if len(set(slaves)) != len(slaves):
        raise ValueError("one constraint per slave DOF")
    msc[slaves] = 0
    meqn, neq = equation_numbers(msc)
    mpmceq, mmceq, ttcc = [1], [], []
    for k, (sd, c0, masters) in enumerate(cons):
        meqn[sd] = -(k + 1)
        mmceq.append(sd + 1)
        ttcc.append(float(c0))
        for md, cm in masters:
            if fixed0[md]:
                continue
            if md in slaves:
                raise ValueError("master DOFs must be free (resolve chains first)")
            mmceq.append(md + 1)
            ttcc.append(float(cm))
        mpmceq.append(len(mmceq) + 1)
    return SamArrays(meqn, neq, np.array(mpmceq, dtype=np.int32),
                     np.array(mmceq, dtype=np.int32), np.array(ttcc, dtype=np.float64))


def periodic_constraints(patch: Patch, ncomp: int, direction: int = 0):
    """Constraints tying each DOF on the face at the end of `direction` to
    the DOF of the node at its start (u_end = u_start), as periodicity
    conditions are given to the reference (ASMs3D::addPeriodicity)."""
    n = patch.n
    k, j, i = [a.ravel() for a in np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")]
    idx = (i, j, k)[direction]
    end = np.nonzero(idx == n[direction] - 1)[0]
    shift = [1, n[0], n[0] * n[1]][direction] * (n[direction] - 1)
    return [(int(nd) * ncomp + c, 0.0, [((int(nd) - shift) * ncomp + c, 1.0)])
            for nd in end for c in range(ncomp)]


# ---- integrand descriptions (same numbering as the C ABI) -----------------
FN_ZERO, FN_CONST, FN_SINPROD, FN_FOURIER, FN_EXP_FOURIER = range(5)
POISSON, ELASTICITY = 1, 2


@dataclasses.dataclass
class Function:
    """Scalar field: CONST scale; SINPROD scale*sin(pi x)sin(pi y)sin(pi z);
    FOURIER mean + scale*sum_m a_m cos(k_m.x + phi_m); EXP_FOURIER exp(...)."""
    type: int = FN_ZERO
    scale: float = 0.0
    mean: float = 0.0
    modes: Optional[np.ndarray] = None  # (nmodes, 5): kx ky kz phase amp

    @property
    def nmodes(self) -> int:
        return 0 if self.modes is None else int(self.modes.shape[0])


@dataclasses.dataclass
class Integrand:
    kind: int = POISSON
    E: float = 1.0
    nu: float = 0.3
    kappa: Function = dataclasses.field(default_factory=Function)
    source: Function = dataclasses.field(default_factory=Function)
    body: Sequence[float] = (0.0, 0.0, 0.0)

    @property
    def ncomp(self) -> int:
        return 3 if self.kind == ELASTICITY else 1


def gaussian_field_modes(seed: int, ell: float, nmodes: int = 128) -> np.ndarray:
    """Random Fourier features of a squared-exponential Gaussian field with
    correlation length ``ell``: G(x) = sqrt(2/M) sum cos(k_m.x + phi_m),
    k_m ~ N(0, I/ell^2), phi_m ~ U(0, 2 pi).  Unit variance; scale by sigma."""
    k = normal(seed, 3 * nmodes).reshape(nmodes, 3) / ell
    phi = 2.0 * np.pi * uniform(seed ^ 0x5DEECE66D, nmodes)
    amp = np.full(nmodes, math.sqrt(2.0 / nmodes))
    return np.ascontiguousarray(np.column_stack([k, phi, amp]))


def lognormal_conductivity(seed: int, sigma: float, ell: float, nmodes: int = 128) -> Function:
    return Function(FN_EXP_FOURIER, sigma, 0.0, gaussian_field_modes(seed, ell, nmodes))


def smooth_source(seed: int, nmodes: int = 16, kmax: int = 3) -> Function:
    """Random smooth source: sum of low-frequency cosines with N(0,1)/(1+|k|^2)
    amplitudes, integer wave numbers times pi up to kmax."""
    u = uniform(seed, 4 * nmodes).reshape(nmodes, 4)
    k = np.floor(u[:, :3] * (kmax + 1)) * np.pi
    phi = 2.0 * np.pi * u[:, 3]
    amp = normal(seed ^ 0xA5A5A5A5, nmodes) / (1.0 + (k ** 2).sum(1) / np.pi ** 2)
    return Function(FN_FOURIER, 1.0, 0.0, np.ascontiguousarray(np.column_stack([k, phi, amp])))


def lame(E: float, nu: float):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
