"""The benchmark problems as concrete model inputs.

Each function returns a ``Problem``: patch, integrand, Gauss count,
equation numbers and an optional traction load, built by ``ifem_amd.model``.

* ``poisson_cube``      -Laplace u = 3 pi^2 sin(pi x)sin(pi y)sin(pi z) on the unit
                        cube, u = 0 on all faces (manufactured solution).
* ``elasticity_cube``   isotropic elasticity (E, nu) on a jittered unit cube,
                        clamped at x = 0, uniform traction (0,0,-1) on x = 1.
* ``diffusion_cube``    -div(kappa grad u) = f with log-normal kappa (random
                        Fourier Gaussian field) and a random smooth f, u = 0.
* ``diffusion_batch``   many independent diffusion patches of uneven size and
                        degree, each with its own kappa and f.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple

import numpy as np

from . import model as M


@dataclasses.dataclass
class Problem:
    name: str
    patch: M.Patch
    integrand: M.Integrand
    ng: Tuple[int, int, int]
    meqn: np.ndarray
    neq: int
    traction: Optional[Tuple[int, Tuple[float, float, float], int]] = None
    # SAM arrays with prescribed values (meqn < 0 at their DOFs), or None
    sam: Optional[M.SamArrays] = None

    def prescribed_values(self) -> np.ndarray:
        """Per DOF prescribed value (0 where none)."""
        return self.sam.prescribed_values() if self.sam is not None else np.zeros(self.ndof)

    @property
    def ncomp(self) -> int:
        return self.integrand.ncomp

    @property
    def ndof(self) -> int:
        return self.patch.nnod * self.ncomp


@dataclasses.dataclass
class MultiProblem:
    """Several patches sharing the nodes of their connected faces: mlgn[q]
    maps patch q's local nodes (GoTools order) to global nodes (1-based);
    meqn holds ncomp equation numbers per global node; tractions are
    (patch, face 0..5, t, ng)."""
    name: str
    patches: list
    mlgn: list
    integrand: M.Integrand
    ng: Tuple[int, int, int]
    meqn: np.ndarray
    neq: int
    tractions: list
    sam: Optional[M.SamArrays] = None

    @property
    def ncomp(self) -> int:
        return self.integrand.ncomp

    @property
    def nnod(self) -> int:
        return int(max(int(m.max()) for m in self.mlgn))

    @property
    def ndof(self) -> int:
        return self.nnod * self.ncomp


def poisson_cube(nel: int = 16, p: int = 2, ng: Optional[int] = None, jitter: float = 0.0,
                 seed: int = 0) -> Problem:
    P = M.cube_patch(p, nel, jitter, seed)
    msc = M.face_status(P, 1, [1] * 6)
    meqn, neq = M.equation_numbers(msc)
    itg = M.Integrand(M.POISSON, source=M.Function(M.FN_SINPROD, 3.0 * math.pi ** 2))
    g = ng or p + 1
    return Problem(f"poisson{nel}p{p}", P, itg, (g, g, g), meqn, neq)


def elasticity_cube(nel: int = 64, p: int = 2, ng: Optional[int] = None, jitter: float = 0.15,
                    seed: int = 2024, E: float = 1.0, nu: float = 0.3) -> Problem:
    P = M.cube_patch(p, nel, jitter, seed)
    fix = [0b111, 0, 0, 0, 0, 0]  # clamped at x = 0 (face u=0)
    msc = M.face_status(P, 3, fix)
    meqn, neq = M.equation_numbers(msc)
    itg = M.Integrand(M.ELASTICITY, E=E, nu=nu)
    g = ng or p + 1
    return Problem(f"elasticity{nel}p{p}", P, itg, (g, g, g), meqn, neq,
                   traction=(M.FACE_U1, (0.0, 0.0, -1.0), g))


def diffusion_cube(nel: int = 48, p: int = 2, ng: Optional[int] = None, sigma: float = 1.5,
                   ell: float = 0.1, seed: int = 7, nmodes: int = 128,
                   source_seed: Optional[int] = None) -> Problem:
    P = M.cube_patch(p, nel)
    msc = M.face_status(P, 1, [1] * 6)
    meqn, neq = M.equation_numbers(msc)
    kappa = M.lognormal_conductivity(seed, sigma, ell, nmodes)
    src = M.smooth_source(seed + 1 if source_seed is None else source_seed)
    itg = M.Integrand(M.POISSON, kappa=kappa, source=src)
    g = ng or p + 1
    return Problem(f"diffusion{nel}p{p}", P, itg, (g, g, g), meqn, neq)


def diffusion_batch(npatch: int = 256, nel_range: Tuple[int, int] = (8, 48),
                    degrees: Tuple[int, ...] = (2, 3), sigma: float = 1.0, ell: float = 0.2,
                    seed: int = 11, nmodes: int = 128):
    """``npatch`` unit-cube diffusion patches; patch q draws its element count
    per direction uniformly from ``nel_range`` (inclusive) and its degree from
    ``degrees`` (splitmix64 stream ``seed``), with a log-normal conductivity
    (seed ``seed*100003 + 2q``) and a smooth source (seed ``seed*100003 + 2q + 1``),
    u = 0 on all faces, ng = p+1."""
    lo, hi = nel_range
    u = M.uniform(seed, 4 * npatch).reshape(npatch, 4)
    out = []
    for q in range(npatch):
        nel = [lo + min(int(u[q, d] * (hi - lo + 1)), hi - lo) for d in range(3)]
        p = degrees[min(int(u[q, 3] * len(degrees)), len(degrees) - 1)]
        P = M.cube_patch(p, nel)
        msc = M.face_status(P, 1, [1] * 6)
        meqn, neq = M.equation_numbers(msc)
        base = seed * 100003 + 2 * q
        kappa = M.lognormal_conductivity(base, sigma, ell, nmodes)
        itg = M.Integrand(M.POISSON, kappa=kappa, source=M.smooth_source(base + 1))
        g = p + 1
        out.append(Problem(f"batch{q}:{nel[0]}x{nel[1]}x{nel[2]}p{p}", P, itg, (g, g, g), meqn, neq))
    return out


def dof_counts(pb: Problem):
    """(total DOFs incl. eliminated ones, free equations)."""
    return pb.ndof, pb.neq
