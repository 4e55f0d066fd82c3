"""Spline multigrid preconditioner (pc = PC_MG).

The hierarchy drops every second distinct interior knot per level, so the
coarse spline space is a subspace of the fine one.  The prolongation is
checked here against an independent construction (collocation at the fine
Greville points with scipy's B-spline design matrices, P = B_f^-1 B_c), the
Galerkin operators against P^T K P formed with scipy from the oracle's K, and
the preconditioned solves against the oracle's solution.
"""
import numpy as np
import pytest
import scipy.interpolate as si
import scipy.sparse as sp

import oracle
from ifem_amd import FP32, FP64, PC_MG, model as M, problems, setup_problem


def coarse_knots(t, p):
    """Every second distinct interior breakpoint of t removed."""
    t = np.asarray(t)
    inner = t[p + 1:len(t) - p - 1]
    br, mult = np.unique(inner, return_counts=True)
    keep = np.repeat(br[1::2], mult[1::2])
    return np.concatenate([t[:p + 1], keep, t[len(t) - p - 1:]])


def prolong_1d(tc, tf, p):
    """P with B_c = B_f P, by collocation at the fine Greville points."""
    g = M.greville(p, np.asarray(tf))
    g[-1] = np.nextafter(g[-1], -np.inf)  # stay inside the last span
    Bf = si.BSpline.design_matrix(g, tf, p).toarray()
    Bc = si.BSpline.design_matrix(g, tc, p).toarray()
    P = np.linalg.solve(Bf, Bc)
    P[np.abs(P) < 1e-13] = 0.0
    return P


def test_prolongation_reference_properties():
    """Partition of unity carries over (P 1 = 1), P >= 0, the dyadic p=2
    interior columns are (1/4, 3/4, 3/4, 1/4), and coarse B-splines equal
    their fine expansions at random points."""
    t = M.open_uniform_knots(2, 8)
    tc = coarse_knots(t, 2)
    assert np.allclose(tc, M.open_uniform_knots(2, 4))
    P = prolong_1d(tc, t, 2)
    assert np.allclose(P.sum(axis=1), 1.0, atol=1e-14) and P.min() >= -1e-15
    assert np.allclose(P[2:6, 2], [0.25, 0.75, 0.75, 0.25], atol=1e-14)
    x = np.random.default_rng(0).random(50)
    Bf = si.BSpline.design_matrix(x, t, 2).toarray()
    Bc = si.BSpline.design_matrix(x, tc, 2).toarray()
    assert np.max(np.abs(Bf @ P - Bc)) < 1e-14
    # odd element count: the last coarse span covers one fine span
    t5 = M.open_uniform_knots(3, 5)
    tc5 = coarse_knots(t5, 3)
    assert len(tc5) - 3 - 1 == 6  # 3 elements
    P5 = prolong_1d(tc5, t5, 3)
    assert np.allclose(P5.sum(axis=1), 1.0, atol=1e-14)


def _dof_prolongation(patch, ncomp):
    """Node-major / component-minor DOF prolongation and the coarse knots."""
    Ps, tcs = [], []
    for d in range(3):
        tc = coarse_knots(patch.knots[d], patch.p[d])
        tcs.append(tc)
        Ps.append(prolong_1d(tc, patch.knots[d], patch.p[d]))
    Pn = sp.kron(sp.csr_matrix(Ps[2]), sp.kron(sp.csr_matrix(Ps[1]), sp.csr_matrix(Ps[0])))
    return sp.kron(Pn, sp.identity(ncomp)).tocsr(), tcs, [P.shape[1] for P in Ps]


def _stencil_to_dense(S, n, p, nc):
    nnode = n[0] * n[1] * n[2]
    W1 = 2 * p + 1
    S = S.reshape(nnode, W1 ** 3, nc, nc).transpose(1, 2, 3, 0)
    A = np.zeros((nnode * nc, nnode * nc))
    k, j, i = [a.ravel() for a in np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")]
    node = np.arange(nnode)
    for off in range(W1 ** 3):
        e0, e1, e2 = off % W1 - p, (off // W1) % W1 - p, off // (W1 * W1) - p
        ii, jj, kk = i + e0, j + e1, k + e2
        ok = (ii >= 0) & (ii < n[0]) & (jj >= 0) & (jj < n[1]) & (kk >= 0) & (kk < n[2])
        nb = ii + n[0] * (jj + n[1] * kk)
        for ci in range(nc):
            for cj in range(nc):
                A[node[ok] * nc + ci, nb[ok] * nc + cj] = S[off, ci, cj, ok]
    return A


def _coarse_free(free_f, P):
    """Coarse DOF kept iff every fine DOF in its column of P is free."""
    bad = (abs(P).T @ (~free_f).astype(float)) > 0
    return ~bad


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", [
    ("elast6p2", lambda: problems.elasticity_cube(6, 2, seed=3)),
    ("diffusion6p3", lambda: problems.diffusion_cube(6, 3, sigma=1.0, ell=0.2, seed=2)),
    ("poisson9p1", lambda: problems.poisson_cube(9, 1, jitter=0.1, seed=4)),
])
def test_galerkin_operators_match_scipy(name, make):
    """Levels 1 and 2 of the hierarchy equal P^T A P (scipy, from the oracle's
    K embedded at the free DOFs) restricted to the kept coarse DOFs, 1e-12."""
    pb = make()
    print(f"[gpu] multigrid Galerkin {name}", flush=True)
    rp, ci, A, b = oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq)
    nc = pb.ncomp
    free = pb.meqn > 0
    idx = np.nonzero(free)[0]
    K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq))
    E = sp.csr_matrix((np.ones(len(idx)), (idx, np.arange(len(idx)))), shape=(pb.ndof, pb.neq))
    Af = (E @ K @ E.T).tocsr()
    patch = pb.patch
    with setup_problem(pb) as s:
        nl, ms = s.mg_setup()
        assert nl >= 2
        for level in (1, 2):
            if level >= nl:
                break
            P, tcs, ncs = _dof_prolongation(patch, nc)
            free = _coarse_free(free, P)
            Ac = (P.T @ Af @ P).toarray()
            Ac[~free, :] = 0.0
            Ac[:, ~free] = 0.0
            n, nfree, lam, S = s.mg_level(level, operator=True)
            assert n == ncs and nfree == int(free.sum())
            G = _stencil_to_dense(S, n, patch.p[0], nc)
            assert np.max(np.abs(G - Ac)) <= 1e-12 * np.max(np.abs(Ac)), level
            # next level from this one
            patch = M.Patch(patch.p, ncs, tcs, np.zeros((ncs[0] * ncs[1] * ncs[2], 3)))
            Af = sp.csr_matrix(Ac)


@pytest.mark.gpu
def test_vcycle_is_symmetric_positive_and_reduces_the_error():
    """<M r1, r2> = <r1, M r2> to FP32 rounding, <r, M r> > 0, and the
    preconditioned Richardson error contracts: ||e - M A e||_A < 0.5 ||e||_A."""
    pb = problems.elasticity_cube(8, 2, seed=5)
    print("[gpu] multigrid V-cycle properties elasticity 8^3", flush=True)
    rp, ci, A, b = oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq)
    K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq))
    rng = np.random.default_rng(3)
    r1, r2, e = (rng.standard_normal(pb.neq) for _ in range(3))
    with setup_problem(pb) as s:
        s.mg_setup()
        z1, z2 = s.mg_apply(r1), s.mg_apply(r2)
        w = s.mg_apply(K @ e)
    assert abs(z1 @ r2 - r1 @ z2) <= 1e-5 * np.sqrt((z1 @ r1) * (z2 @ r2))
    assert z1 @ r1 > 0 and z2 @ r2 > 0
    d = e - w
    assert np.sqrt(d @ (K @ d)) < 0.5 * np.sqrt(e @ (K @ e))


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,prec,smooth,gain", [
    ("elast16", lambda: problems.elasticity_cube(16, 2, seed=6), FP64, 2, 5),
    ("elast16_mixed", lambda: problems.elasticity_cube(16, 2, seed=6), FP32, 2, 5),
    ("poisson16p3", lambda: problems.poisson_cube(16, 3, jitter=0.1, seed=1), FP64, 2, 2),
    ("poisson16p3s4", lambda: problems.poisson_cube(16, 3, jitter=0.1, seed=1), FP64, 4, 3),
    ("diffusion12p1", lambda: problems.diffusion_cube(12, 1, sigma=1.0, ell=0.2, seed=8), FP64, 2,
     3),
])
def test_multigrid_pcg_matches_oracle(name, make, prec, smooth, gain):
    """Multigrid-preconditioned flexible CG reaches the oracle's solution
    (1e-9 relative at rtol 1e-11) in at least `gain` times fewer iterations
    than Jacobi (point smoothers lose strength with the degree: p = 3 gains
    less, and more from more smoothing steps)."""
    pb = make()
    print(f"[gpu] multigrid solve {name}", flush=True)
    rp, ci, A, b = oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, pb.traction)
    xo, ito, _ = oracle.pcg(rp, ci, A, b, rtol=1e-12)
    with setup_problem(pb, precision=prec) as s:
        st = s.solve(rtol=1e-11, maxits=2000, pc=PC_MG, precision=prec, mg_smooth=smooth)
        x = s.solution()
        stj = s.solve(rtol=1e-11, maxits=20000)
    print(f"   iterations multigrid {st['iterations']} Jacobi {stj['iterations']}", flush=True)
    assert st["converged"] and st["rel_residual"] <= 1e-11
    assert np.max(np.abs(x - xo)) <= 1e-9 * np.max(np.abs(xo))
    assert st["iterations"] * gain < stj["iterations"], (st["iterations"], stj["iterations"])


@pytest.mark.gpu
def test_multigrid_with_prescribed_values():
    """Prescribed boundary values: eliminated DOFs carry over to the coarse
    levels (coarse DOFs touching them are dropped); solution equals the
    oracle's."""
    pb = problems.elasticity_cube(10, 2, seed=4)
    P, nc = pb.patch, 3
    n = P.n
    k, j, i = [a.ravel() for a in np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")]
    vals = np.full((P.nnod, 3), np.nan)
    vals[i == 0] = 0.0
    vals[i == n[0] - 1, 0] = 0.02
    vals[k == n[2] - 1, 2] = np.linspace(-0.01, 0.01, int((k == n[2] - 1).sum()))
    sam = M.prescribed_equations(np.ones(P.nnod * 3, dtype=np.int32), vals.ravel())
    pb = problems.Problem("pres10", P, pb.integrand, pb.ng, sam.meqn, sam.neq, None, sam)
    print("[gpu] multigrid prescribed values", flush=True)
    g = sam.prescribed_values()
    rp, ci = oracle.pattern(P, nc, sam.meqn, sam.neq)
    A, b = oracle.assemble(P, pb.integrand, pb.ng, sam.meqn, sam.neq, rp, ci, gdof=g)
    xo, _, _ = oracle.pcg(rp, ci, A, b, rtol=1e-12)
    with setup_problem(pb) as s:
        st = s.solve(rtol=1e-11, maxits=2000, pc=PC_MG)
        x = s.solution()
    assert st["converged"]
    assert np.max(np.abs(x - xo)) <= 1e-9 * np.max(np.abs(xo))
