"""Several patches in one system, sharing the nodes of their common faces.

The oracle assembles patch by patch into the global CSR (each patch with the
global equation numbers of its local DOFs, as SAM::assembleSystem does over
the patches of a model); it is pinned by a p=1 split, which spans the same
space as one patch and must give the same matrix, and by linear patch tests
across the interfaces.  The GPU path is compared with it.
"""
import types

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from ifem_amd import FP32, FP64, model as M


def _global_status(mp, nc, fixmask):
    return M.face_status(types.SimpleNamespace(n=mp.dims), nc, fixmask)


def _local(arr, mlgn, nc):
    """Global per-DOF array gathered to a patch's local DOFs."""
    idx = ((np.asarray(mlgn, dtype=np.int64) - 1)[:, None] * nc + np.arange(nc)).ravel()
    return np.asarray(arr)[idx]


def _union_pattern(mp, nc, meqn, neq):
    rows = [set() for _ in range(neq)]
    for P, ml in zip(mp.patches, mp.mlgn):
        rp, ci = oracle.pattern(P, nc, _local(meqn, ml, nc), neq)
        for r in range(neq):
            rows[r].update(ci[rp[r]:rp[r + 1]].tolist())
    rowptr = np.zeros(neq + 1, dtype=np.int64)
    for r in range(neq):
        rowptr[r + 1] = rowptr[r] + len(rows[r])
    colind = np.concatenate([np.array(sorted(s), dtype=np.int32) for s in rows]) if neq else \
        np.zeros(0, dtype=np.int32)
    return rowptr, colind


def _oracle_system(mp, itg, ng, meqn, neq, gdof=None):
    nc = itg.ncomp
    rp, ci = _union_pattern(mp, nc, meqn, neq)
    A = np.zeros(len(ci))
    b = np.zeros(neq)
    for P, ml in zip(mp.patches, mp.mlgn):
        oracle.assemble_into(P, itg, ng, _local(meqn, ml, nc), neq, rp, ci, A, b,
                             None if gdof is None else _local(gdof, ml, nc))
    return rp, ci, A, b


def test_p1_split_equals_one_patch():
    """Linear splines are C0 anyway: two p=1 patches side by side span the
    same space as one patch over both, so K and b agree (1e-13)."""
    mp = M.x_split_patches(1, [3, 4], 4)
    itg = M.Integrand(M.POISSON, source=M.Function(M.FN_SINPROD, 2.0))
    msc = _global_status(mp, 1, [1] * 6)
    meqn, neq = M.equation_numbers(msc)
    rp, ci, A, b = _oracle_system(mp, itg, [2, 2, 2], meqn, neq)
    P = M.cube_patch(1, [7, 4, 4])
    rp1, ci1, A1, b1 = oracle.system(P, itg, [2, 2, 2], meqn, neq)
    assert np.array_equal(rp, rp1) and np.array_equal(ci, ci1)
    assert np.max(np.abs(A - A1)) <= 1e-13 * np.max(np.abs(A1))
    assert np.max(np.abs(b - b1)) <= 1e-13 * np.max(np.abs(b1))


def _linear_patch_test(mp, p, ng):
    itg = M.Integrand(M.POISSON)
    X = np.zeros((mp.nnod, 3))
    for P, ml in zip(mp.patches, mp.mlgn):
        X[ml - 1] = P.coefs
    u = 0.3 - 1.2 * X[:, 0] + 0.7 * X[:, 1] + 2.0 * X[:, 2]
    gx, ny, nz = mp.dims
    k, j, i = [a.ravel() for a in np.meshgrid(range(nz), range(ny), range(gx), indexing="ij")]
    bnd = (i == 0) | (i == gx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
    sam = M.prescribed_equations(np.ones(mp.nnod, dtype=np.int32), np.where(bnd, u, np.nan))
    return itg, sam, u


def test_linear_patch_test_across_interfaces_oracle():
    """-Laplace u = 0, u linear on the outer boundary, three jittered p=2
    patches: the solution is u at every control point."""
    mp = M.x_split_patches(2, [3, 2, 4], 4, jitter=0.2, seed=6)
    itg, sam, u = _linear_patch_test(mp, 2, 3)
    g = sam.prescribed_values()
    rp, ci, A, b = _oracle_system(mp, itg, [3, 3, 3], sam.meqn, sam.neq, g)
    x, it, res = oracle.pcg(rp, ci, A, b, rtol=1e-14)
    xd = oracle.expand_solution(sam.meqn, x, g)
    assert np.max(np.abs(xd - u)) < 1e-10


def _gpu_setup(mp, itg, ng, meqn, neq, sam=None, precision=FP64):
    from ifem_amd import Session
    s = Session()
    pids = [s.add_patch(P) for P in mp.patches]
    s.bind_patches(pids, mp.mlgn, itg.ncomp, meqn, neq, sam=sam)
    s.init_assembly()
    s.assemble(itg, ng, FP64)
    if precision == FP32:
        s.assemble(itg, ng, FP32)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("name,p,nels,nyz,kind", [
    ("poisson2p", 2, [3, 4], 5, M.POISSON),
    ("diffusion3p3", 3, [2, 3, 2], 3, M.POISSON),
    ("elast2p", 2, [4, 3], 4, M.ELASTICITY),
])
def test_multipatch_assembly_and_solve_match_oracle(name, p, nels, nyz, kind):
    """K, b (1e-12), the PCG solution (1e-12) and the norms of the solution
    (1e-12) of a jittered multi-patch model against the oracle."""
    mp = M.x_split_patches(p, nels, nyz, jitter=0.15, seed=len(name))
    nc = 3 if kind == M.ELASTICITY else 1
    if kind == M.ELASTICITY:
        itg = M.Integrand(M.ELASTICITY, E=1.0, nu=0.3, body=(0.0, 0.0, -1.0))
        msc = _global_status(mp, 3, [7, 0, 0, 0, 0, 0])
    else:
        itg = M.Integrand(M.POISSON, kappa=M.lognormal_conductivity(3, 1.0, 0.3, 16),
                          source=M.smooth_source(4))
        msc = _global_status(mp, 1, [1] * 6)
    meqn, neq = M.equation_numbers(msc)
    ng = [p + 1] * 3
    print(f"[gpu] multi-patch {name} neq={neq}", flush=True)
    rp, ci, A, b = _oracle_system(mp, itg, ng, meqn, neq)
    # solved to where both lie within ~2e-14 of the direct solution (FP64
    # CG stagnates just above 1e-14 on the elasticity model)
    rtol = 1e-13 if kind == M.ELASTICITY else 1e-14
    with _gpu_setup(mp, itg, ng, meqn, neq) as s:
        grp, gci, gA, gb = s.system()
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A))
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
        st = s.solve(rtol=rtol, maxits=20000)
        assert st["converged"]
        x = s.solution()
        xd = s.expanded_solution(len(meqn))
        n = s.norms(itg, ng)
    xo, _, _ = oracle.pcg(rp, ci, A, b, rtol=rtol)
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))
    xdir = spla.spsolve(sp.csr_matrix((A, ci, rp), shape=(neq, neq)).tocsc(), b)
    assert np.max(np.abs(x - xdir)) <= 1e-12 * np.max(np.abs(xdir))
    sq = np.zeros(6)
    for P, ml in zip(mp.patches, mp.mlgn):
        sq += oracle.norms(P, itg, ng, _local(xd, ml, nc)) ** 2
    assert np.all(np.abs(n - np.sqrt(sq)) <= 1e-12 * np.maximum(np.sqrt(sq), 1e-300)), (n, sq)


@pytest.mark.gpu
def test_multipatch_linear_patch_test_and_fp32():
    """The linear patch test across three jittered p=2 patches on the GPU
    (prescribed boundary values), FP64 and mixed precision."""
    mp = M.x_split_patches(2, [3, 2, 4], 4, jitter=0.2, seed=6)
    itg, sam, u = _linear_patch_test(mp, 2, 3)
    print("[gpu] multi-patch linear patch test", flush=True)
    with _gpu_setup(mp, itg, [3, 3, 3], sam.meqn, sam.neq, sam=sam, precision=FP32) as s:
        st = s.solve(rtol=1e-13, maxits=20000)
        assert st["converged"]
        x64 = s.expanded_solution(len(sam.meqn))
        st = s.solve(rtol=1e-12, maxits=20000, precision=FP32)
        assert st["converged"]
        x32 = s.expanded_solution(len(sam.meqn))
    assert np.max(np.abs(x64 - u)) < 1e-10
    assert np.max(np.abs(x32 - u)) < 1e-9


@pytest.mark.gpu
def test_multipatch_traction_on_last_patch():
    """Clamped at x = 0 (patch 0), traction (0,0,-1) on x = 1 (last patch):
    b against the oracle's face integration through the patch's local
    numbering, resultant -1."""
    mp = M.x_split_patches(2, [3, 3], 4, jitter=0.1, seed=2)
    itg = M.Integrand(M.ELASTICITY, E=1.0, nu=0.3)
    msc = _global_status(mp, 3, [7, 0, 0, 0, 0, 0])
    meqn, neq = M.equation_numbers(msc)
    print("[gpu] multi-patch traction", flush=True)
    b = np.zeros(neq)
    last = len(mp.patches) - 1
    oracle.add_traction(mp.patches[last], M.FACE_U1, (0.0, 0.0, -1.0), 3, 3,
                        _local(meqn, mp.mlgn[last], 3), b)
    with _gpu_setup(mp, itg, [3, 3, 3], meqn, neq) as s:
        s.add_traction_patch(last, M.FACE_U1, (0.0, 0.0, -1.0), 3)
        _, _, _, gb = s.system(values=False)
    assert abs(gb[2::3].sum() + 1.0) < 1e-12
    assert np.max(np.abs(gb - b)) <= 1e-13


@pytest.mark.gpu
def test_multipatch_rejects_multigrid_and_bad_maps():
    from ifem_amd import GpuError, PC_MG, Session
    mp = M.x_split_patches(1, [2, 2], 2)
    itg = M.Integrand(M.POISSON, source=M.Function(M.FN_CONST, 1.0))
    meqn, neq = M.equation_numbers(_global_status(mp, 1, [1] * 6))
    with _gpu_setup(mp, itg, [2, 2, 2], meqn, neq) as s:
        with pytest.raises(GpuError):
            s.solve(pc=PC_MG)
    with Session() as s:
        pids = [s.add_patch(P) for P in mp.patches]
        bad = [mp.mlgn[0].copy(), mp.mlgn[1].copy()]
        bad[1][1] = bad[1][0]  # two local nodes on one global node
        with pytest.raises(GpuError):
            s.bind_patches(pids, bad, 1, meqn, neq)
