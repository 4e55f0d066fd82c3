"""Constraint equations with master DOFs on the GPU (u_s = c0 + sum c_m u_m).

The GPU assembles an extended system with the slaves as extra unknowns and
condenses it (K = T^T K_x T, b = T^T (b_x - K_x t0)); the oracle restates
the reference's element-by-element elimination (assemSparse with
constrained columns, src/LinAlg/SparseMatrix.C:612-688; SAM::assembleRHS,
src/LinAlg/SAM.C:605-617) and is pinned by the reference SAM fixtures and by
condensation of the all-free system.  Here the GPU pattern, K, b, solution,
expansion and norms are compared with the oracle.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from ifem_amd import FP32, FP64, model as M

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import mpc_cases  # noqa: E402


def _cases():
    out = []
    for k, (P, ncomp, msc, sam) in enumerate(mpc_cases()):
        P = M.cube_patch(P.p[0], P.nel, jitter=0.1, seed=3 + k)
        itg = (M.Integrand(M.POISSON, source=M.smooth_source(2)) if ncomp == 1 else
               M.Integrand(M.ELASTICITY, body=(0.2, -0.1, 0.3)))
        out.append((f"case{k}", P, ncomp, msc, sam, itg, [P.p[0] + 1] * 3))
    return out


def _transfer(msc, sam):
    """T and t0 with u_all = T u + t0 over the all-free numbering of msc."""
    meq_all, neq_all = M.equation_numbers(np.asarray(msc))
    T = sp.lil_matrix((neq_all, sam.neq))
    t0 = np.zeros(neq_all)
    for d in range(len(msc)):
        ea = meq_all[d] - 1
        if ea < 0:
            continue
        if sam.meqn[d] > 0:
            T[ea, sam.meqn[d] - 1] = 1.0
        elif sam.meqn[d] < 0:
            k = -sam.meqn[d]
            lo, hi = sam.mpmceq[k - 1] - 1, sam.mpmceq[k] - 1
            t0[ea] = sam.ttcc[lo]
            for q in range(lo + 1, hi):
                T[ea, sam.meqn[sam.mmceq[q] - 1] - 1] += sam.ttcc[q]
    return meq_all, T.tocsr(), t0


def _session(P, ncomp, sam, itg, ng, precision=FP64):
    from ifem_amd import Session
    s = Session()
    pid = s.add_patch(P)
    s.bind_constraints(pid, ncomp, sam)
    s.init_assembly()
    s.assemble(itg, ng, FP64)
    if precision == FP32:
        s.assemble(itg, ng, FP32)
    return s


def test_transfer_matrix_reproduces_the_oracle_expansion():
    """CPU check of the helper the GPU tests use: T x + t0 equals the
    oracle's SAM expansion at the non-fixed DOFs."""
    for name, P, ncomp, msc, sam, itg, ng in _cases():
        meq_all, T, t0 = _transfer(msc, sam)
        x = np.random.default_rng(1).standard_normal(sam.neq)
        xd = oracle.expand_sam(sam, x)
        free = meq_all > 0
        assert np.allclose((T @ x + t0), xd[free][np.argsort(meq_all[free])], atol=1e-14), name


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(3))
def test_constrained_assembly_solve_and_norms_match_oracle(k):
    """Pattern (exact), K and b (1e-12), the PCG solution (1e-10), the
    expanded solution (slaves = c0 + sum c_m u_m) and the norms (1e-12)."""
    name, P, ncomp, msc, sam, itg, ng = _cases()[k]
    print(f"[gpu] constraints with masters {name} neq={sam.neq}", flush=True)
    rp, ci = oracle.pattern_sam(P, ncomp, sam)
    A, b = oracle.assemble_sam(P, itg, ng, sam, rp, ci)
    xo, _, _ = oracle.pcg(rp, ci, A, b, rtol=1e-13)
    with _session(P, ncomp, sam, itg, ng) as s:
        grp, gci, gA, gb = s.system()
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A))
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
        st = s.solve(rtol=1e-13, maxits=20000)
        assert st["converged"]
        x = s.solution()
        xd = s.expanded_solution(len(sam.meqn))
        n = s.norms(itg, ng)
    assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))
    xdo = oracle.expand_sam(sam, x)
    assert np.max(np.abs(xd - xdo)) <= 1e-14 * np.max(np.abs(xdo))
    no = oracle.norms(P, itg, ng, xdo)
    assert np.all(np.abs(n - no) <= 1e-12 * np.maximum(no, 1e-300)), (n, no)


@pytest.mark.gpu
def test_constrained_fp32_matrix_and_mixed_precision_solve():
    """FP32 condensed K equals the oracle's to FP32 rounding; the mixed
    precision solve reaches the FP64 solution."""
    name, P, ncomp, msc, sam, itg, ng = _cases()[2]
    print(f"[gpu] constraints with masters FP32 {name}", flush=True)
    rp, ci = oracle.pattern_sam(P, ncomp, sam)
    A, b = oracle.assemble_sam(P, itg, ng, sam, rp, ci)
    xo, _, _ = oracle.pcg(rp, ci, A, b, rtol=1e-13)
    with _session(P, ncomp, sam, itg, ng, precision=FP32) as s:
        A32 = s.values_fp32()
        st = s.solve(rtol=1e-11, maxits=20000, precision=FP32)
        assert st["converged"]
        x = s.solution()
    assert np.max(np.abs(A32 - A)) <= 1e-6 * np.max(np.abs(A))
    assert np.max(np.abs(x - xo)) <= 1e-9 * np.max(np.abs(xo))


@pytest.mark.gpu
def test_constrained_traction_on_the_slave_face():
    """Traction on x = 1, whose DOFs are slaves of x = 0 (periodic in x):
    the face load reaches the masters, b = b_sam + T^T f_all."""
    name, P, ncomp, msc, sam, itg, ng = _cases()[2]
    print(f"[gpu] constraints with masters traction {name}", flush=True)
    rp, ci = oracle.pattern_sam(P, ncomp, sam)
    A, b = oracle.assemble_sam(P, itg, ng, sam, rp, ci)
    meq_all, T, t0 = _transfer(msc, sam)
    f = np.zeros(T.shape[0])
    oracle.add_traction(P, M.FACE_U1, (0.1, 0.0, -1.0), 3, ncomp, meq_all, f)
    expect = b + T.T @ f
    with _session(P, ncomp, sam, itg, ng) as s:
        s.add_traction(M.FACE_U1, (0.1, 0.0, -1.0), 3)
        _, _, _, gb = s.system(values=False)
    assert np.max(np.abs(gb - expect)) <= 1e-12 * np.max(np.abs(expect))


def _periodic_poisson(nel, p=2):
    P = M.cube_patch(p, nel)
    modes = np.array([[2 * np.pi, np.pi, -np.pi, 0.0, 0.25], [2 * np.pi, -np.pi, np.pi, 0.0, 0.25],
                      [2 * np.pi, np.pi, np.pi, 0.0, -0.25], [2 * np.pi, -np.pi, -np.pi, 0.0, -0.25]])
    exact = M.Function(M.FN_FOURIER, 1.0, 0.0, modes)
    itg = M.Integrand(M.POISSON, source=M.Function(M.FN_FOURIER, 6 * np.pi ** 2, 0.0, modes))
    msc = M.face_status(P, 1, [0, 0, 1, 1, 1, 1])
    sam = M.constraint_equations(msc, M.periodic_constraints(P, 1, 0))
    return P, itg, exact, sam


@pytest.mark.gpu
def test_periodic_poisson_on_the_gpu_converges():
    """u = cos(2 pi x) sin(pi y) sin(pi z), periodic in x: GPU L2 error
    norms at nel 8 and 16 equal the oracle's (1e-8) and fall at O(h^3)."""
    errs = []
    for nel in (8, 16):
        P, itg, exact, sam = _periodic_poisson(nel)
        print(f"[gpu] periodic Poisson nel={nel}", flush=True)
        with _session(P, 1, sam, itg, [3, 3, 3]) as s:
            st = s.solve(rtol=1e-13, maxits=20000)
            assert st["converged"]
            xd = s.expanded_solution(len(sam.meqn))
            n = s.norms(itg, [3, 3, 3], exact)
        xd3 = xd.reshape(P.n[2], P.n[1], P.n[0])
        assert np.array_equal(xd3[:, :, 0], xd3[:, :, -1])
        no = oracle.norms(P, itg, [3, 3, 3], xd, exact)
        assert abs(n[4] - no[4]) <= 1e-8 * no[4]
        errs.append(n[4])
    assert 6.0 < errs[0] / errs[1] < 10.0, errs


@pytest.mark.gpu
def test_constrained_system_rejects_multigrid_and_multipatch():
    from ifem_amd import GpuError, PC_MG, Session
    name, P, ncomp, msc, sam, itg, ng = _cases()[0]
    with _session(P, ncomp, sam, itg, ng) as s:
        with pytest.raises(GpuError):
            s.solve(pc=PC_MG)
    mp = M.x_split_patches(1, [2, 2], 2)
    nd = mp.nnod
    msc2 = np.ones(nd, dtype=np.int32)
    sam2 = M.constraint_equations(msc2, [(nd - 1, 0.0, [(0, 1.0)])])
    with Session() as s:
        pids = [s.add_patch(Q) for Q in mp.patches]
        with pytest.raises(GpuError):
            s.bind_patches(pids, mp.mlgn, 1, None, 0, sam=sam2)
