"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs, plus size-independent properties at the full benchmark sizes.

Tolerances: FP64 assembly and solution 1e-12 relative to max|.| (the parity
bar of the FP64 path against the reference's assembly and solution); FP32
assembly 1e-5 relative; mixed-precision solution: energy norm error vs the
FP64 solution <= 1e-6.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from ifem_amd import FP32, FP64, PC_MG, model as M, problems, setup_problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_system(pb, nthreads=0):
    return oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, pb.traction, nthreads)


def _compare_system(s, pb, tol=1e-12):
    rp, ci, A, b = _oracle_system(pb)
    grp, gci, gA, gb = s.system()
    assert np.array_equal(grp, rp)
    assert np.array_equal(gci, ci)
    errA = np.max(np.abs(gA - A)) / np.max(np.abs(A))
    errb = np.max(np.abs(gb - b)) / max(np.max(np.abs(b)), 1e-300)
    assert errA <= tol, errA
    assert errb <= tol, errb
    return rp, ci, A, b


def _with_body(pb, body=(0.2, -0.1, 0.3)):
    """The same problem with a body force (the right-hand-side variant of the
    line kernel)."""
    import dataclasses
    return dataclasses.replace(pb, integrand=dataclasses.replace(pb.integrand, body=body))


CASES = [
    ("poisson16p2", lambda: problems.poisson_cube(16, 2)),
    ("poisson6p1", lambda: problems.poisson_cube(6, 1)),
    ("poisson5p3", lambda: problems.poisson_cube(5, 3)),
    ("poisson6p2ng4", lambda: problems.poisson_cube(6, 2, ng=4, jitter=0.1, seed=5)),
    ("elast6p2", lambda: problems.elasticity_cube(6, 2, seed=3)),
    ("elast4p1", lambda: problems.elasticity_cube(4, 1, seed=4)),
    ("elast3p3", lambda: problems.elasticity_cube(3, 3, seed=5)),
    # full 16-element bundles on the interior lines (one block per line)
    ("elast5p3", lambda: problems.elasticity_cube(5, 3, seed=6)),
    # 5x5x5 Gauss points: three row-component blocks per line
    ("elast3p3ng5", lambda: problems.elasticity_cube(3, 3, ng=5, seed=7)),
    ("elast4p3body", lambda: _with_body(problems.elasticity_cube(4, 3, seed=11))),
    ("elast5p2body", lambda: _with_body(problems.elasticity_cube(5, 2, seed=12))),
    ("diffusion8p2", lambda: problems.diffusion_cube(8, 2, sigma=1.0, ell=0.2, seed=9)),
    ("diffusion4p3", lambda: problems.diffusion_cube(4, 3, sigma=1.0, ell=0.2, seed=10)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_assembly_matches_oracle(name, make):
    pb = make()
    _say(f"[gpu] assembly parity {name} neq={pb.neq}")
    with setup_problem(pb) as s:
        _compare_system(s, pb)


def test_assembly_is_bitwise_deterministic():
    pb = problems.elasticity_cube(8, 2, seed=21)
    _say("[gpu] determinism elasticity 8^3")
    with setup_problem(pb) as s:
        _, _, A1, b1 = s.system()
        s.assemble(pb.integrand, pb.ng, FP64)
        s.add_traction(*pb.traction)
        _, _, A2, b2 = s.system()
    assert np.array_equal(A1, A2) and np.array_equal(b1, b2)


def test_autotuned_assembly_keeps_parity():
    """The element-chunk variants of the line kernel give the oracle's system
    (the autotuner leaves the fastest one assembled)."""
    for pb in (problems.elasticity_cube(6, 2, seed=3), problems.poisson_cube(5, 3),
               problems.diffusion_cube(6, 2, sigma=1.0, ell=0.2, seed=4),
               problems.elasticity_cube(4, 3, seed=9)):
        _say(f"[gpu] autotuned assembly {pb.name}")
        with setup_problem(pb) as s:
            for prec in (FP64, FP32):
                assert s.autotune_assembly(pb.integrand, pb.ng, prec) in (0, 1, 2)
            if pb.traction is not None:
                s.add_traction(*pb.traction)
            _, _, A, _ = _compare_system(s, pb)
            A32 = s.values_fp32().astype(np.float64)
            assert np.max(np.abs(A32 - A)) <= 1e-5 * np.max(np.abs(A))


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("kind", ["elasticity", "elasticity_body", "poisson", "poisson_p1"])
def test_each_assembly_variant_matches_oracle(kind, variant):
    """Every element-chunk variant (full, half, Gauss-point slabs) forced in
    turn: K and b against the oracle (1e-12), bitwise equal on a rerun, FP32
    values within 1e-5."""
    pb = {"elasticity": lambda: problems.elasticity_cube(6, 2, seed=13),
          "elasticity_body": lambda: _with_body(problems.elasticity_cube(5, 2, seed=14)),
          "poisson": lambda: problems.poisson_cube(6, 2),
          "poisson_p1": lambda: problems.poisson_cube(7, 1)}[kind]()
    p, ng, nc = pb.patch.p[0], pb.ng[0], pb.integrand.ncomp
    _say(f"[gpu] assembly variant {variant} {pb.name}")
    with setup_problem(pb) as s:
        for prec in (FP64, FP32):
            s.set_assembly_variant(p, ng, nc, prec, variant)
        s.assemble(pb.integrand, pb.ng, FP64)
        if pb.traction is not None:
            s.add_traction(*pb.traction)
        _compare_system(s, pb)
        _, _, A, b = s.system()
        s.assemble(pb.integrand, pb.ng, FP64)
        if pb.traction is not None:
            s.add_traction(*pb.traction)
        _, _, A2, b2 = s.system()
        assert np.array_equal(A, A2) and np.array_equal(b, b2)
        s.assemble(pb.integrand, pb.ng, FP32)
        A32 = s.values_fp32().astype(np.float64)
        assert np.max(np.abs(A32 - A)) <= 1e-5 * np.max(np.abs(A))


def test_assembly_variant_refusals():
    from ifem_amd._native import GpuError
    pb = problems.elasticity_cube(3, 3, seed=1)
    with setup_problem(pb) as s:
        with pytest.raises(GpuError):
            s.set_assembly_variant(3, 4, 3, FP64, 2)  # no slab variant at 4x4x4 points
        with pytest.raises(GpuError):
            s.set_assembly_variant(2, 7, 3, FP64, 0)


def test_fp32_assembly_close_to_fp64():
    pb = problems.elasticity_cube(6, 2, seed=8)
    _say("[gpu] fp32 assembly elasticity 6^3")
    with setup_problem(pb, precision=FP32) as s:
        _, _, A, _ = s.system()
        A32 = s.values_fp32().astype(np.float64)
    assert np.max(np.abs(A32 - A)) <= 1e-5 * np.max(np.abs(A))


def test_spmv_matches_oracle():
    pb = problems.elasticity_cube(6, 2, seed=3)
    _say("[gpu] spmv")
    with setup_problem(pb) as s:
        rp, ci, A, b = s.system()
        x = np.random.default_rng(0).standard_normal(pb.neq)
        y = s.spmv(x)
    yo = oracle.spmv(rp, ci, A, x)
    assert np.max(np.abs(y - yo)) <= 1e-13 * np.max(np.abs(yo))


def test_pcg_iterates_match_oracle():
    """k Jacobi-PCG iterations from x = 0: GPU iterate vs oracle iterate."""
    pb = problems.poisson_cube(16)
    _say("[gpu] pcg iterate parity")
    with setup_problem(pb) as s:
        rp, ci, A, b = s.system()
        st = s.solve(rtol=1e-30, maxits=20, check_every=5)
        x = s.solution()
    assert st["iterations"] == 20 and not st["converged"]
    xo = oracle.pcg_fixed(rp, ci, A, b, 20)
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))


@pytest.mark.parametrize("name,make", [
    ("poisson16", lambda: problems.poisson_cube(16)),
    ("elast8", lambda: problems.elasticity_cube(8, 2, seed=6)),
])
def test_pcg_solution_matches_oracle(name, make):
    pb = make()
    _say(f"[gpu] pcg solution {name}")
    with setup_problem(pb) as s:
        rp, ci, A, b = _compare_system(s, pb)
        st = s.solve(rtol=1e-13, maxits=20000)
        x = s.solution()
    assert st["converged"] and st["rel_residual"] <= 1e-13
    xo, it, res = oracle.pcg(rp, ci, A, b, rtol=1e-13)
    assert it > 0
    # both solves stop at a relative residual of 1e-13; with the condition
    # numbers of these systems (21 for Poisson 16^3) both lie within ~2e-14
    # of the exact solution, so they agree to 1e-12
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))
    r = b - oracle.spmv(rp, ci, A, x)
    assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(b) * 1.01
    # and against a direct solve of the same system (scipy's sparse LU; for
    # Poisson 16^3 the committed fixture of tests/golden/make_golden.py)
    if name == "poisson16":
        ref = np.load(os.path.join(GOLDEN, "scipy_poisson16.npz"))
        assert np.max(np.abs(b - ref["b"])) <= 1e-14 * np.max(np.abs(ref["b"]))
        xd = ref["x"]
    else:
        K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq)).tocsc()
        xd = spla.spsolve(K, b)
    assert np.max(np.abs(x - xd)) <= 1e-12 * np.max(np.abs(xd))


def test_mixed_precision_solution():
    pb = problems.elasticity_cube(10, 2, seed=12)
    _say("[gpu] mixed precision elasticity 10^3")
    with setup_problem(pb, precision=FP32) as s:
        rp, ci, A, b = s.system()
        st64 = s.solve(rtol=1e-10, maxits=50000)
        x64 = s.solution()
        st = s.solve(rtol=1e-8, maxits=50000, precision=FP32, inner_rtol=1e-4)
        x = s.solution()
    assert st64["converged"] and st["converged"] and st["rel_residual"] <= 1e-8
    K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq))
    e = x - x64
    en = np.sqrt(e @ (K @ e)) / np.sqrt(x64 @ (K @ x64))
    assert en <= 1e-6, en


def test_solver_reports_non_convergence():
    pb = problems.poisson_cube(8)
    _say("[gpu] maxits exhaustion")
    with setup_problem(pb) as s:
        st = s.solve(rtol=1e-14, maxits=3, check_every=1)
    assert st["rc"] == 4 and not st["converged"] and st["iterations"] == 3


def test_errors_are_reported_not_thrown():
    from ifem_amd import GpuError, Session
    pb = problems.poisson_cube(4)
    with Session() as s:
        pid = s.add_patch(pb.patch)
        bad = pb.meqn.copy()
        bad[bad > 0] = bad[bad > 0][::-1]  # not monotone
        with pytest.raises(GpuError):
            s.bind(pid, 1, bad, pb.neq)
        s.bind(pid, 1, pb.meqn, pb.neq)
        with pytest.raises(GpuError):
            s.assemble(pb.integrand, pb.ng)  # before init_assembly
        s.init_assembly()
        with pytest.raises(GpuError):
            s.assemble(M.Integrand(M.ELASTICITY), pb.ng)  # ncomp mismatch


# ---- full benchmark size: size-independent properties -------------------------
def test_elasticity64_full_size_properties():
    """64^3 p=2 jittered elasticity (862,488 DOFs): closed-form nnz, traction
    resultant, rows of a bottom slab vs the oracle, FP64 solve to 1e-8 on the
    true residual."""
    pb = problems.elasticity_cube(64, 2)
    _say("[gpu] full-size elasticity 64^3")
    with setup_problem(pb) as s:
        assert s.nnz == 301386096
        rp, ci, A, b = s.system()
        # traction resultant: sum of z-loads = -1 (face x=1 has area 1)
        assert abs(b[2::3].sum() + 1.0) < 1e-12
        # oracle on elements with e3 in [0,2) gives complete rows for nodes k < 2
        Ao, bo = oracle.assemble(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, rp, ci,
                                 e3_range=(0, 2))
        n1, n2 = pb.patch.n[0], pb.patch.n[1]
        last_row = int(pb.meqn[: 2 * n1 * n2 * 3].max())
        hi = rp[last_row]
        err = np.max(np.abs(A[:hi] - Ao[:hi])) / np.max(np.abs(A[:hi]))
        assert err <= 1e-12, err
        del Ao
        st = s.solve(rtol=1e-8, maxits=20000)
        assert st["converged"] and st["rel_residual"] <= 1e-8


def _free_count(n, lo, p):
    """Per index i >= lo of 0..n-1: free neighbours within +-p (indices >= lo)."""
    i = np.arange(n)
    return np.where(i >= lo, np.minimum(n - 1, i + p) - np.maximum(lo, i - p) + 1, 0)


def test_elasticity96_p3_full_size_properties():
    """Largest single-GPU case: 96^3 p=3 jittered elasticity, 4x4x4 Gauss
    points (2,910,897 DOFs, nnz 2.81e9 > 2^31): closed-form nnz, traction
    resultant, the rows of the bottom three node layers against the oracle,
    FP64 Jacobi-PCG to 1e-8 on the true residual, FP32 assembly with the
    mixed-precision solve (energy-norm error vs FP64 within 1e-6) and the
    multigrid-preconditioned solve."""
    pb = problems.elasticity_cube(96, 3)
    n, p = pb.patch.n[0], 3
    nnz = 9 * int(_free_count(n, 1, p).sum()) * int(_free_count(n, 0, p).sum()) ** 2
    assert nnz == 2813174226  # 312,574,914 3x3 blocks
    _say("[gpu] full-size elasticity 96^3 p=3: set-up and FP64 assembly")
    with setup_problem(pb) as s:
        assert s.nnz == nnz and s.neq == pb.neq == 2881494
        _say("[gpu]   system to host")
        rp, ci, A, b = s.system()
        assert abs(b[2::3].sum() + 1.0) < 1e-12
        _say("[gpu]   oracle rows of node layers k < 3")
        Ao, bo = oracle.assemble(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, rp, ci,
                                 e3_range=(0, 3))
        n1, n2 = pb.patch.n[0], pb.patch.n[1]
        last_row = int(pb.meqn[: 3 * n1 * n2 * 3].max())
        hi = rp[last_row]
        err = np.max(np.abs(A[:hi] - Ao[:hi])) / np.max(np.abs(A[:hi]))
        assert err <= 1e-12, err
        del Ao, A, ci
        _say("[gpu]   FP64 Jacobi-PCG")
        st = s.solve(rtol=1e-8, maxits=50000)
        assert st["converged"] and st["rel_residual"] <= 1e-8
        x64 = s.solution()
        xk = s.spmv(x64)
        _say("[gpu]   FP32 assembly and mixed-precision solve")
        s.assemble(pb.integrand, pb.ng, FP32)
        stm = s.solve(rtol=1e-8, maxits=50000, precision=FP32)
        assert stm["converged"] and stm["rel_residual"] <= 1e-8
        xm = s.solution()
        d = xm - x64
        assert np.sqrt(d @ s.spmv(d)) <= 1e-6 * np.sqrt(x64 @ xk)
        _say("[gpu]   multigrid-preconditioned solve")
        stg = s.solve(rtol=1e-8, maxits=2000, pc=PC_MG)
        assert stg["converged"] and stg["rel_residual"] <= 1e-8
        d = s.solution() - x64
        assert np.sqrt(d @ s.spmv(d)) <= 1e-6 * np.sqrt(x64 @ xk)
        _say(f"[gpu]   iterations FP64 {st['iterations']} mixed {stm['iterations']} "
             f"multigrid {stg['iterations']}")


def test_norms_match_oracle():
    """GPU norm integration of the GPU solution against the oracle's on the same
    coefficients: Poisson with the manufactured exact solution (config 1 at
    16^3) and jittered elasticity."""
    from ifem_amd import model as M
    exact = M.Function(M.FN_SINPROD, 1.0)
    for pb, ex in ((problems.poisson_cube(16, 2), exact),
                   (problems.diffusion_cube(6, 3, sigma=1.0, ell=0.2, seed=4), None),
                   (problems.elasticity_cube(6, 2, seed=9), None)):
        _say(f"[gpu] norms {pb.name}")
        with setup_problem(pb) as s:
            st = s.solve(rtol=1e-12, maxits=20000)
            assert st["converged"]
            n = s.norms(pb.integrand, pb.ng, ex)
            x = s.solution()
        xd = np.zeros(pb.ndof)
        m = pb.meqn > 0
        xd[m] = x[pb.meqn[m] - 1]
        no = oracle.norms(pb.patch, pb.integrand, pb.ng, xd, ex)
        assert np.all(np.abs(n - no) <= 1e-12 * np.maximum(np.abs(no), 1e-300)), (n, no)
        if ex is not None:
            assert abs(n[4] - 2.26239829e-05) <= 1e-6 * n[4]  # O(h^3) error of u_h at 16^3


# ---- prescribed values (constraint equations without masters) -----------------
def _prescribed_case(pb, seed, frac=0.6):
    """Random values on a random part of the boundary, the rest fixed at 0."""
    P, nc = pb.patch, pb.ncomp
    n = P.n
    k, j, i = np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")
    bnd = ((i == 0) | (i == n[0] - 1) | (j == 0) | (j == n[1] - 1) | (k == 0)
           | (k == n[2] - 1)).ravel()
    bnd = np.repeat(bnd, nc)
    rng = np.random.default_rng(seed)
    pick = bnd & (rng.random(P.nnod * nc) < frac)
    vals = np.where(pick, rng.standard_normal(P.nnod * nc), np.nan)
    msc = np.where(bnd & ~pick, 0, 1).astype(np.int32)
    return M.prescribed_equations(msc, vals)


def _setup_prescribed(pb, sam, precision=FP64):
    from ifem_amd import Session
    s = Session()
    pid = s.add_patch(pb.patch)
    s.bind_constraints(pid, pb.ncomp, sam)
    s.init_assembly()
    s.assemble(pb.integrand, pb.ng, FP64)
    if precision == FP32:
        s.assemble(pb.integrand, pb.ng, FP32)
    return s


@pytest.mark.parametrize("name,make", [
    ("diffusion6p2", lambda: problems.diffusion_cube(6, 2, sigma=1.0, ell=0.2, seed=3)),
    ("poisson5p3", lambda: problems.poisson_cube(5, 3, jitter=0.1, seed=2)),
    ("elast5p2", lambda: problems.elasticity_cube(5, 2, seed=8)),
    ("elast4p1", lambda: problems.elasticity_cube(4, 1, seed=1)),
])
def test_prescribed_values_match_oracle(name, make):
    """K, b with prescribed boundary values (b_i -= K_ij g_j for the eliminated
    columns), the FP64 solution, its expansion (g at prescribed DOFs) and the
    norms of the expanded solution, against the oracle at 1e-12."""
    pb = make()
    sam = _prescribed_case(pb, seed=len(name))
    g = sam.prescribed_values()
    _say(f"[gpu] prescribed values {name} neq={sam.neq} nceq={sam.nceq}")
    rp, ci = oracle.pattern(pb.patch, pb.ncomp, sam.meqn, sam.neq)
    A, b = oracle.assemble(pb.patch, pb.integrand, pb.ng, sam.meqn, sam.neq, rp, ci, gdof=g)
    with _setup_prescribed(pb, sam) as s:
        grp, gci, gA, gb = s.system()
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A))
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
        st = s.solve(rtol=1e-13, maxits=20000)
        assert st["converged"]
        xd = s.expanded_solution(pb.ndof)
        n = s.norms(pb.integrand, pb.ng)
    xo, it, res = oracle.pcg(rp, ci, A, b, rtol=1e-13)
    xdo = oracle.expand_solution(sam.meqn, xo, g)
    assert np.max(np.abs(xd - xdo)) <= 1e-10 * np.max(np.abs(xdo))
    assert np.array_equal(xd[sam.meqn < 0], g[sam.meqn < 0])
    no = oracle.norms(pb.patch, pb.integrand, pb.ng, xd)
    assert np.all(np.abs(n - no) <= 1e-12 * np.maximum(np.abs(no), 1e-300)), (n, no)


def test_prescribed_uniaxial_elasticity_closed_form_mixed_precision():
    """u_x = 0 / d on x = 0 / 1, symmetry planes y = 0, z = 0, traction free
    elsewhere: the exact u = (d x, -nu d y, -nu d z) is in the discrete space.
    FP64 PCG and FP32-inner mixed precision both reproduce it at every
    control point (16^3 p=2 jittered)."""
    E, nu, d = 1.0, 0.3, 0.01
    P = M.cube_patch(2, 16, jitter=0.15, seed=7)
    itg = M.Integrand(M.ELASTICITY, E=E, nu=nu)
    n = P.n
    k, j, i = [a.ravel() for a in np.meshgrid(range(n[2]), range(n[1]), range(n[0]), indexing="ij")]
    vals = np.full((P.nnod, 3), np.nan)
    vals[i == 0, 0] = 0.0
    vals[i == n[0] - 1, 0] = d
    vals[j == 0, 1] = 0.0
    vals[k == 0, 2] = 0.0
    sam = M.prescribed_equations(np.ones(P.nnod * 3, dtype=np.int32), vals.ravel())
    pb = problems.Problem("uniaxial16", P, itg, (3, 3, 3), sam.meqn, sam.neq)
    X = P.coefs
    ue = np.stack([d * X[:, 0], -nu * d * X[:, 1], -nu * d * X[:, 2]], axis=1).ravel()
    _say(f"[gpu] uniaxial elasticity 16^3 neq={sam.neq}")
    with _setup_prescribed(pb, sam, precision=FP32) as s:
        st = s.solve(rtol=1e-12, maxits=20000)
        assert st["converged"]
        x64 = s.expanded_solution(pb.ndof)
        st = s.solve(rtol=1e-12, maxits=20000, precision=FP32, inner_rtol=1e-4)
        assert st["converged"]
        x32 = s.expanded_solution(pb.ndof)
    assert np.max(np.abs(x64 - ue)) <= 1e-10 * d
    assert np.max(np.abs(x32 - ue)) <= 1e-10 * d


def test_constraint_chains_are_rejected():
    """A master that is itself a slave (a chain the SAM set-up resolves
    first) is an error; negative equation numbers need the arrays."""
    from ifem_amd import GpuError, Session
    pb = problems.poisson_cube(3)
    sam = _prescribed_case(pb, seed=5)
    # the first constraint's master is the slave of the last one
    master = int(sam.mmceq[sam.mpmceq[-2] - 1])
    bad = M.SamArrays(sam.meqn, sam.neq, np.concatenate([[1], sam.mpmceq[1:] + 1]).astype(np.int32),
                      np.insert(sam.mmceq, 1, master).astype(np.int32),
                      np.insert(sam.ttcc, 1, 1.0))
    with Session() as s:
        pid = s.add_patch(pb.patch)
        with pytest.raises(GpuError):
            s.bind_constraints(pid, 1, bad)
        with pytest.raises(GpuError):
            s.bind(pid, 1, sam.meqn, sam.neq)  # negative meqn without the arrays
        s.bind_constraints(pid, 1, sam)
