"""The reference's own multi-patch inputs (tests/golden/refdata, copied
unchanged from src/ASM/Test/refdata): the 13 two-patch 3d_uneven models and
the eight-patch DomainDecomposition_MPI_3D_4_orient{0..7} cubes of
TestASMs3D.ConnectUneven/Connect (src/ASM/Test/ASM3DTests.h:96-125), read
through the .xinp front-end with <topology> connections and orientation
flags (ASMs3D::connectBasis, src/ASM/ASMs3D.C:721-808).

Known answers: every model connects with the coordinate check (the reference
test's REQUIRE(createFEMmodel())); the eight refined p=1 cubes of the
DomainDecomposition models form one 4x4x4-element cube (5^3 = 125 nodes), and
because linear splines are C0 anyway their Poisson matrix equals the single
4x4x4 p=1 patch's matrix after matching nodes by coordinates, whatever the
patches' orientations.  On the GPU, K and b of every model and the PCG
solution of the elasticity models match the oracle assembled patch by patch.
"""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from ifem_amd import FP64, inputs, model as M

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "refdata")
UNEVEN = sorted(glob.glob(os.path.join(REF, "3d_uneven*.xinp")))
ORIENT = [os.path.join(REF, f"DomainDecomposition_MPI_3D_4_orient{k}.xinp") for k in range(8)]


def _local(arr, mlgn, nc):
    idx = ((np.asarray(mlgn, dtype=np.int64) - 1)[:, None] * nc + np.arange(nc)).ravel()
    return np.asarray(arr)[idx]


def _oracle_system(mp, itg, ng, meqn, neq):
    nc = itg.ncomp
    rows = [set() for _ in range(neq)]
    for P, ml in zip(mp.patches, mp.mlgn):
        rp, ci = oracle.pattern(P, nc, _local(meqn, ml, nc), neq)
        for r in range(neq):
            rows[r].update(ci[rp[r]:rp[r + 1]].tolist())
    rp = np.zeros(neq + 1, dtype=np.int64)
    for r in range(neq):
        rp[r + 1] = rp[r] + len(rows[r])
    ci = np.concatenate([np.array(sorted(s), dtype=np.int32) for s in rows])
    A = np.zeros(len(ci))
    b = np.zeros(neq)
    for P, ml in zip(mp.patches, mp.mlgn):
        oracle.assemble_into(P, itg, ng, _local(meqn, ml, nc), neq, rp, ci, A, b, None)
    return rp, ci, A, b


def test_reference_topology_inputs_all_connect():
    assert len(UNEVEN) == 13 and all(os.path.exists(f) for f in ORIENT)
    for f in UNEVEN:
        mp, _ = inputs.read_xinp(f)
        assert len(mp.patches) == 2
        # the two patches share one face of control points
        (m, k), = [(int(c.get("slave")), int(c.get("sface")))
                   for c in __import__("xml.etree.ElementTree", fromlist=["x"]).parse(f)
                   .getroot().iter("connection")]
        shared = np.prod(inputs.face_size(mp.patches[m - 1], k))
        assert mp.nnod == sum(P.nnod for P in mp.patches) - shared, f
        # every global node's control points coincide
        X = np.full((mp.nnod, 3), np.nan)
        for P, ml in zip(mp.patches, mp.mlgn):
            seen = ~np.isnan(X[ml - 1, 0])
            assert np.all(np.abs(X[ml - 1][seen] - P.coefs[seen]) <= 1e-4)
            X[ml - 1] = P.coefs
    for f in ORIENT:
        mp, _ = inputs.read_xinp(f)
        assert len(mp.patches) == 8 and mp.nnod == 125


def test_wrong_orientation_fails_the_coordinate_check():
    text = open(os.path.join(REF, "3d_uneven.xinp")).read()
    assert 'orient="4"' in text
    with pytest.raises(ValueError, match="non-matching"):
        inputs.read_xinp(text.replace('orient="4"', 'orient="0"'), base_dir=REF)


@pytest.mark.parametrize("k", range(8))
def test_eight_patch_cube_equals_one_patch(k):
    mp, _ = inputs.read_xinp(ORIENT[k])
    itg = M.Integrand(M.POISSON, source=M.Function(M.FN_CONST, 1.0))
    n = mp.nnod
    meqn = np.arange(1, n + 1, dtype=np.int32)  # all free: the bare stiffness
    rp, ci, A, b = _oracle_system(mp, itg, (2, 2, 2), meqn, n)
    K = sp.csr_matrix((A, ci, rp), shape=(n, n)).toarray()
    X = np.zeros((n, 3))
    for Q, ml in zip(mp.patches, mp.mlgn):
        X[ml - 1] = Q.coefs
    L = X.max()  # the cube [0, L]^3 (L = 1, or 2 for orient7)
    P = M.cube_patch(1, 4)
    P.coefs *= L
    rp1, ci1, A1, b1 = oracle.system(P, itg, (2, 2, 2), meqn, n)
    K1 = sp.csr_matrix((A1, ci1, rp1), shape=(n, n)).toarray()
    # global node g <-> single-patch node at the same point
    idx = np.rint(X / L * 4).astype(int)
    perm = idx[:, 0] + 5 * (idx[:, 1] + 5 * idx[:, 2])
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.max(np.abs(K - K1[np.ix_(perm, perm)])) <= 1e-13 * np.max(np.abs(K1))
    assert np.max(np.abs(b - b1[perm])) <= 1e-14 * L ** 3


def _gpu(mp, itg, meqn, neq):
    from ifem_amd import Session
    s = Session()
    pids = [s.add_patch(P) for P in mp.patches]
    s.bind_patches(pids, mp.mlgn, itg.ncomp, meqn, neq)
    s.init_assembly()
    s.assemble(itg, mp.ng, FP64)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(8))
def test_orient_models_elasticity_on_gpu(k):
    """The DomainDecomposition cube with its own symmetry conditions (fixX
    comp 1, fixY comp 2, fixZ comp 3) as an elasticity model with a body
    force: K, b (1e-12) and the FP64 PCG solution (1e-12, against the
    oracle's PCG and a direct solve)."""
    text = open(ORIENT[k]).read().replace(
        "</simulation>", "<elasticity><isotropic E='1' nu='0.3'/></elasticity>"
        "<discretization><nGauss>2</nGauss></discretization></simulation>")
    mp, _ = inputs.read_xinp(text, base_dir=REF)
    itg = M.Integrand(M.ELASTICITY, E=1.0, nu=0.3, body=(0.3, -0.2, -1.0))
    print(f"[gpu] reference orient{k} model, 8 patches, neq={mp.neq}", flush=True)
    rp, ci, A, b = _oracle_system(mp, itg, mp.ng, mp.meqn, mp.neq)
    with _gpu(mp, itg, mp.meqn, mp.neq) as s:
        grp, gci, gA, gb = s.system()
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A))
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
        st = s.solve(rtol=1e-13, maxits=20000)
        x = s.solution()
    assert st["converged"]
    # at a relative residual of 1e-13 both lie within ~5e-15 of the direct solve
    xo, _, _ = oracle.pcg(rp, ci, A, b, rtol=1e-13)
    xd = spla.spsolve(sp.csr_matrix((A, ci, rp), shape=(mp.neq, mp.neq)).tocsc(), b)
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))
    assert np.max(np.abs(x - xd)) <= 1e-12 * np.max(np.abs(xd))


@pytest.mark.gpu
def test_uneven_models_assembly_on_gpu():
    """K and b of the 13 two-patch p=2 models (Poisson with a source, no
    boundary conditions: the bare stiffness) against the oracle (1e-12)."""
    itg = M.Integrand(M.POISSON, kappa=M.Function(M.FN_CONST, 2.0),
                      source=M.Function(M.FN_SINPROD, 1.0))
    for f in UNEVEN:
        mp, _ = inputs.read_xinp(f)
        print(f"[gpu] reference {os.path.basename(f)} neq={mp.neq}", flush=True)
        rp, ci, A, b = _oracle_system(mp, itg, mp.ng, mp.meqn, mp.neq)
        with _gpu(mp, itg, mp.meqn, mp.neq) as s:
            grp, gci, gA, gb = s.system()
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A))
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
