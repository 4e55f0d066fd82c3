"""Writes the golden fixtures in this folder.

* reference_gauss.json: the Gauss-Legendre tables G1..G10 exactly as the
  reference returns them (GaussQuadrature::getCoord/getWeight,
  src/Utility/GaussQuadrature.C:17-60), read through oracle/_ref/libifemref.so
  built from the reference sources by oracle/Makefile.
* reference_jacobian.json: utl::Jacobian (src/Utility/CoordinateMapping.C:24-63)
  on seeded inputs, same route.
* reference_sam.json: SAM equation numbers and DOF couplings
  (src/LinAlg/SAM.C:225-427) for small 3D patches with Dirichlet faces.
* scipy_poisson16.npz: scipy.sparse.linalg.spsolve solution of the 16^3 p=2
  Poisson system assembled by the oracle (a library solve, to pin the PCG).
* reference_sam_mpc.json: SAM equation numbers, DOF couplings and solution
  expansion with constraint equations that have master DOFs
  (src/LinAlg/SAM.C:225-427, 783-807): periodicity in x plus one general
  constraint with two masters and a constant, through oracle/_ref.

Run from the repository root:  python tests/golden/make_golden.py
(python tests/golden/make_golden.py mpc writes only reference_sam_mpc.json)
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ifem_amd import model as M  # noqa: E402
from ifem_amd import problems  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def element_connectivity(P):
    nel = P.nel
    mpmnpc = [1]
    mmnpc = []
    for e3 in range(nel[2]):
        for e2 in range(nel[1]):
            for e1 in range(nel[0]):
                for a3 in range(P.p[2] + 1):
                    for a2 in range(P.p[1] + 1):
                        for a1 in range(P.p[0] + 1):
                            mmnpc.append(1 + (e1 + a1) + P.n[0] * ((e2 + a2) + P.n[1] * (e3 + a3)))
                mpmnpc.append(len(mmnpc) + 1)
    return np.array(mpmnpc, dtype=np.int32), np.array(mmnpc, dtype=np.int32)


def main():
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libifemref.so not built (make -C oracle)")
    out = {}
    for n in range(1, 11):
        xi = np.zeros(n); w = np.zeros(n)
        R.ref_gauss(n, xi, w)
        out[str(n)] = {"xi": [float(v).hex() for v in xi], "w": [float(v).hex() for v in w]}
    with open(os.path.join(HERE, "reference_gauss.json"), "w") as f:
        json.dump(out, f, indent=1)

    rng = np.random.default_rng(1234)
    cases = []
    for nen in (8, 27, 64):
        X = rng.uniform(-1, 1, (3, nen)) + np.repeat(np.eye(3)[:, :1], nen, axis=1)
        D = rng.uniform(-1, 1, (nen, 3))
        Jinv = np.zeros(9); dNdX = np.zeros(nen * 3)
        det = R.ref_jacobian(nen, np.ascontiguousarray(X.ravel()), np.ascontiguousarray(D.ravel()),
                             Jinv, dNdX)
        cases.append({"nen": nen, "X": [v.hex() for v in X.ravel()], "dNdu": [v.hex() for v in D.ravel()],
                      "det": det.hex(), "Jinv": [v.hex() for v in Jinv],
                      "dNdX": [v.hex() for v in dNdX]})
    with open(os.path.join(HERE, "reference_jacobian.json"), "w") as f:
        json.dump(cases, f, indent=1)

    sams = []
    for (p, nel, ncomp, fix) in [(1, (2, 3, 2), 1, [1, 1, 1, 1, 1, 1]),
                                 (2, (3, 2, 2), 1, [1, 0, 1, 0, 0, 1]),
                                 (2, (2, 2, 3), 3, [7, 0, 0, 0, 0, 0]),
                                 (3, (2, 2, 2), 3, [7, 2, 0, 4, 1, 0])]:
        P = M.cube_patch(p, list(nel))
        msc = M.face_status(P, ncomp, fix)
        mpmnpc, mmnpc = element_connectivity(P)
        ndof = P.nnod * ncomp
        meqn = np.zeros(ndof, dtype=np.int32)
        neq = (np.zeros(1, dtype=np.int32))
        cap = ndof * ndof
        rowptr = np.zeros(ndof + 1, dtype=np.int64)
        colind = np.zeros(cap, dtype=np.int32)
        import ctypes as C
        nq = C.c_int(0)
        nnz = R.ref_sam_couplings(P.nnod, ncomp, int(np.prod(P.nel)), mpmnpc, mmnpc,
                                  np.ascontiguousarray(msc, dtype=np.int32), meqn, C.byref(nq),
                                  rowptr, colind, cap)
        assert nnz >= 0
        sams.append({"p": p, "nel": list(nel), "ncomp": ncomp, "fix": fix, "neq": nq.value,
                     "meqn": meqn.tolist(), "rowptr": rowptr[:nq.value + 1].tolist(),
                     "colind": colind[:nnz].tolist()})
    with open(os.path.join(HERE, "reference_sam.json"), "w") as f:
        json.dump(sams, f)

    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    pb = problems.poisson_cube(16)
    rp, ci, A, b = oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, nthreads=1)
    K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq))
    x = spl.spsolve(K.tocsc(), b)
    np.savez_compressed(os.path.join(HERE, "scipy_poisson16.npz"), x=x, b=b)
    print("golden fixtures written")


def mpc_cases():
    """Small patches with periodicity in x and one extra constraint with two
    masters and a constant c0 (the SAM arrays come from
    model.constraint_equations)."""
    out = []
    for (p, nel, ncomp, fix) in [(2, (3, 2, 2), 1, [0, 0, 1, 1, 0, 0]),
                                 (1, (3, 2, 2), 3, [0, 0, 7, 0, 0, 0]),
                                 (2, (2, 2, 3), 3, [0, 0, 0, 0, 7, 0])]:
        P = M.cube_patch(p, list(nel))
        msc = M.face_status(P, ncomp, fix)
        cons = M.periodic_constraints(P, ncomp, 0)
        # an interior DOF tied to two others with a constant
        interior = [d for d in range(P.nnod * ncomp) if msc[d] == 1
                    and d not in {c[0] for c in cons} and d not in
                    {m for c in cons for m, _ in c[2]}]
        s_, m1, m2 = interior[len(interior) // 2], interior[0], interior[-1]
        cons.append((s_, 0.25, [(m1, 0.5), (m2, -1.5)]))
        out.append((P, ncomp, msc, M.constraint_equations(msc, cons)))
    return out


def write_mpc():
    import ctypes as C
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libifemref.so not built (make -C oracle)")
    cases = []
    for P, ncomp, msc, sam in mpc_cases():
        mpmnpc, mmnpc = element_connectivity(P)
        ndof = P.nnod * ncomp
        msc_ref = np.where(sam.meqn < 0, 0, msc).astype(np.int32)
        meqn = np.zeros(ndof, dtype=np.int32)
        nq = C.c_int(0)
        cap = ndof * ndof
        rowptr = np.zeros(ndof + 1, dtype=np.int64)
        colind = np.zeros(cap, dtype=np.int32)
        xeq = 1.0 + 0.1 * np.arange(ndof, dtype=np.float64)
        xdof = np.zeros(ndof)
        nnz = R.ref_sam_couplings_mpc(P.nnod, ncomp, int(np.prod(P.nel)), mpmnpc, mmnpc, msc_ref,
                                      sam.nceq, np.ascontiguousarray(sam.mpmceq, dtype=np.int32),
                                      np.ascontiguousarray(sam.mmceq, dtype=np.int32),
                                      np.ascontiguousarray(sam.ttcc, dtype=np.float64), meqn,
                                      C.byref(nq), rowptr, colind, cap, xeq, xdof)
        assert nnz >= 0, nnz
        cases.append({"p": P.p[0], "nel": P.nel, "ncomp": ncomp, "msc": [int(v) for v in msc],
                      "mpmceq": sam.mpmceq.tolist(), "mmceq": sam.mmceq.tolist(),
                      "ttcc": [float(v).hex() for v in sam.ttcc], "neq": nq.value,
                      "meqn": meqn.tolist(), "rowptr": rowptr[:nq.value + 1].tolist(),
                      "colind": colind[:nnz].tolist(),
                      "xeq": [float(v).hex() for v in xeq[:nq.value]],
                      "xdof": [float(v).hex() for v in xdof]})
    with open(os.path.join(HERE, "reference_sam_mpc.json"), "w") as f:
        json.dump(cases, f)
    print("reference_sam_mpc.json written")


if __name__ == "__main__":
    if sys.argv[1:] == ["mpc"]:
        write_mpc()
    else:
        main()
        write_mpc()
