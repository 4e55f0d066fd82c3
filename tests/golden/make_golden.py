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

Run from the repository root:  python tests/golden/make_golden.py
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


if __name__ == "__main__":
    main()
