"""The reference-interface facade (ifem_amd/facade) driven like a SIM: the C++
driver facade_test runs Poisson and elasticity through ASMs3D::integrate,
AlgEqSystem, GPUMatrix::solve and SAM::expandSolution; its results are
checked here against the oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from ifem_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ifem_amd", "facade_test")


def test_facade_built():
    assert os.path.exists(EXE), "build with __graft_entry__.build()"


def test_facade_gauss_rule_known_answers():
    """ASMs3D::getNoGaussPt of the facade against the known answers read off
    src/ASM/ASMbase.C:1788-1798 (the table in test_inputs)."""
    from test_inputs import GAUSS_KAT
    args = []
    for order, nGauss, neumann, _ in GAUSS_KAT:
        args += [str(order), str(nGauss), "1" if neumann else "0"]
    r = subprocess.run([EXE, "gauss", *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["ng"] == [k[3] for k in GAUSS_KAT]


@pytest.mark.gpu
def test_facade_driver_matches_oracle(tmp_path):
    print("[gpu] facade driver", flush=True)
    h5name = str(tmp_path / "facade_poisson")
    env = dict(os.environ, IFEM_FACADE_HDF5=h5name)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            d = json.loads(line)
            out[(d["case"], d.get("precision", ""))] = d
    pz = out[("poisson", "")]
    pb = problems.poisson_cube(16)
    rp, ci, A, b = oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq)
    assert pz["neq"] == pb.neq and pz["nnz"] == len(ci)
    assert abs(pz["bsum"] - b.sum()) <= 1e-12 * abs(b.sum())
    assert pz["rel_residual"] <= 1e-10
    x, it, res = oracle.pcg(rp, ci, A, b, rtol=1e-12)
    xd = np.zeros(pb.ndof); xd[pb.meqn > 0] = x
    rms = np.sqrt(np.mean(xd ** 2))  # SAM::normL2 is an RMS
    assert abs(pz["normL2"] - rms) <= 1e-9 * rms
    assert pz["max_nodal_err"] < 0.02
    # result export through the facade's HDF5Writer: the expanded solution
    # of the one patch, and its basis
    assert pz["hdf5"] == 1
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        raw = str(tmp_path / "u.bin")
        subprocess.run([h5dump, "-d", "/0/Poisson-1/fields/u/1", "-b", "LE", "-o", raw,
                        h5name + ".hdf5"], capture_output=True, check=True)
        u = np.fromfile(raw, dtype="<f8")
        assert len(u) == pb.ndof
        assert abs(np.sqrt(np.mean(u ** 2)) - pz["normL2"]) <= 1e-13 * pz["normL2"]
        assert np.max(np.abs(u - xd)) <= 1e-8 * np.max(np.abs(xd))
    e64, e32 = out[("elasticity", "fp64")], out[("elasticity", "fp32")]
    assert abs(e64["force_z"] + 1.0) < 1e-12
    assert e64["rel_residual"] <= 1e-8 and e32["rel_residual"] <= 1e-8
    assert abs(e32["min_uz"] - e64["min_uz"]) <= 1e-6 * abs(e64["min_uz"])
    assert out[("gmres_refused", "")]["ok"] is True
    ua = out[("uniaxial", "")]
    n = 8  # control points per direction at nel = 6, p = 2
    assert ua["nceq"] == 2 * n * n and ua["neq"] == 3 * n ** 3 - 4 * n * n
    assert ua["max_err"] <= 1e-11 and ua["rel_residual"] <= 1e-12
    g2 = out[("g2", "")]
    assert g2["n"] == 18 and g2["knot_diff"] <= 1e-15 and g2["coef_diff"] <= 1e-14
