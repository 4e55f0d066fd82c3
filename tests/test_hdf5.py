"""Result export to HDF5 in the reference's layout (ifem_amd.hdf5 over
libifem_hdf5.so).

Pinned by the reference's own HDF5 files (tests/golden/reference_hdf5.json,
from src/Utility/Test/refdata/Field3D-{2P,1P}.hdf5, written by its
HDF5Writer): the same content written here gives the same objects, the same
datatypes and the same bytes.  Every file is read back with h5dump, a reader
independent of the writer.
"""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from ifem_amd import inputs, model as M
from ifem_amd import hdf5 as H

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_hdf5.json")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")


def _gold():
    with open(GOLD) as f:
        return json.load(f)


def _listing(path):
    out = subprocess.run([H5DUMP, "-n", path], capture_output=True, text=True, check=True).stdout
    return [[m.group(1), m.group(2).strip()] for m in
            (re.match(r"\s*(group|dataset)\s+(\S.*)$", ln) for ln in out.splitlines()) if m]


def _dtype(path, ds):
    out = subprocess.run([H5DUMP, "-H", "-d", ds, path], capture_output=True, text=True,
                         check=True).stdout
    return re.search(r"DATATYPE\s+(\S+)", out).group(1)


def _raw(path, ds, tmp):
    f = os.path.join(tmp, "raw.bin")
    subprocess.run([H5DUMP, "-d", ds, "-b", "LE", "-o", f, path], capture_output=True, check=True)
    with open(f, "rb") as fh:
        return fh.read()


def _doubles(path, ds, tmp):
    return np.frombuffer(_raw(path, ds, tmp), dtype="<f8")


def test_basis_text_is_byte_identical_to_the_reference():
    """A patch read from the reference's basis text and written again gives
    the same bytes (GoTools layout, %.15g numbers, empty last line)."""
    g = _gold()
    texts = list(g["Field3D-2P"]["basis"].values()) + list(g["Field3D-1P"]["basis"].values())
    for text in texts:
        (P,) = inputs.read_g2(text)
        assert H.g2_text(P) == text


def test_reference_layout_datatypes_and_bytes(tmp_path):
    """The two-patch reference file's basis and "u" datasets, written from
    their own content: same listing (for those objects), datatypes
    (H5T_STD_I8LE text, H5T_IEEE_F64LE values) and bytes."""
    g = _gold()["Field3D-2P"]
    patches = [inputs.read_g2(g["basis"][k])[0] for k in ("1", "2")]
    u = {k: np.array([float.fromhex(v) for v in g["u"][k]]) for k in ("1", "2")}
    w = H.HDF5Writer(str(tmp_path / "f"))
    w.open_file(0)
    w.write_basis(0, "Stokes-1", patches)
    for k in ("1", "2"):
        w.write_field(0, "Stokes-1", "u", int(k), u[k])
    w.close_file(0)
    path = w.path
    mine = _listing(path)
    keep = {"/", "/0", "/0/Stokes-1", "/0/Stokes-1/basis", "/0/Stokes-1/fields",
            "/0/Stokes-1/fields/u"}
    ref = [[k, p] for k, p in g["listing"] if p in keep or p.startswith("/0/Stokes-1/basis/")
           or p.startswith("/0/Stokes-1/fields/u/")]
    assert mine == ref
    for k, p in mine:
        if k == "dataset":
            assert _dtype(path, p) == g["types"][p], p
    for k in ("1", "2"):
        assert _raw(path, f"/0/Stokes-1/basis/{k}", str(tmp_path)) == g["basis"][k].encode()
        assert np.array_equal(_doubles(path, f"/0/Stokes-1/fields/u/{k}", str(tmp_path)), u[k])


def test_multipatch_solution_export_reads_back(tmp_path):
    """A solved three-patch model (oracle solution of the linear patch test):
    each patch's field holds the solution at its own nodes, the basis text
    reads back as the patch (15 significant digits) and the time is kept."""
    mp = M.x_split_patches(2, [2, 3, 2], 3, jitter=0.2, seed=4)
    itg = M.Integrand(M.POISSON)
    X = np.zeros((mp.nnod, 3))
    for P, ml in zip(mp.patches, mp.mlgn):
        X[ml - 1] = P.coefs
    u = 0.3 - 1.2 * X[:, 0] + 0.7 * X[:, 1] + 2.0 * X[:, 2]
    path = H.write_solution(str(tmp_path / "poisson"), mp.patches, u, 1, mlgns=mp.mlgn,
                            sim_name="Poisson", time=0.25)
    names = [p for k, p in _listing(path) if k == "dataset"]
    assert "/0/timeinfo/level" in names
    assert _doubles(path, "/0/timeinfo/level", str(tmp_path))[0] == 0.25
    for i, (P, ml) in enumerate(zip(mp.patches, mp.mlgn), 1):
        v = _doubles(path, f"/0/Poisson-1/fields/u/{i}", str(tmp_path))
        assert np.array_equal(v, u[ml - 1])
        (Q,) = inputs.read_g2(_raw(path, f"/0/Poisson-1/basis/{i}", str(tmp_path)).decode())
        assert Q.p == P.p and list(Q.n) == list(P.n)
        assert np.allclose(Q.coefs, P.coefs, rtol=1e-14, atol=1e-15)
        for d in range(3):
            assert np.allclose(Q.knots[d], P.knots[d], rtol=1e-14, atol=1e-15)


def test_time_levels_append_and_log(tmp_path):
    """A writer that truncates reports -1; levels written one after the
    other are counted when appending (the reference's getLastTimeLevel
    count); a log text reads back."""
    name = str(tmp_path / "steps")
    w = H.HDF5Writer(name)
    assert w.get_last_time_level() == -1
    for level, t in enumerate((0.0, 0.5, 1.0)):
        w.open_file(level)
        w.write_time_info(level, t)
        if level == 0:
            w.write_log("solver converged\n", "Stokes")
        w.close_file(level)
    assert H.HDF5Writer(name, append=True).get_last_time_level() == 3
    assert H.HDF5Writer(name).get_last_time_level() == -1
    for level, t in enumerate((0.0, 0.5, 1.0)):
        assert _doubles(w.path, f"/{level}/timeinfo/level", str(tmp_path))[0] == t
    assert _raw(w.path, "/log/Stokes/0", str(tmp_path)) == b"solver converged\n"
    # writing a level again replaces its datasets
    w2 = H.HDF5Writer(name, append=True)
    w2.open_file(1)
    w2.write_time_info(1, 0.75)
    w2.close_file(1)
    assert _doubles(w.path, "/1/timeinfo/level", str(tmp_path))[0] == 0.75
    assert H.HDF5Writer(name, append=True).get_last_time_level() == 3


def test_export_errors(tmp_path):
    w = H.HDF5Writer(str(tmp_path / "e"))
    with pytest.raises(RuntimeError):
        w.write_time_info(0, 0.0)  # not open
    w.open_file(0)
    with pytest.raises(RuntimeError):
        w.write_field(0, "M-1", "", 1, np.zeros(3))
    w.close_file(0)
    bad = H.HDF5Writer(str(tmp_path / "no_such_dir" / "x"))
    with pytest.raises(RuntimeError):
        bad.open_file(0)


@pytest.mark.gpu
def test_gpu_solution_export(tmp_path):
    """Elasticity solved on the GPU, exported and read back: the field is
    the expanded solution of the single patch."""
    from ifem_amd import problems, setup_problem
    pb = problems.elasticity_cube(6, 2, seed=2)
    print("[gpu] HDF5 export of a GPU solution", flush=True)
    with setup_problem(pb) as s:
        st = s.solve(rtol=1e-10, maxits=5000)
        assert st["converged"]
        xd = s.expanded_solution(pb.ndof)
    rp, ci, A, b = oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, pb.traction)
    xo, _, _ = oracle.pcg(rp, ci, A, b, rtol=1e-12)
    assert np.max(np.abs(xd - oracle.expand_solution(pb.meqn, xo))) <= 1e-8 * np.max(np.abs(xo))
    path = H.write_solution(str(tmp_path / "elast"), [pb.patch], xd, 3, sim_name="Elasticity")
    v = _doubles(path, "/0/Elasticity-1/fields/u/1", str(tmp_path))
    assert np.array_equal(v, xd)
