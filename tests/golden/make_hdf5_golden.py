"""Writes reference_hdf5.json from the reference's own HDF5 test files
(src/Utility/Test/refdata/Field3D-2P.hdf5 and Field3D-1P.hdf5, written by
the reference's HDF5Writer for its field-function tests).

Kept: the object listing of Field3D-2P, the datatype of each of its
datasets, the basis texts of its two patches and their "u" fields (as hex
floats), and the basis text of Field3D-1P.  The files are read with h5dump,
which executes nothing from them.

    python tests/golden/make_hdf5_golden.py
"""
import json
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/src/Utility/Test/refdata"
H5DUMP = "/opt/conda/bin/h5dump"


def listing(path):
    out = subprocess.run([H5DUMP, "-n", path], capture_output=True, text=True, check=True).stdout
    items = []
    for line in out.splitlines():
        m = re.match(r"\s*(group|dataset)\s+(\S.*)$", line)
        if m:
            items.append([m.group(1), m.group(2).strip()])
    return items


def dtype(path, ds):
    out = subprocess.run([H5DUMP, "-H", "-d", ds, path], capture_output=True, text=True,
                         check=True).stdout
    return re.search(r"DATATYPE\s+(\S+)", out).group(1)


def raw(path, ds):
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "x.bin")
        subprocess.run([H5DUMP, "-d", ds, "-b", "LE", "-o", f, path], capture_output=True,
                       check=True)
        with open(f, "rb") as fh:
            return fh.read()


def main():
    f2 = os.path.join(REF, "Field3D-2P.hdf5")
    f1 = os.path.join(REF, "Field3D-1P.hdf5")
    lst = listing(f2)
    types = {p: dtype(f2, p) for k, p in lst if k == "dataset"}
    basis = {str(i): raw(f2, f"/0/Stokes-1/basis/{i}").decode() for i in (1, 2)}
    u = {str(i): [float(v).hex() for v in np.frombuffer(raw(f2, f"/0/Stokes-1/fields/u/{i}"),
                                                         dtype="<f8")] for i in (1, 2)}
    out = {"Field3D-2P": {"listing": lst, "types": types, "basis": basis, "u": u},
           "Field3D-1P": {"basis": {"1": raw(f1, "/0/Stokes-1/basis/1").decode()},
                          "basis_type": dtype(f1, "/0/Stokes-1/basis/1")}}
    with open(os.path.join(HERE, "reference_hdf5.json"), "w") as fh:
        json.dump(out, fh)
    print("reference_hdf5.json written:", len(lst), "objects")


if __name__ == "__main__":
    main()
