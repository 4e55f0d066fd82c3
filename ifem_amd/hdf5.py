"""Result export to HDF5 in the reference's file layout.

The reference writes a model's patches (their GoTools text) and its
per-patch solution fields with HDF5Writer (src/Utility/HDF5Writer.C):

    /<level>/<sim>-1/basis/<patch>              GoTools text, char
    /<level>/<sim>-1/fields/<field>/<patch>     control-point values, double
    /<level>/<sim>-1/knotspan/<field>/<patch>   element values, double
    /<level>/timeinfo/level                     time, double
    /log/<name>/0                               text, char

with patches numbered from 1.  This module writes the same layout through
libifem_hdf5.so (include/ifem_hdf5.h), so that a GPU solution is read by
the tools that read the reference's .hdf5 files.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import model as M

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """The export library; raises if it was not built (no HDF5 found)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libifem_hdf5.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (HDF5 export needs the HDF5 C library)")
        L = C.CDLL(path)
        vp, ip, sz, dp, cp = C.c_void_p, C.c_int, C.c_size_t, C.POINTER(C.c_double), C.c_char_p
        L.ifem_h5_open.argtypes = [cp, ip, ip, C.POINTER(vp)]
        L.ifem_h5_close.argtypes = [vp]
        L.ifem_h5_last_level.argtypes = [cp]
        L.ifem_h5_write_basis.argtypes = [vp, ip, cp, ip, cp, sz]
        L.ifem_h5_write_field.argtypes = [vp, ip, cp, cp, ip, dp, sz]
        L.ifem_h5_write_knotspan.argtypes = [vp, ip, cp, cp, ip, dp, sz]
        L.ifem_h5_write_time.argtypes = [vp, ip, C.c_double]
        L.ifem_h5_write_log.argtypes = [vp, cp, cp, sz]
        L.ifem_g2_format.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.POINTER(C.c_double)), dp, C.c_char_p, sz,
                                     C.POINTER(sz)]
        L.ifem_h5_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError("HDF5 export: " + lib().ifem_h5_last_error().decode())


def g2_text(P: M.Patch) -> str:
    """GoTools text of a patch as the reference writes a basis (ASMs3D::write,
    src/ASM/ASMs3D.C:163-173): numbers in %.15g, empty last line."""
    order = (C.c_int * 3)(*[p + 1 for p in P.p])
    ncoef = (C.c_int * 3)(*P.n)
    ks = [np.ascontiguousarray(k, dtype=np.float64) for k in P.knots]
    kp = (C.POINTER(C.c_double) * 3)(*[k.ctypes.data_as(C.POINTER(C.c_double)) for k in ks])
    co = np.ascontiguousarray(P.coefs, dtype=np.float64).reshape(-1)
    n = C.c_size_t(0)
    cdp = co.ctypes.data_as(C.POINTER(C.c_double))
    _check(lib().ifem_g2_format(order, ncoef, kp, cdp, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    _check(lib().ifem_g2_format(order, ncoef, kp, cdp, buf, n.value, C.byref(n)))
    return buf.raw[:n.value].decode()


def patch_solutions(xdof: np.ndarray, ncomp: int, mlgns: Optional[Sequence] = None):
    """Per-patch control-point vectors of a global DOF vector
    (SIMbase::extractPatchSolution): the whole vector for one patch, else
    the values at each patch's global nodes (1-based maps)."""
    xdof = np.asarray(xdof, dtype=np.float64)
    if mlgns is None:
        return [xdof]
    out = []
    for ml in mlgns:
        idx = ((np.asarray(ml, dtype=np.int64) - 1)[:, None] * ncomp + np.arange(ncomp)).ravel()
        out.append(xdof[idx])
    return out


class HDF5Writer:
    """HDF5Writer(name, append): writes `name`.hdf5 (src/Utility/HDF5Writer.C:37-49).
    A file not appended to is truncated at the first open_file; later opens
    of the same writer append (closeFile switches to read-write)."""

    def __init__(self, name: str, append: bool = False):
        self.path = name + ".hdf5"
        self.append = append and os.path.exists(self.path)
        self._f = C.c_void_p()

    def get_last_time_level(self) -> int:
        """-1 for a writer that truncates; else the count of levels stored."""
        if not self.append:
            return -1
        return lib().ifem_h5_last_level(self.path.encode())

    def open_file(self, level: int) -> None:
        if self._f.value:
            return
        _check(lib().ifem_h5_open(self.path.encode(), int(self.append), level, C.byref(self._f)))

    def close_file(self, level: int = 0) -> None:
        if self._f.value:
            _check(lib().ifem_h5_close(self._f))
        self._f = C.c_void_p()
        self.append = True

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close_file()

    def _need(self):
        if not self._f.value:
            raise RuntimeError("HDF5Writer: open_file first")

    def write_time_info(self, level: int, time: float) -> None:
        self._need()
        _check(lib().ifem_h5_write_time(self._f, level, float(time)))

    def write_log(self, data: str, name: str) -> None:
        self._need()
        b = data.encode()
        _check(lib().ifem_h5_write_log(self._f, name.encode(), b, len(b)))

    def write_basis(self, level: int, name: str, patches: Sequence[M.Patch]) -> None:
        """/<level>/<name>/basis/<i> for patches i = 1, 2, ..."""
        self._need()
        for i, P in enumerate(patches, 1):
            b = g2_text(P).encode()
            _check(lib().ifem_h5_write_basis(self._f, level, name.encode(), i, b, len(b)))

    def _array(self, fn, level, basis, field, patch, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _check(fn(self._f, level, basis.encode(), field.encode(), patch,
                  v.ctypes.data_as(C.POINTER(C.c_double)), v.size))

    def write_field(self, level: int, basis: str, field: str, patch: int, values) -> None:
        self._need()
        self._array(lib().ifem_h5_write_field, level, basis, field, patch, values)

    def write_knotspan(self, level: int, basis: str, field: str, patch: int, values) -> None:
        self._need()
        self._array(lib().ifem_h5_write_knotspan, level, basis, field, patch, values)

    def write_sim(self, level: int, sim_name: str, patches: Sequence[M.Patch], solutions,
                  field: str = "u", geometry: bool = True) -> None:
        """Basis (at level 0 or when the geometry changed) and the primary
        solution of every patch, as HDF5Writer::writeSIM does for a
        single-basis model (src/Utility/HDF5Writer.C:231-384)."""
        basis = sim_name + "-1"
        if level == 0 or geometry:
            self.write_basis(level, basis, patches)
        for i, s in enumerate(solutions, 1):
            self.write_field(level, basis, field, i, s)


def write_solution(name: str, patches: Sequence[M.Patch], xdof, ncomp: int,
                   mlgns: Optional[Sequence] = None, sim_name: str = "Model", field: str = "u",
                   level: int = 0, time: float = 0.0, append: bool = False) -> str:
    """One call: basis, per-patch solution and time of one level; returns the
    file path."""
    w = HDF5Writer(name, append)
    w.open_file(level)
    try:
        w.write_sim(level, sim_name, patches, patch_solutions(xdof, ncomp, mlgns), field)
        w.write_time_info(level, time)
    finally:
        w.close_file(level)
    return w.path
