"""Reference input formats for the GPU path: GoTools ``.g2`` spline volumes and
the ``.xinp`` subset that describes the hot path's models.

* ``read_g2`` / ``write_g2``: SplineVolume records (class 700, GoTools ASCII:
  header ``700 1 0 0``, ``dim rational``, per direction ``n order`` and the
  n+order knots, then the control points with i fastest), as read by
  ASMs3D::read (src/ASM/ASMs3D.C:117-160) and written by GoTools.
* ``uniform_refine``: ``nins`` equally spaced knots inserted in every
  non-empty knot span of one direction, ASM3D::uniformRefine
  (src/SIM/SIM3D.C:104-128), by Boehm knot insertion, which leaves the
  geometry unchanged.
* ``raise_order``: degree elevation of a patch without interior knots
  (Bernstein degree elevation), ASM3D::raiseOrder (src/SIM/SIM3D.C:152-170).
  With interior knots, elevation would double them and the GPU path needs
  simple interior knots; such patches are refused (raise before refining, as
  IFEM inputs do).
* ``read_xinp``: ``<geometry>`` (``<patch>``/``<patchfile>``, ``<refine>``,
  ``<raiseorder>``, ``<topologysets>`` with face sets),
  ``<boundaryconditions>`` (``<dirichlet>`` on face sets, homogeneous or
  with a constant value,
  constant ``<neumann>`` traction), ``<discretization><nGauss>``,
  ``<linearsolver>`` keys, and the physics: ``<poisson>`` (conductivity,
  source) or ``<elasticity>`` (``<isotropic E nu>``), following the tag
  handling of SIMinput::parse (src/SIM/SIMinput.C:92-200, 410-700, 894-953)
  and SIM3D::parseGeometryTag (src/SIM/SIM3D.C:100-200).
"""
from __future__ import annotations

import math
import os
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import model as M
from .problems import Problem


# ---- .g2 --------------------------------------------------------------------
def read_g2(text: str) -> List[M.Patch]:
    """All SplineVolume records of a .g2 text (non-rational only)."""
    tok = text.split()
    pos = 0
    out = []

    def nxt():
        nonlocal pos
        v = tok[pos]
        pos += 1
        return v

    while pos < len(tok):
        cls = int(nxt())
        for _ in range(3):
            nxt()  # version and reserved fields
        if cls != 700:
            raise ValueError(f"unsupported GoTools class {cls} (only 700, SplineVolume)")
        dim, rational = int(nxt()), int(nxt())
        if dim != 3:
            raise ValueError("only 3D spline volumes are supported")
        if rational:
            raise ValueError("rational spline volumes are not supported by the GPU path")
        n, p, knots = [], [], []
        for _ in range(3):
            nd, order = int(nxt()), int(nxt())
            kn = np.array([float(nxt()) for _ in range(nd + order)])
            n.append(nd)
            p.append(order - 1)
            knots.append(kn)
        cnt = n[0] * n[1] * n[2]
        coefs = np.array([float(nxt()) for _ in range(3 * cnt)]).reshape(cnt, 3)
        out.append(M.Patch(p, n, knots, coefs))
    return out


def write_g2(P: M.Patch) -> str:
    lines = ["700 1 0 0", "3 0"]
    for d in range(3):
        lines.append(f"{P.n[d]} {P.p[d] + 1}")
        lines.append(" ".join(repr(float(v)) for v in P.knots[d]))
    for x in P.coefs:
        lines.append(" ".join(repr(float(v)) for v in x))
    return "\n".join(lines) + "\n"


def _net(P: M.Patch) -> np.ndarray:
    """Control net as (n3, n2, n1, 3)."""
    return P.coefs.reshape(P.n[2], P.n[1], P.n[0], 3)


def _axis(d: int) -> int:
    """Array axis of parametric direction d in the (n3, n2, n1, 3) net."""
    return 2 - d


def insert_knot(P: M.Patch, d: int, u: float) -> M.Patch:
    """Boehm insertion of knot u (inside the parameter range) in direction d."""
    p, t = P.p[d], P.knots[d]
    k = int(np.searchsorted(t, u, side="right")) - 1  # t[k] <= u < t[k+1]
    net = np.moveaxis(_net(P), _axis(d), 0)          # (n_d, ..., 3)
    nd = net.shape[0]
    new = np.empty((nd + 1,) + net.shape[1:])
    new[:k - p + 1] = net[:k - p + 1]
    for i in range(k - p + 1, k + 1):
        a = (u - t[i]) / (t[i + p] - t[i])
        new[i] = a * net[i] + (1.0 - a) * net[i - 1]
    new[k + 1:] = net[k:]
    knots = list(P.knots)
    knots[d] = np.insert(t, k + 1, u)
    n = list(P.n)
    n[d] += 1
    net2 = np.moveaxis(new, 0, _axis(d))
    return M.Patch(list(P.p), n, knots, np.ascontiguousarray(net2.reshape(-1, 3)))


def uniform_refine(P: M.Patch, d: int, nins: int) -> M.Patch:
    """nins equally spaced knots in every non-empty span of direction d."""
    if nins <= 0:
        return P
    t = P.knots[d]
    spans = [(t[i], t[i + 1]) for i in range(len(t) - 1) if t[i + 1] > t[i]]
    for a, b in spans:
        for j in range(1, nins + 1):
            P = insert_knot(P, d, a + (b - a) * j / (nins + 1))
    return P


def raise_order(P: M.Patch, ru: int, rv: int, rw: int) -> M.Patch:
    """Degree elevation by (ru, rv, rw) of a patch without interior knots."""
    for d, r in enumerate((ru, rv, rw)):
        for _ in range(r):
            p, t = P.p[d], P.knots[d]
            if P.n[d] != p + 1:
                raise ValueError("raise_order: only patches without interior knots "
                                 "(raise the order before refining)")
            net = np.moveaxis(_net(P), _axis(d), 0)
            new = np.zeros((p + 2,) + net.shape[1:])
            for i in range(p + 2):
                a = i / (p + 1)
                if i > 0:
                    new[i] += a * net[i - 1]
                if i <= p:
                    new[i] += (1.0 - a) * net[i]
            knots = list(P.knots)
            knots[d] = np.array([t[0]] * (p + 2) + [t[-1]] * (p + 2))
            n = list(P.n)
            n[d] = p + 2
            pp = list(P.p)
            pp[d] = p + 1
            net2 = np.moveaxis(new, 0, _axis(d))
            P = M.Patch(pp, n, knots, np.ascontiguousarray(net2.reshape(-1, 3)))
    return P


def evaluate(P: M.Patch, u: float, v: float, w: float) -> np.ndarray:
    """Point of the volume at (u, v, w) (de Boor's recurrence per direction)."""
    def basis(p, t, x):
        n = len(t) - p - 1
        k = min(max(int(np.searchsorted(t, x, side="right")) - 1, p), n - 1)
        N = np.zeros(p + 1)
        N[0] = 1.0
        left, right = np.zeros(p + 1), np.zeros(p + 1)
        for j in range(1, p + 1):
            left[j] = x - t[k + 1 - j]
            right[j] = t[k + j] - x
            saved = 0.0
            for r in range(j):
                tmp = N[r] / (right[r + 1] + left[j - r])
                N[r] = saved + right[r + 1] * tmp
                saved = left[j - r] * tmp
            N[j] = saved
        return k - p, N
    i0, N1 = basis(P.p[0], P.knots[0], u)
    j0, N2 = basis(P.p[1], P.knots[1], v)
    k0, N3 = basis(P.p[2], P.knots[2], w)
    net = _net(P)
    sub = net[k0:k0 + len(N3), j0:j0 + len(N2), i0:i0 + len(N1)]
    return np.einsum("k,j,i,kjid->d", N3, N2, N1, sub)


# ---- .xinp ------------------------------------------------------------------
_FN_TYPES = {"constant": M.FN_CONST, "sinprod": M.FN_SINPROD}


def _items(text: Optional[str]) -> List[int]:
    return [int(x) for x in (text or "").split()]


def _atoi(text: str) -> int:
    """C atoi: optional sign and leading digits, 0 when there are none."""
    s = text.strip()
    k = 1 if s[:1] in ("+", "-") else 0
    while k < len(s) and s[k].isdigit():
        k += 1
    digits = s[:k]
    return int(digits) if digits not in ("", "+", "-") else 0


# SIMoptions::nGauss default (src/SIM/SIMoptions.C:58)
NGAUSS_DEFAULT = 4


def parse_ngauss(el: Optional[ET.Element]) -> Tuple[int, int]:
    """(main, recovery) integration rules of <discretization><nGauss>, as
    SIMoptions::parseDiscretizationTag reads them (src/SIM/SIMoptions.C:
    139-150): default="N" means 10+N for N > 0 (order + N points) and N
    otherwise; a text value sets both rules, a second value the recovery
    rule; no tag keeps the default 4."""
    ng = [NGAUSS_DEFAULT, NGAUSS_DEFAULT]
    if el is None:
        return ng[0], ng[1]
    dflt = el.get("default")
    if dflt is not None:
        g = _atoi(dflt)
        ng[0] = ng[1] = 10 + g if g > 0 else g
    elif el.text is not None:
        for i, tok in enumerate(el.text.split()[:2]):
            for j in range(i, 2):
                ng[j] = _atoi(tok)
    return ng[0], ng[1]


def _function(el: Optional[ET.Element], default: M.Function) -> M.Function:
    """<conductivity>/<source>: a number, or type="sinprod" with an amplitude
    (scale * sin(pi x) sin(pi y) sin(pi z))."""
    if el is None:
        return default
    typ = (el.get("type") or "constant").lower()
    val = float((el.text or "0").strip() or 0.0)
    if typ not in _FN_TYPES:
        raise ValueError(f"unsupported field type {typ!r} (constant, sinprod)")
    return M.Function(_FN_TYPES[typ], val)


def read_topologysets(geo: ET.Element) -> Dict[str, List[Tuple[int, int]]]:
    """Face sets of <topologysets> as {name: [(patch, face), ...]}
    (1-based patch and IFEM face numbers).  SIMinput::parse
    (src/SIM/SIMinput.C:229-252) reads the set's dimension from its type:
    volume, face/surface, edge/curve, vertex/point, nodes, elements, bbox, and
    a set without a type is a whole volume.  Only face sets are read here;
    every other dimension, the untyped set included, raises ValueError rather
    than being taken for faces."""
    sets: Dict[str, List[Tuple[int, int]]] = {}
    ts = geo.find("topologysets")
    if ts is None:
        return sets
    for st in ts.findall("set"):
        name = st.get("name")
        if name is None:
            continue
        typ = (st.get("type") or "").lower()
        if typ not in ("face", "surface"):
            what = typ or f"untyped (volume, dimension {st.get('dimension', '3')})"
            raise ValueError(f"topology set {name!r}: {what} sets are not supported "
                             "(face sets only)")
        if (st.get("closure") or "").lower() == "open":
            raise ValueError(f"topology set {name!r}: open face sets are not supported")
        items = sets.setdefault(name, [])
        for it in st.findall("item"):
            pq = int(it.get("patch", 1))
            for f in _items(it.text):
                if not 1 <= f <= 6:
                    raise ValueError(f"topology set {name!r}: face {f} is not 1..6")
                items.append((pq, f))
    return sets


def read_xinp(source: str, base_dir: Optional[str] = None) -> Tuple[Problem, Dict[str, str]]:
    """Problem and linear-solver keys of an .xinp file (path or XML text)."""
    if os.path.exists(source):
        base_dir = base_dir or os.path.dirname(os.path.abspath(source))
        root = ET.parse(source).getroot()
    else:
        root = ET.fromstring(source)
    base_dir = base_dir or os.getcwd()
    geo = root.find("geometry")
    if geo is None:
        raise ValueError("no <geometry>")
    patches: List[M.Patch] = []
    for el in geo:
        tag = el.tag.lower()
        if tag == "patch":
            patches += read_g2(el.text or "")
        elif tag == "patchfile":
            with open(os.path.join(base_dir, (el.text or "").strip())) as f:
                patches += read_g2(f.read())
    if len(patches) != 1:
        raise ValueError("the GPU path takes exactly one patch per model")
    P = patches[0]
    for el in geo:
        tag = el.tag.lower()
        if tag == "raiseorder":
            P = raise_order(P, int(el.get("u", 0)), int(el.get("v", 0)), int(el.get("w", 0)))
        elif tag == "refine":
            for d, key in enumerate("uvw"):
                P = uniform_refine(P, d, int(el.get(key, 0)))
    sets = read_topologysets(geo)
    for name, items in sets.items():
        if any(pq != 1 for pq, _ in items):
            raise ValueError(f"topology set {name!r}: only patch 1")
    sets = {name: [f for _, f in items] for name, items in sets.items()}
    # physics
    if root.find("elasticity") is not None:
        el = root.find("elasticity")
        iso = el.find("isotropic")
        E = float(iso.get("E", 1.0)) if iso is not None else 1.0
        nu = float(iso.get("nu", 0.3)) if iso is not None else 0.3
        itg = M.Integrand(M.ELASTICITY, E=E, nu=nu)
    else:
        el = root.find("poisson")
        kappa = _function(el.find("conductivity") if el is not None else None,
                          M.Function(M.FN_ZERO, 0.0))
        src = _function(el.find("source") if el is not None else None,
                        M.Function(M.FN_ZERO, 0.0))
        itg = M.Integrand(M.POISSON, kappa=kappa, source=src)
    nc = itg.ncomp
    # boundary conditions: homogeneous Dirichlet on faces, constant traction
    fix = [0] * 6
    inhom = []  # (face, mask, value) of constant inhomogeneous conditions
    traction = None
    bc = root.find("boundaryconditions")
    if bc is not None:
        for d in bc.findall("dirichlet"):
            faces = sets.get(d.get("set"), [])
            comp = str(d.get("comp", d.get("component", "1" if nc == 1 else "123")))
            mask = 0
            for ch in comp:
                c = int(ch) - 1
                if 0 <= c < nc:
                    mask |= 1 << c
            # a nonzero constant value makes the condition inhomogeneous
            # (Property::DIRICHLET_INHOM, src/SIM/SIMinput.C:576-618)
            typ = (d.get("type") or "").lower()
            text = (d.text or "").strip()
            if typ in ("anasol", "expression", "field") or typ.startswith("symm"):
                raise ValueError(f"<dirichlet type='{typ}'> is not supported")
            value = float(text) if text else 0.0
            for f in faces:
                if value != 0.0:
                    inhom.append((f - 1, mask, value))
                else:
                    fix[f - 1] |= mask
        for nm in bc.findall("neumann"):
            faces = sets.get(nm.get("set"), [])
            if len(faces) != 1 or traction is not None:
                raise ValueError("one <neumann> face is supported")
            direction = int(nm.get("direction", 3))
            t = [0.0, 0.0, 0.0]
            t[direction - 1] = float((nm.text or "0").strip())
            traction = (faces[0] - 1, tuple(t), None)
    msc = M.face_status(P, nc, fix)
    sam = None
    if inhom:
        vals = np.full(P.nnod * nc, np.nan)
        for f, mask, value in inhom:  # later conditions win on shared DOFs
            on = M.face_status(P, nc, [mask if g == f else 0 for g in range(6)]) == 0
            vals[on] = value
        sam = M.prescribed_equations(msc, vals)
        meqn, neq = sam.meqn, sam.neq
    else:
        meqn, neq = M.equation_numbers(msc)
    disc = root.find("discretization")
    nGauss, _ = parse_ngauss(disc.find("nGauss") if disc is not None else None)
    # ASMbase::getNoGaussPt per direction with the spline order p+1; the GPU
    # kernels take the counts as given and refuse those they do not support
    ng = tuple(M.gauss_points(P.p[d] + 1, nGauss) for d in range(3))
    if traction is not None:
        f = traction[0]
        tdirs = [d for d in range(3) if d != f // 2]
        ngf = {M.gauss_points(P.p[d] + 1, nGauss, neumann=True) for d in tdirs}
        if len(ngf) != 1:
            raise ValueError("the traction face needs equal Gauss counts in its two directions")
        traction = (traction[0], traction[1], ngf.pop())
    solver: Dict[str, str] = {}
    ls = root.find("linearsolver")
    if ls is not None:
        for ch in ls:
            solver[ch.tag.lower()] = (ch.text or "").strip()
    return Problem("xinp", P, itg, ng, meqn, neq, traction, sam), solver


def solver_kwargs(solver: Dict[str, str]) -> dict:
    """ifem_amd Session.solve keyword arguments from <linearsolver> keys
    (LinSolParams, src/LinAlg/LinSolParams.C:68-77,158-195)."""
    if solver.get("type", "cg").lower() != "cg":
        raise ValueError("the GPU path solves with CG only")
    pc = solver.get("pc", "jacobi").lower()
    kw = {}
    if pc in ("gamg", "ml", "mg", "multigrid"):
        from ._native import PC_MG
        kw["pc"] = PC_MG
        if "multigrid_no_smooth" in solver:
            kw["mg_smooth"] = int(solver["multigrid_no_smooth"])
    elif pc not in ("jacobi", "default"):
        raise ValueError("the GPU path preconditions with Jacobi or multigrid")
    for key, name, conv in (("rtol", "rtol", float), ("atol", "atol", float),
                            ("dtol", "dtol", float), ("maxits", "maxits", int)):
        if key in solver:
            kw[name] = conv(solver[key])
    return kw
