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
  ``<raiseorder>`` per patch or patch range, ``<topology>`` connections with
  orientation flags, ``<topologysets>`` with face sets),
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


# ---- multi-patch topology ------------------------------------------------------
def face_size(P: M.Patch, face: int) -> Tuple[int, int]:
    """Control points of face 1..6 in its two directions, ASMs3D::getFaceSize
    (src/ASM/ASMs3D.C:3796-3810)."""
    n1, n2, n3 = P.n
    if face in (1, 2):
        return n2, n3
    if face in (3, 4):
        return n1, n3
    return n1, n2


def face_nodes(P: M.Patch, face: int) -> np.ndarray:
    """0-based local nodes of face 1..6 in the order of
    ASMs3D::getBoundaryNodes (src/ASM/ASMs3D.C:1770-1818): faces 1/2 with j
    fastest then k, faces 3/4 i then k, faces 5/6 i then j."""
    n1, n2, n3 = P.n
    if face in (1, 2):
        i = 0 if face == 1 else n1 - 1
        k, j = np.meshgrid(np.arange(n3), np.arange(n2), indexing="ij")
        return (i + n1 * (j + n2 * k)).ravel()
    if face in (3, 4):
        j = 0 if face == 3 else n2 - 1
        k, i = np.meshgrid(np.arange(n3), np.arange(n1), indexing="ij")
        return (i + n1 * (j + n2 * k)).ravel()
    k = 0 if face == 5 else n3 - 1
    j, i = np.meshgrid(np.arange(n2), np.arange(n1), indexing="ij")
    return (i + n1 * (j + n2 * k)).ravel()


def connect_patches(patches: List[M.Patch], connections, xtol: float = 1.0e-4) -> List[np.ndarray]:
    """Global node numbers (MLGN, 1-based) of every patch's local nodes after
    collapsing the nodes of each connected face pair, as
    ASMs3D::connectBasis does (src/ASM/ASMs3D.C:721-808): the slave face's
    node (k, l) is matched to the master face's node (i, j) by the
    orientation flag 0..7, and the two must coincide within xtol per
    coordinate (Vec3::equal, src/Utility/Vec3.h:133-141), else ValueError.
    connections: (master, mface, slave, sface, orient), 1-based patches.
    Global nodes are numbered by first appearance, patch by patch."""
    off = np.concatenate([[0], np.cumsum([P.nnod for P in patches])]).astype(np.int64)
    parent = np.arange(off[-1], dtype=np.int64)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for master, mface, slave, sface, orient in connections:
        if master == slave or not (1 <= master <= len(patches) and 1 <= slave <= len(patches)):
            raise ValueError(f"connection: invalid patch indices {master} {slave}")
        if not 0 <= orient <= 7:
            raise ValueError(f"connection: orientation flag {orient} out of range [0,7]")
        Ps, Pm = patches[slave - 1], patches[master - 1]
        m1, m2 = face_size(Ps, sface)
        n1, n2 = face_size(Pm, mface)
        if orient > 3:
            n1, n2 = n2, n1
        if n1 != m1 or n2 != m2:
            raise ValueError(f"connection P{slave} F{sface} - P{master} F{mface}: non-matching "
                             f"faces, sizes {n1},{n2} and {m1},{m2}")
        if orient > 3:
            n1, n2 = n2, n1
        snodes = face_nodes(Ps, sface)
        mnodes = face_nodes(Pm, mface)
        node = 0
        for j in range(n2):
            for i in range(n1):
                k, l = {0: (i, j), 1: (i, n2 - j - 1), 2: (n1 - i - 1, j),
                        3: (n1 - i - 1, n2 - j - 1), 4: (j, i), 5: (j, m2 - i - 1),
                        6: (m1 - j - 1, i), 7: (m1 - j - 1, m2 - i - 1)}[orient]
                sn = int(snodes[l * m1 + k])
                mn = int(mnodes[node])
                node += 1
                if np.any(np.abs(Pm.coefs[mn] - Ps.coefs[sn]) > xtol):
                    raise ValueError(f"connection P{slave} F{sface} - P{master} F{mface}: "
                                     f"non-matching nodes {mn + 1}: {Pm.coefs[mn]} and "
                                     f"{sn + 1}: {Ps.coefs[sn]}")
                a, b = find(off[master - 1] + mn), find(off[slave - 1] + sn)
                if a != b:
                    parent[max(a, b)] = min(a, b)
    roots = np.array([find(a) for a in range(off[-1])], dtype=np.int64)
    _, first, inv = np.unique(roots, return_index=True, return_inverse=True)
    # number the global nodes in order of their first appearance
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    gid = rank[inv] + 1
    return [gid[off[q]:off[q + 1]].astype(np.int32) for q in range(len(patches))]


def _patch_list(el: ET.Element, npatch: int) -> List[int]:
    """Patches of a <refine>/<raiseorder> tag: patch="n", or lowerpatch ..
    upperpatch (SIMinput::parsePatchList, src/SIM/SIMinput.C:1036-1050);
    default patch 1."""
    lo = up = 1
    if el.get("patch") is not None:
        lo = up = int(el.get("patch"))
    if el.get("lowerpatch") is not None:
        lo = int(el.get("lowerpatch"))
        up = npatch
    if el.get("upperpatch") is not None:
        up = int(el.get("upperpatch"))
    if lo < 1 or up > npatch:
        raise ValueError(f"<{el.tag}>: patch range {lo}..{up} outside 1..{npatch}")
    return list(range(lo, up + 1))


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


def _read_patchfile(base_dir: str, name: str) -> List[M.Patch]:
    """A <patchfile>: relative to the input's folder, or (for inputs copied
    out of their tree) the file of that name beside the input."""
    path = os.path.join(base_dir, name)
    if not os.path.exists(path):
        path = os.path.join(base_dir, os.path.basename(name))
    with open(path) as f:
        return read_g2(f.read())


def read_xinp(source: str, base_dir: Optional[str] = None):
    """Problem (one patch) or MultiProblem (several patches, <topology>) and
    the linear-solver keys of an .xinp file (path or XML text)."""
    from .problems import MultiProblem
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
            patches += _read_patchfile(base_dir, (el.text or "").strip())
    if not patches:
        raise ValueError("no patches in <geometry>")
    for el in geo:
        tag = el.tag.lower()
        if tag == "raiseorder":
            for q in _patch_list(el, len(patches)):
                patches[q - 1] = raise_order(patches[q - 1], int(el.get("u", 0)),
                                             int(el.get("v", 0)), int(el.get("w", 0)))
        elif tag == "refine":
            if (el.text or "").strip():
                raise ValueError("<refine> with explicit knots is not supported")
            for q in _patch_list(el, len(patches)):
                for d, key in enumerate("uvw"):
                    patches[q - 1] = uniform_refine(patches[q - 1], d, int(el.get(key, 0)))
    connections = []
    top = geo.find("topology")
    if top is not None:
        offset = int(top.get("offset", 0))
        for c in top.findall("connection"):
            if int(c.get("dim", 2)) != 2:
                continue  # edge/vertex connections are ignored in serial (SIM3D.C:67)
            if c.get("periodic") not in (None, "false", "0"):
                raise ValueError("periodic connections are not supported")
            if c.get("basis") not in (None, "0", "1"):
                raise ValueError("mixed-basis connections are not supported")
            connections.append((int(c.get("master", 0)) + offset,
                                int(c.get("midx", c.get("mface", 0))),
                                int(c.get("slave", 0)) + offset,
                                int(c.get("sidx", c.get("sface", 0))),
                                int(c.get("orient", 0))))
    multi = len(patches) > 1 or bool(connections)
    mlgn = connect_patches(patches, connections) if multi else \
        [np.arange(1, patches[0].nnod + 1, dtype=np.int32)]
    nnod = int(max(int(m.max()) for m in mlgn))
    sets = read_topologysets(geo)
    for name, items in sets.items():
        for pq, _ in items:
            if not 1 <= pq <= len(patches):
                raise ValueError(f"topology set {name!r}: patch {pq} out of range")
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

    def set_dofs(items, mask):
        """Global DOFs (0-based) of the masked components on the set's faces."""
        out = []
        for pq, f in items:
            nodes = mlgn[pq - 1][face_nodes(patches[pq - 1], f)].astype(np.int64) - 1
            for c in range(nc):
                if (mask >> c) & 1:
                    out.append(nodes * nc + c)
        return np.unique(np.concatenate(out)) if out else np.zeros(0, dtype=np.int64)

    # boundary conditions: Dirichlet on face sets (homogeneous or a constant
    # value), constant tractions
    msc = np.ones(nnod * nc, dtype=np.int32)
    vals = np.full(nnod * nc, np.nan)
    inhom = False
    tractions = []
    bc = root.find("boundaryconditions")
    if bc is not None:
        for d in bc.findall("dirichlet"):
            items = sets.get(d.get("set"), [])
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
            dofs = set_dofs(items, mask)
            if value != 0.0:
                vals[dofs] = value  # later conditions win on shared DOFs
                inhom = True
            else:
                msc[dofs] = 0
        for nm in bc.findall("neumann"):
            items = sets.get(nm.get("set"), [])
            if not items:
                raise ValueError("<neumann> on an empty set")
            direction = int(nm.get("direction", 3))
            t = [0.0, 0.0, 0.0]
            t[direction - 1] = float((nm.text or "0").strip())
            for pq, f in items:
                tractions.append((pq - 1, f - 1, tuple(t[:nc] if nc < 3 else t)))
    sam = None
    if inhom:
        vals[msc == 0] = np.nan  # a homogeneously fixed DOF stays fixed
        sam = M.prescribed_equations(msc, vals)
        meqn, neq = sam.meqn, sam.neq
    else:
        meqn, neq = M.equation_numbers(msc)
    disc = root.find("discretization")
    nGauss, _ = parse_ngauss(disc.find("nGauss") if disc is not None else None)
    # ASMbase::getNoGaussPt per direction with the spline order p+1; the GPU
    # kernels take the counts as given and refuse those they do not support
    ngs = {tuple(M.gauss_points(P.p[d] + 1, nGauss) for d in range(3)) for P in patches}
    if len(ngs) != 1:
        raise ValueError("the patches need equal Gauss counts (equal degrees)")
    ng = ngs.pop()
    tr = []
    for q, f, t in tractions:
        P = patches[q]
        tdirs = [d for d in range(3) if d != f // 2]
        ngf = {M.gauss_points(P.p[d] + 1, nGauss, neumann=True) for d in tdirs}
        if len(ngf) != 1:
            raise ValueError("a traction face needs equal Gauss counts in its two directions")
        tr.append((q, f, t, ngf.pop()))
    solver: Dict[str, str] = {}
    ls = root.find("linearsolver")
    if ls is not None:
        for ch in ls:
            solver[ch.tag.lower()] = (ch.text or "").strip()
    if not multi:
        if len(tr) > 1:
            raise ValueError("one <neumann> face is supported on a single patch")
        traction = (tr[0][1], tr[0][2], tr[0][3]) if tr else None
        return Problem("xinp", patches[0], itg, ng, meqn, neq, traction, sam), solver
    return MultiProblem("xinp", patches, mlgn, itg, ng, meqn, neq, tr, sam), solver


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
