"""Input front-end: GoTools .g2 volumes, knot insertion and degree elevation
(geometry-preserving), and the .xinp subset of the hot path's models, checked
against the generated benchmark problems and against evaluation of the
spline map."""
import math

import numpy as np
import pytest

import oracle
from ifem_amd import inputs, model as M, problems

TRILINEAR_CUBE = """700 1 0 0
3 0
2 2
0 0 1 1
2 2
0 0 1 1
2 2
0 0 1 1
0 0 0
1 0 0
0 1 0
1 1 0
0 0 1
1 0 1
0 1 1
1 1 1
"""


def test_g2_round_trip():
    P = M.cube_patch(2, [3, 4, 5], jitter=0.2, seed=3)
    Q = inputs.read_g2(inputs.write_g2(P))[0]
    assert Q.p == P.p and Q.n == P.n
    for d in range(3):
        assert np.array_equal(Q.knots[d], P.knots[d])
    assert np.array_equal(Q.coefs, P.coefs)


def test_g2_refuses_rational_and_other_classes():
    with pytest.raises(ValueError):
        inputs.read_g2(TRILINEAR_CUBE.replace("3 0\n", "3 1\n", 1))
    with pytest.raises(ValueError):
        inputs.read_g2("200 1 0 0\n" + TRILINEAR_CUBE.split("\n", 1)[1])


def test_raise_then_refine_gives_the_benchmark_cube():
    """A trilinear cube raised to p=2 and refined with 15 knots per span is
    the 16^3 p=2 cube of config 1 (identity map: control points at the
    Greville abscissae)."""
    P = inputs.read_g2(TRILINEAR_CUBE)[0]
    P = inputs.raise_order(P, 1, 1, 1)
    for d in range(3):
        P = inputs.uniform_refine(P, d, 15)
    R = M.cube_patch(2, 16)
    assert P.p == R.p and P.n == R.n
    for d in range(3):
        assert np.max(np.abs(P.knots[d] - R.knots[d])) <= 1e-15
    assert np.max(np.abs(P.coefs - R.coefs)) <= 1e-14


def test_refinement_and_elevation_preserve_the_geometry():
    rng = np.random.default_rng(5)
    P = M.cube_patch(2, [3, 2, 4], jitter=0.25, seed=7)
    R = P
    for d in range(3):
        R = inputs.uniform_refine(R, d, 2)
    assert R.n == [P.n[0] + 2 * 3, P.n[1] + 2 * 2, P.n[2] + 2 * 4]
    # a curved trilinear block, elevated twice
    B = inputs.read_g2(TRILINEAR_CUBE)[0]
    B.coefs[7] += [0.3, -0.2, 0.25]
    B.coefs[2] += [-0.1, 0.15, 0.05]
    E = inputs.raise_order(B, 2, 1, 2)
    assert E.p == [3, 2, 3]
    for _ in range(20):
        u, v, w = rng.random(3)
        assert np.max(np.abs(inputs.evaluate(R, u, v, w) - inputs.evaluate(P, u, v, w))) <= 1e-13
        assert np.max(np.abs(inputs.evaluate(E, u, v, w) - inputs.evaluate(B, u, v, w))) <= 1e-13
    with pytest.raises(ValueError):
        inputs.raise_order(R, 1, 0, 0)  # interior knots


POISSON_XINP = f"""<?xml version="1.0"?>
<simulation>
  <geometry>
    <patch>
{TRILINEAR_CUBE}
    </patch>
    <raiseorder patch="1" u="1" v="1" w="1"/>
    <refine patch="1" u="15" v="15" w="15"/>
    <topologysets>
      <set name="Boundary" type="face"><item patch="1">1 2 3 4 5 6</item></set>
    </topologysets>
  </geometry>
  <boundaryconditions>
    <dirichlet set="Boundary" comp="1"/>
  </boundaryconditions>
  <poisson>
    <source type="sinprod">{3 * math.pi ** 2!r}</source>
  </poisson>
  <discretization><nGauss>3</nGauss></discretization>
  <linearsolver><type>cg</type><pc>jacobi</pc><rtol>1e-10</rtol><maxits>5000</maxits></linearsolver>
</simulation>
"""

ELASTICITY_XINP = f"""<simulation>
  <geometry>
    <patch>{TRILINEAR_CUBE}</patch>
    <raiseorder patch="1" u="1" v="1" w="1"/>
    <refine patch="1" u="3" v="3" w="3"/>
    <topologysets>
      <set name="Left" type="face"><item patch="1">1</item></set>
      <set name="Right" type="face"><item patch="1">2</item></set>
    </topologysets>
  </geometry>
  <boundaryconditions>
    <dirichlet set="Left" comp="123"/>
    <neumann set="Right" direction="3">-1</neumann>
  </boundaryconditions>
  <elasticity><isotropic E="1" nu="0.3"/></elasticity>
</simulation>
"""


def test_xinp_poisson_is_config_1():
    pb, solver = inputs.read_xinp(POISSON_XINP)
    ref = problems.poisson_cube(16, 2)
    assert pb.neq == ref.neq and np.array_equal(pb.meqn, ref.meqn)
    assert np.max(np.abs(pb.patch.coefs - ref.patch.coefs)) <= 1e-14
    assert pb.ng == ref.ng
    assert pb.integrand.source.type == M.FN_SINPROD
    assert abs(pb.integrand.source.scale - ref.integrand.source.scale) <= 1e-15
    assert inputs.solver_kwargs(solver) == {"rtol": 1e-10, "maxits": 5000}


def test_xinp_elasticity_matches_generated_problem():
    pb, solver = inputs.read_xinp(ELASTICITY_XINP)
    ref = problems.elasticity_cube(4, 2, jitter=0.0)
    assert pb.neq == ref.neq and np.array_equal(pb.meqn, ref.meqn)
    # no <nGauss>: SIMoptions' default 4 (a fixed count, also on the face)
    assert pb.ng == (4, 4, 4)
    assert pb.traction[:2] == ref.traction[:2] and pb.traction[2] == 4
    pb3, _ = inputs.read_xinp(ELASTICITY_XINP.replace(
        "</elasticity>", "</elasticity><discretization><nGauss>3</nGauss></discretization>"))
    assert pb3.traction == ref.traction and pb3.ng == ref.ng
    assert pb.integrand.kind == M.ELASTICITY and pb.integrand.nu == 0.3
    assert solver == {}
    with pytest.raises(ValueError):
        inputs.solver_kwargs({"type": "gmres"})


@pytest.mark.gpu
def test_xinp_model_solves_to_the_manufactured_error():
    """Config 1 read from .xinp, assembled and solved on the GPU: the L2 error
    against sin(pi x) sin(pi y) sin(pi z) equals the oracle's 2.2624e-5."""
    from ifem_amd import setup_problem
    pb, solver = inputs.read_xinp(POISSON_XINP)
    print("[gpu] xinp config 1", flush=True)
    with setup_problem(pb) as s:
        st = s.solve(**inputs.solver_kwargs(solver))
        assert st["converged"]
        n = s.norms(pb.integrand, pb.ng, M.Function(M.FN_SINPROD, 1.0))
    assert abs(n[4] - 2.26239829e-05) <= 1e-5 * n[4]


UNIAXIAL_XINP = f"""<simulation>
  <geometry>
    <patch>{TRILINEAR_CUBE}</patch>
    <raiseorder patch="1" u="1" v="1" w="1"/>
    <refine patch="1" u="2" v="2" w="2"/>
    <topologysets>
      <set name="Left" type="face"><item patch="1">1</item></set>
      <set name="Right" type="face"><item patch="1">2</item></set>
      <set name="Front" type="face"><item patch="1">3</item></set>
      <set name="Bottom" type="face"><item patch="1">5</item></set>
    </topologysets>
  </geometry>
  <boundaryconditions>
    <dirichlet set="Left" comp="1"/>
    <dirichlet set="Right" comp="1">0.01</dirichlet>
    <dirichlet set="Front" comp="2"/>
    <dirichlet set="Bottom" comp="3"/>
  </boundaryconditions>
  <elasticity><isotropic E="1" nu="0.3"/></elasticity>
</simulation>
"""


def _uniaxial_exact(pb, d=0.01, nu=0.3):
    X = pb.patch.coefs
    return np.stack([d * X[:, 0], -nu * d * X[:, 1], -nu * d * X[:, 2]], axis=1).ravel()


def test_xinp_inhomogeneous_dirichlet_uniaxial_strain():
    """A nonzero constant <dirichlet> value becomes one constraint equation
    per face DOF (u_x = 0.01 on x = 1); with the symmetry planes the exact
    uniaxial-strain field is linear, so the oracle's solution equals it."""
    pb, _ = inputs.read_xinp(UNIAXIAL_XINP)
    n = pb.patch.n[0]
    assert pb.sam is not None and pb.sam.nceq == n * n
    assert np.all(pb.sam.ttcc == 0.01)
    assert pb.neq == 3 * n ** 3 - 4 * n * n
    rp, ci = oracle.pattern(pb.patch, 3, pb.meqn, pb.neq)
    g = pb.prescribed_values()
    A, b = oracle.assemble(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, rp, ci, gdof=g)
    x, it, res = oracle.pcg(rp, ci, A, b, rtol=1e-14)
    xd = oracle.expand_solution(pb.meqn, x, g)
    assert np.max(np.abs(xd - _uniaxial_exact(pb))) <= 1e-13
    with pytest.raises(ValueError):
        inputs.read_xinp(UNIAXIAL_XINP.replace('comp="1">0.01', 'comp="1" type="expression">x'))


@pytest.mark.gpu
def test_xinp_inhomogeneous_dirichlet_on_gpu():
    from ifem_amd import setup_problem
    pb, _ = inputs.read_xinp(UNIAXIAL_XINP)
    print("[gpu] xinp uniaxial (prescribed values)", flush=True)
    with setup_problem(pb) as s:
        st = s.solve(rtol=1e-13, maxits=5000)
        assert st["converged"]
        xd = s.expanded_solution(pb.ndof)
    assert np.max(np.abs(xd - _uniaxial_exact(pb))) <= 1e-12


# Known answers of ASMbase::getNoGaussPt (src/ASM/ASMbase.C:1788-1798), read
# off the reference with the spline order as its argument (ASMs3D::getOrder
# returns svol->order(d), src/ASM/ASMs3D.C:1898-1906): (order, nGauss,
# neumann) -> points.
GAUSS_KAT = [
    (3, 4, False, 4),     # SIMoptions default 4 (SIMoptions.C:58): fixed count
    (3, 4, True, 4),      # fixed counts hold on faces too
    (2, 10, False, 10),   # 10 is still a fixed count (nGauss < 11)
    (3, 0, False, 3),     # ASMbase's own initial nGauss = 0 (ASMbase.C:65): order
    (3, 0, True, 3),
    (3, 11, False, 4),    # <nGauss default="1"/> = 11: order + 1
    (4, 11, False, 5),
    (3, 13, False, 6),    # order + 13 % 10
    (3, 13, True, 3),     # faces: the order
    (2, 20, False, 2),    # 20 % 10 = 0
    (3, -1, False, 2),    # order + (-1 % 10) with C's remainder = order - 1
    (3, -1, True, 3),
    (3, -2, False, 1),
    (3, -3, False, 0),    # nGauss <= -order: none
    (3, -3, True, 0),     # checked before the Neumann branch
    (2, -13, False, 0),
    (4, -13, False, 0),
    (4, -3, True, 4),     # above -order and not fixed: faces take the order
    (4, -3, False, 1),
]


@pytest.mark.parametrize("order,nGauss,neumann,expect", GAUSS_KAT)
def test_gauss_point_rule_known_answers(order, nGauss, neumann, expect):
    assert M.gauss_points(order, nGauss, neumann) == expect
    assert oracle.num_gauss(order, nGauss, neumann) == expect


@pytest.mark.parametrize("tag,expect", [
    (None, (4, 4)),                                    # SIMoptions.C:58
    ('<nGauss default="1"/>', (11, 11)),               # SIMoptions.C:140-142
    ('<nGauss default="0"/>', (0, 0)),
    ('<nGauss default="-1"/>', (-1, -1)),
    ("<nGauss>5</nGauss>", (5, 5)),                    # SIMoptions.C:143-149
    ("<nGauss>3 2</nGauss>", (3, 2)),
    ("<nGauss> 13 </nGauss>", (13, 13)),
    ("<nGauss>x</nGauss>", (0, 0)),                    # atoi of a non-number
])
def test_ngauss_tag_forms(tag, expect):
    import xml.etree.ElementTree as ET
    el = ET.fromstring(tag) if tag else None
    assert inputs.parse_ngauss(el) == expect


def test_xinp_gauss_rule_applied_per_direction():
    """<nGauss default="1"/> on a p = 2 patch integrates with order + 1 = 4
    points; 13 gives 6, which the GPU kernels refuse when asked, and the
    traction face always takes the order (3) for nGauss > 10."""
    x = POISSON_XINP.replace("<nGauss>3</nGauss>", '<nGauss default="1"/>')
    pb, _ = inputs.read_xinp(x)
    assert pb.ng == (4, 4, 4)
    pb, _ = inputs.read_xinp(POISSON_XINP.replace("<nGauss>3</nGauss>", "<nGauss>13</nGauss>"))
    assert pb.ng == (6, 6, 6)
    e = ELASTICITY_XINP.replace(
        "</elasticity>", "</elasticity><discretization><nGauss>13</nGauss></discretization>")
    pb, _ = inputs.read_xinp(e)
    assert pb.ng == (6, 6, 6) and pb.traction[2] == 3


@pytest.mark.parametrize("settype", ["edge", "vertex", "volume", "nodes", None])
def test_topology_sets_other_than_faces_are_refused(settype):
    """An edge, vertex or untyped (whole-volume, SIMinput.C:235-252) set must
    not be read as a list of faces."""
    attr = f' type="{settype}"' if settype else ""
    x = ELASTICITY_XINP.replace('<set name="Left" type="face">', f'<set name="Left"{attr}>')
    with pytest.raises(ValueError, match="not supported"):
        inputs.read_xinp(x)
