"""Oracle against the reference's RotatingBeamTest constraint cases
(tests/regression_tests/regression/test_rotating_beam.cpp, fixtures in
tests/golden/rotating_beam_constraints.json): rigid joint to a prescribed massless hub, pitch
rotation control on a fixed and on a rotating hub, a free hub on a revolute joint, and a
generator torque on a tilted revolute shaft, each at the test's own tolerance; plus the
kinematic relations each converged joint fixes (rigid carry, root = hub * exp(pitch x), rotation
about the revolute axis only) at 1e-12."""

import numpy as np
import pytest

import rotating_beam_cases as RB
from golden_data import load
from oracle_lib import OracleLib, OracleSystem


@pytest.fixture(scope="module")
def orc():
    return OracleLib()


@pytest.fixture(scope="module")
def fx():
    return load("rotating_beam_constraints.json")


@pytest.mark.parametrize("name", RB.CASES)
def test_rotating_beam_case(orc, fx, name):
    spec = RB.case_spec(fx, name)
    o = OracleSystem(orc, spec)
    for i in range(spec.num_steps):
        assert o.step() == 0, (name, i)
    worst = RB.check_case(fx, name, o.get, lambda c: o.constraint(c, 2))
    inv = RB.check_invariants(fx, name, o.get)
    print(f"{name}: worst deviation from the reference {worst:.2e}, joint invariants {inv:.1e}")


def test_hinged_beam_pendulum(orc, fx):
    """The beam hinged at its root through a three-dof ground node (RigidJoint6DOFsTo3DOFs +
    PrescribedBC3DOFs) swings under gravity with its root fixed in space."""
    spec = RB.hinged_beam_spec(fx)
    o = OracleSystem(orc, spec)
    assert o.num_dofs == 6 * 6 + 3 + 3 + 3
    for i in range(spec.num_steps):
        assert o.step() == 0, i
    q = o.get("q")
    assert np.max(np.abs(q[0, :3])) < 1e-12
    assert np.max(np.abs(q[spec.meta["ground"], :3])) < 1e-14
    assert q[5, 2] < -0.1  # the tip has fallen
