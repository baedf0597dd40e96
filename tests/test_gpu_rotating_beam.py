"""GPU joints and controls against the reference's RotatingBeamTest constraint cases
(tests/regression_tests/regression/test_rotating_beam.cpp; fixtures in
tests/golden/rotating_beam_constraints.json) and against the oracle on the same models, in the
fused and the plain form: rigid joint to a prescribed massless hub, pitch rotation control on a
fixed and on a rotating hub (control signal set every step), a free hub on a revolute joint,
and a generator torque on a tilted revolute shaft.

Bars: the reference's values at its own tolerance, the joint relations at 1e-12, the state
within 1e-12 relative of the FP64 oracle (rates 1e-10 / 1e-9), identical Newton iteration
counts every step, revolute joint output within 1e-10 relative."""

import numpy as np
import pytest

import openturbine_amd as ota
import rotating_beam_cases as RB
from golden_data import load
from oracle_lib import OracleLib, OracleSystem

pytestmark = pytest.mark.gpu

MODES = [pytest.param(ota.EXEC_FUSED, id="fused"), pytest.param(ota.EXEC_PLAIN, id="plain")]
STATE_TOL = {"q": 1e-12, "x": 1e-12, "q_delta": 1e-10, "v": 1e-10, "vd": 1e-9, "a": 1e-9}


@pytest.fixture(scope="module")
def kyn():
    return ota.KynLib()


@pytest.fixture(scope="module")
def orc():
    return OracleLib()


@pytest.fixture(scope="module")
def fx():
    return load("rotating_beam_constraints.json")


def rel(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / (scale if scale > 0 else 1.0))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", RB.CASES)
def test_rotating_beam_case(kyn, orc, fx, name, mode):
    print(f"case {name} mode {mode}", flush=True)
    spec = RB.case_spec(fx, name)
    g = ota.KynSystem(kyn, spec, exec_mode=mode)
    o = OracleSystem(orc, spec)
    for k in range(spec.num_steps):
        inputs = spec.step_inputs(k)
        assert g.step(inputs) == 0 and o.step(inputs) == 0, k
        assert len(g.convergence()) == len(o.convergence()), k
    worst = RB.check_case(fx, name, g.get, lambda c: g.constraint(c, 2))
    inv = RB.check_invariants(fx, name, g.get)
    print(f"{name}: worst deviation from the reference {worst:.2e}, joint invariants {inv:.1e}", flush=True)
    for f, tol in STATE_TOL.items():
        assert rel(g.get(f), o.get(f)) < tol, f
    for c in range(len(spec.constraints)):
        if spec.constraints[c].type == 4:
            assert rel(g.constraint(c, 2)[:3], o.constraint(c, 2)[:3]) < 1e-10


@pytest.mark.parametrize("mode", MODES)
def test_load_on_massless_hub(kyn, orc, fx, mode):
    """A moment and a force on the massless hub of the revolute case (a node outside every
    element: its rows hold only its load and constraint terms) against the oracle."""
    spec = RB.case_spec(fx, "RevoluteJointConstraint")
    hub = 6
    f = np.zeros((spec.num_nodes, 6))
    f[hub] = [0.0, 50.0, 0.0, 0.0, 0.0, 200.0]
    g = ota.KynSystem(kyn, spec, exec_mode=mode)
    o = OracleSystem(orc, spec)
    for k in range(spec.num_steps):
        inputs = dict(spec.step_inputs(k), nodal_loads=f)
        assert g.step(inputs) == 0 and o.step(inputs) == 0, k
        assert len(g.convergence()) == len(o.convergence()), k
    assert abs(o.get("q_delta")[hub, 5] - 0.1) > 1e-3  # the moment changes the spin
    for fld, tol in STATE_TOL.items():
        assert rel(g.get(fld), o.get(fld)) < tol, fld
