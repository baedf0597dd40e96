"""GPU BladeInterface (kyn_blade_interface_*: the C++ beam builder, loads in, one device step,
packed node motion out) against the reference's BladeInterface regression tests
(tests/regression_tests/interfaces/test_blade_interface.cpp, at its 1e-10) and against the
oracle run on the same model (state at the 1e-12 / rates 1e-10 / 1e-9 bars)."""

import math

import numpy as np
import pytest

import blade_builder as BB
import openturbine_amd as ota
from golden_data import load
from oracle_lib import OracleLib, OracleSystem
from openturbine_amd import workloads as W

pytestmark = pytest.mark.gpu

MODES = [pytest.param(ota.EXEC_FUSED, id="fused"), pytest.param(ota.EXEC_PLAIN, id="plain")]
MOTION = {"position": 0, "displacement": 1, "velocity": 2, "acceleration": 3}
ORACLE_FIELD = ("x", "q", "v", "vd")
TOL = (1e-12, 1e-12, 1e-10, 1e-9)


@pytest.fixture(scope="module")
def kyn():
    return ota.KynLib()


@pytest.fixture(scope="module")
def orc():
    return OracleLib()


@pytest.fixture(scope="module")
def fx():
    return load("blade_interface.json")


def solution(d):
    return {"time_step": d["time_step"], "rho_inf": d["rho_inf"], "max_iter": int(d["max_iter"]),
            "atol": d["atol"], "rtol": d["rtol"]}


def root_disp(d, step):
    t = step * d["time_step"]
    u_rot = W.quat_from_rotvec(np.array(d["omega"]) * t)
    x0r = np.array(d["x0_root"])
    return np.concatenate([W.quat_rotate(u_rot, x0r) - x0r, u_rot])


def check_golden(bi, expect, tip):
    worst = 0.0
    for e in expect:
        node = int(e["expr"].split("nodes[")[1].split("]")[0]) if "nodes[" in e["expr"] else tip
        field = e["expr"].split(".")[-1].split("[")[0]
        comp = int(e["expr"].split("[")[-1].rstrip("]"))
        val = bi.motion(node)[MOTION[field]][comp]
        worst = max(worst, abs(val - e["value"]))
        assert abs(val - e["value"]) <= e["tol"], (e, val)
    return worst


def check_oracle(bi, o):
    for f, (name, tol) in enumerate(zip(ORACLE_FIELD, TOL)):
        ref = o.get(name)
        got = np.array([bi.motion(i)[f] for i in range(bi.n)])
        scale = max(np.max(np.abs(ref)), 1e-300)
        assert np.max(np.abs(got - ref)) / scale < tol, name


@pytest.mark.parametrize("mode", MODES)
def test_rotating_beam(kyn, orc, fx, mode):
    d = fx["rotating_beam"]
    bi = ota.KynBladeInterface(kyn, solution(d), BB.rotating_beam_kyn_input(d), exec_mode=mode)
    spec = BB.build_blade_spec(BB.rotating_beam_input(d), h=d["time_step"], rho_inf=d["rho_inf"],
                               max_iter=d["max_iter"], atol=d["atol"], rtol=d["rtol"])
    o = OracleSystem(orc, spec)
    for i in range(1, int(d["num_steps"]) + 1):
        u = root_disp(d, i)
        bi.set_root_displacement(u)
        assert bi.step() == ota.KYN_OK
        assert o.step({"constraint_disp": {0: u}}) == 0
    worst = check_golden(bi, d["expect"], bi.n - 1)
    print(f"rotating beam ({mode}): worst deviation from the reference {worst:.2e}", flush=True)
    check_oracle(bi, o)


@pytest.mark.parametrize("mode", MODES)
def test_blade_windio(kyn, orc, fx, mode):
    d = fx["windio"]
    bi = ota.KynBladeInterface(kyn, solution(d), BB.windio_kyn_input(d), exec_mode=mode)
    spec = BB.build_blade_spec(BB.windio_input(d), h=d["time_step"], rho_inf=d["rho_inf"], max_iter=d["max_iter"],
                               atol=d["atol"], rtol=d["rtol"])
    o = OracleSystem(orc, spec)
    tip = bi.n - 1
    for i in range(1, int(d["num_steps"]) + 1):
        t = i * d["time_step"]
        loads = np.zeros(6)
        loads[1] = d["tip_moment_amplitude"] * math.sin(2.0 * math.pi * d["tip_moment_frequency"] * t)
        bi.set_node_loads(tip, loads)
        f = np.zeros((spec.num_nodes, 6))
        f[tip] = loads
        assert bi.step() == ota.KYN_OK, i
        assert o.step({"nodal_loads": f}) == 0
    worst = check_golden(bi, d["expect"], tip)
    print(f"windIO blade ({mode}): worst deviation from the reference {worst:.2e}", flush=True)
    check_oracle(bi, o)


def test_save_restore_and_point_loads(kyn, fx):
    d = fx["rotating_beam"]
    bi = ota.KynBladeInterface(kyn, solution(d), BB.rotating_beam_kyn_input(d))
    bi.set_root_displacement(root_disp(d, 1))
    assert bi.step() == ota.KYN_OK
    saved = [m.copy() for m in bi.motion(bi.n - 1)]
    bi.save_state()
    bi.set_root_displacement(root_disp(d, 2))
    assert bi.step() == ota.KYN_OK
    after = [m.copy() for m in bi.motion(bi.n - 1)]
    bi.restore_state()  # node motion refreshed from the restored state
    for a, b in zip(saved, bi.motion(bi.n - 1)):
        np.testing.assert_array_equal(a, b)
    assert bi.step() == ota.KYN_OK  # the predictor-corrector pattern: same inputs, same step
    for a, b in zip(after, bi.motion(bi.n - 1)):
        np.testing.assert_array_equal(a, b)
    # Beam::AddPointLoad spreads a load with the node weights at s (beam.cpp:26-45)
    bi.clear_loads()
    load6 = np.array([1.0, -2.0, 3.0, 0.5, 0.0, -0.25])
    bi.add_point_load(0.3, load6)
    w = BB.lagrange_deriv_weights(2 * 0.3 - 1.0, BB.gll_points(int(d["element_order"])))
    for k in range(bi.n):
        np.testing.assert_allclose(bi.node_loads(k), w[k] * load6, rtol=1e-13, atol=1e-13)
    with pytest.raises(RuntimeError, match="Invalid position"):
        bi.add_point_load(1.5, load6)
