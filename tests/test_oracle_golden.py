"""The CPU oracle against the reference's own golden vectors (CPU-only, no GPU).

Fixtures come from the reference's test sources (tests/golden/extract_reference_vectors.py):
  test_solver.cpp         assembled b, solution x, convergence_err[0], state after one step
  test_cantilever_beam.cpp tip displacement over 3 steps
  test_rotor.cpp          IEA-15 3-blade rotor, 11 steps, node q to 1e-13
  test_calculate_strain.cpp strain known answer
Tolerances are the reference's own (1e-6 absolute by default, test_utilities.hpp:21-35;
1e-13 for the rotor, test_rotor.cpp:140-188).
"""

import numpy as np
import pytest

from golden_data import B_VIEW, X_VIEW, load, solver_case
from openturbine_amd import workloads as W
from oracle_lib import OracleLib, OracleSystem


@pytest.fixture(scope="module")
def lib():
    return OracleLib()


def test_solver_assembly_matches_reference(lib):
    d, exp = solver_case("SetUpSolverAndAssemble")
    spec = W.solver_regression_spec(d["inputs"])
    s = OracleSystem(lib, spec)
    s.debug_assemble(spec.step_inputs(0))
    np.testing.assert_allclose(s.rhs(), exp[B_VIEW], atol=1e-6, rtol=0)
    np.testing.assert_allclose(s.get("q").ravel(), exp["state.q"], atol=1e-6)
    np.testing.assert_allclose(s.get("q_delta").ravel(), exp["state.q_delta"], atol=1e-6)
    np.testing.assert_allclose(s.get("v").ravel(), exp["state.v"], atol=1e-6)
    np.testing.assert_allclose(s.get("vd").ravel(), exp["state.vd"], atol=1e-6)
    np.testing.assert_allclose(s.get("a").ravel(), exp["state.a"], atol=1e-6)
    np.testing.assert_allclose(s.get("q_prev").ravel(), exp["state.q_prev"], atol=1e-6)
    np.testing.assert_allclose(s.constraint(0, 1), exp["constraints.residual_terms"], atol=1e-6)
    np.testing.assert_allclose(s.constraint(0, 0), exp["constraints.lambda"], atol=1e-6)
    # much tighter than the reference's tolerance: the dominant entries agree to ~1e-10
    assert np.max(np.abs(s.rhs() - exp[B_VIEW])) < 1e-9


def test_solver_one_iteration_matches_reference(lib):
    d, exp = solver_case("SetupAndTakeNoSteps")
    s = OracleSystem(lib, W.solver_regression_spec(d["inputs"]))
    assert s.step() == 1  # max_iter = 1: Step returns false after one iteration
    np.testing.assert_allclose(s.convergence()[0], d["convergence_err0"], atol=1e-6, rtol=0)
    np.testing.assert_allclose(s.solution(), exp[X_VIEW], atol=1e-6)
    assert np.max(np.abs(s.solution() - exp[X_VIEW])) < 1e-10
    for f in ("q_delta", "v", "vd", "q"):
        np.testing.assert_allclose(s.get(f).ravel(), exp["state." + f], atol=1e-6)
    np.testing.assert_allclose(s.constraint(0, 0), exp["constraints.lambda"], atol=1e-6)


def test_solver_two_iterations_matches_reference(lib):
    d, exp = solver_case("SetupAndTakeTwoSteps")
    s = OracleSystem(lib, W.solver_regression_spec(d["inputs"]))
    s.step()
    np.testing.assert_allclose(s.constraint(0, 1), exp["constraints.residual_terms"], atol=1e-6)
    np.testing.assert_allclose(s.rhs(), exp[B_VIEW], atol=1e-6)
    np.testing.assert_allclose(s.solution(), exp[X_VIEW], atol=1e-6)


def test_cantilever_sine_load_matches_reference(lib):
    fx = load("cantilever_sine.json")
    s = OracleSystem(lib, W.cantilever_sine_spec(fx))
    for k in range(3):
        assert s.step() == 0
        np.testing.assert_allclose(s.get("q")[fx["tip_node"], :3], fx["tip_q_translation"][k], atol=1e-6)


def test_iea15_rotor_matches_reference(lib):
    fx = load("iea15_rotor.json")
    spec = W.iea15_rotor_spec(fx)
    s = OracleSystem(lib, spec)
    for _ in range(spec.num_steps):
        assert s.step() == 0
    q = s.get("q")
    for ch in fx["checks"]:
        ref = q[ch["equals"][0], ch["equals"][1]] if "equals" in ch else ch["value"]
        assert abs(q[ch["node"], ch["comp"]] - ref) <= ch["tol"], ch


def test_strain_known_answer(lib):
    import ctypes as C

    k = load("functor_known_answers.json")["strain"]
    dp = C.POINTER(C.c_double)
    arrs = [np.array(k["inputs"][n], dtype=np.float64) for n in ("x0_prime", "u_prime", "r", "r_prime")]
    out = np.zeros(6)
    lib.L.orc_fn_strain(*[a.ctypes.data_as(dp) for a in arrs], out.ctypes.data_as(dp))
    np.testing.assert_allclose(out, k["expected"]["strain"], atol=1e-15)
