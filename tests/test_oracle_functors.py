"""The oracle's per-functor entry points against the reference's unit-test known answers
(tests/golden/functor_known_answers.json and integrate_known_answers.json, extracted as data from
tests/unit_tests/system/{beams,masses}/*.cpp and tests/unit_tests/state/*.cpp of the reference),
at the reference's own tolerances, plus two functors the reference has no unit test for:
the tangent operator (checked against a central difference of the rotation map) and the dense
LU (checked against numpy's solver)."""

import ctypes as C

import numpy as np
import pytest

from golden_data import load
from oracle_lib import OracleLib

DP = C.POINTER(C.c_double)
FUNCTORS = load("functor_known_answers.json")
INTEGRATE = load("integrate_known_answers.json")


@pytest.fixture(scope="module")
def lib():
    return OracleLib()


def arr(values, n=None):
    a = np.zeros(n if n is not None else len(values))
    a[: len(values)] = values
    return a


def ptr(a):
    return a.ctypes.data_as(DP)


def call(lib, name, *args):
    """Calls orc_fn_<name> with float scalars and numpy arrays (passed as double*)."""
    f = getattr(lib.L, "orc_fn_" + name)
    f.restype = None
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(ptr(a))
        elif isinstance(a, int):
            conv.append(C.c_size_t(a))
        else:
            conv.append(C.c_double(a))
    f(*conv)


def check(out, expected, tol):
    np.testing.assert_allclose(out, np.asarray(expected, dtype=np.float64), rtol=0, atol=tol)


def test_strain(lib):
    k = FUNCTORS["strain"]
    i = k["inputs"]
    out = np.zeros(6)
    call(lib, "strain", arr(i["x0_prime"]), arr(i["u_prime"]), arr(i["r"]), arr(i["r_prime"]), out)
    check(out, k["expected"]["strain"], 1e-15)


def test_Ouu_Puu_Quu(lib):
    for key, name in (("Ouu", "ouu"), ("Puu", "puu"), ("Quu", "quu")):
        k = FUNCTORS[key]
        i = k["inputs"]
        out = np.zeros(36)
        if key == "Ouu":
            call(lib, name, arr(i["Cuu"]), arr(i["x0pupSS"]), arr(i["M_tilde"]), arr(i["N_tilde"]), out)
        else:
            call(lib, name, arr(i["Cuu"]), arr(i["x0pupSS"]), arr(i["N_tilde"]), out)
        check(out, k["expected"][key], 1e-15)


def test_forces_and_temporary_variables(lib):
    k = FUNCTORS["force_FC"]
    out = np.zeros(6)
    call(lib, "fc", arr(k["inputs"]["Cuu"]), arr(k["inputs"]["strain"]), out)
    check(out, k["expected"]["FC"], 1e-15)
    k = FUNCTORS["force_FD"]
    call(lib, "fd", arr(k["inputs"]["x0pupSS"]), arr(k["inputs"]["FC"]), out)
    check(out, k["expected"]["FD"], 1e-15)
    k = FUNCTORS["temporary_variables"]
    out9 = np.zeros(9)
    call(lib, "temporary_variables", arr(k["inputs"]["x0_prime"]), arr(k["inputs"]["u_prime"]), out9)
    check(out9, k["expected"]["x0pupSS"], 1e-15)


def test_mass_functors(lib):
    k = FUNCTORS["mass_matrix_components"]
    eta, rho = np.zeros(3), np.zeros(9)
    call(lib, "eta_rho", arr(k["inputs"]["Muu"]), eta, rho)
    check(eta, k["expected"]["eta"], 1e-15)
    check(rho, k["expected"]["rho"], 1e-15)

    k = FUNCTORS["rotate_section_matrix"]
    out = np.zeros(36)
    call(lib, "rotate_section", arr(k["inputs"]["xr"]), arr(k["inputs"]["Cstar"]), out)
    check(out, k["expected"]["Cuu"], 1e-15)

    k = FUNCTORS["gravity_force"]
    i = k["inputs"]
    out6 = np.zeros(6)
    call(lib, "gravity_force", k["scalars"]["mass"], arr(i["gravity"]), arr(i["eta_tilde"]), out6)
    check(out6, k["expected"]["FG"], 1e-15)

    k = FUNCTORS["inertial_forces"]
    i = k["inputs"]
    call(lib, "inertial_force", k["scalars"]["mass"], arr(i["u_ddot"]), arr(i["omega"]), arr(i["omega_dot"]),
         arr(i["eta"]), arr(i["eta_tilde"]), arr(i["rho"]), arr(i["omega_tilde"]), arr(i["omega_dot_tilde"]), out6)
    check(out6, k["expected"]["FI"], 1e-15)

    k = FUNCTORS["gyroscopic_matrix"]
    i = k["inputs"]
    call(lib, "gyroscopic_matrix", k["scalars"]["mass"], arr(i["omega"]), arr(i["eta"]), arr(i["rho"]),
         arr(i["omega_tilde"]), out)
    check(out, k["expected"]["Guu"], 1e-15)

    k = FUNCTORS["inertia_stiffness_matrix"]
    i = k["inputs"]
    call(lib, "inertia_stiffness_matrix", k["scalars"]["mass"], arr(i["u_ddot"]), arr(i["omega"]),
         arr(i["omega_dot"]), arr(i["eta"]), arr(i["rho"]), arr(i["omega_tilde"]), arr(i["omega_dot_tilde"]), out)
    check(out, k["expected"]["Kuu"], 1e-15)


def test_state_kernels(lib):
    # the reference holds these at 1e-14 (EXPECT_NEAR), algorithmic acceleration exactly
    k = FUNCTORS["next_state"]
    s, i = k["scalars"], k["inputs"]
    qd, v, vd, a = (arr(i[n]) for n in ("q_delta", "v", "vd", "a"))
    call(lib, "next_state", s["h"], s["alpha_f"], s["alpha_m"], s["beta"], s["gamma"], qd, v, vd, a)
    for name, got in (("q_delta", qd), ("v", v), ("vd", vd), ("a", a)):
        check(got, k["expected"][name], 1e-14)

    k = FUNCTORS["displacement"]
    q = np.zeros(7)
    call(lib, "displacement", k["scalars"]["h"], arr(k["inputs"]["q_delta"]), arr(k["inputs"]["q_prev"]), q)
    check(q, k["expected"]["q"], 1e-14)

    k = FUNCTORS["dynamic_prediction"]
    s = k["scalars"]
    qd, v, vd = np.zeros(6), np.zeros(6), np.zeros(6)
    call(lib, "dynamic_prediction", s["h"], s["beta_prime"], s["gamma_prime"], arr(k["inputs"]["x_delta"]), qd, v, vd)
    for name, got in (("q_delta", qd), ("v", v), ("vd", vd)):
        check(got, k["expected"][name], 1e-14)

    k = FUNCTORS["static_prediction"]  # two nodes, six dofs each, in the reference's dof order
    x = arr(k["inputs"]["x_delta"])
    qd = np.zeros(12)
    for node in range(2):
        xn, qn = x[6 * node: 6 * node + 6].copy(), np.zeros(6)
        call(lib, "static_prediction", k["scalars"]["h"], xn, qn)
        qd[6 * node: 6 * node + 6] = qn
    check(qd, k["expected"]["q_delta"], 1e-14)

    k = FUNCTORS["algorithmic_acceleration"]
    acc = np.zeros(6)
    call(lib, "algorithmic_acceleration", k["scalars"]["alpha_f"], k["scalars"]["alpha_m"], arr(k["inputs"]["vd"]), acc)
    check(acc, k["expected"]["acceleration"], 0.0)

    k = FUNCTORS["global_position"]
    xg = np.zeros(7)
    call(lib, "global_position", arr(k["inputs"]["q"]), arr(k["inputs"]["x0"]), xg)
    check(xg, k["expected"]["x"], 1e-14)


def _integrate(lib, kind, case, args):
    """One reference functor call (positional arguments by name) through the oracle."""
    A, S = case["arrays"], case["scalars"]

    def val(name):
        return S[name] if name in S else float(name.rstrip("U").rstrip("."))

    N = int(val(args[1])) if kind != "integrate_residual" else int(S["number_of_nodes"])
    Q = int(val(args[2] if kind != "integrate_residual" else args[1]))

    def table(name):  # [N][Q] shape table: the first N*Q entries of the reference's row-major data
        return arr(A.get(name, [])[: N * Q], N * Q)

    def qp(name, width):
        return arr(A.get(name, []), Q * width)

    if kind == "integrate_stiffness":
        w, J, phi, dphi = qp(args[3], 1), qp(args[4], 1), table(args[5]), table(args[6])
        mats = [qp(n, 36) for n in args[7:12]]
        out = np.zeros(N * N * 36)
        call(lib, "integrate_stiffness", N, Q, w, J, phi, dphi, *mats, out)
        return out
    if kind == "integrate_inertia":
        w, J, phi = qp(args[3], 1), qp(args[4], 1), table(args[5])
        out = np.zeros(N * N * 36)
        call(lib, "integrate_inertia", N, Q, w, J, phi, val(args[8]), val(args[9]), qp(args[6], 36), qp(args[7], 36), out)
        return out
    w, J, phi, dphi = qp(args[2], 1), qp(args[3], 1), table(args[4]), table(args[5])
    FX = arr(A.get(args[6], []), N * 6)
    forces = [qp(n, 6) for n in args[7:12]]
    out = np.zeros(N * 6)
    call(lib, "integrate_residual", N, Q, w, J, phi, dphi, *forces, FX, out)
    return out


@pytest.mark.parametrize("kind,name", [(k, n) for k in INTEGRATE for n in INTEGRATE[k]])
def test_integrate_known_answers(lib, kind, name):
    case = INTEGRATE[kind][name]
    out = np.concatenate([_integrate(lib, kind, case, args) for args in case["calls"]])
    check(out, case["expected"], case["tolerance"])


def _rotmat(rv):
    rv = np.asarray(rv, dtype=np.float64)
    a = np.linalg.norm(rv)
    K = np.array([[0, -rv[2], rv[1]], [rv[2], 0, -rv[0]], [-rv[1], rv[0], 0]])
    if a < 1e-300:
        return np.eye(3)
    return np.eye(3) + np.sin(a) / a * K + (1 - np.cos(a)) / a**2 * K @ K


def test_tangent_operator_against_the_rotation_map(lib):
    """calculate_tangent_operator.hpp: T(psi) with psi = h q_delta_rot is the right Jacobian of
    the rotation-vector map, R(psi + d) = R(psi) exp(T(psi) d) + O(|d|^2); the oracle stores T^T.
    Checked column by column against a central difference of Rodrigues' formula; T(0) = I."""
    rng = np.random.default_rng(7)
    h = 0.01
    for trial in range(6):
        qd = np.zeros(6)
        qd[:3] = rng.normal(size=3)
        qd[3:] = rng.normal(size=3) * (0.0 if trial == 0 else 40.0 * trial)
        out = np.zeros(36)
        call(lib, "tangent_operator", h, qd, out)
        T = out.reshape(6, 6).T  # stored transposed
        np.testing.assert_array_equal(T[:3, :3], np.eye(3))
        assert np.all(T[:3, 3:] == 0) and np.all(T[3:, :3] == 0)
        psi = h * qd[3:]
        R0 = _rotmat(psi)
        eps = 1e-6
        for j in range(3):
            d = np.zeros(3)
            d[j] = eps
            dR = R0.T @ (_rotmat(psi + d) - _rotmat(psi - d)) / (2 * eps)
            col = np.array([dR[2, 1], dR[0, 2], dR[1, 0]])
            np.testing.assert_allclose(T[3:, 3 + j], col, atol=1e-8)


def test_dense_lu_against_numpy(lib):
    """The oracle's pivoted LU (the KLU stand-in) solves random and zero-diagonal systems like
    numpy.linalg.solve."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 33, 80):
        A = rng.normal(size=(n, n)) + n * np.eye(n) * (0.0 if n == 7 else 0.1)
        if n == 7:
            A[np.arange(n), np.arange(n)] = 0.0  # every diagonal entry zero: pivoting required
        b = rng.normal(size=n)
        Ac, bc = A.copy(), b.copy()
        f = lib.L.orc_fn_lu_solve
        f.restype = C.c_int
        rc = f(C.c_size_t(n), ptr(Ac), ptr(bc))
        assert rc == 0
        np.testing.assert_allclose(bc, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


# ---- curved beam (reference unit tests test_curved_beam.cpp with test_curved_beam.hpp: a 3-node
#      second-order element on GLL nodes, 7-point Gauss rule, values from the reference's
#      Mathematica notebooks), at the reference's tolerances
CB = load("curved_beam.json")
CB_N, CB_Q = 3, 7


def test_curved_beam_shape_functions_and_jacobian(lib):
    f = lib.L.orc_fn_lagrange
    f.restype = None
    xs = arr(CB["kGLLNodes"])
    phi, dphi = np.zeros((CB_Q, CB_N)), np.zeros((CB_Q, CB_N))
    for q, x in enumerate(CB["kGaussQuadraturePoints"]):
        w, dw = np.zeros(CB_N), np.zeros(CB_N)
        f(C.c_double(x), C.c_size_t(CB_N), ptr(xs), ptr(w), ptr(dw))
        phi[q], dphi[q] = w, dw
    check(phi.ravel(), CB["kExpectedInterpWeights"], 1e-12)
    check(dphi.ravel(), CB["kExpectedDerivWeights"], 1e-12)
    J, x0p = np.zeros(CB_Q), np.zeros(CB_Q * 3)
    call(lib, "jacobian", CB_N, CB_Q, arr(CB["kCurvedBeamNodes_data"]), arr(CB["kDerivWeightsFlat"][: CB_N * CB_Q]), J, x0p)
    check(J, CB["kExpectedJacobians"], 1e-12)
    np.testing.assert_allclose(np.linalg.norm(x0p.reshape(CB_Q, 3), axis=1), 1.0, atol=1e-12)


def test_curved_beam_section_terms(lib):
    out6, out36 = np.zeros(6), np.zeros(36)
    call(lib, "fc", arr(CB["kCurvedBeamCuu_data"]), arr(CB["kCurvedBeamStrain_data"]), out6)
    check(out6, CB["kExpectedFc_data"], 1e-6)
    call(lib, "fd", arr(CB["kX0pupSS_data"]), arr(CB["kExpectedFc_data"]), out6)
    check(out6, CB["kExpectedFd_data"], 1e-6)
    call(lib, "rotate_section", arr(CB["xr_data"]), arr(CB["Cstar_data"]), out36)
    check(out36, CB["kExpectedCuu_data"], 1e-6)
    Cuu, xs = arr(CB["kCurvedBeamCuu_data"]), arr(CB["kX0pupSS_data"])
    Mt, Nt = arr(CB["kCurvedBeamM_tilde_data"]), arr(CB["kCurvedBeamN_tilde_data"])
    call(lib, "ouu", Cuu, xs, Mt, Nt, out36)
    check(out36, CB["kExpectedOuu_data"], 1e-6)
    call(lib, "puu", Cuu, xs, Nt, out36)
    check(out36, CB["kExpectedPuu_data"], 1e-6)
    call(lib, "quu", Cuu, xs, Nt, out36)
    check(out36, CB["kExpectedQuu_data"], 1e-6)


def test_curved_beam_residual_and_stiffness_integration(lib):
    w, J = arr(CB["kGaussQuadratureWeights"]), arr(CB["kExpectedJacobians"])
    phi, dphi = arr(CB["kInterpWeightsFlat"][: CB_N * CB_Q]), arr(CB["kDerivWeightsFlat"][: CB_N * CB_Q])
    zero6 = np.zeros(CB_Q * 6)
    res = np.zeros(CB_N * 6)
    call(lib, "integrate_residual", CB_N, CB_Q, w, J, phi, dphi, arr(CB["kFc"]), arr(CB["kFd"]), arr(CB["kFi"]),
         zero6, zero6, np.zeros(CB_N * 6), res)
    check(res, CB["kExpectedResidualVector"], 1e-8)
    K = np.zeros(CB_N * CB_N * 36)
    call(lib, "integrate_stiffness", CB_N, CB_Q, w, J, phi, dphi, arr(CB["kKuu"]), arr(CB["kPuu"]), arr(CB["kCuu"]),
         arr(CB["kOuu"]), arr(CB["kQuu"]), K)
    check(K, CB["kExpectedStiffnessMatrix_data"], 1e-7)
