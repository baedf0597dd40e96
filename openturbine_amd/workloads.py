"""Synthetic and reference-test workloads for the beam Newton path.

Each builder returns a ``ModelSpec``: plain numpy arrays describing nodes, beam elements,
sections, quadrature and constraints, plus the step parameters and a per-step driver that
produces the time-dependent inputs (prescribed root motion, nodal loads).  ``build_model``
feeds a spec through any library exposing the model entry points of include/kynema_amd.h
(``kyn_model_*``), so the same inputs reach every consumer bit for bit.

Random inputs use a counter-based SplitMix64 (seed 20261202 by default): the value for key
``k`` is the SplitMix64 finaliser applied to ``seed + (k + 1) * 0x9E3779B97F4A7C15``;
uniform doubles take the top 53 bits.  No std::*_distribution is involved, so every
consumer sees the same numbers.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

SEED = 20261202
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)

# Constraint type codes (src/constraints/constraint_type.hpp:9-21)
FIXED_BC = 1
PRESCRIBED_BC = 2
RIGID_JOINT = 3
REVOLUTE_JOINT = 4
ROTATION_CONTROL = 5
FIXED_BC_3DOFS = 6
PRESCRIBED_BC_3DOFS = 7
RIGID_JOINT_6DOFS_TO_3DOFS = 8


# --------------------------------------------------------------------------------------
# random numbers
# --------------------------------------------------------------------------------------
def splitmix64(seed: int, keys) -> np.ndarray:
    k = np.asarray(keys, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (k + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, keys) -> np.ndarray:
    return (splitmix64(seed, keys) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, keys) -> np.ndarray:
    """Box-Muller on two SplitMix64 draws per key (keys 2k and 2k+1)."""
    k = np.asarray(keys, dtype=np.uint64)
    u1 = uniform01(seed, k * np.uint64(2))
    u2 = uniform01(seed, k * np.uint64(2) + np.uint64(1))
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# --------------------------------------------------------------------------------------
# quadrature / node locations (input generation only)
# --------------------------------------------------------------------------------------
def gll_points(order: int) -> np.ndarray:
    """Gauss-Lobatto-Legendre points on [-1, 1] (order+1 points)."""
    if order == 1:
        return np.array([-1.0, 1.0])
    c = np.zeros(order + 1)
    c[-1] = 1.0
    interior = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    pts = np.concatenate([[-1.0], np.sort(interior), [1.0]])
    # polish interior points with Newton on P'_order
    d1 = np.polynomial.legendre.legder(c)
    d2 = np.polynomial.legendre.legder(d1)
    for _ in range(3):
        pts[1:-1] -= np.polynomial.legendre.legval(pts[1:-1], d1) / np.polynomial.legendre.legval(pts[1:-1], d2)
    return pts


def gauss_legendre(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


# --------------------------------------------------------------------------------------
# quaternions (input generation only)
# --------------------------------------------------------------------------------------
def quat_from_rotvec(rv) -> np.ndarray:
    rv = np.asarray(rv, dtype=np.float64)
    a = float(np.sqrt(rv @ rv))
    if abs(a) < 1e-12:
        return np.array([1.0, 0.0, 0.0, 0.0])
    s = math.sin(a / 2.0) / a
    return np.array([math.cos(a / 2.0), rv[0] * s, rv[1] * s, rv[2] * s])


def rotvec_from_quat(q) -> np.ndarray:
    """Rotation vector (axis times angle in [0, pi]) of a unit quaternion {w, x, y, z}."""
    q = np.asarray(q, dtype=np.float64)
    if q[0] < 0.0:
        q = -q
    s = float(np.sqrt(q[1:] @ q[1:]))
    if s < 1e-300:
        return 2.0 * q[1:]
    return (2.0 * math.atan2(s, q[0]) / s) * q[1:]


def quat_compose(a, b) -> np.ndarray:
    return np.array([
        a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
        a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
        a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
        a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0],
    ])


def quat_rotate(q, v) -> np.ndarray:
    q0, q1, q2, q3 = q
    R = np.array([
        [q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
        [2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1)],
        [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3],
    ])
    return R @ np.asarray(v, dtype=np.float64)


# --------------------------------------------------------------------------------------
# model spec
# --------------------------------------------------------------------------------------
@dataclass
class BeamSpec:
    nodes: np.ndarray          # [N] node ids
    section_s: np.ndarray      # [S]
    section_M: np.ndarray      # [S, 36]
    section_C: np.ndarray      # [S, 36]
    quad: np.ndarray           # [Q, 2] (xi, w)


@dataclass
class ConstraintSpec:
    type: int
    base: int
    target: int
    axis: np.ndarray = field(default_factory=lambda: np.zeros(3))
    init: np.ndarray = field(default_factory=lambda: np.array([0.0, 0, 0, 1, 0, 0, 0]))


@dataclass
class MassSpec:
    node: int
    M: np.ndarray              # [36] mass matrix in the material frame


@dataclass
class SpringSpec:
    n1: int
    n2: int
    k: float
    l0: float


@dataclass
class ModelSpec:
    name: str
    x0: np.ndarray             # [n, 7]
    u: np.ndarray              # [n, 7]
    v: np.ndarray              # [n, 6]
    vd: np.ndarray             # [n, 6]
    s: np.ndarray              # [n]
    beams: list
    constraints: list
    gravity: np.ndarray
    is_dynamic: bool = True
    max_iter: int = 10
    h: float = 0.01
    rho_inf: float = 0.9
    atol: float = 1e-5
    rtol: float = 1e-3
    num_steps: int = 1
    # step_inputs(step_index, read_state) -> dict with optional keys
    #   "constraint_disp": {constraint_id: [7]}, "nodal_loads": [n, 6] array,
    #   "element_node_loads": {elem: [N, 6]}, "control": {constraint_id: value}
    step_inputs: Optional[Callable] = None
    meta: dict = field(default_factory=dict)
    masses: list = field(default_factory=list)
    springs: list = field(default_factory=list)

    @property
    def num_nodes(self) -> int:
        return int(self.x0.shape[0])


class _NodeTable:
    def __init__(self):
        self.x0, self.u, self.v, self.vd, self.s = [], [], [], [], []

    def add(self, x0, u=None, v=None, vd=None, s=0.0) -> int:
        self.x0.append(np.asarray(x0, dtype=np.float64))
        self.u.append(np.asarray(u if u is not None else [0, 0, 0, 1, 0, 0, 0], dtype=np.float64))
        self.v.append(np.asarray(v if v is not None else np.zeros(6), dtype=np.float64))
        self.vd.append(np.asarray(vd if vd is not None else np.zeros(6), dtype=np.float64))
        self.s.append(float(s))
        return len(self.x0) - 1

    def arrays(self):
        return (np.array(self.x0).reshape(-1, 7), np.array(self.u).reshape(-1, 7),
                np.array(self.v).reshape(-1, 6), np.array(self.vd).reshape(-1, 6), np.array(self.s))


def build_model(api, spec: ModelSpec):
    """Feed a spec through a library's model entry points; returns the model handle."""
    m = api.model_create()
    api.model_set_gravity(m, spec.gravity)
    for i in range(spec.num_nodes):
        api.model_add_node(m, spec.x0[i], spec.u[i], spec.v[i], spec.vd[i], spec.s[i])
    for b in spec.beams:
        api.model_add_beam_element(m, b.nodes, b.section_s, b.section_M, b.section_C, b.quad)
    for e in spec.masses:
        api.model_add_mass_element(m, e.node, e.M)
    for e in spec.springs:
        api.model_add_spring_element(m, e.n1, e.n2, e.k, e.l0)
    for c in spec.constraints:
        api.model_add_constraint(m, c.type, c.base, c.target, c.axis, c.init)
    return m


# --------------------------------------------------------------------------------------
# reference point-mass / spring regression tests (inputs in tests/golden/masses_springs.json)
# --------------------------------------------------------------------------------------
def masses_external_force_spec(fx: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_masses.cpp (MassesTest.ExternalForce): one free
    point mass, constant force and torque, rho_inf = 0."""
    nt = _NodeTable()
    nt.add([0, 0, 0, 1, 0, 0, 0])
    x0, u, v, vd, s = nt.arrays()
    f = np.zeros((1, 6))
    f[0, 0] = fx["force_x"]
    f[0, 4] = fx["torque_y"]
    max_iter, rho = fx["step_parameters"]

    def step_inputs(step, _state=None):
        return {"nodal_loads": f}

    return ModelSpec("masses_external_force", x0, u, v, vd, s, [], [], np.zeros(3), max_iter=int(max_iter),
                     h=fx["time_step"], rho_inf=rho, num_steps=int(fx["n_steps"]), step_inputs=step_inputs,
                     masses=[MassSpec(0, np.array(fx["mass_matrix"]))])


def precession_spec(fx: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_precession.cpp: a free spinning point mass."""
    nt = _NodeTable()
    nt.add([0, 0, 0, 1, 0, 0, 0], v=fx["velocity"])
    x0, u, v, vd, s = nt.arrays()
    return ModelSpec("precession", x0, u, v, vd, s, [], [], np.zeros(3), max_iter=int(fx["max_iter"]),
                     h=fx["step_size"], rho_inf=fx["rho_inf"], num_steps=int(fx["num_steps"]),
                     masses=[MassSpec(0, np.array(fx["mass_matrix"]))])


def spring_mass_spec(fx: dict, anchor_type: int = FIXED_BC) -> ModelSpec:
    """tests/regression_tests/regression/test_spring_mass_system.cpp (SpringMassSystemTest): a
    point mass on a spring to a clamped node, released 2 m from equilibrium; h = T / steps.
    anchor_type FIXED_BC_3DOFS holds the anchor, which only the spring touches, as a three-dof
    node (the same motion of the mass)."""
    nt = _NodeTable()
    for p in fx["positions"]:
        nt.add(p)
    x0, u, v, vd, s = nt.arrays()
    m, j, k = fx["m"], fx["j"], fx["k"]
    period = 2.0 * np.pi * np.sqrt(m / k)
    M = _diag6([m, m, m, j, j, j])
    return ModelSpec("spring_mass", x0, u, v, vd, s, [], [ConstraintSpec(anchor_type, 0, 0)], np.zeros(3),
                     max_iter=int(fx["max_iter"]), h=period / fx["num_steps"], rho_inf=fx["rho_inf"],
                     num_steps=int(fx["num_steps"]), masses=[MassSpec(1, M)],
                     springs=[SpringSpec(0, 1, k, fx["l0"])])


def spring_mass_chain_spec(fx: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_spring_mass_system.cpp (SpringMassChainSystemTest):
    equidistant point masses joined by zero-length springs between two clamped anchors, in
    static equilibrium for one period of a single mass on one spring."""
    nt = _NodeTable()
    nm = int(fx["number_of_masses"])
    for i in range(nm + 2):
        nt.add([i * fx["spacing"], 0, 0, 1, 0, 0, 0])
    x0, u, v, vd, s = nt.arrays()
    m, j, k = fx["m"], fx["j"], fx["k"]
    M = _diag6([m, m, m, j, j, j])
    return ModelSpec("spring_mass_chain", x0, u, v, vd, s, [],
                     [ConstraintSpec(FIXED_BC, 0, 0), ConstraintSpec(FIXED_BC, 0, nm + 1)], np.zeros(3),
                     max_iter=int(fx["max_iter"]), h=2.0 * np.pi * np.sqrt(m / k) / fx["num_steps"],
                     rho_inf=fx["rho_inf"], num_steps=int(fx["num_steps"]),
                     masses=[MassSpec(i + 1, M) for i in range(nm)],
                     springs=[SpringSpec(i, i + 1, k, fx["l0"]) for i in range(nm + 1)])


def heavy_top_spec(fx: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_heavy_top.cpp: a spinning top (point mass with
    inertia) held at a fixed pivot by a rigid joint to a three-dof ground node under gravity."""
    nt = _NodeTable()
    x = np.array(fx["x"])
    om = np.array(fx["omega"])
    nt.add(np.concatenate([x, [1, 0, 0, 0]]), v=np.concatenate([np.cross(om, x), om]),
           vd=np.concatenate([fx["x_ddot"], fx["omega_dot"]]))
    ground = nt.add([0, 0, 0, 1, 0, 0, 0])
    x0, u, v, vd, s = nt.arrays()
    M = _diag6([fx["mass"]] * 3 + list(fx["inertia"]))
    cons = [ConstraintSpec(RIGID_JOINT_6DOFS_TO_3DOFS, 0, ground), ConstraintSpec(PRESCRIBED_BC_3DOFS, 0, ground)]
    return ModelSpec("heavy_top", x0, u, v, vd, s, [], cons, np.array(fx["gravity"]), max_iter=int(fx["max_iter"]),
                     h=fx["step_size"], rho_inf=fx["rho_inf"], atol=fx["atol"], rtol=fx["rtol"],
                     num_steps=int(fx["num_steps"]), masses=[MassSpec(0, M)])


def _diag6(vals):
    M = np.zeros((6, 6))
    np.fill_diagonal(M, vals)
    return M.reshape(36)


# --------------------------------------------------------------------------------------
# reference regression tests (inputs transcribed into tests/golden fixtures)
# --------------------------------------------------------------------------------------
def solver_regression_spec(inp: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_solver.cpp:18-100 (one 6-node element,
    prescribed root, 0.1 rad/s about z)."""
    nt = _NodeTable()
    omega = inp["omega"]
    off = inp["x_root_offset"]
    ids = []
    for s in inp["node_s"]:
        x = 10.0 * s + off
        ids.append(nt.add([x, 0, 0, 1, 0, 0, 0], v=[0, x * omega, 0, 0, 0, omega], s=s))
    x0, u, v, vd, s = nt.arrays()
    beam = BeamSpec(np.array(ids), np.array([0.0, 1.0]), np.array([inp["mass_matrix"]] * 2),
                    np.array([inp["stiffness_matrix"]] * 2), np.array(inp["quadrature"]).reshape(-1, 2))
    cons = [ConstraintSpec(PRESCRIBED_BC, 0, ids[0])]
    h = inp["step_size"]
    x0_root = np.array([off, 0.0, 0.0])

    def step_inputs(step, _state=None):
        q = quat_from_rotvec([0.0, 0.0, omega * h])
        if inp["bc_translation_follows_rotation"]:
            xr = quat_rotate(q, x0_root)
            ut = xr - x0_root
        else:
            ut = np.zeros(3)
        return {"constraint_disp": {0: np.concatenate([ut, q])}}

    return ModelSpec("solver_regression", x0, u, v, vd, s, [beam], cons, np.zeros(3),
                     max_iter=int(inp["max_iter"]), h=h, rho_inf=inp["rho_inf"],
                     step_inputs=step_inputs)


def cantilever_sine_spec(fx: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_cantilever_beam.cpp:31-138."""
    nt = _NodeTable()
    ids = [nt.add([10.0 * s, 0, 0, 1, 0, 0, 0], s=s) for s in fx["node_s"]]
    x0, u, v, vd, s = nt.arrays()
    beam = BeamSpec(np.array(ids), np.array([0.0, 1.0]), np.array([fx["mass_matrix"]] * 2),
                    np.array([fx["stiffness_matrix"]] * 2), np.array(fx["quadrature"]).reshape(-1, 2))
    cons = [ConstraintSpec(FIXED_BC, 0, ids[0])]
    h = fx["step_size"]
    n_nodes = len(ids)

    def step_inputs(step, _state=None):
        fxl = np.zeros((n_nodes, 6))
        fxl[fx["tip_node"], 2] = 100.0 * math.sin(10.0 * h * (step + 1))
        return {"element_node_loads": {0: fxl}}

    return ModelSpec("cantilever_sine", x0, u, v, vd, s, [beam], cons, np.zeros(3),
                     max_iter=int(fx["max_iter"]), h=h, rho_inf=fx["rho_inf"], num_steps=3,
                     step_inputs=step_inputs)


def iea15_rotor_spec(fx: dict) -> ModelSpec:
    """tests/regression_tests/regression/test_rotor.cpp:25-135 (IEA-15 3-blade rotor)."""
    nt = _NodeTable()
    node_loc = [(xi + 1.0) / 2.0 for xi in fx["node_xi"]]
    coords = np.array(fx["node_coords"]).reshape(-1, 7)
    omega = np.array(fx["omega"])
    hub = np.array(fx["hub_radius"])
    nb = fx["num_blades"]
    sec_s = np.array([sec["s"] for sec in fx["sections"]])
    sec_M = np.array([sec["M"] for sec in fx["sections"]])
    sec_C = np.array([sec["C"] for sec in fx["sections"]])
    quad = np.array(fx["quadrature"]).reshape(-1, 2)
    beams, roots = [], []
    for b in range(nb):
        q_rot = quat_from_rotvec([0.0, 0.0, -2.0 * math.pi * b / nb])
        ids = []
        for j, sl in enumerate(node_loc):
            p = coords[j, :3] + hub                      # Model::TranslateBeam
            p = quat_rotate(q_rot, p)                    # Node::RotateAboutPoint (origin)
            r = quat_compose(q_rot, coords[j, 3:])
            vel = np.concatenate([np.cross(omega, p), omega])  # Node::SetVelocityAboutPoint
            ids.append(nt.add(np.concatenate([p, r]), v=vel, s=sl))
        beams.append(BeamSpec(np.array(ids), sec_s, sec_M, sec_C, quad))
        roots.append(ids[0])
    x0, u, v, vd, s = nt.arrays()
    cons = [ConstraintSpec(PRESCRIBED_BC, 0, r) for r in roots]
    h = fx["step_size"]
    num_steps = int(fx["t_end"] / h + 1.0)

    def step_inputs(step, _state=None):
        q_hub = quat_from_rotvec(omega * h * (step + 1))
        d = np.concatenate([np.zeros(3), q_hub])
        return {"constraint_disp": {i: d for i in range(nb)}}

    return ModelSpec("iea15_rotor", x0, u, v, vd, s, beams, cons, np.array(fx["gravity"]),
                     max_iter=int(fx["max_iter"]), h=h, rho_inf=fx["rho_inf"], num_steps=num_steps,
                     step_inputs=step_inputs)


# --------------------------------------------------------------------------------------
# synthetic workloads
# --------------------------------------------------------------------------------------
C1_STIFF = np.array([1e9, 4e8, 4e8, 1e7, 2e7, 2e7])
C1_MASS = np.array([10.0, 10.0, 10.0, 0.5, 0.25, 0.25])


def cantilever_config_spec(num_nodes=8, num_qps=11, length=10.0, tip_force=1e5, ramp=0.1,
                           h=0.005, num_steps=200, rho_inf=0.9, max_iter=50, atol=1e-10, rtol=1e-6):
    """Straight cantilever along x: one element, GLL nodes, Gauss quadrature, clamped root,
    tip follower force in local z ramped over `ramp` seconds (re-oriented once per step
    from the start-of-step tip rotation)."""
    nt = _NodeTable()
    xi = gll_points(num_nodes - 1)
    ids = [nt.add([length * (x + 1) / 2, 0, 0, 1, 0, 0, 0], s=(x + 1) / 2) for x in xi]
    x0, u, v, vd, s = nt.arrays()
    gx, gw = gauss_legendre(num_qps)
    beam = BeamSpec(np.array(ids), np.array([0.0, 1.0]), np.array([_diag6(C1_MASS)] * 2),
                    np.array([_diag6(C1_STIFF)] * 2), np.stack([gx, gw], axis=1))
    cons = [ConstraintSpec(FIXED_BC, 0, ids[0])]
    tip = ids[-1]
    n = len(ids)

    def step_inputs(step, read_state=None):
        t = h * (step + 1)
        mag = tip_force * min(t / ramp, 1.0)
        q_tip = read_state("q")[tip, 3:] if read_state is not None else np.array([1.0, 0, 0, 0])
        f = np.zeros((n, 6))
        f[tip, :3] = quat_rotate(q_tip, [0.0, 0.0, mag])
        return {"nodal_loads": f}

    return ModelSpec("cantilever", x0, u, v, vd, s, [beam], cons, np.zeros(3), max_iter=max_iter,
                     h=h, rho_inf=rho_inf, atol=atol, rtol=rtol, num_steps=num_steps,
                     step_inputs=step_inputs, meta={"tip_node": tip})


def beam_mass_spring_spec(num_steps=20, h=0.005, rho_inf=0.9, max_iter=50):
    """The cantilever above (tip follower force, clamped root) with a point mass at the tip and a
    second, offset point mass hung from the tip by a spring: node `extra` sits 1 m below the tip
    with a rotated reference frame and a mass matrix whose centre of mass is off the node
    (nonzero eta), under gravity, so every term of the mass and spring elements is exercised in
    one KKT block with the beam."""
    base = cantilever_config_spec(num_steps=num_steps, h=h, rho_inf=rho_inf, max_iter=max_iter)
    tip = base.meta["tip_node"]
    n0 = base.num_nodes
    half = 0.3
    r_extra = np.array([np.cos(half), np.sin(half) * 0.6, 0.0, np.sin(half) * 0.8])
    x0 = np.vstack([base.x0, np.concatenate([base.x0[tip, :3] + [0.0, 0.0, -1.0], r_extra])])
    u = np.vstack([base.u, [0, 0, 0, 1, 0, 0, 0]])
    v = np.vstack([base.v, np.zeros(6)])
    vd = np.vstack([base.vd, np.zeros(6)])
    s = np.concatenate([base.s, [1.0]])

    def rigid_mass(m, eta, J):
        et = np.array([[0, -eta[2], eta[1]], [eta[2], 0, -eta[0]], [-eta[1], eta[0], 0]])
        M = np.zeros((6, 6))
        M[:3, :3] = m * np.eye(3)
        M[3:, :3] = m * et
        M[:3, 3:] = (m * et).T
        M[3:, 3:] = J
        return M.reshape(36)

    J = np.array([[4.0, 0.3, -0.2], [0.3, 3.0, 0.1], [-0.2, 0.1, 5.0]])
    masses = [MassSpec(tip, rigid_mass(50.0, [0.0, 0.0, 0.0], np.eye(3) * 2.0)),
              MassSpec(n0, rigid_mass(120.0, [0.05, -0.02, 0.1], J))]
    springs = [SpringSpec(tip, n0, 4.0e5, 0.8)]

    def step_inputs(step, read_state=None):
        f = base.step_inputs(step, read_state)["nodal_loads"]
        return {"nodal_loads": np.vstack([f, np.zeros(6)])}

    return ModelSpec("beam_mass_spring", x0, u, v, vd, s, base.beams, base.constraints,
                     np.array([0.0, 0.0, -9.81]), max_iter=max_iter, h=h, rho_inf=rho_inf, atol=base.atol,
                     rtol=base.rtol, num_steps=num_steps, step_inputs=step_inputs,
                     meta={"tip_node": tip, "extra_node": n0}, masses=masses, springs=springs)


def _spd_project(A: np.ndarray) -> np.ndarray:
    """Symmetric part, eigenvalues clipped to >= 1e-6 * lambda_max (stacked [..., 6, 6])."""
    A = 0.5 * (A + np.swapaxes(A, -1, -2))
    w, V = np.linalg.eigh(A)
    w = np.maximum(w, 1e-6 * w.max(axis=-1, keepdims=True))
    B = (V * w[..., None, :]) @ np.swapaxes(V, -1, -2)
    return 0.5 * (B + np.swapaxes(B, -1, -2))


SCALE_EXPONENTS = np.array([2.0, 2.0, 2.0, 4.0, 4.0, 4.0])


def section_properties(seed: int, blade_key: int, num_stations: int = 21, size_factor: float = 1.0):
    """Tapered, noisy, coupled SPD sections: diagonal = base * 20^-s * exp(U(-ln1.1, ln1.1)),
    off-diagonal = 0.05 * sqrt(C_ii C_jj) * U(-1, 1), projected to SPD.
    Random key of station i, matrix m (0 mass, 1 stiffness), entry k:
    (blade_key * num_stations * 128) + i * 128 + m * 64 + k (k < 6 diagonal, 6..20 upper).
    size_factor s scales the base section geometrically: mass / axial / shear entries by s^2,
    rotary inertias / torsion / bending by s^4."""
    s = np.linspace(0.0, 1.0, num_stations)
    iu = np.triu_indices(6, 1)
    base_key = np.uint64(blade_key) * np.uint64(num_stations * 128)
    st = np.arange(num_stations, dtype=np.uint64) * np.uint64(128)
    out = []
    for mat_id, base in enumerate((C1_MASS * size_factor**SCALE_EXPONENTS, C1_STIFF * size_factor**SCALE_EXPONENTS)):
        keys = base_key + st[:, None] + np.uint64(mat_id * 64) + np.arange(21, dtype=np.uint64)[None, :]
        u = uniform01(seed, keys)                                        # [S, 21]
        d = base[None, :] * (20.0 ** (-s))[:, None] * np.exp((2.0 * u[:, :6] - 1.0) * math.log(1.1))
        A = np.zeros((num_stations, 6, 6))
        A[:, np.arange(6), np.arange(6)] = d
        off = 0.05 * np.sqrt(d[:, iu[0]] * d[:, iu[1]]) * (2.0 * u[:, 6:21] - 1.0)
        A[:, iu[0], iu[1]] = off
        A[:, iu[1], iu[0]] = off
        out.append(_spd_project(A).reshape(num_stations, 36))
    return s, out[0], out[1]


def chain_section_properties(seed: int, chain_key: int, element: int, num_elements: int, num_stations: int = 5,
                             size_factor: float = 1.0):
    """Sections of element `element` of a chain of `num_elements`: the generator above evaluated
    on the chain coordinate (the 20x taper runs root to tip over the whole chain, noise keyed by
    chain and station), returned at element-local positions s in [0, 1]."""
    s_loc = np.linspace(0.0, 1.0, num_stations)
    s_chain = (element + s_loc) / num_elements
    iu = np.triu_indices(6, 1)
    out = []
    for mat_id, base in enumerate((C1_MASS * size_factor**SCALE_EXPONENTS, C1_STIFF * size_factor**SCALE_EXPONENTS)):
        st = (np.uint64(chain_key) * np.uint64(1 << 20) + np.uint64(element * num_stations)
              + np.arange(num_stations, dtype=np.uint64)) * np.uint64(128)
        keys = st[:, None] + np.uint64(mat_id * 64) + np.arange(21, dtype=np.uint64)[None, :]
        u = uniform01(seed, keys)
        d = base[None, :] * (20.0 ** (-s_chain))[:, None] * np.exp((2.0 * u[:, :6] - 1.0) * math.log(1.1))
        A = np.zeros((num_stations, 6, 6))
        A[:, np.arange(6), np.arange(6)] = d
        off = 0.05 * np.sqrt(d[:, iu[0]] * d[:, iu[1]]) * (2.0 * u[:, 6:21] - 1.0)
        A[:, iu[0], iu[1]] = off
        A[:, iu[1], iu[0]] = off
        out.append(_spd_project(A).reshape(num_stations, 36))
    return s_loc, out[0], out[1]


def _blade_frame(azimuth: float) -> np.ndarray:
    """Orientation taking material x onto the blade axis (0, cos a, sin a)."""
    return quat_compose(quat_from_rotvec([azimuth, 0.0, 0.0]), quat_from_rotvec([0.0, 0.0, math.pi / 2]))


def _add_rotor(nt, beams, cons, roots_info, *, num_nodes_per_blade, length_per_blade, omega,
               hub_radius, sec_seed, blade_keys, length_scaled_stiffness=False):
    """Append one 3-blade rotor (hub axis x, blades in the y-z plane)."""
    w = np.array([omega, 0.0, 0.0])
    for b in range(3):
        az = 2.0 * math.pi * b / 3.0
        d = np.array([0.0, math.cos(az), math.sin(az)])
        qb = _blade_frame(az)
        N = int(num_nodes_per_blade[b])
        L = float(length_per_blade[b])
        xi = gll_points(N - 1)
        ids = []
        for x in xi:
            sl = (x + 1) / 2
            p = (hub_radius + sl * L) * d
            ids.append(nt.add(np.concatenate([p, qb]), v=np.concatenate([np.cross(w, p), w]), s=sl))
        gx, gw = gauss_legendre(2 * N - 2)
        # optional geometric scaling of the 10 m base section to the blade length
        ss, Ms, Cs = section_properties(sec_seed, blade_keys[b],
                                        size_factor=(L / 10.0) if length_scaled_stiffness else 1.0)
        beams.append(BeamSpec(np.array(ids), ss, Ms, Cs, np.stack([gx, gw], axis=1)))
        cons.append(ConstraintSpec(PRESCRIBED_BC, 0, ids[0]))
        roots_info.append((len(cons) - 1, ids[0], ids[-1], omega))


def rotor_config_spec(num_nodes=11, num_qps=20, length=60.0, omega=1.0, hub_radius=2.0, h=0.01,
                      num_steps=500, rho_inf=0.9, max_iter=50, atol=1e-10, rtol=1e-6, seed=SEED,
                      gravity=(0.0, 0.0, -9.81)):
    """3 blades, one element each, prescribed root motion of a hub spinning about x."""
    assert num_qps == 2 * num_nodes - 2 or num_qps > 0
    nt = _NodeTable()
    beams, cons, roots = [], [], []
    _add_rotor(nt, beams, cons, roots, num_nodes_per_blade=[num_nodes] * 3,
               length_per_blade=[length] * 3, omega=omega, hub_radius=hub_radius, sec_seed=seed,
               blade_keys=[0, 1, 2])
    if num_qps != 2 * num_nodes - 2:
        for b in beams:
            gx, gw = gauss_legendre(num_qps)
            b.quad = np.stack([gx, gw], axis=1)
    x0, u, v, vd, s = nt.arrays()
    return _rotor_spec("rotor", x0, u, v, vd, s, beams, cons, roots, np.array(gravity), h, num_steps,
                       rho_inf, max_iter, atol, rtol, tip_load_sigma=0.0, seed=seed)


def _rotor_spec(name, x0, u, v, vd, s, beams, cons, roots, gravity, h, num_steps, rho_inf,
                max_iter, atol, rtol, tip_load_sigma, seed, rotor_of_root=None, tip_load_tau=0.0):
    root_pos = {cid: x0[node, :3].copy() for cid, node, _, _ in roots}
    n = x0.shape[0]
    tips = np.array([tip for _, _, tip, _ in roots])
    omegas = np.array([om for _, _, _, om in roots])

    def constraint_disp(step):
        t = h * (step + 1)
        out = {}
        for cid, _node, _tip, om in roots:
            q = quat_from_rotvec([om * t, 0.0, 0.0])
            p = root_pos[cid]
            out[cid] = np.concatenate([quat_rotate(q, p) - p, q])
        return out

    def white(step):
        nb = len(tips)
        keys = (np.uint64(step) * np.uint64(nb) + np.arange(nb, dtype=np.uint64)) * np.uint64(3)
        return tip_load_sigma * normal(seed + 7, keys[:, None] + np.arange(3, dtype=np.uint64)[None, :])

    # tip_load_tau > 0: f_k = rho f_(k-1) + sqrt(1 - rho^2) w_k with rho = exp(-h / tau) and w_k
    # white normal(0, sigma): every step's load is still normal(0, sigma), successive steps are
    # correlated over tau seconds.  Values are generated once, in step order, and cached.
    ar = {"f": np.zeros((0, len(tips), 3))}
    rho = math.exp(-h / tip_load_tau) if tip_load_tau > 0.0 else 0.0

    def tip_values(step):
        if tip_load_tau <= 0.0:
            return white(step)
        f = ar["f"]
        if step >= f.shape[0]:
            k0 = f.shape[0]
            ext = np.empty((max(step + 1, 2 * k0), len(tips), 3))
            ext[:k0] = f
            for k in range(k0, ext.shape[0]):
                ext[k] = white(k) if k == 0 else rho * ext[k - 1] + math.sqrt(1.0 - rho * rho) * white(k)
            ar["f"] = f = ext
        return f[step]

    def tip_loads(step):
        f = np.zeros((n, 6))
        if tip_load_sigma > 0.0:
            f[tips, :3] = tip_values(step)
        return f

    def step_inputs(step, _state=None):
        d = {"constraint_disp": constraint_disp(step)}
        if tip_load_sigma > 0.0:
            d["nodal_loads"] = tip_loads(step)
        return d

    spec = ModelSpec(name, x0, u, v, vd, s, beams, cons, gravity, max_iter=max_iter, h=h,
                     rho_inf=rho_inf, atol=atol, rtol=rtol, num_steps=num_steps,
                     step_inputs=step_inputs,
                     meta={"tips": tips, "omegas": omegas, "constraint_disp": constraint_disp,
                           "tip_loads": tip_loads, "rotor_of_root": rotor_of_root})
    return spec


def ensemble_config_specs(num_rotors=4096, h=0.01, num_steps=100, rho_inf=0.9, max_iter=50,
                          atol=1e-10, rtol=1e-6, seed=SEED, tip_load_sigma=1e4, hub_radius=2.0,
                          gravity=(0.0, 0.0, -9.81), first_rotor=0, tip_load_tau=0.1, fixed_nodes=None):
    """The batched ensemble: independent 3-blade rotors, per-blade N uniform in
    {6, 11, 16, 21}, Q = 2N-2, L uniform in [30, 120] m, omega uniform in [0.5, 1.5] rad/s,
    per-step tip loads normal(0, tip_load_sigma) N.  Sections are the config-1 base section
    scaled geometrically to the blade (s = L / 10 m: mass, EA, GA x s^2; rotary inertia, GJ,
    EI x s^4): with unscaled base sections a 120 m blade sags ~127 m under gravity and the
    reference's Newton iteration diverges.  The tip loads are correlated from step to step over
    tip_load_tau seconds (an AR(1) sequence whose every value is normal(0, tip_load_sigma)):
    loads drawn afresh each step (tip_load_tau = 0) keep pumping energy into the high modes of
    the light GLL tip nodes until the reference's Newton iteration diverges (the oracle: 7 of
    the first 64 rotors within 150 steps, the first at step 75; none with tau = 0.1 s).
    fixed_nodes: every blade with that many nodes (single-size ensembles for tuning).
    Returns one ModelSpec per rotor."""
    specs = []
    for r in range(first_rotor, first_rotor + num_rotors):
        keys = np.uint64(r) * np.uint64(8) + np.arange(8, dtype=np.uint64)
        u = uniform01(seed + 1, keys)
        Ns = [6 + 5 * min(int(u[b] * 4.0), 3) for b in range(3)] if fixed_nodes is None else [int(fixed_nodes)] * 3
        Ls = [30.0 + 90.0 * u[3 + b] for b in range(3)]
        om = 0.5 + float(u[6])
        nt = _NodeTable()
        beams, cons, roots = [], [], []
        _add_rotor(nt, beams, cons, roots, num_nodes_per_blade=Ns, length_per_blade=Ls, omega=om,
                   hub_radius=hub_radius, sec_seed=seed, blade_keys=[3 * r, 3 * r + 1, 3 * r + 2],
                   length_scaled_stiffness=True)
        x0, uu, v, vd, s = nt.arrays()
        spec = _rotor_spec(f"ensemble_rotor_{r}", x0, uu, v, vd, s, beams, cons, roots,
                           np.array(gravity), h, num_steps, rho_inf, max_iter, atol, rtol,
                           tip_load_sigma, seed=seed + 1000003 * (r + 1), tip_load_tau=tip_load_tau)
        spec.meta["rotor"] = r
        spec.meta["blade_nodes"] = Ns
        specs.append(spec)
    return specs


def ensemble_schedule(specs, num_steps, first_step=0):
    """Device-schedule arrays for a list of rotor specs advanced as one system (models in
    order): constraint displacements [steps, total_constraints, 7], the loaded (tip) node ids
    [total_tips] and their loads [steps, total_tips, 6].  Values equal the specs' own
    step_inputs (prescribed rigid rotation of each root about x, normal tip loads)."""
    node_off = np.cumsum([0] + [sp.num_nodes for sp in specs[:-1]])
    roots, omegas, tips, seeds = [], [], [], []
    for sp, off in zip(specs, node_off):
        for c in sp.constraints:
            roots.append(sp.x0[c.target, :3])
        omegas.extend(sp.meta["omegas"].tolist())
        tips.extend((off + sp.meta["tips"]).tolist())
    roots = np.array(roots)
    omegas = np.array(omegas)
    steps = np.arange(first_step, first_step + num_steps)
    t = (steps[:, None] * 0 + steps[:, None] + 1) * specs[0].h                     # [S,1]
    ang = omegas[None, :] * t                                                     # [S,C]
    a = np.abs(ang)
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(a < 1e-12, 0.0, np.sin(a / 2.0) / np.where(a < 1e-12, 1.0, a))
    q0 = np.where(a < 1e-12, 1.0, np.cos(a / 2.0))
    q1 = ang * f
    c2 = q0 * q0 - q1 * q1
    cs = 2.0 * (0.0 - q0 * q1)
    py, pz = roots[None, :, 1], roots[None, :, 2]
    ux = (q0 * q0 + q1 * q1) * roots[None, :, 0] - roots[None, :, 0]
    uy = (c2 * py + cs * pz) - py
    uz = (-cs * py + c2 * pz) - pz
    cdisp = np.stack([ux, uy, uz, q0, q1, np.zeros_like(q0), np.zeros_like(q0)], axis=-1)
    loads = np.zeros((num_steps, len(tips), 6))
    k = 0
    for sp in specs:
        nb = len(sp.meta["tips"])
        for si, st in enumerate(steps):
            f = sp.meta["tip_loads"](int(st))
            loads[si, k:k + nb] = f[sp.meta["tips"]]
        k += nb
    return np.ascontiguousarray(cdisp), np.array(tips, dtype=np.uint64), loads


def structure_config_spec(num_nodes=21, num_qps=40, tower_elems=8, shaft_elems=8, blade_elems=16,
                          tower_height=80.0, shaft_length=8.0, blade_length=60.0, hub_radius=2.0,
                          h=0.005, num_steps=1000, rho_inf=0.9, max_iter=50, atol=1e-10, rtol=1e-6,
                          seed=SEED + 2, tip_load_sigma=1e4, gravity=(0.0, 0.0, -9.81)):
    """One large structure: a tower of `tower_elems` elements along z (FixedBC at the base), a
    shaft of `shaft_elems` elements along x from the tower top, a revolute joint (axis x, zero
    torque) from the shaft end to a hub node, and three blades of `blade_elems` elements in the
    y-z plane tied to the hub by rigid joints at `hub_radius`.  Every element has its own
    `num_nodes` GLL nodes (a node's element location is unique, as in the reference's Node);
    consecutive elements of a chain are joined by rigid joints between coincident end nodes.
    Sections: the seeded generator along each chain, the 10 m base section scaled geometrically
    to the chain (tower height / 10, shaft 4, blade length / 10; as the ensemble's blades).  Blade
    tips take normal(0, sigma) loads per step.  All blocks form one KKT system (about 8k dofs at the default sizes)."""
    nt = _NodeTable()
    beams, cons = [], []
    xi = gll_points(num_nodes - 1)
    gx, gw = gauss_legendre(num_qps)
    quad = np.stack([gx, gw], axis=1)

    def chain(start, direction, frame, length, ne, key, size_factor):
        ends = []
        for k in range(ne):
            ids = []
            for x in xi:
                sl = (x + 1) / 2
                p = start + direction * length * (k + sl) / ne
                ids.append(nt.add(np.concatenate([p, frame]), s=sl))
            ss, Ms, Cs = chain_section_properties(seed, key, k, ne, size_factor=size_factor)
            beams.append(BeamSpec(np.array(ids), ss, Ms, Cs, quad))
            ends.append((ids[0], ids[-1]))
        for k in range(1, ne):
            cons.append(ConstraintSpec(RIGID_JOINT, ends[k - 1][1], ends[k][0]))
        return ends

    tower_frame = quat_from_rotvec([0.0, -math.pi / 2, 0.0])  # material x -> +z
    tower = chain(np.zeros(3), np.array([0.0, 0.0, 1.0]), tower_frame, tower_height, tower_elems, 0, tower_height / 10.0)
    cons.insert(0, ConstraintSpec(FIXED_BC, 0, tower[0][0]))
    top = np.array([0.0, 0.0, tower_height])
    shaft = chain(top, np.array([1.0, 0.0, 0.0]), np.array([1.0, 0, 0, 0]), shaft_length, shaft_elems, 1, 4.0)
    cons.append(ConstraintSpec(RIGID_JOINT, tower[-1][1], shaft[0][0]))
    hub_pos = top + np.array([shaft_length, 0.0, 0.0])
    hub = nt.add(np.concatenate([hub_pos, [1.0, 0, 0, 0]]))
    cons.append(ConstraintSpec(REVOLUTE_JOINT, shaft[-1][1], hub, axis=np.array([1.0, 0.0, 0.0])))
    tips = []
    for b in range(3):
        az = 2.0 * math.pi * b / 3.0
        d = np.array([0.0, math.cos(az), math.sin(az)])
        blade = chain(hub_pos + hub_radius * d, d, _blade_frame(az), blade_length, blade_elems, 2 + b,
                      blade_length / 10.0)
        cons.append(ConstraintSpec(RIGID_JOINT, hub, blade[0][0]))
        tips.append(blade[-1][1])
    x0, u, v, vd, s = nt.arrays()
    n = x0.shape[0]
    tips = np.array(tips)

    def step_inputs(step, _state=None):
        f = np.zeros((n, 6))
        keys = (np.uint64(step) * np.uint64(3) + np.arange(3, dtype=np.uint64)) * np.uint64(3)
        f[tips, :3] = tip_load_sigma * normal(seed + 7, keys[:, None] + np.arange(3, dtype=np.uint64)[None, :])
        return {"nodal_loads": f}

    return ModelSpec("structure", x0, u, v, vd, s, beams, cons, np.array(gravity), max_iter=max_iter, h=h,
                     rho_inf=rho_inf, atol=atol, rtol=rtol, num_steps=num_steps, step_inputs=step_inputs,
                     meta={"tips": tips, "hub": hub})
