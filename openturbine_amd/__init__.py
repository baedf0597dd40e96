"""openturbine_amd — MI355X (gfx950) beam Newton path behind a C-ABI.

The product is the HIP library ``libkynema_amd.so`` built in-tree from ``csrc/`` (see
include/kynema_amd.h).  This module is a thin ctypes binding used by tests and the bench;
it never falls back to a CPU implementation: if the library is missing it raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import workloads  # noqa: F401

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libkynema_amd.so"

KYN_OK = 0
KYN_NOT_CONVERGED = 1
KYN_ERR_INVALID = -1
KYN_ERR_HIP = -2
KYN_ERR_UNSUPPORTED = -3
EXEC_PLAIN = 0
EXEC_FUSED = 1
TANGENT_FP64 = 0
TANGENT_FP32 = 1

FIELDS = {"x0": (0, 7), "x": (1, 7), "q_prev": (2, 7), "q": (3, 7), "q_delta": (4, 6), "v": (5, 6),
          "vd": (6, 6), "a": (7, 6), "f": (8, 6), "tangent": (9, 36)}

_dp = C.POINTER(C.c_double)
_sp = C.POINTER(C.c_size_t)
_ip = C.POINTER(C.c_int)


class StepParams(C.Structure):
    _fields_ = [("is_dynamic_solve", C.c_int), ("max_iter", C.c_size_t), ("h", C.c_double),
                ("rho_inf", C.c_double), ("atol", C.c_double), ("rtol", C.c_double)]


class ExecOptions(C.Structure):
    _fields_ = [("exec_mode", C.c_int), ("precision", C.c_int), ("refine_sweeps", C.c_int),
                ("refine_err_bound", C.c_double)]


def build(force: bool = True) -> None:
    """Compile libkynema_amd.so for gfx950 (hipcc, in-tree)."""
    import subprocess

    subprocess.run(["make", "-B" if force else "-s", "-C", str(PKG)], check=True)


class SolutionInput(C.Structure):
    """kyn_solution_input (components::SolutionInput)."""
    _fields_ = [("gravity", C.c_double * 3), ("dynamic_solve", C.c_int), ("time_step", C.c_double),
                ("rho_inf", C.c_double), ("max_iter", C.c_size_t), ("absolute_error_tolerance", C.c_double),
                ("relative_error_tolerance", C.c_double)]


class BeamInput(C.Structure):
    """kyn_beam_input (components::BeamInput); pointer fields refer to arrays kept by the caller."""
    _fields_ = [("element_order", C.c_size_t), ("section_refinement", C.c_size_t),
                ("reference_axis_orientation", C.c_int), ("num_ref_axis_points", C.c_size_t),
                ("ref_axis_grid", C.POINTER(C.c_double)), ("ref_axis_coordinates", C.POINTER(C.c_double)),
                ("num_twist_points", C.c_size_t), ("twist_grid", C.POINTER(C.c_double)),
                ("twist", C.POINTER(C.c_double)), ("num_sections", C.c_size_t),
                ("section_grid", C.POINTER(C.c_double)), ("section_mass", C.POINTER(C.c_double)),
                ("section_stiffness", C.POINTER(C.c_double)), ("prescribe_root_motion", C.c_int),
                ("root_position", C.c_double * 7), ("root_velocity", C.c_double * 6),
                ("root_acceleration", C.c_double * 6)]


REF_AXIS_X = 0
REF_AXIS_Z = 1


def beam_input(*, element_order, ref_grid, ref_coords, twist_grid, twist, section_grid, section_mass,
               section_stiffness, orientation=REF_AXIS_X, section_refinement=0, prescribe_root_motion=False,
               root_position=(0.0, 0, 0, 1, 0, 0, 0), root_velocity=(0.0,) * 6, root_acceleration=(0.0,) * 6):
    """A BeamInput plus the arrays it points to (keep the returned tuple alive while in use)."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (ref_grid, np.asarray(ref_coords).reshape(-1, 3), twist_grid, twist, section_grid,
             np.asarray(section_mass).reshape(-1, 36), np.asarray(section_stiffness).reshape(-1, 36))]
    b = BeamInput()
    b.element_order = int(element_order)
    b.section_refinement = int(section_refinement)
    b.reference_axis_orientation = int(orientation)
    b.num_ref_axis_points = len(keep[0])
    b.ref_axis_grid = keep[0].ctypes.data_as(_dp)
    b.ref_axis_coordinates = keep[1].ctypes.data_as(_dp)
    b.num_twist_points = len(keep[2])
    b.twist_grid = keep[2].ctypes.data_as(_dp)
    b.twist = keep[3].ctypes.data_as(_dp)
    b.num_sections = len(keep[4])
    b.section_grid = keep[4].ctypes.data_as(_dp)
    b.section_mass = keep[5].ctypes.data_as(_dp)
    b.section_stiffness = keep[6].ctypes.data_as(_dp)
    b.prescribe_root_motion = int(bool(prescribe_root_motion))
    b.root_position[:] = list(root_position)
    b.root_velocity[:] = list(root_velocity)
    b.root_acceleration[:] = list(root_acceleration)
    return b, keep


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class KynLib:
    """ctypes binding of include/kynema_amd.h."""

    def __init__(self, path=LIB_PATH):
        if not Path(path).exists():
            raise RuntimeError(f"HIP library {path} is missing: run __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = C.CDLL(str(path))
        self.L = L
        L.kyn_last_error.restype = C.c_char_p
        L.kyn_device_info.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]
        L.kyn_model_create.restype = C.c_void_p
        L.kyn_model_destroy.argtypes = [C.c_void_p]
        L.kyn_model_set_gravity.argtypes = [C.c_void_p, _dp]
        L.kyn_model_add_node.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double]
        L.kyn_model_add_node.restype = C.c_int64
        L.kyn_model_add_beam_element.argtypes = [C.c_void_p, C.c_size_t, _sp, C.c_size_t, _dp, _dp, _dp,
                                                 C.c_size_t, _dp]
        L.kyn_model_add_beam_element.restype = C.c_int64
        L.kyn_model_add_constraint.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_size_t, _dp, _dp]
        L.kyn_model_add_constraint.restype = C.c_int64
        L.kyn_model_add_mass_element.argtypes = [C.c_void_p, C.c_size_t, _dp]
        L.kyn_model_add_mass_element.restype = C.c_int64
        L.kyn_model_add_spring_element.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_double, C.c_double]
        L.kyn_model_add_spring_element.restype = C.c_int64
        L.kyn_system_create.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.POINTER(StepParams),
                                        C.POINTER(ExecOptions)]
        L.kyn_system_create.restype = C.c_void_p
        L.kyn_system_destroy.argtypes = [C.c_void_p]
        for f in ("kyn_system_num_models", "kyn_system_num_nodes", "kyn_system_num_constraints"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_size_t
        for f in ("kyn_system_num_dofs", "kyn_system_node_offset", "kyn_system_constraint_offset",
                  "kyn_restart_size"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_size_t]
            getattr(L, f).restype = C.c_size_t
        L.kyn_step.argtypes = [C.c_void_p]
        L.kyn_get_step_status.argtypes = [C.c_void_p, _ip, _ip]
        L.kyn_step_many.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
        L.kyn_get_totals.argtypes = [C.c_void_p, _ip, _ip]
        L.kyn_get_convergence_history.argtypes = [C.c_void_p, C.c_size_t, _dp, C.c_size_t]
        L.kyn_get_convergence_history.restype = C.c_size_t
        L.kyn_last_step_ms.argtypes = [C.c_void_p]
        L.kyn_last_step_ms.restype = C.c_double
        L.kyn_last_kernel_ms.argtypes = [C.c_void_p]
        L.kyn_last_kernel_ms.restype = C.c_double
        L.kyn_set_constraint_displacement.argtypes = [C.c_void_p, C.c_size_t, _dp]
        L.kyn_set_constraint_displacements.argtypes = [C.c_void_p, C.c_size_t, _sp, _dp]
        L.kyn_set_constraint_control.argtypes = [C.c_void_p, C.c_size_t, C.c_double]
        L.kyn_set_nodal_loads.argtypes = [C.c_void_p, _dp]
        L.kyn_set_element_node_loads.argtypes = [C.c_void_p, C.c_size_t, _dp]
        L.kyn_set_input_schedule.argtypes = [C.c_void_p, C.c_size_t, _dp, C.c_size_t, _sp, _dp]
        L.kyn_get_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.kyn_set_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.kyn_get_constraint_values.argtypes = [C.c_void_p, C.c_size_t, C.c_int, _dp]
        L.kyn_get_rhs.argtypes = [C.c_void_p, C.c_size_t, _dp]
        L.kyn_get_solution.argtypes = [C.c_void_p, C.c_size_t, _dp]
        L.kyn_save_state.argtypes = [C.c_void_p]
        L.kyn_restore_state.argtypes = [C.c_void_p]
        L.kyn_write_restart.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_ubyte)]
        L.kyn_read_restart.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_ubyte), C.c_size_t]
        L.kyn_debug_assemble.argtypes = [C.c_void_p]
        L.kyn_debug_set_wait_limit.argtypes = [C.c_void_p, C.c_int]
        L.kyn_debug_set_contraction_mask.argtypes = [C.c_void_p, C.c_uint]
        L.kyn_get_element_terms.argtypes = [C.c_void_p, C.c_size_t, _dp, _dp]
        L.kyn_element_terms_noalloc.argtypes = [C.c_void_p, C.c_size_t, _dp, _dp]
        L.kyn_debug_get_crs.argtypes = [C.c_void_p, C.c_size_t, _sp, _sp, _dp, C.c_size_t]
        L.kyn_debug_get_crs.restype = C.c_size_t
        L.kyn_get_motion.argtypes = [C.c_void_p, _dp]
        L.kyn_model_num_nodes.argtypes = [C.c_void_p]
        L.kyn_model_num_nodes.restype = C.c_size_t
        L.kyn_model_get_node.argtypes = [C.c_void_p, C.c_size_t, _dp, _dp, _dp, _dp, _dp]
        L.kyn_model_beam_element_info.argtypes = [C.c_void_p, C.c_size_t, _sp, _sp, _sp]
        L.kyn_model_get_beam_element.argtypes = [C.c_void_p, C.c_size_t, _sp, _dp, _dp, _dp, _dp]
        L.kyn_beam_build_model.argtypes = [C.c_void_p, C.POINTER(BeamInput), C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int64)]
        L.kyn_blade_interface_create.argtypes = [C.POINTER(SolutionInput), C.POINTER(BeamInput),
                                                 C.POINTER(ExecOptions)]
        L.kyn_blade_interface_create.restype = C.c_void_p
        L.kyn_blade_interface_destroy.argtypes = [C.c_void_p]
        L.kyn_blade_interface_num_nodes.argtypes = [C.c_void_p]
        L.kyn_blade_interface_num_nodes.restype = C.c_size_t
        for f in ("kyn_blade_interface_set_node_loads", "kyn_blade_interface_get_node_loads"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_size_t, _dp]
        L.kyn_blade_interface_add_point_load.argtypes = [C.c_void_p, C.c_double, _dp]
        for f in ("kyn_blade_interface_clear_loads", "kyn_blade_interface_step", "kyn_blade_interface_save_state",
                  "kyn_blade_interface_restore_state"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.kyn_blade_interface_get_node_motion.argtypes = [C.c_void_p, C.c_size_t, _dp, _dp, _dp, _dp]
        L.kyn_blade_interface_set_root_displacement.argtypes = [C.c_void_p, _dp]
        L.kyn_blade_interface_system.argtypes = [C.c_void_p]
        L.kyn_blade_interface_system.restype = C.c_void_p

    def error(self) -> str:
        return self.L.kyn_last_error().decode()

    def device_info(self):
        name = C.create_string_buffer(256)
        arch = C.create_string_buffer(64)
        rc = self.L.kyn_device_info(name, 256, arch, 64)
        if rc != 0:
            raise RuntimeError(self.error())
        return name.value.decode(), arch.value.decode()

    # ---- model entry points (same names as the oracle adapter used by the tests)
    def model_create(self):
        return self.L.kyn_model_create()

    def model_destroy(self, m):
        self.L.kyn_model_destroy(m)

    def model_set_gravity(self, m, g):
        a, p = _d(g)
        self.L.kyn_model_set_gravity(m, p)

    def model_add_node(self, m, x0, u, v, vd, s):
        a1, p1 = _d(x0)
        a2, p2 = _d(u)
        a3, p3 = _d(v)
        a4, p4 = _d(vd)
        return self.L.kyn_model_add_node(m, p1, p2, p3, p4, float(s))

    def model_add_beam_element(self, m, nodes, sec_s, sec_M, sec_C, quad):
        nodes = np.ascontiguousarray(nodes, dtype=np.uintp)
        a1, p1 = _d(sec_s)
        a2, p2 = _d(sec_M)
        a3, p3 = _d(sec_C)
        a4, p4 = _d(quad)
        return self.L.kyn_model_add_beam_element(m, len(nodes), nodes.ctypes.data_as(_sp), len(a1), p1, p2, p3,
                                                 a4.size // 2, p4)

    def model_add_constraint(self, m, ctype, base, target, axis, init):
        a1, p1 = _d(axis)
        a2, p2 = _d(init)
        return self.L.kyn_model_add_constraint(m, int(ctype), int(base), int(target), p1, p2)

    def model_add_mass_element(self, m, node, mass_matrix):
        a, p = _d(np.asarray(mass_matrix, dtype=np.float64).reshape(36))
        return self.L.kyn_model_add_mass_element(m, int(node), p)

    def model_add_spring_element(self, m, n1, n2, k, l0):
        return self.L.kyn_model_add_spring_element(m, int(n1), int(n2), float(k), float(l0))


class KynBladeInterface:
    """interfaces::BladeInterface through the C-ABI (kyn_blade_interface_*)."""

    def __init__(self, lib: KynLib, solution: dict, beam, *, exec_mode=EXEC_FUSED, precision=TANGENT_FP64,
                 refine_sweeps=0):
        self.lib = lib
        sol = SolutionInput()
        sol.gravity[:] = list(solution.get("gravity", (0.0, 0.0, 0.0)))
        sol.dynamic_solve = int(solution.get("dynamic_solve", True))
        sol.time_step = float(solution["time_step"])
        sol.rho_inf = float(solution.get("rho_inf", 0.0))
        sol.max_iter = int(solution.get("max_iter", 5))
        sol.absolute_error_tolerance = float(solution.get("atol", 1e-5))
        sol.relative_error_tolerance = float(solution.get("rtol", 1e-3))
        b, self._keep = beam
        opts = ExecOptions(exec_mode, precision, refine_sweeps, 0.0)
        self.h = lib.L.kyn_blade_interface_create(C.byref(sol), C.byref(b), C.byref(opts))
        if not self.h:
            raise RuntimeError("kyn_blade_interface_create failed: " + lib.error())
        self.n = lib.L.kyn_blade_interface_num_nodes(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.L.kyn_blade_interface_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(self.lib.error())
        return rc

    def set_node_loads(self, node, loads):
        a, p = _d(loads)
        self._check(self.lib.L.kyn_blade_interface_set_node_loads(self.h, int(node), p))

    def node_loads(self, node):
        out = np.zeros(6)
        self._check(self.lib.L.kyn_blade_interface_get_node_loads(self.h, int(node), out.ctypes.data_as(_dp)))
        return out

    def add_point_load(self, s, loads):
        a, p = _d(loads)
        self._check(self.lib.L.kyn_blade_interface_add_point_load(self.h, float(s), p))

    def clear_loads(self):
        self._check(self.lib.L.kyn_blade_interface_clear_loads(self.h))

    def step(self):
        return self._check(self.lib.L.kyn_blade_interface_step(self.h))

    def motion(self, node):
        """(position, displacement, velocity, acceleration) of one blade node."""
        out = [np.zeros(7), np.zeros(7), np.zeros(6), np.zeros(6)]
        self._check(self.lib.L.kyn_blade_interface_get_node_motion(self.h, int(node), *[o.ctypes.data_as(_dp) for o in out]))
        return out

    def save_state(self):
        self._check(self.lib.L.kyn_blade_interface_save_state(self.h))

    def restore_state(self):
        self._check(self.lib.L.kyn_blade_interface_restore_state(self.h))

    def set_root_displacement(self, u):
        a, p = _d(u)
        self._check(self.lib.L.kyn_blade_interface_set_root_displacement(self.h, p))

    def system_handle(self):
        return self.lib.L.kyn_blade_interface_system(self.h)


def built_beam_model(lib: KynLib, beam):
    """Runs the C-ABI beam builder on a host-only model; returns (nodes, beam element) as arrays:
    nodes {x0, u, v, vd, s}, element {node_ids, section_s, M, C, quad}."""
    b, keep = beam
    m = lib.L.kyn_model_create()
    try:
        be = C.c_int64(-1)
        rc_ = C.c_int64(-1)
        rc = lib.L.kyn_beam_build_model(m, C.byref(b), C.byref(be), C.byref(rc_))
        if rc != 0:
            raise RuntimeError(lib.error())
        n = lib.L.kyn_model_num_nodes(m)
        x0, u, v, vd, s = np.zeros((n, 7)), np.zeros((n, 7)), np.zeros((n, 6)), np.zeros((n, 6)), np.zeros(n)
        for i in range(n):
            sv = C.c_double(0.0)
            lib.L.kyn_model_get_node(m, i, x0[i].ctypes.data_as(_dp), u[i].ctypes.data_as(_dp),
                                     v[i].ctypes.data_as(_dp), vd[i].ctypes.data_as(_dp), C.byref(sv))
            s[i] = sv.value
        nn, ns, nq = C.c_size_t(), C.c_size_t(), C.c_size_t()
        lib.L.kyn_model_beam_element_info(m, be.value, C.byref(nn), C.byref(ns), C.byref(nq))
        ids = np.zeros(nn.value, dtype=np.uintp)
        ss, M, Cm, q = np.zeros(ns.value), np.zeros((ns.value, 36)), np.zeros((ns.value, 36)), np.zeros((nq.value, 2))
        lib.L.kyn_model_get_beam_element(m, be.value, ids.ctypes.data_as(_sp), ss.ctypes.data_as(_dp),
                                         M.ctypes.data_as(_dp), Cm.ctypes.data_as(_dp), q.ctypes.data_as(_dp))
        return ({"x0": x0, "u": u, "v": v, "vd": vd, "s": s},
                {"node_ids": ids, "section_s": ss, "M": M, "C": Cm, "quad": q, "root_constraint": rc_.value})
    finally:
        lib.L.kyn_model_destroy(m)


class KynSystem:
    """One system advancing one or more ModelSpecs (each keeps its own Step semantics)."""

    def __init__(self, lib: KynLib, specs, *, exec_mode=EXEC_FUSED, precision=TANGENT_FP64, refine_sweeps=0,
                 refine_err_bound=0.0, max_iter=None, atol=None, rtol=None):
        from .workloads import build_model

        if not isinstance(specs, (list, tuple)):
            specs = [specs]
        self.lib = lib
        self.specs = list(specs)
        s0 = self.specs[0]
        models = [build_model(lib, sp) for sp in self.specs]
        arr = (C.c_void_p * len(models))(*models)
        params = StepParams(int(s0.is_dynamic), int(max_iter or s0.max_iter), s0.h, s0.rho_inf,
                            s0.atol if atol is None else atol, s0.rtol if rtol is None else rtol)
        opts = ExecOptions(exec_mode, precision, refine_sweeps, float(refine_err_bound))
        self.h = lib.L.kyn_system_create(arr, len(models), C.byref(params), C.byref(opts))
        for m in models:
            lib.model_destroy(m)
        if not self.h:
            raise RuntimeError("kyn_system_create failed: " + lib.error())
        self.n = lib.L.kyn_system_num_nodes(self.h)
        self.num_models = len(models)
        self.num_cons = lib.L.kyn_system_num_constraints(self.h)
        self.max_iter = params.max_iter
        self.node_off = [lib.L.kyn_system_node_offset(self.h, k) for k in range(self.num_models)]
        self.con_off = [lib.L.kyn_system_constraint_offset(self.h, k) for k in range(self.num_models)]
        self.step_index = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.L.kyn_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(self.lib.error())
        return rc

    # ---- state
    def get(self, field):
        fid, w = FIELDS[field]
        out = np.zeros((self.n, w))
        self._check(self.lib.L.kyn_get_state(self.h, fid, out.ctypes.data_as(_dp)))
        return out

    def set(self, field, arr):
        fid, w = FIELDS[field]
        a, p = _d(np.asarray(arr).reshape(self.n, w))
        self._check(self.lib.L.kyn_set_state(self.h, fid, p))

    def model_slice(self, k):
        end = self.node_off[k + 1] if k + 1 < self.num_models else self.n
        return slice(self.node_off[k], end)

    # ---- inputs
    def apply_inputs(self, inputs, model=0):
        """Apply one model's step inputs (ModelSpec.step_inputs format)."""
        cd = inputs.get("constraint_disp", {})
        if cd:
            ids = np.array([self.con_off[model] + c for c in cd.keys()], dtype=np.uintp)
            vals = np.ascontiguousarray(np.array(list(cd.values())), dtype=np.float64)
            self._check(self.lib.L.kyn_set_constraint_displacements(self.h, len(ids), ids.ctypes.data_as(_sp),
                                                                    vals.ctypes.data_as(_dp)))
        if "nodal_loads" in inputs:
            f = self.get("f")
            f[self.model_slice(model)] = inputs["nodal_loads"]
            a, p = _d(f)
            self._check(self.lib.L.kyn_set_nodal_loads(self.h, p))
        for e, fx in inputs.get("element_node_loads", {}).items():
            a, p = _d(fx)
            self._check(self.lib.L.kyn_set_element_node_loads(self.h, e, p))
        for cid, val in inputs.get("control", {}).items():
            self._check(self.lib.L.kyn_set_constraint_control(self.h, self.con_off[model] + cid, float(val)))

    def set_schedule(self, num_steps, cdisp=None, load_nodes=None, loads=None):
        cp = None
        if cdisp is not None:
            cdisp = np.ascontiguousarray(cdisp, dtype=np.float64)
            cp = cdisp.ctypes.data_as(_dp)
        nl = 0
        np_ = None
        lp = None
        if loads is not None:
            load_nodes = np.ascontiguousarray(load_nodes, dtype=np.uintp)
            loads = np.ascontiguousarray(loads, dtype=np.float64)
            nl = len(load_nodes)
            np_ = load_nodes.ctypes.data_as(_sp)
            lp = loads.ctypes.data_as(_dp)
        self._check(self.lib.L.kyn_set_input_schedule(self.h, num_steps, cp, nl, np_, lp))
        self._sched_keep = (cdisp, load_nodes, loads)

    # ---- stepping
    def step(self, inputs=None):
        """One step; with inputs=None the spec's step_inputs drive every model."""
        if inputs is None:
            for k, sp in enumerate(self.specs):
                if sp.step_inputs is not None:
                    sl = self.model_slice(k)
                    self.apply_inputs(sp.step_inputs(self.step_index, lambda f, sl=sl: self.get(f)[sl]), k)
        elif inputs:
            self.apply_inputs(inputs)
        rc = self._check(self.lib.L.kyn_step(self.h))
        self.step_index += 1
        return rc

    def step_many(self, num_steps, use_graph=True):
        rc = self._check(self.lib.L.kyn_step_many(self.h, num_steps, int(use_graph)))
        self.step_index += num_steps
        return rc

    def totals(self):
        it = np.zeros(self.num_models, dtype=np.int32)
        fl = np.zeros(self.num_models, dtype=np.int32)
        self._check(self.lib.L.kyn_get_totals(self.h, it.ctypes.data_as(_ip), fl.ctypes.data_as(_ip)))
        return it, fl

    def status(self):
        st = np.zeros(self.num_models, dtype=np.int32)
        it = np.zeros(self.num_models, dtype=np.int32)
        self._check(self.lib.L.kyn_get_step_status(self.h, st.ctypes.data_as(_ip), it.ctypes.data_as(_ip)))
        return st, it

    def convergence(self, model=0):
        buf = np.zeros(256)
        n = self.lib.L.kyn_get_convergence_history(self.h, model, buf.ctypes.data_as(_dp), 256)
        return buf[: min(n, 256)].copy()

    def last_step_ms(self):
        return self.lib.L.kyn_last_step_ms(self.h)

    def last_kernel_ms(self):
        return self.lib.L.kyn_last_kernel_ms(self.h)

    def rhs(self, model=0):
        out = np.zeros(self.lib.L.kyn_system_num_dofs(self.h, model))
        self._check(self.lib.L.kyn_get_rhs(self.h, model, out.ctypes.data_as(_dp)))
        return out

    def solution(self, model=0):
        out = np.zeros(self.lib.L.kyn_system_num_dofs(self.h, model))
        self._check(self.lib.L.kyn_get_solution(self.h, model, out.ctypes.data_as(_dp)))
        return out

    def constraint(self, cid, which):
        out = np.zeros(6)
        self._check(self.lib.L.kyn_get_constraint_values(self.h, cid, which, out.ctypes.data_as(_dp)))
        return out[:3] if which == 2 else out

    def debug_assemble(self, inputs=None):
        if inputs:
            self.apply_inputs(inputs)
        self._check(self.lib.L.kyn_debug_assemble(self.h))

    def crs_matrix(self, model=0):
        """(row_ptrs, col_inds, values) of a model that forms one large block (plain form), in the
        reference's dof order: the graph and the values of the last assembly."""
        L = self.lib.L
        nnz = L.kyn_debug_get_crs(self.h, model, None, None, None, 0)
        if nnz == 0:
            raise RuntimeError(self.lib.error())
        ndofs = int(L.kyn_system_num_dofs(self.h, model))
        rp = np.zeros(ndofs + 1, dtype=np.uint64)
        ci = np.zeros(nnz, dtype=np.uint64)
        v = np.zeros(nnz)
        got = L.kyn_debug_get_crs(self.h, model, rp.ctypes.data_as(_sp), ci.ctypes.data_as(_sp), v.ctypes.data_as(_dp),
                                  nnz)
        assert got == nnz
        return rp.astype(np.int64), ci.astype(np.int64), v

    def element_terms(self, elem, N):
        r = np.zeros((N, 6))
        m = np.zeros((N, N, 36))
        self._check(self.lib.L.kyn_get_element_terms(self.h, elem, r.ctypes.data_as(_dp), m.ctypes.data_as(_dp)))
        return r, m

    def save_state(self):
        self._check(self.lib.L.kyn_save_state(self.h))

    def restore_state(self):
        self._check(self.lib.L.kyn_restore_state(self.h))

    def write_restart(self, model=0) -> bytes:
        n = self.lib.L.kyn_restart_size(self.h, model)
        buf = (C.c_ubyte * n)()
        self._check(self.lib.L.kyn_write_restart(self.h, model, buf))
        return bytes(buf)

    def read_restart(self, data: bytes, model=0):
        buf = (C.c_ubyte * len(data)).from_buffer_copy(data)
        self._check(self.lib.L.kyn_read_restart(self.h, model, buf, len(data)))
