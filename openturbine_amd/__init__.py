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
    _fields_ = [("exec_mode", C.c_int), ("precision", C.c_int), ("refine_sweeps", C.c_int)]


def build(force: bool = True) -> None:
    """Compile libkynema_amd.so for gfx950 (hipcc, in-tree)."""
    import subprocess

    subprocess.run(["make", "-B" if force else "-s", "-C", str(PKG)], check=True)


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
        L.kyn_get_element_terms.argtypes = [C.c_void_p, C.c_size_t, _dp, _dp]

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


class KynSystem:
    """One system advancing one or more ModelSpecs (each keeps its own Step semantics)."""

    def __init__(self, lib: KynLib, specs, *, exec_mode=EXEC_FUSED, precision=TANGENT_FP64, refine_sweeps=0,
                 max_iter=None, atol=None, rtol=None):
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
        opts = ExecOptions(exec_mode, precision, refine_sweeps)
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
