"""The HIP library loads on a CPU-only host and exports every entry point that
include/kynema_amd.h declares (no compute call is made here)."""

import ctypes
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "kynema_amd.h"
LIB = ROOT / "openturbine_amd" / "libkynema_amd.so"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(kyn_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("kyn_model_create", "kyn_model_add_beam_element", "kyn_system_create", "kyn_step",
                     "kyn_get_state", "kyn_set_constraint_displacement", "kyn_get_convergence_history",
                     "kyn_write_restart", "kyn_model_add_mass_element", "kyn_model_add_spring_element",
                     "kyn_get_motion", "kyn_beam_build_model", "kyn_blade_interface_create",
                     "kyn_blade_interface_step", "kyn_blade_interface_get_node_motion",
                     "kyn_blade_interface_set_root_displacement"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "openturbine_amd")], check=True)
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (kyn_[a-z_0-9]+)\b", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_without_gpu():
    lib = ctypes.CDLL(str(LIB))
    for n in declared_functions():
        assert hasattr(lib, n), n
    lib.kyn_model_create.restype = ctypes.c_void_p
    lib.kyn_model_destroy.argtypes = [ctypes.c_void_p]
    m = lib.kyn_model_create()  # host-only object
    assert m
    lib.kyn_model_destroy(m)


def test_python_binding_refuses_missing_library(tmp_path):
    import pytest

    from openturbine_amd import KynLib

    with pytest.raises(RuntimeError, match="no CPU fallback"):
        KynLib(tmp_path / "missing.so")
