"""GPU: kyn_element_terms_noalloc -- one beam element's residual and iteration matrix at the
current state, computed into the library's file-scope device scratch with no allocation --
against the allocating path it stands in for (kyn_debug_assemble, whose element kernel writes
into library-allocated element buffers, then kyn_get_element_terms), through the C-ABI, on two
calls in a row: first the rotor's 21-node blade, then a smaller one."""

import ctypes as C

import numpy as np
import pytest

import openturbine_amd as ota
from openturbine_amd import workloads as W

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def kyn():
    return ota.KynLib()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rotor_with_mixed_blades():
    """First ensemble rotor with a 21-node blade and a smaller one (elements = blades, in order)."""
    for sp in W.ensemble_config_specs(num_rotors=32, num_steps=3):
        ns = sp.meta["blade_nodes"]
        if 21 in ns and min(ns) < 21:
            return sp, ns.index(21), ns.index(min(ns)), ns
    raise AssertionError("no rotor with mixed blades")


def noalloc(kyn, g, e, N):
    r = np.full((N, 6), np.nan)
    m = np.full((N, N, 36), np.nan)
    rc = kyn.L.kyn_element_terms_noalloc(g.h, e, r.ctypes.data_as(_dp), m.ctypes.data_as(_dp))
    return rc, r, m


@pytest.mark.parametrize("mode", [ota.EXEC_PLAIN, ota.EXEC_FUSED], ids=["plain", "fused"])
def test_noalloc_element_terms_match_the_allocating_path(kyn, mode):
    spec, big, small, ns = rotor_with_mixed_blades()
    print(f"no-alloc element terms: mode {mode}, blades N = {ns}, elements {big} then {small}", flush=True)
    g = ota.KynSystem(kyn, spec, exec_mode=mode)
    assert g.step() == ota.KYN_OK  # off the initial state
    g.debug_assemble(spec.step_inputs(1))
    for e in (big, small):  # the second call is the smaller element
        N = ns[e]
        ref_r, ref_m = g.element_terms(e, N)
        rc, r, m = noalloc(kyn, g, e, N)
        assert rc == ota.KYN_OK, kyn.error()
        dr, dm = rel(r, ref_r), rel(m, ref_m)
        print(f"  element {e} (N = {N}): residual rel diff {dr:.1e}, iteration matrix rel diff {dm:.1e}", flush=True)
        assert dr <= 1e-13 and dm <= 1e-13


def test_noalloc_element_terms_errors(kyn):
    spec, _, _, _ = rotor_with_mixed_blades()
    g = ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_FUSED)
    assert kyn.L.kyn_element_terms_noalloc(g.h, 3, None, None) == ota.KYN_ERR_INVALID  # 3 elements: 0..2
    assert "bad element" in kyn.error()
    assert kyn.L.kyn_element_terms_noalloc(None, 0, None, None) == ota.KYN_ERR_INVALID
