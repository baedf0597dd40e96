"""The C-ABI beam component builder (kyn_beam_build_model, host only) against the numpy
restatement the oracle tests use (tests/blade_builder.py, pinned by the reference's
BladeInterface regression values): same nodes, sections and quadrature for the rotating beam
and the windIO IEA-15 blade (reference axis and sections given along z)."""

import numpy as np
import pytest

import blade_builder as BB
import openturbine_amd as ota
from golden_data import load


@pytest.fixture(scope="module")
def lib():
    return ota.KynLib()


@pytest.fixture(scope="module")
def fx():
    return load("blade_interface.json")


def compare(lib, kyn_in, py_in):
    nodes, elem = ota.built_beam_model(lib, kyn_in)
    spec = BB.build_blade_spec(py_in, h=0.01, rho_inf=0.0, max_iter=6, atol=1e-6, rtol=1e-4)
    b = spec.beams[0]
    np.testing.assert_allclose(nodes["x0"], spec.x0, rtol=0, atol=1e-12 * np.max(np.abs(spec.x0)))
    np.testing.assert_allclose(nodes["v"], spec.v, rtol=0, atol=1e-13)
    np.testing.assert_allclose(nodes["vd"], spec.vd, rtol=0, atol=1e-13)
    np.testing.assert_allclose(nodes["s"], spec.s, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(elem["node_ids"], b.nodes)
    np.testing.assert_allclose(elem["section_s"], b.section_s, rtol=0, atol=1e-15)
    np.testing.assert_allclose(elem["M"], b.section_M, rtol=0, atol=1e-12 * np.max(np.abs(b.section_M)))
    np.testing.assert_allclose(elem["C"], b.section_C, rtol=0, atol=1e-12 * np.max(np.abs(b.section_C)))
    np.testing.assert_allclose(elem["quad"], b.quad, rtol=0, atol=1e-15)
    assert elem["root_constraint"] == 0


def test_rotating_beam_builder(lib, fx):
    d = fx["rotating_beam"]
    compare(lib, BB.rotating_beam_kyn_input(d), BB.rotating_beam_input(d))


def test_windio_blade_builder(lib, fx):
    d = fx["windio"]
    compare(lib, BB.windio_kyn_input(d), BB.windio_input(d))


def test_builder_rejects_bad_input(lib):
    bad = ota.beam_input(element_order=0, ref_grid=[0.0, 1.0], ref_coords=[[0, 0, 0], [1, 0, 0]], twist_grid=[0.0],
                         twist=[0.0], section_grid=[0.0], section_mass=[np.eye(6).ravel()],
                         section_stiffness=[np.eye(6).ravel()])
    with pytest.raises(RuntimeError, match="Element order"):
        ota.built_beam_model(lib, bad)
