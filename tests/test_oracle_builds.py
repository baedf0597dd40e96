"""The oracle's two builds -- with FMA contraction (the checker) and without it -- restate the
same algorithm and differ only in rounding: on the config-1 cantilever and the config-2 rotor
they take the same Newton iteration counts in every step and end within rounding of each other.
The GPU parity tests scale their Newton-history bars by the distance between the two builds, so
this test also checks that the distance is a rounding-level one (nonzero, and small)."""

import numpy as np
import pytest

from openturbine_amd import workloads as W
from oracle_lib import NC_LIB_PATH, OracleLib, OracleSystem


@pytest.fixture(scope="module")
def libs():
    return OracleLib(), OracleLib(NC_LIB_PATH)


@pytest.mark.parametrize("which", ["config1", "config2"])
def test_two_oracle_builds_agree_to_rounding(libs, which):
    spec = W.cantilever_config_spec(num_steps=30) if which == "config1" else W.rotor_config_spec(num_steps=30)
    a, b = OracleSystem(libs[0], spec), OracleSystem(libs[1], spec)
    worst = 0.0
    for k in range(30):
        assert a.step(spec.step_inputs(k, a.get)) == 0
        assert b.step(spec.step_inputs(k, b.get)) == 0
        ha, hb = np.asarray(a.convergence()), np.asarray(b.convergence())
        assert len(ha) == len(hb), k
        worst = max(worst, float(np.max(np.abs(ha[:-1] - hb[:-1]) / ha[:-1])))
    qa, qb = a.get("q"), b.get("q")
    dq = float(np.max(np.abs(qa - qb)) / np.max(np.abs(qa)))
    assert 0.0 < dq < 1e-13, dq        # different builds, same algorithm
    assert 0.0 < worst < 1e-4, worst   # histories: rounding-level, not identical
