"""Loaders for the golden fixtures under tests/golden (data extracted from the reference's
own test sources by tests/golden/extract_reference_vectors.py)."""

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return json.loads((GOLDEN / name).read_text())


def solver_case(fn):
    d = load("solver_regression.json")[fn]
    exp = {e["view"]: np.array(e["values"]) for e in d["expect"]}
    return d, exp


B_VIEW = "Kokkos::subview(solver.b,Kokkos::ALL,0)"
X_VIEW = "Kokkos::subview(solver.x,Kokkos::ALL,0)"
