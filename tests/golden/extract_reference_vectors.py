"""Extract the reference's golden vectors into JSON fixtures under tests/golden/.

The reference keeps its known answers as numeric literals inside C++ test sources.  This
script reads those sources as text under /root/reference (it is run in the build container
only; the fixtures it writes are what the tests load) and stores the numbers, with the
inputs they belong to, as plain JSON data.  Nothing of the reference's code is kept: only
the inputs and the expected outputs.

Sources (paths relative to the reference root):
  tests/regression_tests/regression/test_solver.cpp         -> solver_regression.json
  tests/regression_tests/regression/test_cantilever_beam.cpp -> cantilever_sine.json
  tests/regression_tests/regression/test_rotor.cpp
  tests/regression_tests/regression/iea15_rotor_data.hpp    -> iea15_rotor.json
  tests/unit_tests/system/beams/test_calculate_strain.cpp   -> functor_known_answers.json
  tests/unit_tests/system/beams/test_calculate_{Ouu,Puu,Quu,force_FC,force_FD,
    temporary_variables}.cpp,
  tests/unit_tests/system/masses/*.cpp,
  tests/unit_tests/state/test_{calculate_next_state,calculate_displacement,
    update_dynamic_prediction,update_static_prediction,update_algorithmic_acceleration,
    update_global_position}.cpp                             -> functor_known_answers.json
  tests/unit_tests/system/beams/test_integrate_{stiffness_matrix,inertia_matrix,
    residual_vector}.cpp                                    -> integrate_known_answers.json
  tests/unit_tests/system/beams/test_curved_beam.hpp         -> curved_beam.json
  tests/unit_tests/solver/test_compute_{row_ptrs,col_inds}.cpp -> crs_known_answers.json
  tests/regression_tests/regression/test_masses.cpp,
  tests/regression_tests/regression/test_precession.cpp,
  tests/regression_tests/regression/test_spring_mass_system.cpp,
  tests/regression_tests/regression/test_heavy_top.cpp        -> masses_springs.json
  tests/regression_tests/interfaces/test_blade_interface.cpp,
  tests/regression_tests/interfaces/interfaces_test_files/IEA-15-240-RWT.yaml
                                                             -> blade_interface.json
  tests/regression_tests/regression/test_rotating_beam.cpp  -> rotating_beam_constraints.json

Usage: python tests/golden/extract_reference_vectors.py [/root/reference]
"""

import json
import math
import re
import sys
from pathlib import Path

NUM = re.compile(r"(?<![\w.])-?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")
HERE = Path(__file__).resolve().parent


def numbers(text):
    return [float(t) for t in NUM.findall(text)]


def arith(expr, env=None):
    """Value of a numeric literal expression (+ - * / and parentheses, e.g. `1. / 3.`; `NAME[i]`
    for an element of an array parsed earlier, given in `env`)."""
    import ast

    def ev(n):
        if isinstance(n, ast.Subscript) and env is not None and isinstance(n.value, ast.Name) and n.value.id in env:
            return float(env[n.value.id][int(ev(n.slice))])
        if isinstance(n, ast.Call) and env is not None and isinstance(n.func, ast.Name) and n.func.id in env \
                and len(n.args) == 1:  # a view read `name(i)`
            return float(env[n.func.id][int(ev(n.args[0]))])
        if isinstance(n, ast.Name) and env is not None and n.id in env and not isinstance(env[n.id], list):
            return float(env[n.id])
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            return -ev(n.operand) if isinstance(n.op, ast.USub) else ev(n.operand)
        if isinstance(n, ast.BinOp) and isinstance(n.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, ast.Add):
                return a + b
            if isinstance(n.op, ast.Sub):
                return a - b
            return a * b if isinstance(n.op, ast.Mult) else a / b
        raise ValueError("not a numeric literal expression: " + expr)

    e = re.sub(r"(\d)\.(?![\d])", r"\1.0", expr.strip())  # C++ `1.` -> `1.0`
    e = re.sub(r"(?<=\d)[uUlLfF]+\b", "", e)
    return ev(ast.parse(e, mode="eval"))


def values(braced):
    """Element values of a `{a, b, ...}` initializer whose elements may be literal expressions."""
    inner = braced.strip()
    if inner.startswith("{"):
        inner = balanced(inner, 0)
    return [arith(t) for t in split_top(inner) if t.strip()]


def balanced(text, start):
    """Return the text between the parenthesis/brace opened at text[start] and its match."""
    open_c = text[start]
    close_c = {"(": ")", "{": "}", "[": "]"}[open_c]
    depth = 0
    for i in range(start, len(text)):
        if text[i] == open_c:
            depth += 1
        elif text[i] == close_c:
            depth -= 1
            if depth == 0:
                return text[start + 1 : i]
    raise ValueError("unbalanced")


def split_top(args):
    out, depth, cur = [], 0, []
    for ch in args:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    out.append("".join(cur))
    return out


def strip_comments(text):
    text = re.sub(r"//[^\n]*", "", text)
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def expectations(body):
    """All expect_kokkos_view_{1D,2D}_equal calls in a function body, in order."""
    res = []
    for m in re.finditer(r"expect_kokkos_view_(1D|2D)_equal\s*\(", body):
        args = split_top(balanced(body, m.end() - 1))
        name = re.sub(r"\s+", "", args[0])
        data = numbers(",".join(args[1:]))
        res.append({"view": name, "values": data})
    return res


def function_body(text, name):
    m = re.search(r"\b" + re.escape(name) + r"\s*\(\s*\)\s*\{", text)
    return balanced(text, m.end() - 1)


def solver_regression(root):
    src = strip_comments((root / "tests/regression_tests/regression/test_solver.cpp").read_text())
    out = {"source": "tests/regression_tests/regression/test_solver.cpp"}
    for fn in ("SetUpSolverAndAssemble", "SetupAndTakeNoSteps", "SetupAndTakeTwoSteps"):
        body = function_body(src, fn)
        entry = {"expect": expectations(body)}
        m = re.search(r"EXPECT_NEAR\(solver\.convergence_err\[0\],\s*([^,]+),", body)
        if m:
            entry["convergence_err0"] = float(m.group(1))
        mass = numbers(balanced(body, body.index("{", body.index("mass_matrix"))))
        stiff = numbers(balanced(body, body.index("{", body.index("stiffness_matrix"))))
        node_s = numbers(balanced(body, body.index("{", body.index("node_s ="))))
        quad = numbers(balanced(body, body.index("(", body.index("model.AddBeamElement"))).split("std::array{", 2)[-1])
        # quadrature: the last std::array{...} group inside AddBeamElement holds xi,w pairs
        qtext = balanced(body, body.index("(", body.index("model.AddBeamElement")))
        qblock = qtext[qtext.rfind("std::array{\n") :]
        quad = numbers(qblock)
        entry["inputs"] = {
            "mass_matrix": mass,
            "stiffness_matrix": stiff,
            "node_s": node_s,
            "quadrature": quad,
            "x_root_offset": 2.0 if "x0_root[0]" in body.split("AddNode")[0] + body.split("AddNode")[1][:200] else 0.0,
            "omega": numbers(re.search(r"omega = ([^;]+);", body).group(1))[0],
            "max_iter": numbers(re.search(r"max_iter = ([^;]+);", body).group(1))[0],
            "step_size": numbers(re.search(r"step_size = ([^;]+);", body).group(1))[0],
            "rho_inf": numbers(re.search(r"rho_inf = ([^;]+);", body).group(1))[0],
            "bc_translation_follows_rotation": "u_trans" in body,
        }
        out[fn] = entry
    return out


def cantilever(root):
    src = strip_comments((root / "tests/regression_tests/regression/test_cantilever_beam.cpp").read_text())
    body = src[src.index("CantileverBeamSineLoad") :]
    mass = numbers(balanced(body, body.index("{", body.index("mass_matrix ="))))
    stiff = numbers(balanced(body, body.index("{", body.index("stiffness_matrix ="))))
    node_s = numbers(balanced(body, body.index("{", body.index("node_s ="))))
    qtext = balanced(body, body.index("(", body.index("model.AddBeamElement")))
    quad = numbers(qtext[qtext.rfind("std::array{\n") :])
    loads = [numbers(m.group(1)) for m in re.finditer(r"100\. \* std::sin\(([^)]*)\)", body)]
    tips = [e["values"] for e in expectations(body)]
    return {
        "source": "tests/regression_tests/regression/test_cantilever_beam.cpp",
        "mass_matrix": mass,
        "stiffness_matrix": stiff,
        "node_s": node_s,
        "quadrature": quad,
        "load_sin_args": loads,
        "max_iter": numbers(re.search(r"max_iter\((\d+)\)", body).group(1))[0],
        "step_size": numbers(re.search(r"step_size\(([^)]+)\)", body).group(1))[0],
        "rho_inf": numbers(re.search(r"rho_inf\(([^)]+)\)", body).group(1))[0],
        "tip_node": 4,
        "tip_q_translation": tips,
    }


def iea15(root):
    data = strip_comments((root / "tests/regression_tests/regression/iea15_rotor_data.hpp").read_text())
    node_xi = numbers(balanced(data, data.index("{", data.index("node_xi ="))))
    coords = numbers(balanced(data, data.index("{", data.index("node_coords ="))))
    quad = numbers(balanced(data, data.index("{", data.index("trapz_quadrature ="))))
    sec_text = balanced(data, data.index("{", data.index("material_sections =")))
    sections = []
    for m in re.finditer(r"BeamSection\s*\{", sec_text):
        vals = numbers(balanced(sec_text, m.end() - 1))
        assert len(vals) == 73, len(vals)
        sections.append({"s": vals[0], "M": vals[1:37], "C": vals[37:73]})
    test = strip_comments((root / "tests/regression_tests/regression/test_rotor.cpp").read_text())
    body = test[test.index("TEST(RotorTest, IEA15Rotor)") :]
    body = body[: body.index("TEST(", 10)] if "TEST(" in body[10:] else body
    omega = numbers(re.search(r"omega = std::array\{([^}]*)\}", body).group(1))
    hub = numbers(re.search(r"hub_rad = std::array\{([^}]*)\}", test).group(1))
    gravity = numbers(re.search(r"SetGravity\(([^)]*)\)", test).group(1))
    checks = []
    for m in re.finditer(r"EXPECT_NEAR\(q\((\d+), (\d+)\),\s*(q\(\d+, \d+\)|[^,]+),\s*([^)]+)\)", body):
        node, comp = int(m.group(1)), int(m.group(2))
        rhs = m.group(3).strip()
        tol = float(m.group(4))
        ref = re.match(r"q\((\d+), (\d+)\)", rhs)
        if ref:
            checks.append({"node": node, "comp": comp, "equals": [int(ref.group(1)), int(ref.group(2))], "tol": tol})
        else:
            checks.append({"node": node, "comp": comp, "value": float(rhs), "tol": tol})
    return {
        "source": "tests/regression_tests/regression/test_rotor.cpp + iea15_rotor_data.hpp",
        "node_xi": node_xi,
        "node_coords": coords,
        "quadrature": quad,
        "sections": sections,
        "omega": omega,
        "hub_radius": hub,
        "gravity": gravity,
        "num_blades": 3,
        "max_iter": numbers(re.search(r"max_iter\((\d+)\)", body).group(1))[0],
        "step_size": numbers(re.search(r"step_size\(([^)]+)\)", body).group(1))[0],
        "rho_inf": numbers(re.search(r"rho_inf\(([^)]+)\)", body).group(1))[0],
        "t_end": numbers(re.search(r"t_end\(([^)]+)\)", body).group(1))[0],
        "checks": checks,
    }


FUNCTOR_TESTS = {
    # key: reference unit test (relative path) -- one functor call on named inputs each
    "strain": "tests/unit_tests/system/beams/test_calculate_strain.cpp",
    "Ouu": "tests/unit_tests/system/beams/test_calculate_Ouu.cpp",
    "Puu": "tests/unit_tests/system/beams/test_calculate_Puu.cpp",
    "Quu": "tests/unit_tests/system/beams/test_calculate_Quu.cpp",
    "force_FC": "tests/unit_tests/system/beams/test_calculate_force_FC.cpp",
    "force_FD": "tests/unit_tests/system/beams/test_calculate_force_FD.cpp",
    "temporary_variables": "tests/unit_tests/system/beams/test_calculate_temporary_variables.cpp",
    "gravity_force": "tests/unit_tests/system/masses/test_calculate_gravity_force.cpp",
    "gyroscopic_matrix": "tests/unit_tests/system/masses/test_calculate_gyroscopic_matrix.cpp",
    "inertia_stiffness_matrix": "tests/unit_tests/system/masses/test_calculate_inertia_stiffness_matrix.cpp",
    "inertial_forces": "tests/unit_tests/system/masses/test_calculate_inertial_forces.cpp",
    "mass_matrix_components": "tests/unit_tests/system/masses/test_calculate_mass_matrix_components.cpp",
    "rotate_section_matrix": "tests/unit_tests/system/masses/test_rotate_section_matrix.cpp",
    "next_state": "tests/unit_tests/state/test_calculate_next_state.cpp",
    "displacement": "tests/unit_tests/state/test_calculate_displacement.cpp",
    "dynamic_prediction": "tests/unit_tests/state/test_update_dynamic_prediction.cpp",
    "static_prediction": "tests/unit_tests/state/test_update_static_prediction.cpp",
    "algorithmic_acceleration": "tests/unit_tests/state/test_update_algorithmic_acceleration.cpp",
    "global_position": "tests/unit_tests/state/test_update_global_position.cpp",
}


def functor_case(text):
    """Named inputs (CreateView<...>("name", std::array{...})), scalar constants
    (`auto|double name = value;`) and expected outputs (`name_exact_data = std::array{...}`)."""
    inputs = {}
    for m in re.finditer(r'Create(?:Mutable|Left)?View<[^>]*>\(\s*"(\w+)",\s*std::array(?:<[^>]*>)?\s*\{', text):
        try:
            inputs[m.group(1)] = values(balanced(text, m.end() - 1))
        except (ValueError, SyntaxError):  # a non-numeric table (e.g. freedom signatures)
            continue
    scalars = {m.group(1): float(m.group(2))
               for m in re.finditer(r"(?:constexpr auto|const double|constexpr double)\s+(\w+)\s*=\s*(-?[\d.eE+-]+)\s*;", text)}
    expected = {}
    for m in re.finditer(r"(\w+)_exact_data\s*=\s*std::array(?:<[^>]*>)?\s*\{", text):
        expected[m.group(1)] = values(balanced(text, m.end() - 1))
    return {"inputs": inputs, "scalars": scalars, "expected": expected}


def functor_known_answers(root):
    out = {}
    for key, rel in FUNCTOR_TESTS.items():
        case = functor_case(strip_comments((root / rel).read_text()))
        case["source"] = rel
        out[key] = case
    return out


INTEGRATE_TESTS = {
    "integrate_stiffness": "tests/unit_tests/system/beams/test_integrate_stiffness_matrix.cpp",
    "integrate_inertia": "tests/unit_tests/system/beams/test_integrate_inertia_matrix.cpp",
    "integrate_residual": "tests/unit_tests/system/beams/test_integrate_residual_vector.cpp",
}


def statements(body):
    """Top-level `const auto NAME = ...;` and `constexpr auto NAME = ...;` statements of a body."""
    out = {}
    for m in re.finditer(r"(?:const|constexpr)\s+auto\s+(\w+)\s*=", body):
        depth, i = 0, m.end()
        while i < len(body):
            ch = body[i]
            if ch in "({[":
                depth += 1
            elif ch in ")}]":
                depth -= 1
            elif ch == ";" and depth == 0:
                break
            i += 1
        out[m.group(1)] = body[m.end():i].strip()
    return out


def view_data(expr, src):
    """Numbers a CreateView / CreateLeftView statement fills its view with (None: zero view)."""
    if not re.match(r"Create(?:Left|Mutable)?View", expr):
        return None
    args = split_top(balanced(expr, expr.index("(")))
    data = ",".join(args[1:]).strip()  # template argument lists hold commas too
    call = re.fullmatch(r"(\w+)\(\)", data)
    if call:  # a constexpr data function
        m = re.search(r"\b" + call.group(1) + r"\(\)\s*\{", src)
        body = balanced(src, m.end() - 1).split("return", 1)[1]
        return values(body[body.index("{"):body.rindex("}") + 1])
    return values(data[data.index("{"):])


def scalar_value(expr):
    m = re.fullmatch(r"size_t\{(\d+)U?\}", expr)
    if m:
        return float(m.group(1))
    m = re.fullmatch(r"-?[\d.]+(?:[eE][-+]?\d+)?", expr)
    return float(expr) if m else None


def integrate_cases(src, rel):
    """One entry per TEST: the integrator functor, its positional arguments by name, the arrays
    and scalars those names hold (a helper's parameters share the caller's names) and the
    expected output."""
    cases = {}
    for m in re.finditer(r"TEST\((\w+),\s*(\w+)\)\s*\{", src):
        body = balanced(src, m.end() - 1)
        call = re.fullmatch(r"\s*(\w+)\(\);\s*", body)
        if call:
            body = function_body(src, call.group(1))
        helper_tol = None
        helper = re.search(r"\b(TestIntegrate\w+)\(\s*qp_\w+", body)
        if helper:
            hm = re.search(r"void\s+" + helper.group(1) + r"\s*\(", src)
            params = balanced(src, hm.end() - 1)
            t = re.search(r"tolerance\s*=\s*([\d.eE+-]+)", params)
            helper_tol = float(t.group(1)) if t else None
            call_args = split_top(balanced(body, helper.end() - len(helper.group(0)) + len(helper.group(1))))
            if len(call_args) > 6:
                helper_tol = float(call_args[6])
            hbody = balanced(src, src.index("{", hm.end() + len(params)))
            body = body + "\n" + hbody
        st = statements(body)
        calls = list(re.finditer(r"beams::Integrate(\w+)Element<[^>]*>\s*\{", body))
        fm = calls[0]
        # one functor call per element (element index first); the expected output holds the
        # elements' results in that order
        args = [[a.strip() for a in split_top(balanced(body, c.end() - 1))] for c in calls]
        arrays, scalars = {}, {}
        for name, expr in st.items():
            if name.endswith("_data") or name.startswith("exact") or "exact" in name:
                continue
            sv = scalar_value(expr)
            if sv is not None:
                scalars[name] = sv
            else:
                d = view_data(expr, src)
                if d is not None:
                    arrays[name] = d
        exact = [v for k, v in st.items() if re.search(r"exact\w*_data$", k)]
        cases[m.group(2)] = {
            "functor": "Integrate" + fm.group(1) + "Element",
            "calls": args,
            "arrays": arrays,
            "scalars": scalars,
            "expected": values(exact[0][exact[0].index("{"):]),
            "tolerance": helper_tol if helper_tol is not None else 1e-15,
            "source": rel,
        }
    return cases


def integrate_known_answers(root):
    return {key: integrate_cases(strip_comments((root / rel).read_text()), rel) for key, rel in INTEGRATE_TESTS.items()}


def flat_values(braced, env=None):
    """Flattened element values of a possibly nested `{{...}, {...}}` initializer."""
    inner = balanced(braced.strip(), 0)
    out = []
    for t in split_top(inner):
        t = t.strip()
        if not t:
            continue
        out.extend(flat_values(t, env) if t.startswith("{") else [arith(t, env)])
    return out


def curved_beam(root):
    """Every constexpr data array of the curved-beam unit tests (a 3-node, 7-point element whose
    expected values come from the reference's Mathematica notebooks), by name."""
    rel = "tests/unit_tests/system/beams/test_curved_beam.hpp"
    text = strip_comments((root / rel).read_text())
    out = {"source": rel}
    for m in re.finditer(r"constexpr\s+(?:std::array<[^=]*>|auto)\s+(\w+)\s*=\s*(?:std::array(?:<[^>]*>)?)?\s*\{", text):
        out[m.group(1)] = flat_values(text[m.end() - 1:], out)
    out["kGLLNodes"] = values(re.search(r"kGLLNodes(\{[^}]*\})", text).group(1))
    return out


CRS_TESTS = {
    "row_ptrs": "tests/unit_tests/solver/test_compute_row_ptrs.cpp",
    "col_inds": "tests/unit_tests/solver/test_compute_col_inds.cpp",
}


def expectation_loops(body, env):
    """Evaluates the checks after the call under test: nested `for (auto v : std::views::iota(a,
    b)) { ... }` loops, `auto v = expr;` locals and `EXPECT_EQ(name(index), value);` -> a list of
    [name, index, value] (the loops run over literal ranges; expressions are arithmetic on the
    test's constants, loop variables and input arrays)."""
    checks = []

    def run(text, env):
        i = 0
        while i < len(text):
            m = re.compile(r"\s*for\s*\(\s*auto\s+(\w+)\s*:\s*std::views::iota\(").match(text, i)
            if m:
                rng = split_top(balanced(text, m.end() - 1))
                lo, hi = int(arith(rng[0], env)), int(arith(rng[1], env))
                j = text.index("{", m.end())
                inner = balanced(text, j)
                for v in range(lo, hi):
                    run(inner, dict(env, **{m.group(1): v}))
                i = j + len(inner) + 2
                continue
            m = re.compile(r"\s*(?:const\s+)?auto\s+(\w+)\s*=\s*([^;]+);").match(text, i)
            if m:
                try:
                    env = dict(env, **{m.group(1): arith(m.group(2), env)})
                except (ValueError, SyntaxError, KeyError):  # a mirror view or other non-number
                    pass
                i = m.end()
                continue
            m = re.compile(r"\s*EXPECT_EQ\(\s*(\w+)\(").match(text, i)
            if m:
                args = split_top(balanced(text, text.index("(", m.start() + m.group(0).index("EXPECT_EQ"))))
                idx = args[0][args[0].index("(") + 1: args[0].rindex(")")]
                checks.append([m.group(1), int(arith(idx, env)), arith(args[1], env)])
                i = text.index(";", m.end()) + 1
                continue
            i += 1

    run(body, env)
    return checks


def crs_cases(src, rel):
    out = {}
    for m in re.finditer(r"TEST\((\w+),\s*(\w+)\)\s*\{", src):
        body = balanced(src, m.end() - 1)
        env, arrays = {}, {}
        for c in re.finditer(r"constexpr auto (\w+)\s*=\s*([^;]+);", body):
            env[c.group(1)] = arith(c.group(2), env)
        for c in re.finditer(r'CreateView<(?:[^<>]|<[^<>]*>)*>\(\s*"(\w+)",', body):
            expr = balanced(body, c.end() - len(c.group(0)) + c.group(0).index("("))
            data = ",".join(split_top(expr)[1:])
            if "make_pair" in data:
                arrays[c.group(1)] = [[arith(a), arith(b)] for a, b in re.findall(r"make_pair\(([^,]+),\s*([^)]+)\)", data)]
            else:
                arrays[c.group(1)] = values(data[data.index("{"):])
        tail = body[body.index("::invoke("):]
        tail = tail[tail.index(";") + 1:]
        lookup = dict(env)
        if "row_ptrs" in arrays:
            lookup["row_ptrs_host"] = arrays["row_ptrs"]
        out[m.group(2)] = {"scalars": env, "arrays": arrays, "sorted": "sort_crs_graph" in body,
                           "checks": expectation_loops(tail, lookup), "source": rel}
    return out


def crs_known_answers(root):
    return {key: crs_cases(strip_comments((root / rel).read_text()), rel) for key, rel in CRS_TESTS.items()}


def near_values(body, pattern):
    """EXPECT_NEAR(<expr matching pattern>, value, tol) -> [(expr, value, tol)], in order."""
    out = []
    for m in re.finditer(r"EXPECT_NEAR\s*\(", body):
        args = split_top(balanced(body, m.end() - 1))
        expr = re.sub(r"\s+", "", args[0])
        if re.search(pattern, expr):
            out.append({"expr": expr, "value": numbers(args[1])[0], "tol": numbers(args[2])[0]})
    return out


def symbolic_matrix(text, body):
    """Entries of a brace-initialized matrix whose entries are literals or named constants
    (`const auto m = 1.;` / `constexpr auto j = 1.;` in the same body)."""
    consts = {m.group(1): float(m.group(2)) for m in
              re.finditer(r"(?:const|constexpr) auto (\w+) = (-?[\d.]+(?:[eE][-+]?\d+)?);", body)}
    vals = []
    for tok in re.findall(r"[A-Za-z_]\w*|-?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?", text):
        vals.append(consts[tok] if tok in consts else float(tok))
    return vals


def masses_springs(root):
    reg = root / "tests/regression_tests/regression"
    masses = strip_comments((reg / "test_masses.cpp").read_text())
    body = masses[masses.index("MassesTest, ExternalForce") :]
    ext = {
        "source": "tests/regression_tests/regression/test_masses.cpp (MassesTest.ExternalForce)",
        "mass_matrix": symbolic_matrix(balanced(body, body.index("{", body.index("model.AddMassElement"))), body),
        "time_step": numbers(re.search(r"time_step = ([^;]+);", body).group(1))[0],
        "step_parameters": numbers(re.search(r"StepParameters\(true, ([^)]*)\)", body).group(1)),
        "force_x": numbers(re.search(r"force_x = ([^;]+);", body).group(1))[0],
        "torque_y": numbers(re.search(r"torque_y = ([^;]+);", body).group(1))[0],
        "n_steps": numbers(re.search(r"n_steps = ([^;]+);", body).group(1))[0],
        "expect_v": near_values(body, r"^v_host"),
    }
    prec = strip_comments((reg / "test_precession.cpp").read_text())
    pbody = function_body(prec, "SetUpPrecessionTest")
    precession = {
        "source": "tests/regression_tests/regression/test_precession.cpp",
        "velocity": numbers(re.search(r"SetVelocity\(([^)]*)\)", pbody).group(1)),
        "mass_matrix": symbolic_matrix(balanced(pbody, pbody.index("{", pbody.index("model.AddMassElement"))), pbody),
        "max_iter": numbers(re.search(r"max_iter\((\d+)\)", pbody).group(1))[0],
        "step_size": numbers(re.search(r"step_size\(([^)]+)\)", pbody).group(1))[0],
        "rho_inf": numbers(re.search(r"rho_inf\(([^)]+)\)", pbody).group(1))[0],
        "num_steps": numbers(re.search(r"iota\(0, (\d+)\)", pbody).group(1))[0],
        "expect_q": near_values(pbody, r"^q_host"),
    }
    sm = strip_comments((reg / "test_spring_mass_system.cpp").read_text())
    sbody = sm[sm.index("SetUpSpringMassSystem") : sm.index("SetUpSpringMassChainSystem")]
    positions = [numbers(m.group(1)) for m in re.finditer(r"SetPosition\(([^)]*)\)", sbody)]
    spring_mass = {
        "source": "tests/regression_tests/regression/test_spring_mass_system.cpp (SpringMassSystemTest)",
        "positions": positions,
        "m": numbers(re.search(r"constexpr auto m = ([^;]+);", sbody).group(1))[0],
        "j": numbers(re.search(r"constexpr auto j = ([^;]+);", sbody).group(1))[0],
        "k": numbers(re.search(r"const auto k = ([^;]+);", sbody).group(1))[0],
        "l0": numbers(re.search(r"const auto l0 = ([^;]+);", sbody).group(1))[0],
        "num_steps": numbers(re.search(r"num_steps = ([^;]+);", sbody).group(1))[0],
        "max_iter": numbers(re.search(r"max_iter\((\d+)\)", sbody).group(1))[0],
        "rho_inf": numbers(re.search(r"rho_inf\(([^)]+)\)", sbody).group(1))[0],
        "expect_half": near_values(sbody[: sbody.index("time_step == num_steps)")], r"^q\(1,0\)"),
        "expect_end": near_values(sbody[sbody.index("time_step == num_steps)") :], r"^q\(1,0\)"),
    }
    chain = sm[sm.index("SetUpSpringMassChainSystem") :]
    spring_mass_chain = {
        "source": "tests/regression_tests/regression/test_spring_mass_system.cpp (SpringMassChainSystemTest)",
        "number_of_masses": numbers(re.search(r"number_of_masses = (\d+)U;", chain).group(1))[0],
        "spacing": numbers(re.search(r"constexpr auto displacement = ([^;]+);", chain).group(1))[0],
        "m": numbers(re.search(r"constexpr auto m = ([^;]+);", chain).group(1))[0],
        "j": numbers(re.search(r"constexpr auto j = ([^;]+);", chain).group(1))[0],
        "k": numbers(re.search(r"const auto k = ([^;]+);", chain).group(1))[0],
        "l0": numbers(re.search(r"AddSpringElement\([^,]+,[^,]+, k, ([^)]+)\)", chain).group(1))[0],
        "num_steps": numbers(re.search(r"num_steps = ([^;]+);", chain).group(1))[0],
        "max_iter": numbers(re.search(r"max_iter\((\d+)\)", chain).group(1))[0],
        "rho_inf": numbers(re.search(r"rho_inf\(([^)]+)\)", chain).group(1))[0],
        "expect_q_x": 0.0,
        "tol": numbers(re.search(r"EXPECT_NEAR\(q\(node, 0\), 0\., ([^)]+)\)", chain).group(1))[0],
    }
    ht = strip_comments((reg / "test_heavy_top.cpp").read_text())
    hb = function_body(ht, "SetUpHeavyTopTest")

    def arr3(name):
        return numbers(re.search(name + r" =\s*std::array\{([^}]*)\}", hb).group(1))

    heavy_top = {
        "source": "tests/regression_tests/regression/test_heavy_top.cpp",
        "gravity": numbers(re.search(r"SetGravity\(([^)]*)\)", hb).group(1)),
        "mass": numbers(re.search(r"constexpr auto mass = ([^;]+);", hb).group(1))[0],
        "inertia": arr3("inertia"),
        "x": arr3(r"const auto x"),
        "omega": arr3("omega"),
        "omega_dot": arr3("omega_dot"),
        "x_ddot": arr3("x_ddot"),
        "max_iter": numbers(re.search(r"max_iter\((\d+)\)", hb).group(1))[0],
        "step_size": numbers(re.search(r"step_size\(([^)]+)\)", hb).group(1))[0],
        "rho_inf": numbers(re.search(r"rho_inf\(([^)]+)\)", hb).group(1))[0],
        "atol": numbers(re.search(r"a_tol\(([^)]+)\)", hb).group(1))[0],
        "rtol": numbers(re.search(r"r_tol\(([^)]+)\)", hb).group(1))[0],
        "num_steps": numbers(re.search(r"iota\(0, (\d+)\)", hb).group(1))[0],
        "expect_q": near_values(hb, r"^q\[0\]"),
    }
    return {"masses_external_force": ext, "precession": precession, "spring_mass": spring_mass,
            "spring_mass_chain": spring_mass_chain, "heavy_top": heavy_top}


def blade_interface(root):
    """BladeInterfaceTest.RotatingBeam and .BladeWindIO: builder inputs, solution parameters,
    per-step inputs and the expected node states.  The windIO blade data come from the YAML
    file the test reads (loaded with yaml.safe_load: data only)."""
    import yaml

    src = strip_comments((root / "tests/regression_tests/interfaces/test_blade_interface.cpp").read_text())
    rb = src[src.index("BladeInterfaceTest, RotatingBeam") : src.index("WriteMatrixToFile")]
    sect = rb[rb.index("builder.Blade().AddSection(") :]
    mats = re.findall(r"std::array\{\s*(std::array\{[^;]*?\})\s*,\s*std::array\{\s*(std::array\{[^;]*?\}),\s*interfaces", sect, re.S)
    M = numbers(mats[0][0])
    K = numbers(mats[0][1])
    rotating = {
        "source": "tests/regression_tests/interfaces/test_blade_interface.cpp (BladeInterfaceTest.RotatingBeam)",
        "time_step": numbers(re.search(r"time_step\{([^}]*)\}", rb).group(1))[0],
        "omega": numbers(re.search(r"omega = std::array\{([^}]*)\}", rb).group(1)),
        "x0_root": numbers(re.search(r"x0_root = std::array\{([^}]*)\}", rb).group(1)),
        "max_iter": numbers(re.search(r"SetMaximumNonlinearIterations\(([^)]*)\)", rb).group(1))[0],
        "atol": numbers(re.search(r"SetAbsoluteErrorTolerance\(([^)]*)\)", rb).group(1))[0],
        "rtol": numbers(re.search(r"SetRelativeErrorTolerance\(([^)]*)\)", rb).group(1))[0],
        "rho_inf": numbers(re.search(r"SetDampingFactor\(([^)]*)\)", rb).group(1))[0],
        "element_order": numbers(re.search(r"SetElementOrder\(([^)]*)\)", rb).group(1))[0],
        "node_s": numbers(balanced(rb, rb.index("{", rb.index("node_s = std::vector")))),
        "section_s": numbers(balanced(rb, rb.index("{", rb.index("section_s")))),
        "mass_matrix": M,
        "stiffness_matrix": K,
        "num_steps": numbers(re.search(r"iota\(1U, (\d+)U\)", rb).group(1))[0] - 1,
        "expect": near_values(rb, r"^interface\.Blade\(\)\.nodes"),
    }
    wi = src[src.index("BladeInterfaceTest, BladeWindIO") : src.index("BladeInterfaceTest, RotatingBeam")]
    wio = yaml.safe_load((root / "tests/regression_tests/interfaces/interfaces_test_files/IEA-15-240-RWT.yaml").read_text())
    blade = wio["components"]["blade"]
    ra = blade["outer_shape_bem"]["reference_axis"]
    six = blade["elastic_properties_mb"]["six_x_six"]
    windio = {
        "source": "tests/regression_tests/interfaces/test_blade_interface.cpp (BladeInterfaceTest.BladeWindIO), "
                  "blade data from interfaces_test_files/IEA-15-240-RWT.yaml",
        "time_step": numbers(re.search(r"time_step\{([^}]*)\}", wi).group(1))[0],
        "max_iter": numbers(re.search(r"SetMaximumNonlinearIterations\(([^)]*)\)", wi).group(1))[0],
        "atol": numbers(re.search(r"SetAbsoluteErrorTolerance\(([^)]*)\)", wi).group(1))[0],
        "rtol": numbers(re.search(r"SetRelativeErrorTolerance\(([^)]*)\)", wi).group(1))[0],
        "rho_inf": numbers(re.search(r"SetDampingFactor\(([^)]*)\)", wi).group(1))[0],
        "n_nodes": numbers(re.search(r"n_nodes\{([^}]*)\}", wi).group(1))[0],
        "num_steps": numbers(re.search(r"iota\(1U, (\d+)U\)", wi).group(1))[0] - 1,
        "tip_moment_amplitude": numbers(re.search(r"loads\[1\] = ([^*]+)\*", wi).group(1))[0],
        "tip_moment_frequency": numbers(re.search(r"numbers::pi \* ([\d.]+) \* t", wi).group(1))[0],
        "ref_axis_grid": ra["x"]["grid"],
        "ref_axis_x": ra["x"]["values"], "ref_axis_y": ra["y"]["values"], "ref_axis_z": ra["z"]["values"],
        "twist_grid": blade["outer_shape_bem"]["twist"]["grid"],
        "twist_values_deg": blade["outer_shape_bem"]["twist"]["values"],
        "section_grid": six["stiff_matrix"]["grid"],
        "mass_upper21": six["inertia_matrix"]["values"],
        "stiffness_upper21": six["stiff_matrix"]["values"],
        "expect": near_values(wi, r"^tip_node\.position"),
    }
    return {"rotating_beam": rotating, "windio": windio}


def test_body(src, suite, name):
    m = re.search(r"TEST\(\s*" + suite + r"\s*,\s*" + name + r"\s*\)\s*\{", src)
    return balanced(src, m.end() - 1)


def rotating_beam_constraints(root):
    """RotatingBeamTest cases that join the 6-node beam to hub/ground nodes through constraints:
    MasslessConstraints (rigid joint + prescribed hub), RotationControlConstraint (pitch
    control, fixed hub), CompoundRotationControlConstraint (pitch control on a prescribed
    rotating hub), RevoluteJointConstraint (free hub on a revolute joint to a fixed ground
    node) and GeneratorTorque_Tilt0/_Tilt90 (generator torque on a revolute shaft).  Stored:
    the file's section matrices, node locations and quadrature, each case's scalar inputs and
    its expected values with the test's tolerance (expect_kokkos_view_*_equal: 1e-6)."""
    src = strip_comments((root / "tests/regression_tests/regression/test_rotating_beam.cpp").read_text())
    head = src[: src.index("TEST(")]

    def arr(name):
        i = head.index(name)
        return numbers(balanced(head, head.index("{", i)))

    def scalar(body, pat):
        m = re.search(pat, body)
        return numbers(m.group(1))[0] if m else None

    out = {
        "source": "tests/regression_tests/regression/test_rotating_beam.cpp",
        "mass_matrix": arr("mass_matrix = std::array"),
        "stiffness_matrix": arr("stiffness_matrix = std::array"),
        "node_s": arr("node_s = std::vector"),
        "quadrature": arr("quadrature = std::vector"),
        "view_tol": 1e-6,
    }

    def params(body):
        return {
            "omega": scalar(body, r"omega = ([^;]+);"),
            "hub_size": scalar(body, r"hub_size = ([^;]+);"),
            "max_iter": scalar(body, r"max_iter\((\d+)\)"),
            "step_size": scalar(body, r"step_size\(([^)]+)\)"),
            "rho_inf": scalar(body, r"rho_inf\(([^)]+)\)"),
            "num_steps": scalar(body, r"iota\(0, (\d+)\)"),
        }

    for name in ("MasslessConstraints", "RotationControlConstraint", "RevoluteJointConstraint"):
        body = test_body(src, "RotatingBeamTest", name)
        e = params(body)
        e["expect_q"] = expectations(body)[0]["values"]
        out[name] = e
    body = test_body(src, "RotatingBeamTest", "CompoundRotationControlConstraint")
    e = params(body)
    e["expect_rv"] = near_values(body, r"^rv\[")
    out["CompoundRotationControlConstraint"] = e
    helper = function_body_args(src, "GeneratorTorqueWithAxisTilt")
    gp = params(helper)
    gp["torque"] = scalar(helper, r"auto torque = ([^;]+);")
    out["GeneratorTorque"] = gp
    for name, tilt in (("GeneratorTorque_Tilt0", 0.0), ("GeneratorTorque_Tilt90", math.pi / 2.0)):
        body = test_body(src, "RotatingBeamTest", name)
        groups = [numbers(g) for g in re.findall(r"\{([^{}]*)\}", body[body.index("(") :])]
        out[name] = {"tilt": tilt, "expect_azimuth_q": groups[0], "expect_azimuth_v": groups[1],
                     "expect_revolute_output": groups[2]}
    return out


def function_body_args(text, name):
    """Body of a free function with parameters: `void name(...) {...}`."""
    m = re.search(r"\bvoid\s+" + re.escape(name) + r"\s*\(", text)
    i = m.end() - 1
    balanced(text, i)
    depth, j = 0, i
    for j in range(i, len(text)):
        if text[j] == "(":
            depth += 1
        elif text[j] == ")":
            depth -= 1
            if depth == 0:
                break
    k = text.index("{", j)
    return balanced(text, k)


def main():
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    (HERE / "rotating_beam_constraints.json").write_text(json.dumps(rotating_beam_constraints(root), indent=1))
    (HERE / "blade_interface.json").write_text(json.dumps(blade_interface(root), indent=1))
    (HERE / "solver_regression.json").write_text(json.dumps(solver_regression(root), indent=1))
    (HERE / "cantilever_sine.json").write_text(json.dumps(cantilever(root), indent=1))
    (HERE / "iea15_rotor.json").write_text(json.dumps(iea15(root)))
    (HERE / "functor_known_answers.json").write_text(json.dumps(functor_known_answers(root), indent=1))
    (HERE / "integrate_known_answers.json").write_text(json.dumps(integrate_known_answers(root), indent=1))
    (HERE / "curved_beam.json").write_text(json.dumps(curved_beam(root), indent=1))
    (HERE / "crs_known_answers.json").write_text(json.dumps(crs_known_answers(root), indent=1))
    (HERE / "masses_springs.json").write_text(json.dumps(masses_springs(root), indent=1))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
