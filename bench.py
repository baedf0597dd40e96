"""Bench: batched FP32-tangent time steps/s vs plain FP64 on the 4096-rotor ensemble (1 GPU).

Workload (one "step" = one generalized-alpha time step of every model): 4096 independent
3-blade rotors, one beam element per blade with N uniform in {6, 11, 16, 21} GLL nodes and
Q = 2N - 2 Gauss points, L in [30, 120] m, hub speed in [0.5, 1.5] rad/s, gravity, tip loads
normal(0, 1e4 N) in every step (correlated from step to step over 0.1 s: with loads drawn
afresh each step the reference's Newton iteration diverges on some rotors), h = 0.01 s,
rho_inf = 0.9, Newton with the reference's increment criterion (atol 1e-10, rtol 1e-6).  Inputs are synthetic (SplitMix64 seed
20261202) and resident in HBM (device input schedule) before timing starts.

Timed legs (all on the one GPU; each the median of --repeats runs of --steps steps after the
warm-up, every run timed with a synchronize on both sides of its step loop):
  value            fused FP32-tangent form with --refine FP64 iterative-refinement sweeps per
                   solve in the Newton iterations whose previous convergence error is below
                   --refine-bound (default 10: in practice the iteration that converges, the
                   only one whose solve error survives into the state), one hipGraph-replayed
                   kernel per step; null (with `error` set) if any model-step of this leg failed
                   to converge
  fp32_refine_every_iteration  the same with the sweeps in every Newton iteration
  fp32_without_refinement  the same without refinement
  fused_fp64       fused form, FP64 tangent
  plain_fp64       the reference's launch structure (one kernel per stage, host reads the
                   convergence error every Newton iteration), FP64, --plain-steps per run
Roofline: the fused FP32 step kernel; algorithmic HBM bytes per launch, 8(67N + 87Q) + 36 s N^2
per element and Newton iteration (s = 4), summed over the iterations actually taken, over its
HIP-event-timed duration; the kernel is latency-bound (the `bound` field names the byte
roofline it is measured against), so the matrix-core rate of its MFMA work (contraction
144 N^2 Q of the element sizes on the matrix cores, LU trailing updates) is reported beside it
against the FP32 MFMA peak.
accuracy (extra field, untimed): the FP32-tangent form (the benched refinement policy, and
refinement in every iteration) and the FP64 fused form against the CPU FP64 oracle on the first
--acc-rotors rotors over --acc-steps steps (100 steps, the ensemble's length): final q
per model relative to that model's max |q|, worst and median, models above 1e-9, and Newton
iteration totals that differ from the oracle's.  The FP32 form is an inexact-tangent Newton
method: where its path takes one iteration more or less than the oracle's in some step, the
final state differs by up to a few 1e-9 of the model's max |q| (the figure is reported as
measured, per model, not held to a bar here; tests/test_gpu_parity.py holds smaller cases).
config3 (extra field, untimed): the FP32-tangent form of the config-1 cantilever over 200 steps
against the oracle: max-norm tip displacement and rotation-vector history error, Newton counts,
refinement sweeps per solve.
cpu_baseline: the CPU oracle (same algorithm, dense pivoted LU instead of KLU), rebuilt with
-march=native on the host it runs on, on a bounded sample of the same ensemble, OpenMP over
rotors on all host cores; a single-thread run of a smaller sample beside it.
structure (extra field): the large single structure (tower 8 + shaft 8 + 3 x 16 blade elements
of 21 nodes / 40 points, joints, 8459 dofs) in the plain form (CRS assembly, condensed solve),
FP64 and the FP32-tangent option (FP32 element factors, two FP64 refinement sweeps); steps/s
and the element kernel's algorithmic HBM rate.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import openturbine_amd as ota  # noqa: E402
from openturbine_amd import workloads as W  # noqa: E402

METRIC = "time steps/s + qp-kernel HBM GB/s: batched FP32-tangent vs plain FP64, 1 GPU"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def block_flops(specs):
    """Per model: flops of one Newton iteration summed over its blocks (blades)."""
    out = []
    for sp in specs:
        f = 0.0
        for b in sp.beams:
            N = len(b.nodes)
            Q = len(b.quad)
            n = 6 * N + 6
            f += 144.0 * N * N * Q + (2.0 / 3.0) * n**3 + 2.0 * n * n
        out.append(f)
    return np.array(out)


def block_bytes(specs, s):
    """Per model: algorithmic HBM bytes of one Newton iteration, 8(67N + 87Q) + 36 s N^2 per
    element (element state/section inputs and its tangent at s bytes per entry)."""
    out = []
    for sp in specs:
        b_ = 0.0
        for b in sp.beams:
            N = len(b.nodes)
            Q = len(b.quad)
            b_ += 8.0 * (67 * N + 87 * Q) + 36.0 * s * N * N
        out.append(b_)
    return np.array(out)


def contraction_mfma_mask():
    """Bit N set: elements of N nodes contract on the matrix cores (the tuning table compiled into
    the library, openturbine_amd/csrc/kyn_tuning.inc); all bits set if the table is not found."""
    import re

    path = ROOT / "openturbine_amd" / "csrc" / "kyn_tuning.inc"
    m = re.search(r"kContractMfmaMask\s*=\s*0x([0-9a-fA-F]+)", path.read_text()) if path.exists() else None
    return int(m.group(1), 16) if m else ~0


def contraction_flops(specs, mask=~0):
    """Per model: node-pair contraction flops of one Newton iteration, 144 N^2 Q per element, over
    the elements whose size N has bit N set in `mask`."""
    return np.array([sum(144.0 * len(b.nodes) ** 2 * len(b.quad) for b in sp.beams if (mask >> len(b.nodes)) & 1)
                     for sp in specs])


def lu_update_flops(specs):
    """Per model: flops of the FP32 LU's trailing updates on the matrix cores per Newton
    iteration: sum over 16-column panels of 2 * 16 * m * (m + 1), m the rows after the panel."""
    out = []
    for sp in specs:
        f = 0.0
        for b in sp.beams:
            n = 6 * len(b.nodes) + 6
            for k0 in range(0, n - 16, 16):
                m = n - k0 - 16
                f += 2.0 * 16 * m * (m + 1)
        out.append(f)
    return np.array(out)


def measured_traffic(refine, refine_bound):
    """HBM bytes per launch of the FP32 step kernel from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, copy-calibrated), if it was taken with this
    refinement setting."""
    path = ROOT / "profiles" / "k_fused_step_f32_ensemble4096_traffic.txt"
    if not path.exists():
        return None
    fields = {}
    for line in path.read_text().splitlines():
        if ":" in line and line.split(":", 1)[0] in ("refine_sweeps", "refine_err_bound", "traffic_bytes_per_launch"):
            k, v = line.split(":", 1)
            fields[k] = float(v)
    if fields.get("refine_sweeps") != refine or fields.get("refine_err_bound") != refine_bound:
        return None
    if "traffic_bytes_per_launch" not in fields:
        return None
    return {"bytes_per_launch": fields["traffic_bytes_per_launch"], "source": str(path.relative_to(ROOT))}


def run_leg(lib, specs, sched, *, exec_mode, precision, warmup, steps, repeats=1, refine=0, refine_bound=0.0,
            use_graph=True):
    """Warm-up steps, then `repeats` timed runs of `steps` steps each (synchronized on both sides
    by step_many); the leg's rate is the median run."""
    t0 = time.time()
    sysm = ota.KynSystem(lib, specs, exec_mode=exec_mode, precision=precision, refine_sweeps=refine,
                         refine_err_bound=refine_bound)
    log(f"  system built in {time.time() - t0:.1f}s ({sysm.n} nodes, {sysm.num_models} models)")
    cdisp, tips, loads = sched
    total = warmup + repeats * steps
    sysm.set_schedule(total, cdisp[:total], tips, loads[:total])
    if warmup:
        sysm.step_many(warmup, use_graph=use_graph)
    walls, kerns, rc = [], [], 0
    it0, _ = sysm.totals()
    for _ in range(repeats):
        t0 = time.perf_counter()
        rc |= sysm.step_many(steps, use_graph=use_graph)
        walls.append(time.perf_counter() - t0)
        kerns.append(sysm.last_kernel_ms())
    it1, fl = sysm.totals()
    order = np.argsort(walls)
    med = int(order[len(order) // 2])
    return {"system": sysm, "rc": rc, "wall_s": walls[med], "walls": walls, "kernel_ms": kerns[med],
            "iters": (it1 - it0) / repeats, "fails": int(fl.sum())}


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes (torch's elementwise copy kernel): the reference
    reading for the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) that scale the step
    kernel's HBM traffic."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "read_KiB": nbytes / 1024, "write_KiB": nbytes / 1024}),
          flush=True)


def structure_legs(lib, steps):
    """Large single structure, plain form: host-driven steps (tip loads set each step)."""
    spec = W.structure_config_spec(num_steps=steps + 1)
    out = {"workload": "tower 8 + shaft 8 + 3 blades x 16 elements, N=21, Q=40, rigid joints between "
                       "elements, revolute hub joint, h=0.005, tip loads normal(0, 1e4 N)",
           "dofs": int(spec.num_nodes * 6 + sum(5 if c.type == W.REVOLUTE_JOINT else 6 for c in spec.constraints))}
    elem_bytes = sum(8.0 * (67 * len(b.nodes) + 87 * len(b.quad)) + 36.0 * 8 * len(b.nodes) ** 2 for b in spec.beams)
    for name, prec, refine in (("fp64", ota.TANGENT_FP64, 0), ("fp32", ota.TANGENT_FP32, 2)):
        g = ota.KynSystem(lib, spec, exec_mode=ota.EXEC_PLAIN, precision=prec, refine_sweeps=refine)
        assert g.step(spec.step_inputs(0)) == 0  # warm-up
        iters, kms, fails = 0, 0.0, 0
        t0 = time.perf_counter()
        for k in range(1, steps + 1):
            fails += g.step(spec.step_inputs(k)) != 0
            iters += len(g.convergence())
            kms += g.last_kernel_ms()
        wall = time.perf_counter() - t0
        out[name] = {"steps_per_s": round(steps / wall, 3), "newton_iters_per_step": round(iters / steps, 2),
                     "failed_steps": fails}
        if name == "fp64":
            ms = kms / steps
            out["element_kernel"] = {"kernel": "k_plain_element", "ms_per_launch": round(ms, 4),
                                     "algorithmic_bytes_per_launch": elem_bytes,
                                     "GB_s": round(elem_bytes / (ms / 1e3) / 1e9, 1)}
        log(f"  structure {name}: {steps} steps in {wall:.3f}s, {iters / steps:.1f} iters/step, fails {fails}")
    return out


def native_oracle():
    """The oracle rebuilt from its sources with -march=native for the host this runs on (into a
    temporary directory); the shipped x86-64-v3 build if no compiler is there."""
    import subprocess
    import tempfile

    out = Path(tempfile.gettempdir()) / f"liboracle_kyn_native_{os.getpid()}.so"
    try:
        subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "oracle"), "ARCH=native", f"OUT={out}", str(out)], check=True,
                       capture_output=True, timeout=300)
        return out, "native"
    except (OSError, subprocess.SubprocessError):
        return ROOT / "oracle" / "liboracle_kyn.so", "x86-64-v3"


def cpu_baseline(specs_sample, steps, lib_path, threads=None):
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleLib, OracleSystem, step_many

    if threads is not None:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    lib = OracleLib(lib_path)
    if threads is not None:
        import ctypes as C

        try:  # the OpenMP runtime was already initialized by an earlier load: set it directly
            C.CDLL("libgomp.so.1").omp_set_num_threads(int(threads))
        except OSError:
            pass
    systems = [OracleSystem(lib, sp) for sp in specs_sample]
    t0 = time.perf_counter()
    for k in range(steps):
        inputs = [sp.step_inputs(k) for sp in specs_sample]
        st = step_many(lib, systems, inputs)
        assert all(s == 0 for s in st), "oracle did not converge"
    return time.perf_counter() - t0


def accuracy_leg(lib, lib_path, rotors, steps, refine, refine_bound):
    """FP32-tangent (refine sweeps under the benched bound, and in every iteration) and FP64 fused
    forms against the FP64 oracle, per model."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleLib, OracleSystem, step_many

    specs = W.ensemble_config_specs(num_rotors=rotors, num_steps=steps)
    sched = W.ensemble_schedule(specs, steps)
    out = {"rotors": rotors, "steps": steps}
    res = {}
    forms = [("fp32", ota.TANGENT_FP32, refine, refine_bound)]
    if refine > 0 and refine_bound > 0:
        forms.append(("fp32_refine_every_iteration", ota.TANGENT_FP32, refine, 0.0))
    forms.append(("fp64", ota.TANGENT_FP64, 0, 0.0))
    for name, prec, rf, bound in forms:
        g = ota.KynSystem(lib, specs, exec_mode=ota.EXEC_FUSED, precision=prec, refine_sweeps=rf,
                          refine_err_bound=bound)
        g.set_schedule(steps, *sched)
        g.step_many(steps, use_graph=True)
        it, fl = g.totals()
        res[name] = (g.get("q"), it, int(fl.sum()))
        offs = np.append(np.asarray(g.node_off, dtype=np.int64), g.n)
        del g
    olib = OracleLib(lib_path)
    systems = [OracleSystem(olib, sp) for sp in specs]
    its = np.zeros(rotors, dtype=np.int64)
    t0 = time.perf_counter()
    for k in range(steps):
        st = step_many(olib, systems, [sp.step_inputs(k) for sp in specs])
        assert all(v == 0 for v in st), "oracle did not converge"
        its += np.array([len(s.convergence()) for s in systems])
    log(f"  oracle: {rotors} rotors x {steps} steps in {time.perf_counter() - t0:.1f}s")
    qo = np.concatenate([s.get("q") for s in systems])
    for name, (q, it, fails) in res.items():
        d = np.max(np.abs(q - qo), axis=1)
        m = np.max(np.abs(qo), axis=1)
        per_model = np.maximum.reduceat(d, offs[:-1]) / np.maximum.reduceat(m, offs[:-1])
        out[name] = {"q_rel_worst_model": float(np.max(per_model)), "q_rel_median_model": float(np.median(per_model)),
                     "models_above_1e-9": int(np.sum(per_model > 1e-9)),
                     "models_with_other_newton_total": int(np.sum(it != its)), "failed_model_steps": fails}
        if name.startswith("fp32"):
            out[name]["refine_sweeps"] = refine
            out[name]["refine_err_bound"] = refine_bound if name == "fp32" else 0.0
        log(f"  {name} vs oracle after {steps} steps: worst model {out[name]['q_rel_worst_model']:.2e}, median "
            f"{out[name]['q_rel_median_model']:.2e}, {out[name]['models_above_1e-9']} of {rotors} above 1e-9, "
            f"{out[name]['models_with_other_newton_total']} with other Newton totals")
    return out


def config3_leg(lib, lib_path, refine, refine_bound, steps=200):
    """The FP32-tangent form of the follower-load cantilever (200 steps) against the oracle."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleLib, OracleSystem

    spec = W.cantilever_config_spec(num_steps=steps)
    tip = spec.meta["tip_node"]
    g = ota.KynSystem(lib, spec, exec_mode=ota.EXEC_FUSED, precision=ota.TANGENT_FP32, refine_sweeps=refine,
                      refine_err_bound=refine_bound)
    o = OracleSystem(OracleLib(lib_path), spec)
    du = dr = umax = rmax = 0.0
    it32 = it64 = differ = 0
    for _ in range(steps):
        inputs = spec.step_inputs(g.step_index, o.get)
        g.step(inputs)
        o.step(inputs)
        n32, n64 = len(g.convergence()), len(o.convergence())
        it32, it64, differ = it32 + n32, it64 + n64, differ + int(n32 != n64)
        qg, qo = g.get("q")[tip], o.get("q")[tip]
        du, umax = max(du, float(np.max(np.abs(qg[:3] - qo[:3])))), max(umax, float(np.max(np.abs(qo[:3]))))
        rg, ro = W.rotvec_from_quat(qg[3:]), W.rotvec_from_quat(qo[3:])
        dr, rmax = max(dr, float(np.max(np.abs(rg - ro)))), max(rmax, float(np.max(np.abs(ro))))
    out = {"steps": steps, "refine_sweeps_per_solve": refine, "refine_err_bound": refine_bound, "tip_displacement_history_error": du / umax,
           "tip_rotation_vector_history_error": dr / rmax, "newton_iterations": {"fp32": it32, "oracle_fp64": it64},
           "steps_with_other_newton_count": differ}
    log(f"  config 3: tip displacement {out['tip_displacement_history_error']:.2e}, rotation "
        f"{out['tip_rotation_vector_history_error']:.2e}, iterations {it32} vs {it64}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--repeats", type=int, default=5, help="timed runs per leg (median reported)")
    ap.add_argument("--rotors", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-rotors", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--cpu1-rotors", type=int, default=256, help="single-thread CPU sample (rotors, same steps)")
    ap.add_argument("--acc-rotors", type=int, default=512, help="accuracy leg against the oracle (0: skip)")
    ap.add_argument("--acc-steps", type=int, default=100)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--refine", type=int, default=1, help="FP64 refinement sweeps of the FP32-tangent solve")
    ap.add_argument("--refine-bound", type=float, default=10.0,
                    help="refine only Newton iterations whose previous convergence error is below this (0: every "
                         "iteration, timed beside the default as fp32_refine_every_iteration)")
    ap.add_argument("--skip-unrefined", action="store_true", help="skip the FP32 leg without refinement")
    ap.add_argument("--skip-refine-all", action="store_true",
                    help="skip the FP32 leg with refinement in every Newton iteration")
    ap.add_argument("--structure-steps", type=int, default=10, help="steps of the large-structure legs (0: skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the ensemble "
              "would shard by rotor across GPUs with no per-iteration exchange", file=sys.stderr)
        sys.exit(2)

    lib = ota.KynLib()
    name, arch = lib.device_info()
    log(f"device {name} ({arch})")
    t0 = time.time()
    nsched = args.warmup + args.repeats * max(args.steps, args.plain_steps)
    specs = W.ensemble_config_specs(num_rotors=args.rotors, num_steps=nsched)
    log(f"built {len(specs)} rotor specs in {time.time() - t0:.1f}s")
    t0 = time.time()
    sched = W.ensemble_schedule(specs, nsched)
    log(f"built input schedule ({nsched} steps) in {time.time() - t0:.1f}s")

    flops_iter = block_flops(specs)
    bytes_iter32 = block_bytes(specs, 4)

    def leg(title, **kw):
        log(f"leg: {title}")
        r = run_leg(lib, specs, sched, **kw)
        log(f"  {kw['steps']} steps x {kw.get('repeats', 1)} runs: median {r['wall_s']:.3f}s "
            f"(runs {', '.join(f'{w:.3f}' for w in r['walls'])}), fails {r['fails']}, mean iters/model/step "
            f"{r['iters'].mean() / kw['steps']:.2f}, kernel {r['kernel_ms']:.3f} ms")
        return r

    r32 = leg("fused FP32 tangent (hipGraph)", exec_mode=ota.EXEC_FUSED, precision=ota.TANGENT_FP32,
              warmup=args.warmup, steps=args.steps, repeats=args.repeats, refine=args.refine,
              refine_bound=args.refine_bound)
    del r32["system"]
    rall = None
    if args.refine > 0 and args.refine_bound > 0 and not args.skip_refine_all:
        rall = leg("fused FP32 tangent, refinement in every Newton iteration (hipGraph)", exec_mode=ota.EXEC_FUSED,
                   precision=ota.TANGENT_FP32, warmup=args.warmup, steps=args.steps, repeats=args.repeats,
                   refine=args.refine, refine_bound=0.0)
        del rall["system"]
    r0 = None
    if args.refine > 0 and not args.skip_unrefined:
        r0 = leg("fused FP32 tangent without refinement (hipGraph)", exec_mode=ota.EXEC_FUSED,
                 precision=ota.TANGENT_FP32, warmup=args.warmup, steps=args.steps, repeats=args.repeats, refine=0)
        del r0["system"]
    r64 = None
    if not args.skip_fp64:
        r64 = leg("fused FP64", exec_mode=ota.EXEC_FUSED, precision=ota.TANGENT_FP64, warmup=args.warmup,
                  steps=args.steps, repeats=args.repeats)
        del r64["system"]
    rp = None
    if not args.skip_plain:
        rp = leg("plain FP64 (reference launch structure)", exec_mode=ota.EXEC_PLAIN, precision=ota.TANGENT_FP64,
                 warmup=1, steps=args.plain_steps, repeats=args.repeats, use_graph=False)
        del rp["system"]

    oracle_path, oracle_arch = native_oracle() if not (args.skip_cpu and args.acc_rotors <= 0) else (None, None)
    accuracy = None
    if args.acc_rotors > 0:
        log(f"accuracy: FP32 (refine {args.refine}) and FP64 fused forms vs the oracle, {args.acc_rotors} rotors x "
            f"{args.acc_steps} steps")
        accuracy = accuracy_leg(lib, oracle_path, args.acc_rotors, args.acc_steps, args.refine, args.refine_bound)
        log("config 3: FP32-tangent cantilever vs the oracle, 200 steps")
        config3 = {f"refine{rf}": config3_leg(lib, oracle_path, rf, args.refine_bound) for rf in sorted({0, args.refine})}
    else:
        config3 = None

    structure = None
    if args.structure_steps > 0:
        log("leg: large structure (plain form, condensed solve)")
        structure = structure_legs(lib, args.structure_steps)

    failed = r32["fails"] > 0 or r32["rc"] != 0
    value = None if failed else args.steps / r32["wall_s"]
    ms_per_step = 1000.0 * r32["wall_s"] / args.steps
    # roofline of the fused FP32 step kernel (HIP events around each graph launch of the median
    # run, on the library's stream): algorithmic HBM bytes 8(67N + 87Q) + 36 s N^2 per element
    # and Newton iteration (s = 4), summed over the iterations each model actually took
    iters_per_step = r32["iters"] / args.steps
    bytes_launch = float(np.sum(iters_per_step * bytes_iter32))
    flops_launch = float(np.sum(iters_per_step * flops_iter))
    mfma_flops_launch = float(np.sum(iters_per_step * contraction_flops(specs, contraction_mfma_mask())))
    lu_mfma_flops_launch = float(np.sum(iters_per_step * lu_update_flops(specs)))
    kern_s = (r32["kernel_ms"] if r32["kernel_ms"] > 0 else ms_per_step) / 1000.0
    achieved = bytes_launch / kern_s / 1e9
    traffic = measured_traffic(args.refine, args.refine_bound)
    plain_rate = args.plain_steps / rp["wall_s"] if rp else None
    result = {
        "metric": METRIC,
        "value": round(value, 4) if value is not None else None,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-tangent/f64-residual",
        "data": "synthetic",
        "config": {"workload": f"ensemble of {args.rotors} independent 3-blade rotors, N in {{6,11,16,21}}, "
                               "Q=2N-2, h=0.01, fused one-kernel-per-step, hipGraph",
                   "models": args.rotors, "blades": 3 * args.rotors, "refine_sweeps": args.refine,
                   "refine_err_bound": args.refine_bound,
                   "timing": f"median of {args.repeats} runs of {args.steps} steps after {args.warmup} warm-up"},
        "error": "the FP32-tangent leg had non-converged model-steps" if failed else None,
        "fused_fp64_steps_per_s": round(args.steps / r64["wall_s"], 4) if r64 else None,
        "plain_fp64_steps_per_s": round(plain_rate, 4) if rp else None,
        "speedup_fp32_fused_vs_plain_fp64": round(value / plain_rate, 2) if (rp and value) else None,
        "newton_iters_per_model_step": {"fp32": round(float(r32["iters"].mean()) / args.steps, 3),
                                         "fp64": round(float(r64["iters"].mean()) / args.steps, 3) if r64 else None},
        "failed_model_steps": {"fp32": r32["fails"], "fp64": r64["fails"] if r64 else None},
        "fp32_refine_every_iteration": {"steps_per_s": round(args.steps / rall["wall_s"], 4),
                                        "kernel_ms": round(rall["kernel_ms"], 4),
                                        "failed_model_steps": rall["fails"]} if rall else None,
        "fp32_without_refinement": {"steps_per_s": round(args.steps / r0["wall_s"], 4),
                                    "kernel_ms": round(r0["kernel_ms"], 4)} if r0 else None,
        "accuracy_vs_oracle": accuracy,
        "config3_fp32_cantilever": config3,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 5),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "k_fused_step_f32", "kernel_ms": round(kern_s * 1000.0, 4),
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "note": "latency-bound kernel: HBM and matrix-core rates both far from their peaks",
                     "mfma": {"flops_per_launch": mfma_flops_launch + lu_mfma_flops_launch,
                              "contraction_flops": mfma_flops_launch, "lu_update_flops": lu_mfma_flops_launch,
                              "note": "contraction flops of the element sizes the tuning table puts on the "
                                      "matrix cores (the others contract on the VALU)",
                              "TFLOPs": round((mfma_flops_launch + lu_mfma_flops_launch) / kern_s / 1e12, 3),
                              "peak_TFLOPs": 157.3,
                              "frac": round((mfma_flops_launch + lu_mfma_flops_launch) / kern_s / 1e12 / 157.3, 5)},
                     "all_flops_per_launch": flops_launch,
                     "all_TFLOPs": round(flops_launch / kern_s / 1e12, 3)},
        "cpu_baseline": None,
        "structure": structure,
    }
    if not args.skip_cpu:
        sample = specs[: args.cpu_rotors]
        log(f"cpu baseline ({oracle_arch}): oracle on {len(sample)} rotors x {args.cpu_steps} steps, all cores")
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        wall = cpu_baseline(sample, args.cpu_steps, oracle_path)
        rate = len(sample) * args.cpu_steps / wall / args.rotors
        sample1 = specs[: args.cpu1_rotors]
        log(f"cpu baseline: single thread, {len(sample1)} rotors x {args.cpu_steps} steps")
        wall1 = cpu_baseline(sample1, args.cpu_steps, oracle_path, threads=1)
        rate1 = len(sample1) * args.cpu_steps / wall1 / args.rotors
        result["cpu_baseline"] = {"value": round(rate, 6), "unit": "steps/s", "cores": cores, "kind": "port",
                                  "sample": f"{len(sample)} rotors x {args.cpu_steps} steps of the same ensemble "
                                            f"({wall:.1f}s wall), scaled to {args.rotors} rotors; oracle built "
                                            f"-O3 -march={oracle_arch}",
                                  "single_thread": {"value": round(rate1, 6), "unit": "steps/s", "cores": 1,
                                                    "sample": f"{len(sample1)} rotors x {args.cpu_steps} steps "
                                                              f"({wall1:.1f}s wall), scaled to {args.rotors} rotors"}}
    print(json.dumps(result), flush=True)
    if failed:
        sys.exit(1)


if __name__ == "__main__":
    main()
