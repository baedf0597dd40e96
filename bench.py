"""Bench: batched FP32-tangent time steps/s vs plain FP64 on the 4096-rotor ensemble (1 GPU).

Workload (one "step" = one generalized-alpha time step of every model): 4096 independent
3-blade rotors, one beam element per blade with N uniform in {6, 11, 16, 21} GLL nodes and
Q = 2N - 2 Gauss points, L in [30, 120] m, hub speed in [0.5, 1.5] rad/s, gravity, normal
tip loads (sigma 1e4 N) every step, h = 0.01 s, rho_inf = 0.9, Newton with the reference's
increment criterion (atol 1e-10, rtol 1e-6).  Inputs are synthetic (SplitMix64 seed
20261202) and resident in HBM (device input schedule) before timing starts.

Timed legs (all on the one GPU):
  value            fused FP32-tangent form with one FP64 iterative-refinement sweep per solve
                   (--refine; the setting whose final state meets 1e-9 against FP64), one
                   hipGraph-replayed kernel per step
  fp32_without_refinement  the same without refinement (faster, 5e-9 off FP64 after 22 steps)
  fused_fp64       fused form, FP64 tangent
  plain_fp64       the reference's launch structure (one kernel per stage, host reads the
                   convergence error every Newton iteration), FP64
Roofline: the fused FP32 step kernel; algorithmic HBM bytes per launch, 8(67N + 87Q) + 36 s N^2
per element and Newton iteration (s = 4), summed over the iterations actually taken, over its
HIP-event-timed duration; flop rates (contraction 144 N^2 Q; with LU 2/3 n^3 and solves 2 n^2)
as extra fields; measured traffic from the committed counter summary.
final_state_fp32_vs_fp64 (extra field): the FP32-tangent leg's final q against the FP64 leg's
after the same inputs and steps, relative to max |q| (whole ensemble and worst single model);
the FP64 fused form equals the oracle bit for bit on the parity cases.
cpu_baseline: the CPU oracle (same algorithm, dense pivoted LU instead of KLU) on the same
ensemble (all rotors, 4 steps by default), OpenMP over rotors on the host cores.
structure (extra field): the large single structure (tower 8 + shaft 8 + 3 x 16 blade elements
of 21 nodes / 40 points, joints, 8459 dofs) in the plain form, FP64 and FP32 element factors
with two FP64 refinement sweeps; steps/s and the element kernel's algorithmic HBM rate.
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


def contraction_flops(specs):
    """Per model: node-pair contraction flops of one Newton iteration, 144 N^2 Q per element."""
    return np.array([sum(144.0 * len(b.nodes) ** 2 * len(b.quad) for b in sp.beams) for sp in specs])


def measured_traffic(refine):
    """HBM bytes per launch of the FP32 step kernel from the committed counter summary for this
    refinement setting (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, copy-calibrated), if present."""
    name = f"k_fused_step_f32_refine{refine}_ensemble4096_traffic.txt" if refine else \
        "k_fused_step_f32_ensemble4096_traffic.txt"
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return {"bytes_per_launch": float(line.split(":", 1)[1]), "source": str(path.relative_to(ROOT))}
    return None


def run_leg(lib, specs, sched, *, exec_mode, precision, warmup, steps, refine=0, use_graph=True):
    t0 = time.time()
    sysm = ota.KynSystem(lib, specs, exec_mode=exec_mode, precision=precision, refine_sweeps=refine)
    log(f"  system built in {time.time() - t0:.1f}s ({sysm.n} nodes, {sysm.num_models} models)")
    cdisp, tips, loads = sched
    sysm.set_schedule(warmup + steps, cdisp[: warmup + steps], tips, loads[: warmup + steps])
    if warmup:
        sysm.step_many(warmup, use_graph=use_graph)
    it0, _ = sysm.totals()
    t0 = time.perf_counter()
    rc = sysm.step_many(steps, use_graph=use_graph)
    wall = time.perf_counter() - t0
    it1, fl = sysm.totals()
    dev_ms = sysm.last_step_ms()
    kern_ms = sysm.last_kernel_ms()
    return {"system": sysm, "rc": rc, "wall_s": wall, "device_ms": dev_ms, "kernel_ms": kern_ms,
            "iters": (it1 - it0), "fails": int(fl.sum())}


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


def cpu_baseline(specs_sample, steps):
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OracleLib, OracleSystem, step_many

    lib = OracleLib()
    systems = [OracleSystem(lib, sp) for sp in specs_sample]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    for k in range(steps):
        inputs = [sp.step_inputs(k) for sp in specs_sample]
        st = step_many(lib, systems, inputs)
        assert all(s == 0 for s in st), "oracle did not converge"
    wall = time.perf_counter() - t0
    return wall, cores


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rotors", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-rotors", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--refine", type=int, default=1,
                    help="FP64 refinement sweeps of the FP32-tangent solve (one meets the 1e-9 final-state bar)")
    ap.add_argument("--skip-unrefined", action="store_true", help="skip the FP32 leg without refinement")
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
    specs = W.ensemble_config_specs(num_rotors=args.rotors, num_steps=args.warmup + args.steps)
    log(f"built {len(specs)} rotor specs in {time.time() - t0:.1f}s")
    t0 = time.time()
    nsched = args.warmup + max(args.steps, args.plain_steps)
    sched = W.ensemble_schedule(specs, nsched)
    log(f"built input schedule ({nsched} steps) in {time.time() - t0:.1f}s")

    flops_iter = block_flops(specs)
    bytes_iter32 = block_bytes(specs, 4)

    log("leg: fused FP32 tangent (hipGraph)")
    r32 = run_leg(lib, specs, sched, exec_mode=ota.EXEC_FUSED, precision=ota.TANGENT_FP32,
                  warmup=args.warmup, steps=args.steps, refine=args.refine)
    log(f"  {args.steps} steps in {r32['wall_s']:.3f}s, fails {r32['fails']}, mean iters/model/step "
        f"{r32['iters'].mean() / args.steps:.2f}, step kernel {r32['kernel_ms']:.3f} ms")
    q32 = r32["system"].get("q")
    offs = np.append(np.asarray(r32["system"].node_off, dtype=np.int64), r32["system"].n)
    del r32["system"]

    r0 = None
    if args.refine > 0 and not args.skip_unrefined:
        log("leg: fused FP32 tangent without refinement (hipGraph)")
        r0 = run_leg(lib, specs, sched, exec_mode=ota.EXEC_FUSED, precision=ota.TANGENT_FP32,
                     warmup=args.warmup, steps=args.steps, refine=0)
        log(f"  {args.steps} steps in {r0['wall_s']:.3f}s, fails {r0['fails']}, step kernel {r0['kernel_ms']:.3f} ms")
        q0 = r0["system"].get("q")
        del r0["system"]

    r64 = None
    if not args.skip_fp64:
        log("leg: fused FP64")
        r64 = run_leg(lib, specs, sched, exec_mode=ota.EXEC_FUSED, precision=ota.TANGENT_FP64,
                      warmup=args.warmup, steps=args.steps)
        log(f"  {args.steps} steps in {r64['wall_s']:.3f}s, fails {r64['fails']}")
        q64 = r64["system"].get("q")
        del r64["system"]

    rp = None
    if not args.skip_plain:
        log("leg: plain FP64 (reference launch structure)")
        rp = run_leg(lib, specs, sched, exec_mode=ota.EXEC_PLAIN, precision=ota.TANGENT_FP64,
                     warmup=1, steps=args.plain_steps, use_graph=False)
        log(f"  {args.plain_steps} steps in {rp['wall_s']:.3f}s, fails {rp['fails']}")
        del rp["system"]

    structure = None
    if args.structure_steps > 0:
        log("leg: large structure (plain form, condensed solve)")
        structure = structure_legs(lib, args.structure_steps)

    # final state of the FP32-tangent leg against the FP64 leg (same inputs, same warmup + steps):
    # max |dq| over the ensemble relative to max |q|, and the worst model's own relative deviation
    def state_error(qa, qb):
        d = np.max(np.abs(qa - qb), axis=1)
        m = np.max(np.abs(qb), axis=1)
        per_model = np.maximum.reduceat(d, offs[:-1]) / np.maximum.reduceat(m, offs[:-1])
        return {"q_rel": float(np.max(d) / np.max(m)), "q_rel_worst_model": float(np.max(per_model)),
                "after_steps": args.warmup + args.steps}

    state_err = state_error(q32, q64) if r64 else None
    if state_err:
        log(f"  FP32-tangent final q vs FP64: {state_err['q_rel']:.2e} (worst model {state_err['q_rel_worst_model']:.2e})")
    unrefined = None
    if r0:
        unrefined = {"steps_per_s": round(args.steps / r0["wall_s"], 4), "kernel_ms": round(r0["kernel_ms"], 4),
                     "final_state_vs_fp64": state_error(q0, q64) if r64 else None}
        if r64:
            e = unrefined["final_state_vs_fp64"]
            log(f"  without refinement: final q vs FP64 {e['q_rel']:.2e} (worst model {e['q_rel_worst_model']:.2e})")

    value = args.steps / r32["wall_s"]
    ms_per_step = 1000.0 * r32["wall_s"] / args.steps
    # roofline of the fused FP32 step kernel (HIP events around each graph launch of the timed
    # loop, on the library's stream): algorithmic HBM bytes 8(67N + 87Q) + 36 s N^2 per element
    # and Newton iteration (s = 4), summed over the iterations each model actually took
    iters_per_step = r32["iters"] / args.steps
    bytes_launch = float(np.sum(iters_per_step * bytes_iter32))
    flops_launch = float(np.sum(iters_per_step * flops_iter))
    mfma_flops_launch = float(np.sum(iters_per_step * contraction_flops(specs)))
    kern_s = (r32["kernel_ms"] if r32["kernel_ms"] > 0 else ms_per_step) / 1000.0
    achieved = bytes_launch / kern_s / 1e9
    traffic = measured_traffic(args.refine)
    result = {
        "metric": METRIC,
        "value": round(value, 4),
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
                   "models": args.rotors, "blades": 3 * args.rotors, "refine_sweeps": args.refine},
        "fused_fp64_steps_per_s": round(args.steps / r64["wall_s"], 4) if r64 else None,
        "plain_fp64_steps_per_s": round(args.plain_steps / rp["wall_s"], 4) if rp else None,
        "speedup_fp32_fused_vs_plain_fp64": round(value / (args.plain_steps / rp["wall_s"]), 2) if rp else None,
        "newton_iters_per_model_step": {"fp32": round(float(r32["iters"].mean()) / args.steps, 3),
                                         "fp64": round(float(r64["iters"].mean()) / args.steps, 3) if r64 else None},
        "failed_model_steps": {"fp32": r32["fails"], "fp64": r64["fails"] if r64 else None},
        "final_state_fp32_vs_fp64": state_err,
        "fp32_without_refinement": unrefined,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 5),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "k_fused_step_f32", "kernel_ms": round(kern_s * 1000.0, 4),
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "contraction_flops_per_launch": mfma_flops_launch,
                     "contraction_TFLOPs": round(mfma_flops_launch / kern_s / 1e12, 3),
                     "all_flops_per_launch": flops_launch,
                     "all_TFLOPs": round(flops_launch / kern_s / 1e12, 3)},
        "cpu_baseline": None,
        "structure": structure,
    }
    if not args.skip_cpu:
        sample = specs[: args.cpu_rotors]
        log(f"cpu baseline: oracle on {len(sample)} rotors x {args.cpu_steps} steps")
        wall, cores = cpu_baseline(sample, args.cpu_steps)
        rotor_steps_per_s = len(sample) * args.cpu_steps / wall
        result["cpu_baseline"] = {"value": round(rotor_steps_per_s / args.rotors, 6), "unit": "steps/s",
                                  "cores": cores, "kind": "port",
                                  "sample": f"{len(sample)} rotors x {args.cpu_steps} steps of the same ensemble "
                                            f"({wall:.1f}s wall), scaled to {args.rotors} rotors"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
