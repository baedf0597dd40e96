"""Autotuner of the FP32 step kernel's node-pair contraction engine, per element-size bucket.

For each bucket N in {6, 11, 16, 21} (Q = 2N - 2, the ensemble's element sizes) it times the
fused FP32 step on a single-size ensemble (every blade N nodes) with the contraction on the
matrix cores and on the VALU register tile (kyn_debug_set_contraction_mask), at the bench
default refinement policy (one sweep, in the Newton iterations whose previous convergence error is
below 10), and keeps the faster engine.  Sizes between buckets take the nearest
bucket's choice.  With --write it writes the table the library compiles in
(csrc/kyn_tuning.inc) and a summary under profiles/; rebuild the library afterwards.

    python3 -m openturbine_amd.autotune --rotors 2048 --steps 6 --write
"""

import argparse
import statistics
import time
from pathlib import Path

from . import EXEC_FUSED, TANGENT_FP32, KynLib, KynSystem
from . import workloads as W

BUCKETS = (6, 11, 16, 21)
ROOT = Path(__file__).resolve().parent


def time_engine(lib, specs, sched, mask, steps, repeats, refine, bound):
    s = KynSystem(lib, specs, exec_mode=EXEC_FUSED, precision=TANGENT_FP32, refine_sweeps=refine,
                  refine_err_bound=bound)
    assert lib.L.kyn_debug_set_contraction_mask(s.h, mask) == 0
    total = 1 + repeats * steps
    s.set_schedule(total, sched[0][:total], sched[1], sched[2][:total])
    s.step_many(1, use_graph=True)  # warm-up (and graph capture)
    kms = []
    for _ in range(repeats):
        s.step_many(steps, use_graph=True)
        kms.append(s.last_kernel_ms())
    it, fails = s.totals()
    return statistics.median(kms), float(it.sum()) / (len(specs) * total), int(fails.sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rotors", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--refine-bound", type=float, default=10.0)
    ap.add_argument("--write", action="store_true", help="write csrc/kyn_tuning.inc and the profile summary")
    ap.add_argument("--out", default=None, help="directory for both files instead of their places in the tree")
    args = ap.parse_args()
    lib = KynLib()
    rows, choice = [], {}
    for N in BUCKETS:
        t0 = time.time()
        nsteps = 1 + args.repeats * args.steps
        specs = W.ensemble_config_specs(num_rotors=args.rotors, num_steps=nsteps, fixed_nodes=N)
        sched = W.ensemble_schedule(specs, nsteps)
        mfma = time_engine(lib, specs, sched, 1 << N, args.steps, args.repeats, args.refine, args.refine_bound)
        valu = time_engine(lib, specs, sched, 0, args.steps, args.repeats, args.refine, args.refine_bound)
        choice[N] = mfma[0] <= valu[0]
        rows.append((N, 2 * N - 2, mfma, valu))
        print(f"N = {N:2d} (Q = {2 * N - 2}): matrix cores {mfma[0]:.3f} ms, VALU {valu[0]:.3f} ms per step "
              f"({mfma[1]:.2f} / {valu[1]:.2f} Newton iterations per model-step, fails {mfma[2]} / {valu[2]}) "
              f"-> {'matrix cores' if choice[N] else 'VALU'}  [{time.time() - t0:.0f} s]", flush=True)
    mask = 0
    for n in range(22):
        nearest = min(BUCKETS, key=lambda b: (abs(b - n), b))
        if choice[nearest]:
            mask |= 1 << n
    print(f"kContractMfmaMask = 0x{mask:x}", flush=True)
    if not args.write:
        return
    table = "\n".join(f"//   N = {N:2d}, Q = {Q:2d}: matrix cores {m[0]:7.3f} ms, VALU {v[0]:7.3f} ms -> "
                      f"{'matrix cores' if choice[N] else 'VALU'}" for N, Q, m, v in rows)
    inc = Path(args.out) / "kyn_tuning.inc" if args.out else ROOT / "csrc" / "kyn_tuning.inc"
    prof = Path(args.out) / "autotune_contraction_ensemble.txt" if args.out else \
        ROOT.parent / "profiles" / "autotune_contraction_ensemble.txt"
    inc.write_text(
        "// Tuning table of the FP32 tangent's node-pair contraction: bit N set -> elements of N nodes\n"
        "// contract on the matrix cores, clear -> on the VALU register tile.  Written by\n"
        "// openturbine_amd/autotune.py from fused FP32 step-kernel times of single-size ensembles\n"
        f"// ({args.rotors} rotors of three N-node blades, refine {args.refine}, bound {args.refine_bound:g}, median of\n"
        f"// {args.repeats} runs of {args.steps} steps); sizes between buckets take the nearest bucket's engine.\n"
        f"{table}\n"
        "#pragma once\n"
        "namespace kyn {\n"
        f"constexpr unsigned kContractMfmaMask = 0x{mask:x}u;  // N = 0 .. 21\n"
        "}  // namespace kyn\n")
    prof.write_text(
        "Contraction engine per element-size bucket — fused FP32 step kernel (k_fused_step_f32),\n"
        f"single-size ensembles of {args.rotors} rotors (three N-node blades each, Q = 2N - 2), {args.refine} FP64\n"
        f"refinement sweep(s) in the Newton iterations whose previous convergence error is below "
        f"{args.refine_bound:g} (0: every iteration),\nmedian of {args.repeats} runs of {args.steps} steps "
        f"(HIP events around the graph launches), MI355X.\n\n"
        "Command (repository root):\n"
        f"  python3 -m openturbine_amd.autotune --rotors {args.rotors} --steps {args.steps} --repeats {args.repeats} "
        "--write\n\n"
        "  N   Q   matrix cores ms   VALU ms   iterations/model-step   choice\n" +
        "\n".join(f"  {N:2d}  {Q:2d}  {m[0]:15.3f}  {v[0]:8.3f}   {m[1]:10.2f} / {v[1]:.2f}      "
                  f"{'matrix cores' if choice[N] else 'VALU'}" for N, Q, m, v in rows) +
        f"\n\ntable: kContractMfmaMask = 0x{mask:x} (csrc/kyn_tuning.inc)\n")
    print(f"wrote {inc} and {prof}", flush=True)


if __name__ == "__main__":
    main()
