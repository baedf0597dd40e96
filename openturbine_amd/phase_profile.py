"""Per-phase cycle counts of the fused step kernel on the batched ensemble.

Loads the profiling build (libkynema_amd_phases.so: the same kernels with a timestamp at each
phase boundary of a block-iteration) and prints the mean shader cycles per block and Newton
iteration for each phase.

    python -m openturbine_amd.phase_profile --precision fp32 --rotors 512 --steps 2
"""

import argparse
import ctypes as C
from pathlib import Path

import numpy as np

from . import EXEC_FUSED, TANGENT_FP32, TANGENT_FP64, KynLib, KynSystem
from . import workloads as W

PHASES = ["element: gather", "element: QP values", "element: residual", "element: contraction",
          "element: write into KKT", "constraints + conditioning", "refinement copy + rhs column",
          "LU (blocked, panels)", "back substitution", "refinement + unconditioning", "error + update"]
LIB = Path(__file__).resolve().parent / "libkynema_amd_phases.so"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--lib", default=str(LIB), help="profiling library to load (default: the in-tree build)")
    ap.add_argument("--rotors", type=int, default=512)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--refine", type=int, default=0)
    ap.add_argument("--refine-bound", type=float, default=0.0,
                    help="refine only Newton iterations whose previous convergence error is below this (0: all)")
    args = ap.parse_args()
    lib = KynLib(Path(args.lib))
    lib.L.kyn_phase_read.argtypes = [C.POINTER(C.c_ulonglong)]
    specs = W.ensemble_config_specs(num_rotors=args.rotors, num_steps=args.steps + 1)
    cd, tips, loads = W.ensemble_schedule(specs, args.steps + 1)
    prec = TANGENT_FP32 if args.precision == "fp32" else TANGENT_FP64
    s = KynSystem(lib, specs, exec_mode=EXEC_FUSED, precision=prec, refine_sweeps=args.refine,
                  refine_err_bound=args.refine_bound)
    s.set_schedule(args.steps + 1, cd, tips, loads)
    s.step_many(1, use_graph=False)
    it0, _ = s.totals()
    lib.L.kyn_phase_reset()
    s.step_many(args.steps, use_graph=False)
    it1, _ = s.totals()
    acc = (C.c_ulonglong * 28)()
    lib.L.kyn_phase_read(acc)
    a = [int(v) for v in acc]
    blocks_per_model = np.array([len(sp.beams) for sp in specs])
    count = int(((it1 - it0) * blocks_per_model).sum())
    total = 0.0
    print(f"{args.precision} refine {args.refine} (bound {args.refine_bound:g}), {args.rotors} rotors x {args.steps} steps, "
          f"{count} block-iterations, step kernel {s.last_kernel_ms():.3f} ms")
    for i, name in enumerate(PHASES, start=1):
        c = ((a[i] - a[i - 1]) % (1 << 64)) / count
        total += c
        print(f"  {name:32s} {c:10.0f} cycles")
    print(f"  {'total per block-iteration':32s} {total:10.0f} cycles")
    spans = ["LU wait for look-ahead (wave 0)", "LU wait + panel (wave 0)", "LU look-ahead + trailing (wave 1)",
             "LU panel steps incl. barrier", "refinement residual (HBM copy)", "refinement solve",
             "refinement forward sweep (wave 0)", "unpivoted LU: diagonal tile + L21", "unpivoted LU: U12 rows",
             "unpivoted LU: MFMA trailing update"]
    for i, name in enumerate(spans):
        if i < 4 or a[16 + i]:
            print(f"    {name:34s} {a[16 + i] / count:10.0f} cycles")


if __name__ == "__main__":
    main()
