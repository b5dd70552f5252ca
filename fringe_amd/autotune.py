"""Autotuner of the launch schedules (run on the GPU):

    python -m fringe_amd.autotune [--out fringe_amd/schedules.json] [--quick]

For every key of the schedule table it times each candidate on a resident
synthetic tile, with the engine's own stage timers (HIP events around the
stage's launches), and keeps the fastest:

* Gram variant, per (date class, window): the staged MFMA kernel on 12, 8
  or 4 waves (8 x 16, 4 x 16 or up to 4 x 8 pixel tiles), or the unstaged one.
* Solver launch, per (date class, estimator, precision): one wave per unit,
  or a persistent grid of W waves per compute unit taking G units at a time
  from the work queue.
* Above 32 dates, per (date class, precision): each row in registers (up to
  128 dates; two or, up to 48 dates, four threads per row in FP32, four in
  FP64), or one row per lane (up to 64 dates), or the solver's work matrix
  in LDS (where it fits) or in global memory (above 64 dates).

The chosen table (with every candidate's time) is written as JSON; the
library installs it when it is loaded (fringe_amd._lib.load_schedules).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

MODES = {0: ("phase_link", "MLE"), 1: ("evd", "EVD"), 2: ("evd", "MLE")}


def _time_stage(L, run, stage, reps):
    """Median over reps of the summed stage time (ms) of one call of run()."""
    import torch
    from fringe_amd import _lib
    run()   # warm-up (kernel loading, caches)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        L.fringe_timing_reset()
        L.fringe_timing_enable(1)
        run()
        torch.cuda.synchronize()
        L.fringe_timing_enable(0)
        ts.append(_lib.stage_times().get(stage, (0.0, 0))[0])
    return float(np.median(ts))


def tune(quick=False, log=sys.stderr):
    import torch
    from fringe_amd import _lib, synthetic, tile
    L = _lib.lib()
    L.fringe_schedule_clear()   # tune from the built-in defaults
    dev = torch.device("cuda:0")
    reps = 2 if quick else 3
    # the full 2048 x 2048 tile: on a 256-row tile the persistent solver
    # launch measured 4 % faster than the grid, on the full tile 7 % slower
    rows = 96 if quick else 2048
    cols = 2048
    gram, solve = [], []
    stacks = {}

    def data(n, Nx, Ny, fill=None):
        key = (n, Nx, Ny, fill)
        if key not in stacks:
            slc = synthetic.stack_torch(n, rows, cols, 4000 + n, device=dev, patches=32)
            if fill is None:
                w, _ = tile.nmap(slc, Nx, Ny, 0.05)
            else:
                w = synthetic.random_mask_torch(rows, cols, Nx, Ny, fill, 4100 + n, device=dev)
            stacks[key] = (slc, w)
        return stacks[key]

    # ---- Gram variants (mixed precision, date class <= 32) ----------------
    for (Nx, Ny) in ((7, 5), (5, 5), (6, 5), (10, 5), (3, 2)):
        n = 30
        slc, w = data(n, Nx, Ny)
        p = tile.link_params(rows, cols, n, Nx, Ny, min_neighbors=5, precision="mixed")
        times = {}
        for name, v in _lib.GRAM_VARIANTS.items():
            L.fringe_schedule_set(_lib.FRINGE_SCHED_GRAM_VARIANT, n, Nx, Ny, v)
            r = tile.LinkRunner(p, device=dev)
            times[name] = _time_stage(L, lambda: r.run(slc, w), "gram", reps)
            del r
        L.fringe_schedule_set(_lib.FRINGE_SCHED_GRAM_VARIANT, n, Nx, Ny, 0)
        best = min(times, key=times.get)
        gram.append({"n": n, "Nx": Nx, "Ny": Ny, "variant": best, "ms": times})
        print(f"gram n<=32 window {2 * Nx + 1}x{2 * Ny + 1}: {best} {times}", file=log, flush=True)

    # ---- solver launches ---------------------------------------------------
    cands = [("grid", 0, 0)] + [("persistent", wv, g) for wv in ((8, 16) if quick else (8, 12, 16, 24))
                                for g in ((1, 4) if quick else (1, 2, 4, 8))]
    cases = [(30, 0, "mixed"), (30, 0, "fp64"), (20, 2, "mixed"), (30, 1, "mixed"),
             (48, 0, "mixed"), (64, 1, "mixed"), (96, 0, "mixed")]
    if quick:
        cases = cases[:3]
    for n, mode, prec in cases:
        module, method = MODES[mode]
        Nx, Ny = (10, 5) if n == 64 else (7, 5)
        slc, w = data(n, Nx, Ny, fill=None if n <= 64 else 0.9)
        p = tile.link_params(rows if n <= 64 else 16, cols, n, Nx, Ny, module=module, method=method,
                             min_neighbors=5, precision=prec)
        if n > 64:
            slc, w = slc[:, :16].contiguous(), w[:16].contiguous()
        key = (n, mode, 0 if prec == "fp64" else 1)
        # the launch schedule serves the one-row-per-lane solver and the
        # solver with a work matrix, not the register rows
        ga_default = L.fringe_schedule_get(_lib.FRINGE_SCHED_GENERIC_A, n, key[2], 0)
        if n > 32:
            L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, n, key[2], 0, 0)
        times = {}
        for launch, wv, g in cands:
            L.fringe_schedule_set(_lib.FRINGE_SCHED_SOLVE, *key, 0 if launch == "grid" else (wv << 8) | g)
            r = tile.LinkRunner(p, device=dev)
            times[f"{launch}:{wv}:{g}"] = _time_stage(L, lambda: r.run(slc, w), "solve", reps)
            del r
        L.fringe_schedule_set(_lib.FRINGE_SCHED_SOLVE, *key, 0)
        if n > 32:
            L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, n, key[2], 0, ga_default)
        best = min(times, key=times.get)
        launch, wv, g = best.split(":")
        solve.append({"n": n, "mode": mode, "precision": key[2], "estimator": f"{module}/{method}",
                      "launch": launch, "waves_per_cu": int(wv), "grab": int(g), "ms": times})
        print(f"solve n={n} {module}/{method} {prec}: {best} ({times[best]:.2f} ms, grid {times['grid:0:0']:.2f})",
              file=log, flush=True)
    # ---- above 32 dates: rows in registers, or one row per lane (up to 64
    # dates), or a work matrix in LDS or global memory (above) --------------
    generic_a = []
    for n, prec in ((48, "mixed"), (64, "mixed"), (48, "fp64"), (64, "fp64"), (72, "mixed"), (96, "mixed"),
                    (72, "fp64"), (96, "fp64"), (128, "mixed"), (128, "fp64")):
        if quick and n != 96:
            continue
        prow = 16 if n > 64 else 256
        slc, w = data(n, 7, 5, fill=0.9)
        slc, w = slc[:, :prow].contiguous(), w[:prow].contiguous()
        p = tile.link_params(prow, cols, n, 7, 5, min_neighbors=5, precision=prec)
        pk = 0 if prec == "fp64" else 1
        default = L.fringe_schedule_get(_lib.FRINGE_SCHED_GENERIC_A, n, pk, 0)
        times = {}
        places = ["lanes", "registers"] if n <= 64 else ["lds", "global"] + (["registers"] if n <= 128 else [])
        for place in places:
            v = _lib.GENERIC_A_PLACES[place]
            L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, n, pk, 0, v)
            r = tile.LinkRunner(p, device=dev)
            times[place] = _time_stage(L, lambda: r.run(slc, w), "solve", reps)
            del r
        L.fringe_schedule_set(_lib.FRINGE_SCHED_GENERIC_A, n, pk, 0, default)
        best = min(times, key=times.get)
        generic_a.append({"n": n, "precision": pk, "work_matrix": best, "ms": times})
        print(f"solver layout n={n} {prec}: {best} {times}", file=log, flush=True)
    return {"device": torch.cuda.get_device_name(0), "tile": f"{rows} x {cols}",
            "note": "stage times in ms per call (median); the library installs the chosen entries",
            "gram": gram, "solve": solve, "generic_a": generic_a}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=os.path.join(HERE, "schedules.json"))
    ap.add_argument("--quick", action="store_true", help="fewer candidates and a smaller tile")
    a = ap.parse_args(argv)
    t0 = time.time()
    table = tune(quick=a.quick)
    table["tuning_s"] = round(time.time() - t0, 1)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({"written": a.out, "gram": len(table["gram"]), "solve": len(table["solve"])}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
