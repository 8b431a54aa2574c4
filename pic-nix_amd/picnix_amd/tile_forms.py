"""Time and accuracy of the FP32 current deposit's LDS tile forms.

    python -m picnix_amd.tile_forms --config thermal128 --steps 10 \
        --out ../profiles/fp32_tile_forms_thermal128.txt

The FP32 march deposit sums each claimed cell's particles in registers and adds
the sums to a ring tile in LDS before flushing the tile to J.  The tile can be
  f64     (default) one FP64 value per node (LDS double atomics);
  staged  compensated FP32 pairs (hi, lo), updated without atomics: the cells
          of one phase store their sums in an LDS stage and one thread per
          node adds them in ascending cell order with TwoSum; --stage-cells
          sets the cells per phase (PICNIX_STAGE_CELLS);
  comp    the same pairs updated by LDS float atomics, the error of each add
          recovered from its returned old value (TwoSum);
  plain   one FP32 value per node, no compensation.
PICNIX_FP32_TILE=staged|comp|plain selects a form at pn_create.
Every form is created from the same loader output.  Accuracy: J after the
first step against the FP64 form's J from the same loader output (which
includes the FP32 form's rounding of the particles), and against the f64 tile
form's J (the tile's own summation error; both flush with FP32 global
atomics), as the largest node difference over the largest |J| of each
component.  Time: the deposit phase's HIP-event time per launch over --steps
eager steps after two warm-up steps (pn_timing).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import picnix_amd

FORMS = ["staged", "comp", "f64", "plain"]


def _current(sim):
    """Interior J of every chunk, stacked: [chunks, nz, ny, nx, 3] (FP64)."""
    out = []
    for c in range(sim.nchunks):
        info = sim.chunk_info(c)
        nb = int(info[9])
        nz, ny, nx = (int(v) for v in info[6:9])
        j = sim.field(c, 1)
        z0, z1 = (nb, nz - nb) if nz > 1 else (0, 1)
        y0, y1 = (nb, ny - nb) if ny > 1 else (0, 1)
        out.append(np.array(j[z0:z1, y0:y1, nb:nx - nb, 1:4], dtype=np.float64))
    return np.stack(out)


def _rel(a, b):
    """Largest node difference over the largest |b|, per component (x, y, z)."""
    return [float(np.max(np.abs(a[..., k] - b[..., k])) / max(float(np.max(np.abs(b[..., k]))),
                                                                1e-300)) for k in range(3)]


def _env(form, stage_cells):
    for k in ("PICNIX_FP32_TILE", "PICNIX_STAGE_CELLS"):
        os.environ.pop(k, None)
    os.environ["PICNIX_FP32_TILE"] = form
    if form == "staged" and stage_cells:
        os.environ["PICNIX_STAGE_CELLS"] = str(stage_cells)


def run_form(text32, dt, form, stage_cells, steps):
    _env(form, stage_cells)
    sim = picnix_amd.Simulation(text32)
    for k in ("PICNIX_FP32_TILE", "PICNIX_STAGE_CELLS"):
        os.environ.pop(k, None)
    sim.step(dt)
    sim.synchronize()
    j = _current(sim)
    sim.step(dt, 2)
    sim.synchronize()
    sim.timing(True)
    sim.step(dt, steps)
    sim.synchronize()
    ms, n = sim.timing_report()["deposit"]
    sim.timing(False)
    sim.close()
    return j, ms / max(n, 1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="thermal128")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--stage-cells", default="1,2,4,8",
                    help="cells per phase of the staged form to time (comma list)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    with open(picnix_amd.config_path(args.config)) as fp:
        text64 = fp.read()
    text32 = text64.replace("[application.option]",
                            "[application.option]\n    precision = 'fp32'", 1)
    dt = None
    for line in text64.splitlines():
        s = line.strip()
        if s.startswith("delt"):
            dt = float(s.split("=")[1])
    t0 = time.perf_counter()
    print(f"[tile_forms] {args.config}: FP64 form, one step", flush=True)
    ref = picnix_amd.Simulation(text64)
    ref.step(dt)
    ref.synchronize()
    j64 = _current(ref)
    ref.close()
    runs = [("staged", int(c)) for c in args.stage_cells.split(",") if c] + \
           [(f, 0) for f in FORMS[1:]]
    rows = []
    jf64 = None
    results = {}
    for form, kst in runs:
        j, ms = run_form(text32, dt, form, kst, args.steps)
        results[(form, kst)] = (j, ms)
        if form == "f64":
            jf64 = j
        print(f"[tile_forms] {form}{'/' + str(kst) if kst else ''}: deposit {ms:.3f} ms/launch "
              f"({time.perf_counter() - t0:.0f} s)", flush=True)
    for (form, kst), (j, ms) in results.items():
        rows.append({"form": form, "stage_cells": kst or None, "deposit_ms": ms,
                     "J_vs_fp64_form": _rel(j, j64),
                     "J_vs_f64_tile": _rel(j, jf64) if jf64 is not None else None})
    if args.out:
        cmd = (f"python -m picnix_amd.tile_forms --config {args.config} --steps {args.steps} "
               f"--stage-cells {args.stage_cells} --out <file>")
        out = [f"# FP32 deposit tile forms: {args.config}", "#", "# command (from pic-nix_amd/, "
               "on the GPU):", f"#   {cmd}", "#",
               "# deposit: HIP-event time per launch of the deposit phase (eager steps).",
               "# J errors: max |dJ| / max |J| per component after the first step, against",
               "# the FP64 form and against the FP64-tile FP32 form.", "",
               f"{'form':<10} {'cells':>5} {'deposit ms':>11} {'Jx vs fp64':>11} "
               f"{'Jy vs fp64':>11} {'Jz vs fp64':>11} {'Jx vs f64t':>11} {'Jy vs f64t':>11} "
               f"{'Jz vs f64t':>11}"]
        for r in rows:
            e2 = r["J_vs_f64_tile"] or [float("nan")] * 3
            out.append(f"{r['form']:<10} {str(r['stage_cells'] or '-'):>5} "
                       f"{r['deposit_ms']:>11.3f} " +
                       " ".join(f"{v:>11.2e}" for v in r["J_vs_fp64_form"] + e2))
        with open(args.out, "w") as fp:
            fp.write("\n".join(out) + "\n")
    print(json.dumps({"config": args.config, "steps": args.steps, "rows": rows}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
