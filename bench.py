#!/usr/bin/env python3
"""Throughput of the pic-nix per-timestep hot path on one MI355X.

One "step" is PicApplication::push_openmp's full sequence (B half push, gather +
Boris, position push + count, Esirkepov deposit, B half push, current fold, E push,
particle migration + counting sort) over every chunk of the workload.  The
metric is particle-steps/s = (mobile particles) x steps / wall time of the timed
steps; inputs are resident in HBM before timing starts.

Default workload: the 128^3 thermal plasma (512 chunks of 16^3, electrons +
protons, 32 particles/cell/species = 134M particles, 2nd-order shape, FP64),
the configuration the throughput targets are quoted on; the FP32 batched form and
the plain FP64 path run on the same workload.

The line also carries
  * roofline: the dominant kernel's algorithmic bytes per launch over its mean
    launch time (HIP events on the library's stream), against 8 TB/s;
  * cpu_baseline: the FP64 CPU oracle (a restatement of the reference's OpenMP
    path) timed on this host's cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pic-nix_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "particle-steps/s (push+deposit+sort+FDTD), FP64 plain vs FP32 batched; % HBM peak"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "thermal64": "thermal plasma 3D 64^3 cells, 64 chunks of 16^3, electrons+protons "
                 "(m_p/m_e=100, vth_e=0.1c), 64 particles/cell/species (33.5M), "
                 "2nd-order shape, FP64 plain",
    "thermal128": "thermal plasma 3D 128^3 cells, 512 chunks of 16^3, electrons+protons, "
                  "32 particles/cell/species (134M), 2nd-order shape, FP64 plain",
    "twostream2d": "two-stream 2D 1x16x512, 32 chunks of 1x16x16, two beams +-0.1c, "
                   "64 particles/cell/beam, 1st-order shape, FP64 plain",
    "shock3d": "shock-like 3D 512x64x64 cells, 512 chunks of 16^3, reflecting wall at x=0 and "
               "Poisson injection at x=512, density 1 -> 4 (16 -> 64 particles/cell/species) "
               "with a 32-cell ramp at x=256, drift 0.2c, vth_e=0.05c, m_i/m_e=64, "
               "2nd-order shape, FP64 plain (and the FP32 form)",
}


def algorithmic_bytes(phase, npart, ncell, rho_direct=False, fused=False, deferred=False):
    """Bytes each launch must move at minimum (FP64 operator-boundary model).

    Per particle / per cell, one launch:
      bfd      : read E, B; write B                      72 B/cell
      velocity : read x, u; write u (72 B/particle) + one E, B gather per cell (48 B/cell)
      position : read x, u; write x_old, x, cell index  100 B/particle
      deposit  : read x_old, x (48 B/particle); write J (24 B/cell)
      fold     : zero J and fold the periodic padding (32 B/cell)
      efd      : read E, B, J; write E                   96 B/cell
      sort     : read cell index + record, write record (116 B/particle) + counts (12 B/cell)
      rho      : carried by continuity, rho -= dt div J: read rho, Jx, Jy, Jz and
                 write rho (40 B/cell), no particle reads; a direct deposit instead
                 reads x (24 B/particle) and writes rho (8 B/cell)
    fused (the gather runs the position push): velocity reads x, u once and writes
    u, x_old, x and the cell index, 124 B/particle; deferred (the sort's record
    move runs in the next gather): the gather also reads the source index and
    moves the id, +20 B/particle, and the sort reads the cell index and writes
    the source index, 8 B/particle.
    """
    if phase == "rho" and not rho_direct:
        return 40 * ncell
    per_p = {"bfd": 0, "velocity": 124 if fused else 72, "position": 100, "deposit": 48,
             "fold": 0, "efd": 0, "sort": 8 if deferred else 116, "rho": 24}[phase]
    if phase == "velocity" and deferred:
        per_p += 20
    per_c = {"bfd": 72, "velocity": 48, "position": 0, "deposit": 24, "fold": 32, "efd": 96,
             "sort": 12, "rho": 8}[phase]
    return per_p * npart + per_c * ncell


# kernel that carries each timed phase (its name prefix in rocprofv3 output)
PHASE_KERNEL = {"bfd": "k_push_bfd", "velocity": "k_push_velocity_rows",
                "position": "k_push_position", "deposit": "k_deposit_march",
                "fold": "k_fold_line", "efd": "k_push_efd", "sort": "k_rank_bin",
                "rho": "k_rho_continuity"}


def launch_form(phase, tuning, fused, dim=3, order=2):
    """The template form of the phase's kernel that the tuned FP64 step launches,
    as rocprofv3 names it (None when the phase has no such form)."""
    t = tuning or {}
    b = lambda v: "true" if v else "false"
    if phase == "velocity":
        return (f"k_push_velocity_rows<{dim}, {order}, true, {b(t.get('gather_queue'))}, "
                f"{b(fused)}>")
    if phase == "deposit":
        return f"k_deposit_march<{order}, {b(t.get('deposit_queue'))}>"
    return None


def profile_traffic(config, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    profiles/hbm_traffic_<config>.txt: FETCH_SIZE + WRITE_SIZE per dispatch,
    each scaled by the copy-probe calibration recorded in the same file.
    `kernel` is a full form (matched exactly) or a name prefix (the form with
    the most dispatches).  None when the summary or the kernel is missing."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{config}.txt")
    if not os.path.exists(path):
        return None
    # per counter section, the launch form of `kernel` with the most dispatches
    # (an autotuning session also runs the forms it rejects)
    best = {}
    section = None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts and parts[0] in ("FETCH_SIZE", "WRITE_SIZE"):
                section = parts[0]
            elif section and len(parts) >= 4 and " ".join(parts[:-3]).startswith(kernel):
                if "<" in kernel and " ".join(parts[:-3]) != kernel:
                    continue
                try:
                    n, mb = int(parts[-3]), float(parts[-1])
                except ValueError:
                    continue
                if section not in best or n > best[section][0]:
                    best[section] = (n, mb * 1e6)  # calibrated MB per dispatch
    return sum(v[1] for v in best.values()) if len(best) == 2 else None


def step_bytes(npart, ncell):
    """Whole-step algorithmic bytes: 336 B/particle + 332 B/cell (FP64)."""
    return 336 * npart + 332 * ncell


def host_cpu():
    """The host's CPU model and logical CPU counts (as lscpu / nproc report them)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)),
            "logical_cpus": os.cpu_count()}


def progress(msg):
    """One line on stderr as each part of the run finishes."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(cfg_name, threads, steps):
    """Time the FP64 oracle on the host: `steps` full steps of the same workload."""
    import helpers
    import oracle_lib
    cfg = helpers.load_config(cfg_name)
    oracle_lib.set_threads(threads)
    o = oracle_lib.Oracle(cfg)
    npart = sum(o.np(c, s) for c in range(o.nchunks) for s in range(o.Ns))
    o.step(1)  # first touch of the old-position buffers
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return {
        "value": npart * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} full steps of the same {cfg_name} workload ({npart} particles) after "
                  f"one untimed step; FP64 oracle, C++/OpenMP over chunks, scalar (not xsimd); "
                  f"{threads} threads = this job's CPU share of the host",
        "seconds": dt,
        "host": host_cpu(),
    }


def copy_probe(mb):
    """One device-to-device hipMemcpy of `mb` MiB (ROCm's copy kernel): under
    rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE calibrate the counters, since it
    reads and writes exactly mb MiB."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="thermal128", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    # this job's share of the host: 16 threads on a one-GPU box, whatever nproc says
    ap.add_argument("--cpu-threads", type=int,
                    default=min(16, len(os.sched_getaffinity(0)),
                                int(os.environ.get("OMP_NUM_THREADS", "16"))))
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--autotune", default="on", choices=["on", "off"],
                    help="run an autotuning session (pn_autotune) before the warm-up steps")
    ap.add_argument("--fp32-graph", default="on", choices=["on", "off"],
                    help="time the FP32 form's steps as replays of the captured step graph")
    ap.add_argument("--fp32", default="on", choices=["on", "off"],
                    help="also time the FP32 form on the same workload")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device copy of this size (calibrates HBM counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design "
              "(see DESIGN.md) and --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import picnix_amd
    import tomlmini
    with open(picnix_amd.config_path(args.config)) as f:
        toml_text = f.read()
    cfg = tomlmini.loads(toml_text)
    par = cfg["parameter"]
    dt = float(par["delt"])
    ncell = int(par["Nx"]) * int(par["Ny"]) * int(par["Nz"])

    def run_form(precision, batched=True):
        """Warm-up, then K timed steps of one precision form; returns the wall
        time, the particle counts, the per-phase report and the final history.
        batched=False: the plain form, with the built-in launch shapes (no
        autotuning session, one block per work item, eager steps)."""
        text = toml_text
        if precision == "fp32":
            text = text.replace("[application.option]",
                                "[application.option]\n    precision = 'fp32'", 1)
        sim = picnix_amd.Simulation(text)
        progress(f"{precision} {'batched' if batched else 'plain'}: set up")
        tuning = None
        if args.autotune == "on" and batched:
            sim.autotune(True)
            nt = sim.tuning()["steps_left"]
            sim.step(dt, nt)
            tuning = dict(sim.tuning(), session_steps=nt)
            progress(f"{precision}: tuning session done ({nt} steps)")
        npart = sum(sim.np(c, s) for c in range(sim.nchunks) for s in range(sim.Ns))
        # the FP32 batched form replays the step as a captured hipGraph; the
        # per-phase HIP events then come from eager steps after the timed ones
        graph = precision == "fp32" and args.fp32_graph == "on" and batched
        sim.step_graph(graph)
        sim.step(dt, args.warmup)
        sim.synchronize()
        if not graph:
            sim.timing(True)
        t0 = time.perf_counter()
        sim.step(dt, args.steps)
        sim.synchronize()
        t1 = time.perf_counter()
        progress(f"{precision} {'batched' if batched else 'plain'}: {args.steps} timed steps, "
                 f"{(t1 - t0) / args.steps * 1e3:.3f} ms/step")
        npart_end = sum(sim.np(c, s) for c in range(sim.nchunks) for s in range(sim.Ns))
        hist = sim.history()  # after warm-up + timed steps, as for the other form
        if graph:
            sim.step_graph(False)
            sim.timing(True)
            sim.step(dt, min(args.steps, 5))
            sim.synchronize()
        rep = sim.timing_report()
        sim.timing(False)
        sim.close()
        if batched:
            run_form.tuning[precision] = tuning
        return t1 - t0, npart, npart_end, rep, hist
    run_form.tuning = {}

    elapsed, npart, npart_end, rep, hist64 = run_form("fp64")
    fp32 = None
    plain = None
    if args.fp32 == "on":
        # the plain FP64 path beside the batched forms: built-in shapes, eager steps
        ep, np_p, np_p_end, _, _ = run_form("fp64", batched=False)
        plain = {"value": 0.5 * (np_p + np_p_end) * args.steps / ep,
                 "ms_per_step": ep / args.steps * 1e3,
                 "form": "FP64, built-in launch shapes, no autotuning, no work queue, "
                         "eager steps"}
        e32, n32, n32_end, rep32, hist32 = run_form("fp32")
        # error of the FP32 form against the FP64 form after the same steps
        names = ["div(E)-rho", "div(B)", "E^2/2", "B^2/2"] + \
                [f"kinetic[{k}]" for k in range(len(hist64) - 4)]
        err = {names[k]: (abs(hist32[k] - hist64[k]) / abs(hist64[k]) if hist64[k] != 0 else
                          abs(hist32[k])) for k in range(2, len(hist64))}
        err["div(E)-rho (abs)"] = abs(hist32[0] - hist64[0])
        # the same count as the FP64 lines: the mean of the first and last step
        v32 = 0.5 * (n32 + n32_end) * args.steps / e32
        fp32 = {"value": v32, "ms_per_step": e32 / args.steps * 1e3,
                "speedup_vs_fp64": v32 / (0.5 * (npart + npart_end) * args.steps / elapsed),
                "speedup_vs_fp64_plain": v32 / plain["value"],
                "particles_conserved": n32_end == n32, "dtype": "f32",
                "step_graph": args.fp32_graph == "on",
                "rel_error_vs_fp64_after_steps": args.warmup + args.steps,
                "rel_error_vs_fp64": err,
                "phases_ms_per_launch": {k: (ms / n if n else 0.0) for k, (ms, n) in rep32.items()},
                "form": "FP32 cell-relative particles and fields, FP64 LDS current tiles, "
                        "FP64 energy/charge reductions"}
    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)

    # an injecting boundary changes the particle count during the timed steps:
    # count the mean of the first and last step
    value = 0.5 * (npart + npart_end) * args.steps / elapsed
    phases = {}
    # the gather ran the position push when no position launch was timed; pn_step
    # defers its sort's record move on the periodic workloads (PICNIX_DEFER_SORT)
    fused = rep.get("position", (0.0, 0))[1] == 0
    deferred = (not args.config.startswith("shock") and
                os.environ.get("PICNIX_DEFER_SORT", "1") != "0")
    for ph, (ms, n) in rep.items():
        if n == 0:
            continue
        per = ms / n
        # the shock problem deposits rho every step; the periodic workloads carry it
        b = algorithmic_bytes(ph, npart, ncell, rho_direct=args.config.startswith("shock"),
                              fused=fused, deferred=deferred)
        phases[ph] = {"ms_per_launch": per, "launches_per_step": n / args.steps,
                      "bytes_per_launch": b, "GBps": b / (per * 1e-3) / 1e9}
    dom = max(phases, key=lambda p: phases[p]["ms_per_launch"] * phases[p]["launches_per_step"])
    d = phases[dom]
    # the tuned form of the dominant kernel (thermal workloads: 3D, order 2)
    form = (launch_form(dom, (run_form.tuning.get("fp64") or {}), fused)
            if args.config.startswith("thermal") else None)
    roofline = {
        "bound": "hbm",
        "achieved": d["GBps"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": d["GBps"] / HBM_PEAK_GBS,
        "traffic": profile_traffic(args.config, form or PHASE_KERNEL.get(dom, dom)),
        "kernel": form or PHASE_KERNEL.get(dom, dom),
        "phase": dom,
        "bytes_per_launch": d["bytes_per_launch"],
    }
    step_gbs = step_bytes(npart, ncell) * args.steps / elapsed / 1e9

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded loader of the reference's %s example)" % (
            "shock" if args.config.startswith("shock") else "thermal"),
        "config": {"workload": WORKLOADS[args.config], "config": args.config,
                   "particles": npart, "cells": ncell, "precision": "fp64",
                   "particles_conserved": npart_end == npart},
        "roofline": roofline,
        "step_roofline": {"bytes_per_step": step_bytes(npart, ncell), "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "phases": phases,
        "tuning": run_form.tuning,
        "fp32_form": fp32,
        "fp64_plain": plain,
    }
    if args.cpu_baseline == "on":
        progress("CPU baseline")
        line["cpu_baseline"] = cpu_baseline(args.config, args.cpu_threads, args.cpu_steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
