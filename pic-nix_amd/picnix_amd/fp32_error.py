"""Error of the FP32 form against the FP64 form over a long run.

    python -m picnix_amd.fp32_error --config thermal128 --steps 1000 --every 100 --twin \
        --out ../profiles/fp32_error_thermal128.txt

Both forms start from the same loader output and advance side by side on one
GPU; every `--every` steps the history quantities (div(E) - rho, div(B), E^2/2,
B^2/2, kinetic energy per species, as HistoryDiag writes them) of both are
compared.  The FP64 form is the reference here: at this size the CPU oracle is
too slow for 1000 steps, and the FP64 GPU path is held to it at steps 1 and 10
by the parity tests.  Prints one progress line per checkpoint and a final JSON
object with the relative errors at every checkpoint.

--twin adds a third run: the FP64 form started from exactly the FP32 form's
initial particles (the loader's values rounded to FP32 cell-relative offsets and
momenta).  Its distance from the FP64 run measures how fast this plasma
amplifies an FP32-sized change of the initial state (the trajectories are
chaotic); its distance from the FP32 run is the part that the FP32 arithmetic
of the step adds.

Statistical agreement: the history quantities are also sampled every
`--sample` steps, and over the second half of the run (`--stat-from`, default
steps/2) their time means and standard deviations are compared -- once the
runs have decorrelated, the FP32 form must reproduce the FP64 form's thermal
state (the same mean energies within the fluctuation level), not its
trajectory.  `--out` writes both tables as text.
"""
import argparse
import json
import math
import sys
import time

import picnix_amd


def _mean_std(xs):
    m = sum(xs) / len(xs)
    v = sum((x - m) ** 2 for x in xs) / max(1, len(xs) - 1)
    return m, math.sqrt(v)


def statistics(samples, names):
    """Time mean and standard deviation of every energy over the sampled window,
    per run, and each run's mean relative to the FP64 run's mean."""
    out = {}
    ref = samples["fp64"]
    if not ref:
        return out
    for k in range(2, len(names)):
        q = {}
        m64, s64 = _mean_std([h[k] for h in ref])
        for run, hs in samples.items():
            if not hs:
                continue
            m, sd = _mean_std([h[k] for h in hs])
            q[run] = {"mean": m, "std": sd,
                      "mean_rel_to_fp64": (m - m64) / m64 if m64 != 0 else m - m64,
                      "std_rel": sd / abs(m) if m != 0 else sd}
        q["samples"] = len(ref)
        out[names[k]] = q
    return out


def report(args, dt, stat_from, rows, stats, names):
    cmd = (f"python -m picnix_amd.fp32_error --config {args.config} --steps {args.steps} "
           f"--every {args.every} --sample {args.sample}" + (" --twin" if args.twin else "") +
           (f" --out {args.out}" if args.out else ""))
    out = [f"# FP32 form against the FP64 form over {args.steps} steps: {args.config}",
           "#", "# command (from pic-nix_amd/, on the GPU):", f"#   {cmd}", "#",
           "# Pointwise relative differences of the history energies at each checkpoint",
           "# (twin = the FP64 form started from the FP32 form's initial state).", ""]
    hdr = ["step", "E^2/2 fp32-fp64", "E^2/2 twin-fp64", "E^2/2 fp32-twin",
           "B^2/2 fp32-fp64", "B^2/2 twin-fp64", "kin[0] fp32-fp64", "kin[0] twin-fp64",
           "div(E)-rho fp32 (abs)"]
    out.append(" ".join(f"{h:>17}" for h in hdr))
    for r in rows:
        e = r["rel_error"]
        t = r.get("twin_vs_fp64", {})
        ft = r.get("fp32_vs_twin", {})
        vals = [r["step"], e["E^2/2"], t.get("E^2/2", float("nan")),
                ft.get("E^2/2", float("nan")), e["B^2/2"], t.get("B^2/2", float("nan")),
                e["kinetic[0]"], t.get("kinetic[0]", float("nan")), e["div(E)-rho (abs)"]]
        out.append(f"{vals[0]:>17d} " + " ".join(f"{v:>17.3e}" for v in vals[1:]))
    out += ["", f"# Statistical agreement over steps {stat_from + args.sample}..{args.steps} "
            f"(one sample every {args.sample} steps):",
            "# time mean of each energy, its relative difference from the FP64 run's mean,",
            "# and the relative fluctuation (std/mean) of each run.", ""]
    out.append(f"{'quantity':<14} {'run':<6} {'mean':>14} {'mean vs fp64':>14} {'std/mean':>12}")
    for k in range(2, len(names)):
        q = stats.get(names[k], {})
        for run in ("fp64", "fp32", "twin"):
            if run not in q:
                continue
            v = q[run]
            out.append(f"{names[k]:<14} {run:<6} {v['mean']:>14.6e} "
                       f"{v['mean_rel_to_fp64']:>14.3e} {v['std_rel']:>12.3e}")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="thermal128")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--graph", default="on", choices=["on", "off"])
    ap.add_argument("--twin", action="store_true")
    ap.add_argument("--sample", type=int, default=10)
    ap.add_argument("--stat-from", type=int, default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    if args.every % args.sample != 0:
        ap.error("--every must be a multiple of --sample")
    stat_from = args.steps // 2 if args.stat_from is None else args.stat_from
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
    print(f"[fp32_error] {args.config}: creating the FP64 form", flush=True)
    a = picnix_amd.Simulation(text64)
    print(f"[fp32_error] {args.config}: creating the FP32 form", flush=True)
    b = picnix_amd.Simulation(text32)
    c = None
    if args.twin:
        print(f"[fp32_error] {args.config}: creating the FP64 twin from the FP32 initial state",
              flush=True)
        c = picnix_amd.Simulation(text64, load=False)
        for k in range(c.nchunks):
            for sp in range(c.Ns):
                c.set_particles(k, sp, b.particles(k, sp))
    if args.graph == "on":
        a.step_graph(True)
        b.step_graph(True)
        if c is not None:
            c.step_graph(True)
    print(f"[fp32_error] {args.config}: created both forms in {time.perf_counter() - t0:.1f} s",
          flush=True)
    names = ["div(E)-rho", "div(B)", "E^2/2", "B^2/2"] + \
            [f"kinetic[{k}]" for k in range(a.Ns)]
    def relerr(h, ref):
        rel = {}
        for k in range(2, len(ref)):
            rel[names[k]] = abs(h[k] - ref[k]) / abs(ref[k]) if ref[k] != 0 else abs(h[k])
        rel["div(E)-rho (abs)"] = abs(h[0] - ref[0])
        return rel

    def line(tag, rel):
        return (f"{tag}: E^2/2 {rel['E^2/2']:.3e} B^2/2 {rel['B^2/2']:.3e} "
                f"kinetic {rel['kinetic[0]']:.3e} div(E)-rho {rel['div(E)-rho (abs)']:.3e}")

    rows = []
    samples = {"fp64": [], "fp32": [], "twin": []}
    step = 0
    while step < args.steps:
        n = min(args.sample, args.steps - step)
        a.step(dt, n)
        b.step(dt, n)
        if c is not None:
            c.step(dt, n)
        step += n
        ha, hb = a.history(), b.history()
        hc = c.history() if c is not None else None
        if step > stat_from:
            samples["fp64"].append(list(map(float, ha)))
            samples["fp32"].append(list(map(float, hb)))
            if hc is not None:
                samples["twin"].append(list(map(float, hc)))
        if step % args.every != 0 and step != args.steps:
            continue
        row = {"step": step, "fp64": list(map(float, ha)), "fp32": list(map(float, hb)),
               "rel_error": relerr(hb, ha)}
        msg = f"[fp32_error] step {step}: " + line("fp32 vs fp64", row["rel_error"])
        if hc is not None:
            row["twin"] = list(map(float, hc))
            row["twin_vs_fp64"] = relerr(hc, ha)
            row["fp32_vs_twin"] = relerr(hb, hc)
            msg += "; " + line("twin vs fp64", row["twin_vs_fp64"])
            msg += "; " + line("fp32 vs twin", row["fp32_vs_twin"])
        rows.append(row)
        print(msg, flush=True)
    a.close()
    b.close()
    if c is not None:
        c.close()
    stats = statistics(samples, names)
    if args.out:
        with open(args.out, "w") as fp:
            fp.write(report(args, dt, stat_from, rows, stats, names))
    print(json.dumps({"config": args.config, "steps": args.steps, "dt": dt, "rows": rows,
                      "statistics": stats}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
