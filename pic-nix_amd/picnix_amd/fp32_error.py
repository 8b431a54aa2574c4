"""Error of the FP32 form against the FP64 form over a long run.

    python -m picnix_amd.fp32_error --config thermal128 --steps 1000 --every 100

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
"""
import argparse
import json
import sys
import time

import picnix_amd


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="thermal128")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--graph", default="on", choices=["on", "off"])
    ap.add_argument("--twin", action="store_true")
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
    step = 0
    while step < args.steps:
        n = min(args.every, args.steps - step)
        a.step(dt, n)
        b.step(dt, n)
        if c is not None:
            c.step(dt, n)
        step += n
        ha, hb = a.history(), b.history()
        row = {"step": step, "fp64": list(map(float, ha)), "fp32": list(map(float, hb)),
               "rel_error": relerr(hb, ha)}
        msg = f"[fp32_error] step {step}: " + line("fp32 vs fp64", row["rel_error"])
        if c is not None:
            hc = c.history()
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
    print(json.dumps({"config": args.config, "steps": args.steps, "dt": dt, "rows": rows}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
