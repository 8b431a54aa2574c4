"""Summaries of rocprofv3 result databases (rocpd SQLite) for the bench workload.

    python -m picnix_amd.profsum trace <dir>/..._results.db
    python -m picnix_amd.profsum pmc [--probe-mb=M] FETCH_SIZE <fetch.db> [WRITE_SIZE <write.db>]
    python -m picnix_amd.profsum sq <sq.db> [kernel-prefix ...]

`trace` prints calls, total and average duration per kernel (from a
`--kernel-trace --stats` run), then, for kernels launched with several grids
(an autotuning session tries candidate shapes), the same per launch shape.  `pmc` prints the per-dispatch average of each
counter per kernel (KiB per dispatch for FETCH_SIZE / WRITE_SIZE) and, when the
run held the bench's copy probe (`bench.py --copy-probe-mb M`: one device copy
of M MiB by ROCm's copy kernel), the counter's calibration factor, i.e. the
bytes the probe is known to move divided by what the counter reported for it,
together with the calibrated bytes per dispatch.
"""
import sqlite3
import sys


def _short(name, width=64):
    name = name.replace("picnix::(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    i = name.find("(")
    if i > 0:
        name = name[:i]
    return name[:width]


def trace(db):
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average from top_kernels").fetchall()
    tot = sum(r[2] for r in rows) or 1.0
    out = [f"{'kernel':<48} {'calls':>6} {'total_us':>11} {'avg_us':>10} {'share':>6}"]
    for name, n, t, avg in rows:
        out.append(f"{_short(name, 48):<48} {n:>6} {t:>11.1f} {avg:>10.1f} {100 * t / tot:>5.1f}%")
    return "\n".join(out)


def trace_shapes(db):
    """Kernels launched with more than one grid (the autotuner's candidates): calls
    and average duration per launch shape, the chosen shape being the one with the
    most calls."""
    c = sqlite3.connect(db)
    rows = c.execute(
        "select name, grid_x, grid_y, grid_z, count(*), avg(duration) from kernels "
        "group by name, grid_x, grid_y, grid_z order by name, count(*) desc").fetchall()
    by = {}
    for name, gx, gy, gz, n, avg in rows:
        by.setdefault(name, []).append((gx, gy, gz, n, avg))
    out = [f"{'kernel (launch shapes)':<48} {'grid (threads)':>22} {'calls':>6} {'avg_us':>10}"]
    for name, shapes in by.items():
        if len(shapes) < 2:
            continue
        for gx, gy, gz, n, avg in shapes:
            out.append(f"{_short(name, 48):<48} {f'{gx}x{gy}x{gz}':>22} {n:>6} {avg / 1e3:>10.1f}")
    return "\n".join(out)


def _probe_value(c, counter):
    """Counter value of the largest __amd_rocclr_copyBuffer dispatch: the copy probe."""
    r = c.execute("select max(value) from counters_collection where counter_name=? and "
                  "kernel_name like '%copyBuffer%'", (counter,)).fetchone()
    return r[0] if r else None


def pmc(pairs, probe_mb=None):
    out = []
    for counter, db in pairs:
        c = sqlite3.connect(db)
        rows = c.execute(
            "select kernel_name, count(*), avg(value) from counters_collection where counter_name=? "
            "group by kernel_name order by avg(value) * count(*) desc", (counter,)).fetchall()
        factor = None
        pv = _probe_value(c, counter)
        if probe_mb and pv:
            factor = probe_mb * 1024.0 / pv
        head = f"{counter}" + (f"  (copy probe {probe_mb} MiB, counter {pv:.0f} KiB -> factor {factor:.3f})"
                               if factor else "")
        out.append(head)
        out.append(f"  {'kernel':<48} {'n':>4} {'KiB/dispatch':>14} {'MB/dispatch (cal.)':>19}")
        for name, n, avg in rows:
            cal = f"{avg * factor * 1024 / 1e6:>19.1f}" if factor else f"{'':>19}"
            out.append(f"  {_short(name, 48):<48} {n:>4} {avg:>14.1f} {cal}")
    return "\n".join(out)


def sq(db, prefixes=()):
    """Per-dispatch averages of every counter of a multi-counter --pmc run, one row
    per kernel (the kernels whose names start with one of `prefixes`, or all)."""
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, counter_name, count(*), avg(value) from "
                     "counters_collection group by kernel_name, counter_name").fetchall()
    table, counters = {}, []
    for name, counter, n, avg in rows:
        short = _short(name, 40)
        if prefixes and not any(short.startswith(p) for p in prefixes):
            continue
        table.setdefault(short, {})[counter] = (n, avg)
        if counter not in counters:
            counters.append(counter)
    counters.sort()
    out = [f"{'kernel':<40} {'n':>4} " + " ".join(f"{k:>22}" for k in counters)]
    for name in sorted(table):
        vals = table[name]
        n = max(v[0] for v in vals.values())
        out.append(f"{name:<40} {n:>4} " +
                   " ".join(f"{vals[k][1]:>22.4g}" if k in vals else f"{'':>22}" for k in counters))
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace(argv[1]))
        shapes = trace_shapes(argv[1])
        if shapes.count("\n"):
            print()
            print(shapes)
        return 0
    if len(argv) >= 3 and argv[0] == "pmc":
        probe = None
        args = argv[1:]
        if args[0].startswith("--probe-mb="):
            probe = float(args[0].split("=", 1)[1])
            args = args[1:]
        pairs = [(args[i], args[i + 1]) for i in range(0, len(args) - 1, 2)]
        print(pmc(pairs, probe))
        return 0
    if len(argv) >= 2 and argv[0] == "sq":
        print(sq(argv[1], tuple(argv[2:])))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
