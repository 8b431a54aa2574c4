"""Build the FP64 CPU oracle (test infrastructure) into oracle/liboracle.so.

g++ with -ffp-contract=off so every expression is rounded exactly as written,
and OpenMP over chunks (the reference's `omp parallel for schedule(dynamic)`).
The reference's own hot-path headers cannot be compiled here: nix/nix.hpp
includes <mpi.h>, which this image lacks, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose=True):
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fPIC", "-shared", "-Wall", "-Wno-unused-variable", "-o", out,
           os.path.join(HERE, "picnix_oracle.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
