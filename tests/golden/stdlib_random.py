"""Write stdlib_random.json: reference outputs of this image's libstdc++
std::gamma_distribution<double> and std::poisson_distribution<int> (the
distributions example/shock draws through nix/random.hpp:9-12), each driven by
std::mt19937_64(seed).  The oracle's restatements are checked against them.

    python tests/golden/stdlib_random.py

Compiles a short C++ program with g++ in a temporary directory and runs it.
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

PROGRAM = r"""
#include <cstdio>
#include <random>
int main()
{
  const double gam[][2] = {{1.5, 0.0025}, {2.0, 0.5}, {2.5, 2.0}, {3.0, 1e-4}, {0.7, 1.3}};
  const double mean[]   = {0.3, 1.6, 6.3, 11.5};
  std::printf("{\"gamma\": [");
  for (int k = 0; k < 5; k++) {
    std::mt19937_64 g(1000 + k);
    std::gamma_distribution<double> d(gam[k][0], gam[k][1]);
    std::printf("%s{\"alpha\": %.17g, \"beta\": %.17g, \"seed\": %d, \"values\": [", k ? ", " : "",
                gam[k][0], gam[k][1], 1000 + k);
    for (int i = 0; i < 64; i++)
      std::printf("%s%.17g", i ? ", " : "", d(g));
    std::printf("]}");
  }
  std::printf("], \"poisson\": [");
  for (int k = 0; k < 4; k++) {
    std::mt19937_64 g(2000 + k);
    std::poisson_distribution<int> d(mean[k]);
    std::printf("%s{\"mean\": %.17g, \"seed\": %d, \"values\": [", k ? ", " : "", mean[k], 2000 + k);
    for (int i = 0; i < 256; i++)
      std::printf("%s%d", i ? ", " : "", d(g));
    std::printf("]}");
  }
  std::printf("]}\n");
  return 0;
}
"""


def main():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "gen.cpp")
        exe = os.path.join(tmp, "gen")
        with open(src, "w") as f:
            f.write(PROGRAM)
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    data["generator"] = "g++ " + subprocess.run(["g++", "-dumpfullversion"], capture_output=True,
                                                text=True).stdout.strip() + " libstdc++"
    with open(os.path.join(HERE, "stdlib_random.json"), "w") as f:
        json.dump(data, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
