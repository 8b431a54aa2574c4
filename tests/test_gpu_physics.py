"""Physics pins of the full GPU step (the same set-ups as test_oracle_physics.py).

The two-stream instability of two cold counter-streaming beams (example/beam/
main.cpp:67-116 at a resolved wavelength: 1D, 512 cells, mode 16 the fastest
growing) started from a quiet start grows on the GPU at the rate of the cold
two-beam dispersion relation, and its mode amplitude follows the oracle's run
of the same input.
"""
import sys

import numpy as np
import pytest

import oracle_lib
from test_oracle_physics import growth_rate, mode_amplitude, twostream_quiet, twostream_theory

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2])
def test_gpu_twostream_growth_rate_linear_theory(order):
    import helpers
    import picnix_amd
    cfg, k, ro, fill = twostream_quiet(order=order)
    dt = float(cfg["parameter"]["delt"])
    print(f"[physics] two-stream order {order}: GPU run", flush=True)
    g = picnix_amd.Simulation(helpers.dumps(cfg))
    o = oracle_lib.Oracle(cfg)
    try:
        fill(g)
        fill(o)
        rate = growth_rate(g, lambda n: g.step(dt, n), dt, 512, 16, 800, 1400)
        gt = twostream_theory(k, 0.1, ro)
        print(f"[physics] two-stream order {order}: growth {rate:.6f}, theory {gt:.6f}", flush=True)
        sys.stdout.flush()
        assert abs(rate / gt - 1.0) <= 0.02, (rate, gt)
        o.step(1400)
        a, b = mode_amplitude(g, 512, 16), mode_amplitude(o, 512, 16)
        assert abs(a - b) <= 1e-6 * b, (a, b)
    finally:
        g.close()
        o.close()
