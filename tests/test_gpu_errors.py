"""Error paths of the C ABI on the GPU: a particle segment that overflows.

The reference grows its particle buffers when a chunk receives more particles
than it holds (nix/xtensor_halo3d.hpp:406-418); this library keeps a fixed
capacity per (chunk, species) segment, round_up_alloc(mp * (1 + buffer_ratio))
(nix/particle.hpp:146-153), and reports PN_ERR_CAPACITY instead.  The kernels
must stay inside the segments while it happens, and the error must come back
from pn_synchronize and from the following pn_step.
"""
import sys

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu

PN_ERR_CAPACITY = -4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_capacity_overflow_is_reported(precision):
    import picnix_amd
    cfg = helpers.small_thermal(N=(8, 8, 8), C=(1, 1, 2), np_=1, order=2,
                                buffer_ratio=0.0, precision=precision)
    _say(f"[errors] capacity overflow {precision}: create")
    sim = picnix_amd.Simulation(helpers.dumps(cfg))
    # capacity of a segment: the largest count pn_set_particles accepts
    cap = 256 * 3 // 2
    for c in range(sim.nchunks):
        for s in range(sim.Ns):
            sim.set_particles(c, s, np.zeros((0, 7)))
    with pytest.raises(picnix_amd.PicnixError):
        sim.set_particles(0, 0, np.zeros((cap + 1, 7)))
    rng = np.random.default_rng(7)
    lo = {c: float(sim.chunk_info(c)[5]) for c in range(sim.nchunks)}
    src = min(lo, key=lo.get)           # x in [0, 4)
    dst = max(lo, key=lo.get)           # x in [4, 8)
    # the destination is nearly full and at rest; 300 particles next to its
    # lower x face move into it at 0.9c
    n_dst, n_src = cap - 4, 300
    a = np.zeros((n_dst, 7))
    a[:, 0] = 4.5 + 3.0 * rng.random(n_dst)
    a[:, 1] = 8.0 * rng.random(n_dst)
    a[:, 2] = 8.0 * rng.random(n_dst)
    a[:, 6] = np.arange(n_dst, dtype=np.int64).view(np.float64)
    b = np.zeros((n_src, 7))
    b[:, 0] = 3.99
    b[:, 1] = 8.0 * rng.random(n_src)
    b[:, 2] = 8.0 * rng.random(n_src)
    b[:, 3] = 0.9
    b[:, 6] = np.arange(n_dst, n_dst + n_src, dtype=np.int64).view(np.float64)
    sim.set_particles(dst, 0, a)
    sim.set_particles(src, 0, b)
    dt = float(cfg["parameter"]["delt"])
    _say(f"[errors] capacity overflow {precision}: step")
    sim.step(dt)
    with pytest.raises(picnix_amd.PicnixError) as e:
        sim.synchronize()
    assert e.value.code == PN_ERR_CAPACITY
    # the count was clamped to the capacity, and the next step refuses to run
    assert sim.np(dst, 0) == cap
    with pytest.raises(picnix_amd.PicnixError) as e:
        sim.step(dt)
    assert e.value.code == PN_ERR_CAPACITY
    _say(f"[errors] capacity overflow {precision}: reported")
    sim.close()


def test_capacity_within_limit_is_silent():
    """The same set-up with room to spare runs without an error and keeps
    every particle."""
    import picnix_amd
    cfg = helpers.small_thermal(N=(8, 8, 8), C=(1, 1, 2), np_=1, order=2, buffer_ratio=0.0)
    sim = picnix_amd.Simulation(helpers.dumps(cfg))
    for c in range(sim.nchunks):
        for s in range(sim.Ns):
            sim.set_particles(c, s, np.zeros((0, 7)))
    lo = {c: float(sim.chunk_info(c)[5]) for c in range(sim.nchunks)}
    src = min(lo, key=lo.get)
    b = np.zeros((300, 7))
    b[:, 0] = 3.99
    b[:, 1] = np.linspace(0.1, 7.9, 300)
    b[:, 2] = 4.0
    b[:, 3] = 0.9
    b[:, 6] = np.arange(300, dtype=np.int64).view(np.float64)
    sim.set_particles(src, 0, b)
    sim.step(float(cfg["parameter"]["delt"]))
    sim.synchronize()
    assert sum(sim.np(c, 0) for c in range(sim.nchunks)) == 300
    sim.close()
