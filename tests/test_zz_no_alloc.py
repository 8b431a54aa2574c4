"""pn_get_energy_static: the energy sums with no device allocation.

Checked through the C ABI against pn_get_energy on the same handle: two calls in
a row on a 3D run, then a second, smaller handle (2D, fewer chunks and
particles) on the same GPU, whose call reuses the scratch array the first one
wrote; both handles again afterwards; and the FP32 form.  The two entry points
sum the same terms in different orders, so they agree to round-off (1e-12
relative).  The device's free memory does not change across the calls.
"""
import ctypes
import sys

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(f"[no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _check(sim, label):
    e0, b0, p0 = sim.energy()
    e1, b1, p1 = sim.energy_static()
    _say(f"{label}: E2 {e1:.6e} vs {e0:.6e}, B2 {b1:.6e} vs {b0:.6e}, kinetic {p1} vs {p0}")
    assert abs(e1 - e0) <= TOL * abs(e0), (e0, e1)
    assert abs(b1 - b0) <= TOL * max(abs(b0), 1e-300) or (b0 == 0 and b1 == 0), (b0, b1)
    assert np.allclose(p1, p0, rtol=TOL, atol=0), (p0, p1)
    return e1, b1, p1


def test_energy_static_matches_pn_get_energy():
    import picnix_amd
    big = helpers.small_thermal(N=(32, 32, 32), C=(2, 2, 2), np_=8, order=2)
    small = helpers.small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=4, order=1)
    dt = float(big["parameter"]["delt"])
    a = picnix_amd.Simulation(helpers.dumps(big))
    b = None
    try:
        _say("3D handle: 3 steps")
        a.step(dt, 3)
        a.synchronize()
        ra = _check(a, "3D first call")
        ra2 = _check(a, "3D second call")
        assert ra[0] == ra2[0] and ra[1] == ra2[1] and np.array_equal(ra[2], ra2[2])
        _say("2D handle (smaller): 2 steps")
        b = picnix_amd.Simulation(helpers.dumps(small))
        b.step(float(small["parameter"]["delt"]), 2)
        b.synchronize()
        free0 = _free_bytes()
        rb = _check(b, "2D call after the 3D calls")
        ra3 = _check(a, "3D again after the 2D call")
        assert ra3[0] == ra[0] and np.array_equal(ra3[2], ra[2])
        for _ in range(20):
            a.energy_static()
            b.energy_static()
        assert _free_bytes() == free0
        assert rb[2].shape == (2,)
    finally:
        a.close()
        if b is not None:
            b.close()


def test_energy_static_fp32_and_shock_rejected():
    import picnix_amd
    cfg = helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2, precision="fp32")
    s = picnix_amd.Simulation(helpers.dumps(cfg))
    try:
        s.step(float(cfg["parameter"]["delt"]), 2)
        _check(s, "FP32 3D")
    finally:
        s.close()
    sh = helpers.small_shock(N=(1, 1, 64), C=(1, 1, 4), nppc=4, order=2)
    s = picnix_amd.Simulation(helpers.dumps(sh))
    try:
        with pytest.raises(picnix_amd.PicnixError) as e:
            s.energy_static()
        assert e.value.code == -5
    finally:
        s.close()
