"""Autotuner (pn_autotune): a session picks launch shapes on the run's own steps
and leaves the physics unchanged (to the order of the current sums)."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_session_matches_untuned_run(precision):
    import picnix_amd
    cfg = helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2,
                                precision=precision)
    dt = float(cfg["parameter"]["delt"])
    text = helpers.dumps(cfg)
    a = picnix_amd.Simulation(text)
    b = picnix_amd.Simulation(text)
    try:
        a.autotune(True)
        t0 = a.tuning()
        assert t0["state"] == 1 and t0["steps_left"] >= 4
        n = t0["steps_left"]
        for _ in range(n + 3):
            a.step(dt)
            b.step(dt)
        t1 = a.tuning()
        print(f"[autotune] {precision}: {n} tuning steps -> {t1}", flush=True)
        assert t1["state"] == 2 and t1["steps_left"] == 0
        assert t1["deposit_rows"] >= 1 and t1["rho_rows"] >= 0
        assert t1["gather_rows"] in (-1, 1, 2, 4)
        assert t1["gather_per_species"] in (0, 1)
        assert t1["deposit_queue"] in (0, 1) and t1["gather_queue"] in (0, 1)
        assert t1["fused_push"] in (0, 1)
        tol = 1e-12 if precision == "fp64" else 1e-4
        for c in range(a.nchunks):
            info = a.chunk_info(c)
            fa = helpers.interior(a.field(c, 0), info, 3)
            fb = helpers.interior(b.field(c, 0), info, 3)
            assert helpers.rel_err(fa, fb) <= tol
            if precision == "fp64":
                for s in range(2):
                    pa, pb = a.particles(c, s), b.particles(c, s)
                    assert np.array_equal(pa[:, 6].view(np.int64), pb[:, 6].view(np.int64))
                    assert helpers.rel_err(pa[:, :6], pb[:, :6]) <= tol
        a.autotune(False)
        assert a.tuning()["state"] == 0 and a.tuning()["deposit_rows"] == 0
        assert a.tuning()["gather_rows"] == 0
    finally:
        a.close()
        b.close()
