"""GPU parity of the shock problem (example/shock: [application] problem =
'shock') against the oracle on the same seeded inputs.

The physical x boundaries are on: the conducting wall and particle reflection
at x = 0, the fixed upstream fields and the redraw / Poisson injection of
particles at x = Lx, with the example's own random streams (std::mt19937_64,
libstdc++ distributions, Maxwell-Juttner).  Bars as for the periodic step:
particle order and values bit-exact at step 1, order exact and values within
1e-12 relative at step 10, fields and currents within 1e-12 relative, and the
per-chunk boundary state (random seed, last id, influx, efflux) identical.

1D runs compare with the reference's own operator order.  In 2D and 3D the
oracle resets the wall fields before the EMF pack as well (emf_order = 1): the
reference's order makes a wall chunk's private ghost rows along y and z hold a
neighbour's pre-reset Ex, which the global arrays cannot represent.
"""
import sys

import numpy as np
import pytest

import helpers
import oracle_lib

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = {
    # hot electrons (betae = 8, vth_e = 0.4c) so that some cross x = Lx in 10 steps
    "shock1d_o1": (1, dict(N=(1, 1, 128), C=(1, 1, 4), nppc=64, order=1, emf_order=0, betae=8.0)),
    "shock1d_o2": (1, dict(N=(1, 1, 128), C=(1, 1, 4), nppc=64, order=2, emf_order=0, betae=8.0)),
    "shock1d_o3": (1, dict(N=(1, 1, 96), C=(1, 1, 4), nppc=64, order=3, emf_order=0, betae=8.0)),
    "shock2d_o2": (2, dict(N=(1, 16, 64), C=(1, 2, 4), nppc=4, order=2, emf_order=1)),
    "shock3d_o2": (3, dict(N=(8, 8, 32), C=(2, 2, 2), nppc=2, order=2, emf_order=1)),
    "shock3d_ramp": (3, dict(N=(8, 8, 32), C=(2, 2, 4), nppc=2, order=1, emf_order=1,
                             density=3.0, ramp_x=12.0, ramp_w=8.0)),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(o, g, dim, exact):
    worst = 0.0
    for c in range(o.nchunks):
        for s in range(2):
            a = o.particles(c, s)
            b = g.particles(c, s)
            assert a.shape == b.shape, f"chunk {c} species {s}: np {a.shape[0]} vs {b.shape[0]}"
            assert np.array_equal(a[:, 6].view(np.int64), b[:, 6].view(np.int64)), \
                f"chunk {c} species {s}: particle order differs"
            if exact:
                assert np.array_equal(a[:, :6], b[:, :6]), f"chunk {c} species {s}: values differ"
            else:
                worst = max(worst, helpers.rel_err(b[:, :6], a[:, :6]))
        assert g.boundary_state(c) == o.shock_state(c), f"chunk {c}: boundary state"
    rel = {}
    for which, n in ((0, 6), (1, 4)):
        num = np.zeros(n)
        den = np.zeros(n)
        for c in range(o.nchunks):
            info = o.chunk_info(c)
            fo = helpers.interior(o.field(c, which), info, dim)
            fg = helpers.interior(g.field(c, which), info, dim)
            num = np.maximum(num, np.max(np.abs(fg - fo).reshape(-1, n), axis=0))
            den = np.maximum(den, np.max(np.abs(fo).reshape(-1, n), axis=0))
        for k in range(n):
            rel[f"{'uf' if which == 0 else 'uj'}{k}"] = num[k] / den[k] if den[k] > 0 else num[k]
    return worst, rel


@pytest.mark.parametrize("name", list(CASES))
def test_shock_parity(name):
    _shock_parity(name)


@pytest.mark.parametrize("name", ["shock2d_o2", "shock3d_o2", "shock3d_ramp"])
def test_shock_parity_fused_push(name, monkeypatch):
    """The row gather runs the position push (wall reflection, upstream
    marking, bins and leaver directions) for the particles it holds
    (PICNIX_FUSED_PUSH=1): the same bars as the separate kernels."""
    monkeypatch.setenv("PICNIX_FUSED_PUSH", "1")
    _shock_parity(name)


def _shock_parity(name):
    import picnix_amd
    dim, kw = CASES[name]
    cfg = helpers.small_shock(**kw)
    dt = float(cfg["parameter"]["delt"])
    _say(f"[shock] {name}: oracle + GPU set-up")
    o = oracle_lib.Oracle(cfg)
    g = picnix_amd.Simulation(helpers.dumps(cfg))
    try:
        _, rel = _compare(o, g, dim, exact=True)
        assert max(rel.values()) == 0.0, rel
        _say(f"[shock] {name}: initial state identical; step 1")
        o.step()
        g.step(dt)
        _, rel = _compare(o, g, dim, exact=True)
        assert max(rel.values()) <= TOL, rel
        injected = redrawn = 0
        for _ in range(9):
            o.step()
            g.step(dt)
            for c in range(o.nchunks):
                st = o.shock_state(c)
                injected += st[2] + st[3]
                redrawn += st[4] + st[5]
        _say(f"[shock] {name}: step 10")
        worst, rel = _compare(o, g, dim, exact=False)
        assert worst <= TOL, worst
        assert max(rel.values()) <= TOL, rel
        ho = o.history()
        hg = g.history()
        assert abs(hg[0] - ho[0]) < 1e-10 and abs(hg[1] - ho[1]) < 1e-10, (ho[:2], hg[:2])
        for k in range(2, len(ho)):
            assert abs(hg[k] - ho[k]) <= TOL * abs(ho[k]) + 1e-300, (k, ho[k], hg[k])
        # the boundaries acted: particles were injected and redrawn at x = Lx
        assert injected > 0 and redrawn > 0, (injected, redrawn)
        _say(f"[shock] {name}: ok ({injected} injected, {redrawn} redrawn, "
             f"worst particle {worst:.2e})")
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("name", ["shock1d_o2", "shock2d_o2", "shock3d_ramp"])
def test_shock_fp32_follows_fp64(name):
    """The FP32 form with the same boundaries: the injections (their draws do not
    depend on the particles) are identical, the redraws nearly so, and the
    energies follow the FP64 run to FP32 round-off after 10 steps."""
    import picnix_amd
    dim, kw = CASES[name]
    kw = dict(kw)
    kw.pop("emf_order", None)
    cfg64 = helpers.small_shock(**kw)
    cfg32 = helpers.small_shock(precision="fp32", **kw)
    dt = float(cfg64["parameter"]["delt"])
    a = picnix_amd.Simulation(helpers.dumps(cfg64))
    b = picnix_amd.Simulation(helpers.dumps(cfg32))
    try:
        inj = [0, 0]
        red = [0, 0]
        for _ in range(10):
            a.step(dt)
            b.step(dt)
            for c in range(a.nchunks):
                sa, sb = a.boundary_state(c), b.boundary_state(c)
                assert sa[2:4] == sb[2:4], (c, sa, sb)  # influx
                inj[0] += sa[2] + sa[3]
                inj[1] += sb[2] + sb[3]
                red[0] += sa[4] + sa[5]
                red[1] += sb[4] + sb[5]
        na = sum(a.np(c, s) for c in range(a.nchunks) for s in range(2))
        nb = sum(b.np(c, s) for c in range(b.nchunks) for s in range(2))
        ha, hb = a.history(), b.history()
        _say(f"[shock] {name} fp32 vs fp64: particles {nb} vs {na}, injected {inj}, "
             f"redrawn {red}, E2 {hb[2]:.6e} vs {ha[2]:.6e}, B2 {hb[3]:.6e} vs {ha[3]:.6e}, "
             f"K {hb[4]:.6e} vs {ha[4]:.6e}")
        assert inj[0] == inj[1] and inj[0] > 0
        assert abs(red[0] - red[1]) <= max(2, red[0] // 20)
        assert abs(na - nb) <= max(4, na // 1000)
        for k in (2, 3, 4, 5):
            assert abs(hb[k] - ha[k]) <= 2e-3 * abs(ha[k]), (k, ha[k], hb[k])
    finally:
        a.close()
        b.close()
