"""Physics pins of the oracle's full step (CPU only).

The operator-level checks pin each piece of the step on its own; these run the
whole step (PicApplication::push_openmp, pic/pic_application.cpp:219-292) and
compare it with what the mathematics fixes:

  * the two-stream instability of two cold, counter-streaming electron beams
    (the set-up of example/beam/main.cpp:67-116) grows at the rate of the cold
    two-beam dispersion relation
        1 = w_b^2 / (w - k v0)^2 + w_b^2 / (w + k v0)^2,
    w_b^2 = |qm|^2 ro / gamma0^3 per beam (the loader gives every particle
    m = ro / np and q = qm m, so a beam's plasma frequency squared is qm^2 ro in
    these units, with the longitudinal mass gamma0^3 m of a beam moving at
    u0 = gamma0 v0).  Its unstable root is
        w^2 = k^2 v0^2 + w_b^2 - w_b sqrt(w_b^2 + 4 k^2 v0^2).
    The beams start from a quiet start (evenly spaced particles displaced by a
    small sinusoid of one mode), so the growing eigenmode dominates after a few
    e-foldings and its growth is measured without shot noise.  A wrong charge or
    current normalisation, a wrong sign in the field update or the pusher, or a
    wrong relativistic factor each change the rate by far more than the bar;
  * the plasma frequency of a cold plasma at rest is sqrt(sum qm^2 ro);
  * the total energy of a thermal plasma whose Debye length is about a cell
    (little grid heating) stays within a bound over 200 steps, and its change
    shrinks as the time step is halved.
"""
import numpy as np
import pytest

import helpers
import oracle_lib as ol


def twostream_theory(k, u0, ro):
    """Growth rate of the cold symmetric two-beam instability at wavenumber k."""
    g0 = np.sqrt(1.0 + u0 * u0)
    v0 = u0 / g0
    wb2 = ro / g0 ** 3
    w2 = k * k * v0 * v0 + wb2 - np.sqrt(wb2 * (wb2 + 4.0 * k * k * v0 * v0))
    return np.sqrt(-w2) if w2 < 0 else 0.0


def twostream_quiet(N=512, chunks=8, npc=64, order=1, mode=16, u0=0.1, eps=1e-5):
    """Config and quiet-start particle arrays for a 1D two-stream run whose most
    unstable wavenumber is the given mode: returns (cfg, k, ro, fill) where
    fill(sim) replaces every chunk's particles by the quiet start."""
    g0 = np.sqrt(1.0 + u0 * u0)
    v0 = u0 / g0
    k = 2.0 * np.pi * mode / N
    # k v0 = (sqrt(3) / 2) w_b: the fastest-growing wavenumber of the cold beams
    ro = (2.0 * k * v0 / np.sqrt(3.0)) ** 2 * g0 ** 3
    cfg = helpers.small_twostream(N=(1, 1, N), C=(1, 1, chunks), np_=npc, order=order)
    for s, sgn in zip(cfg["parameter"]["particle"], (1.0, -1.0)):
        s["ro"] = float(ro)
        s["vt"] = 0.0
        s["vx"] = sgn * u0
    L = N // chunks

    def fill(sim):
        for c in range(sim.nchunks):
            x0 = float(sim.chunk_info(c)[5])
            for s in range(2):
                a = sim.particles(c, s).copy()
                n = a.shape[0]
                x = x0 + (np.arange(n) + 0.5) * (L / n)
                x = np.sort(x + eps / k * np.sin(k * x))
                a[:, 0] = x
                a[:, 3] = u0 if s == 0 else -u0
                a[:, 4] = 0.0
                a[:, 5] = 0.0
                sim.set_particles(c, s, a)
        sim.sort_particle()
    return cfg, k, ro, fill


def mode_amplitude(sim, N, mode, dim=1):
    """|Fourier coefficient| of Ex along x (1D runs, chunks ordered by offset)."""
    parts = []
    for c in range(sim.nchunks):
        info = sim.chunk_info(c)
        parts.append((int(info[5]), helpers.interior(sim.field(c, 0), info, dim)[0, 0, :, 0]))
    ex = np.concatenate([e for _, e in sorted(parts, key=lambda t: t[0])])
    return abs(np.fft.rfft(ex)[mode]) / N


def growth_rate(sim, step, dt, N, mode, n0, n1, every=20):
    """Least-squares slope of log|Ex(mode)| over steps [n0, n1]."""
    ts, la = [], []
    done = 0
    for n in range(0, n1 + 1, every):
        if n > done:
            step(n - done)
            done = n
        if n >= n0:
            ts.append(n * dt)
            la.append(np.log(mode_amplitude(sim, N, mode)))
    return np.polyfit(ts, la, 1)[0]


@pytest.mark.parametrize("order", [1, 2])
def test_twostream_growth_rate_linear_theory(order):
    ol.set_threads(8)
    cfg, k, ro, fill = twostream_quiet(order=order)
    o = ol.Oracle(cfg)
    try:
        fill(o)
        dt = float(cfg["parameter"]["delt"])
        g = growth_rate(o, o.step, dt, 512, 16, 800, 1400)
        gt = twostream_theory(k, 0.1, ro)
        assert abs(g / gt - 1.0) <= 0.02, (g, gt)
    finally:
        o.close()


def plasma_quiet(ro, mode=8, N=512, chunks=8, npc=64, eps=1e-4):
    """Cold plasma of two electron species at rest on evenly spaced positions,
    with a velocity perturbation ux = eps sin(k x) of one mode (E starts at zero,
    so a density perturbation alone would never move)."""
    cfg = helpers.small_twostream(N=(1, 1, N), C=(1, 1, chunks), np_=npc, order=1)
    for s in cfg["parameter"]["particle"]:
        s["ro"] = ro
        s["vt"] = 0.0
        s["vx"] = 0.0
    k = 2.0 * np.pi * mode / N
    L = N // chunks

    def fill(sim):
        for c in range(sim.nchunks):
            x0 = float(sim.chunk_info(c)[5])
            for s in range(2):
                a = sim.particles(c, s).copy()
                n = a.shape[0]
                a[:, 0] = x0 + (np.arange(n) + 0.5) * (L / n)
                a[:, 3] = eps * np.sin(k * a[:, 0])
                a[:, 4] = 0.0
                a[:, 5] = 0.0
                sim.set_particles(c, s, a)
        sim.sort_particle()
    return cfg, fill


def test_plasma_frequency():
    """A cold plasma at rest with a small mode-8 velocity perturbation oscillates
    at w_p = sqrt(sum qm^2 ro): the zero crossings of Ex(mode 8) are half a
    plasma period apart."""
    ol.set_threads(8)
    ro = 0.01
    cfg, fill = plasma_quiet(ro)
    o = ol.Oracle(cfg)
    try:
        fill(o)
        dt = float(cfg["parameter"]["delt"])
        sig = []
        for n in range(400):
            parts = []
            for c in range(o.nchunks):
                info = o.chunk_info(c)
                parts.append((int(info[5]), helpers.interior(o.field(c, 0), info, 1)[0, 0, :, 0]))
            ex = np.concatenate([e for _, e in sorted(parts, key=lambda t: t[0])])
            sig.append(np.fft.rfft(ex)[8].imag)
            o.step()
        sig = np.array(sig)
        # linearly interpolated zero crossings
        z = [i + sig[i] / (sig[i] - sig[i + 1]) for i in range(len(sig) - 1)
             if sig[i] == 0.0 or sig[i] * sig[i + 1] < 0]
        half = np.mean(np.diff(z)) * dt
        wp = np.pi / half
        wt = np.sqrt(2 * ro)
        # finite-difference and shape corrections at k dx = 0.1 are ~1e-3
        assert abs(wp / wt - 1.0) <= 5e-3, (wp, wt)
    finally:
        o.close()


def test_energy_heating_bound():
    """Thermal plasma (electrons + protons, order 2) at 1/100 of the benchmark
    density, so the electron Debye length is about one cell: over 100 plasma
    time units the total (field + kinetic) energy changes by < 2e-4 relative,
    and the change converges as dt is halved (successive differences shrink)."""
    ol.set_threads(8)
    drift = []
    for dt in (0.5, 0.25, 0.125):
        cfg = helpers.small_thermal(N=(8, 16, 16), C=(1, 2, 2), np_=16, order=2)
        cfg["parameter"]["delt"] = dt
        for s in cfg["parameter"]["particle"]:
            s["ro"] = s["ro"] * 0.01
        o = ol.Oracle(cfg)
        try:
            h0 = o.history()
            e0 = h0[2] + h0[3] + h0[4:].sum()
            o.step(int(round(100.0 / dt)))
            h = o.history()
            e1 = h[2] + h[3] + h[4:].sum()
        finally:
            o.close()
        drift.append((e1 - e0) / e0)
    assert all(abs(d) < 2e-4 for d in drift), drift
    d1, d2 = abs(drift[0] - drift[1]), abs(drift[1] - drift[2])
    assert d2 < 0.75 * d1, drift
