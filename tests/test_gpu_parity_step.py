"""GPU parity of the full FP64 step against the oracle (same seeded inputs).

Bars (from the task's parity definition):
  * particle order and cell indices: bit-exact;
  * particle positions/momenta: bit-exact at step 1 (identical arithmetic), and
    within 1e-12 relative afterwards (the current is summed in a different
    order, which perturbs the fields at round-off);
  * fields and currents: <= 1e-12 relative (max |diff| / max |value|) after 1 and
    10 steps;
  * div(E) - rho stays at round-off.
"""
import sys

import numpy as np
import pytest

import helpers
import oracle_lib

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _sim(cfg):
    import picnix_amd
    return picnix_amd.Simulation(helpers.dumps(cfg))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare_particles(o, g, exact):
    worst = 0.0
    for c in range(o.nchunks):
        for s in range(o.Ns):
            a = o.particles(c, s)
            b = g.particles(c, s)
            assert a.shape == b.shape, f"chunk {c} species {s}: np {a.shape[0]} vs {b.shape[0]}"
            # ids identify particles: same order <=> identical sort
            assert np.array_equal(a[:, 6].view(np.int64), b[:, 6].view(np.int64)), \
                f"chunk {c} species {s}: particle order differs"
            if exact:
                assert np.array_equal(a[:, :6], b[:, :6]), f"chunk {c} species {s}: values differ"
            else:
                worst = max(worst, helpers.rel_err(b[:, :6], a[:, :6]))
    return worst


def _compare_fields(o, g, dim):
    worst = {}
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        uf_o = helpers.interior(o.field(c, 0), info, dim)
        uf_g = helpers.interior(g.field(c, 0), info, dim)
        uj_o = helpers.interior(o.field(c, 1), info, dim)
        uj_g = helpers.interior(g.field(c, 1), info, dim)
        for k in range(6):
            worst[f"uf{k}"] = max(worst.get(f"uf{k}", 0.0), np.max(np.abs(uf_g[..., k] - uf_o[..., k])))
            worst[f"uf{k}_scale"] = max(worst.get(f"uf{k}_scale", 0.0), np.max(np.abs(uf_o[..., k])))
        for k in range(4):
            worst[f"uj{k}"] = max(worst.get(f"uj{k}", 0.0), np.max(np.abs(uj_g[..., k] - uj_o[..., k])))
            worst[f"uj{k}_scale"] = max(worst.get(f"uj{k}_scale", 0.0), np.max(np.abs(uj_o[..., k])))
    rel = {}
    for k in [f"uf{i}" for i in range(6)] + [f"uj{i}" for i in range(4)]:
        sc = worst[k + "_scale"]
        rel[k] = worst[k] / sc if sc > 0 else worst[k]
    return rel


CASES = {
    "thermal3d_o2": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2)),
    "thermal3d_o1": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=1)),
    "thermal3d_o3": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=4, order=3)),
    "thermal3d_o4": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=4, order=4)),
    "twostream2d_o1": (2, lambda: helpers.small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1)),
    "twostream2d_o2": (2, lambda: helpers.small_twostream(N=(1, 16, 64), C=(1, 2, 4), np_=16, order=2)),
    # 1D (Nz = Ny = 1): Yee 1D, 1D gather, Esirkepov 1D
    "twostream1d_o1": (1, lambda: helpers.small_twostream(N=(1, 1, 128), C=(1, 1, 8), np_=64, order=1)),
    "thermal1d_o2": (1, lambda: helpers.small_thermal(N=(1, 1, 64), C=(1, 1, 4), np_=32, order=2)),
    "thermal1d_o3": (1, lambda: helpers.small_thermal(N=(1, 1, 64), C=(1, 1, 4), np_=32, order=3)),
    "thermal1d_o4": (1, lambda: helpers.small_thermal(N=(1, 1, 64), C=(1, 1, 4), np_=32, order=4)),
}


@pytest.mark.parametrize("kind", list(CASES))
def test_step_parity(kind):
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[parity] {kind}: create oracle + GPU")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    _say(f"[parity] {kind}: compare loader + initial sort")
    _compare_particles(o, g, exact=True)

    _say(f"[parity] {kind}: step 1")
    o.step()
    g.step(dt)
    g.synchronize()
    _compare_particles(o, g, exact=True)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] {kind}: step 1 field rel errors {rel}")
    assert max(rel.values()) <= TOL

    for k in range(9):
        o.step()
        g.step(dt)
    g.synchronize()
    _say(f"[parity] {kind}: step 10")
    perr = _compare_particles(o, g, exact=False)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] {kind}: step 10 particle rel err {perr:.3e}, fields {rel}")
    assert perr <= TOL
    assert max(rel.values()) <= TOL

    ho = o.history()
    hg = g.history()
    _say(f"[parity] {kind}: history oracle {ho} gpu {hg}")
    # div(E) - rho: both sides hold the same value to round-off (it is 0 for the
    # neutral thermal plasma; for the two-stream beams it is -sum(rho) because the
    # immobile ions are a background that is not deposited); div(B) ~ 0
    ncell = np.prod([cfg["parameter"][k] for k in ("Nx", "Ny", "Nz")])
    assert abs(hg[0] - ho[0]) < 1e-11 * ncell
    assert abs(hg[1]) < 1e-11 * ncell
    assert np.allclose(hg[2:4], ho[2:4], rtol=1e-10, atol=1e-14)
    assert np.allclose(hg[4:], ho[4:], rtol=1e-10)
    g.close()
    o.close()


@pytest.mark.parametrize("rows", [2, 1000])
def test_step_parity_march_rows(rows, monkeypatch):
    """The 3D deposit blocks walk `rows` bin rows each (PICNIX_MARCH_ROWS; 1000 =
    whole columns), so the z-ring of the LDS tile wraps and flushes mid-block."""
    monkeypatch.setenv("PICNIX_MARCH_ROWS", str(rows))
    for kind in ("thermal3d_o2", "thermal3d_o3"):
        dim, make = CASES[kind]
        cfg = make()
        dt = float(cfg["parameter"]["delt"])
        _say(f"[parity] march rows {rows} {kind}: create")
        o = oracle_lib.Oracle(cfg)
        g = _sim(cfg)
        for k in range(4):
            o.step()
            g.step(dt)
        g.synchronize()
        perr = _compare_particles(o, g, exact=False)
        rel = _compare_fields(o, g, dim)
        _say(f"[parity] march rows {rows} {kind}: step 4 particle rel err {perr:.3e}, fields {rel}")
        assert perr <= TOL
        assert max(rel.values()) <= TOL
        g.close()
        o.close()


@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o3", "twostream2d_o1"])
def test_step_parity_work_queue(kind, monkeypatch):
    """The march deposit and the row gather as persistent blocks walking the
    device work queue (PICNIX_WORK_QUEUE=1): same steps as one block per item.
    Ten steps also exercise the queue counters' reset between launches."""
    monkeypatch.setenv("PICNIX_WORK_QUEUE", "1")
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[parity] work queue {kind}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    for k in range(10):
        o.step()
        g.step(dt)
    g.synchronize()
    perr = _compare_particles(o, g, exact=False)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] work queue {kind}: step 10 particle rel err {perr:.3e}, fields {rel}")
    assert perr <= TOL
    assert max(rel.values()) <= TOL
    g.close()
    o.close()


@pytest.mark.parametrize("queue", ["0", "1"])
@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o1", "thermal3d_o4",
                                  "twostream2d_o1", "twostream2d_o2"])
def test_step_parity_fused_push(kind, queue, monkeypatch):
    """The row gather also runs each particle's position push, bin count and
    leaver marking (PICNIX_FUSED_PUSH=1), and a separate pass rebuilds the
    per-tile leaver counts from the keys: particles bit-exact at step 1 (same
    arithmetic, same sort), everything within 1e-12 at step 10."""
    monkeypatch.setenv("PICNIX_FUSED_PUSH", "1")
    monkeypatch.setenv("PICNIX_WORK_QUEUE", queue)
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[parity] fused push {kind} queue {queue}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    o.step()
    g.step(dt)
    g.synchronize()
    _compare_particles(o, g, exact=True)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] fused push {kind}: step 1 fields {rel}")
    assert max(rel.values()) <= TOL
    for k in range(9):
        o.step()
        g.step(dt)
    g.synchronize()
    perr = _compare_particles(o, g, exact=False)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] fused push {kind}: step 10 particle rel err {perr:.3e}, fields {rel}")
    assert perr <= TOL
    assert max(rel.values()) <= TOL
    g.close()
    o.close()


@pytest.mark.parametrize("kind", ["thermal3d_o2", "twostream2d_o1"])
def test_step_parity_sort_in_place(kind, monkeypatch):
    """pn_step's sort moves the records itself (PICNIX_DEFER_SORT=0) instead of
    leaving the move to the next step's row gather (the default, which every
    other parity test runs): same bars."""
    monkeypatch.setenv("PICNIX_DEFER_SORT", "0")
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[parity] sort in place {kind}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    for k in range(10):
        o.step()
        g.step(dt)
    g.synchronize()
    perr = _compare_particles(o, g, exact=False)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] sort in place {kind}: step 10 particle rel err {perr:.3e}, fields {rel}")
    assert perr <= TOL
    assert max(rel.values()) <= TOL
    g.close()
    o.close()


def _with_fields(cfg, E=(0.0, 0.0, 0.0), B=(0.0, 0.0, 0.0)):
    p = cfg["parameter"]
    p.update(Ex=E[0], Ey=E[1], Ez=E[2], Bx=B[0], By=B[1], Bz=B[2])
    return cfg


OPTION_CASES = {
    # pushers under a background field strong enough to rotate the electrons
    "vay_3d_o2": (3, lambda: _with_fields(helpers.small_thermal(
        N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2, pusher="Vay"),
        E=(0.002, 0.0, -0.001), B=(0.3, 0.1, 0.5))),
    "hc_2d_o1": (2, lambda: _with_fields(helpers.small_twostream(
        N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1, pusher="HigueraCary"),
        B=(0.0, 0.2, 0.4))),
    # Friedman filter theta > 0 (ff history kept on the GPU)
    "friedman_3d_o1": (3, lambda: helpers.small_thermal(
        N=(16, 16, 16), C=(2, 2, 2), np_=8, order=1, friedman=0.1)),
    "friedman_vay_2d_o2": (2, lambda: _with_fields(helpers.small_twostream(
        N=(1, 16, 64), C=(1, 2, 4), np_=16, order=2, friedman=0.2, pusher="Vay"),
        B=(0.1, 0.0, 0.3))),
    # WT interpolation of the integer-grid weights (nix/primitives.hpp:331-566)
    "wt_3d_o2": (3, lambda: _with_fields(helpers.small_thermal(
        N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2, interpolation="WT"),
        E=(0.001, 0.002, 0.0), B=(0.2, 0.0, 0.1))),
    "wt_hc_2d_o3": (2, lambda: _with_fields(helpers.small_twostream(
        N=(1, 16, 64), C=(1, 1, 4), np_=16, order=3, interpolation="WT", pusher="HigueraCary"),
        E=(0.0, 0.001, 0.0), B=(0.0, 0.1, 0.2))),
    "friedman_wt_vay_1d_o2": (1, lambda: _with_fields(helpers.small_thermal(
        N=(1, 1, 64), C=(1, 1, 4), np_=32, order=2, friedman=0.2, interpolation="WT",
        pusher="Vay"), E=(0.001, 0.0, 0.0), B=(0.1, 0.2, 0.3))),
}


@pytest.mark.parametrize("kind", list(OPTION_CASES))
def test_step_parity_options(kind):
    """Vay / Higuera-Cary pushers (nix/primitives.hpp:191-253), the Friedman
    filter (pic/engine/maxwell.hpp) and WT interpolation against the oracle:
    particles bit-exact at step 1, everything within 1e-12 at step 10."""
    dim, make = OPTION_CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[parity] options {kind}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    _compare_particles(o, g, exact=True)
    o.step()
    g.step(dt)
    g.synchronize()
    _compare_particles(o, g, exact=True)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] options {kind}: step 1 fields {rel}")
    assert max(rel.values()) <= TOL
    for k in range(9):
        o.step()
        g.step(dt)
    g.synchronize()
    perr = _compare_particles(o, g, exact=False)
    rel = _compare_fields(o, g, dim)
    _say(f"[parity] options {kind}: step 10 particle rel err {perr:.3e}, fields {rel}")
    assert perr <= TOL
    assert max(rel.values()) <= TOL
    g.close()
    o.close()


@pytest.mark.parametrize("queue", ["0", "1"])
@pytest.mark.parametrize("kind", ["thermal3d_o2", "twostream2d_o1"])
def test_step_graph_replays(kind, queue, monkeypatch):
    """pn_step replayed from captured hipGraphs (one per particle-buffer parity),
    with eager work between the replays (energy reductions, a particle read-back,
    an eager operator step), checked against the oracle after every replay.
    queue = 1: the graphs hold the persistent work-queue kernels, whose counters
    must be back at zero for every replay."""
    monkeypatch.setenv("PICNIX_WORK_QUEUE", queue)
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[parity] graph {kind} queue {queue}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    g.step_graph(True)
    for it in range(6):
        o.step()
        g.step(dt)
        g.synchronize()
        perr = _compare_particles(o, g, exact=(it == 0))
        rel = _compare_fields(o, g, dim)
        _say(f"[parity] graph {kind}: step {it + 1} particle {perr:.2e} fields max {max(rel.values()):.2e}")
        assert perr <= TOL
        assert max(rel.values()) <= TOL
        g.energy()                      # eager reduction between replays
        if it == 2:
            # an operator-by-operator step in between: eager, then graphs again
            o.step()
            g.push_bfd(0.5 * dt)
            g.push_velocity(dt)
            g.push_position(dt)
            g.deposit_current(dt)
            g.push_bfd(0.5 * dt)
            g.exchange(1)
            g.push_efd(dt)
            g.exchange(0)
            g.exchange(3)
    ho, hg = o.history(), g.history()
    assert np.allclose(hg[2:], ho[2:], rtol=1e-10, atol=1e-14)
    g.close()
    o.close()
