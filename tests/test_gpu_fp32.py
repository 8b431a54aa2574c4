"""The FP32 form (precision = "fp32") against the FP64 oracle.

FP32 cell-relative particles and FP32 fields are compared with the FP64 oracle
on the same seeded inputs; the bars are FP32 error bounds, stated per quantity:
  * after the loader and the initial sort: every particle present with the same
    id; positions within 1e-6 of a cell, momenta within 1e-6 relative (FP32
    rounding of the FP64 loader's values);
  * after 1 and 10 steps: particle positions within 1e-5 of a cell and momenta
    within 1e-4 relative; E, B and J within 1e-4 relative (max |diff| / max |ref|);
  * field and kinetic energies within 1e-4 relative after 10 steps;
  * div(E) - rho: the FP32 charge-conservation residual, summed over all cells,
    below 1e-4 of the summed |rho| after 10 steps.
Particles are matched by id (for odd orders the FP32 form bins by cell, the
reference by the half-shifted grid, so the orders differ).
"""
import copy
import sys

import numpy as np
import pytest

import helpers
import oracle_lib

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fp32(cfg):
    cfg = copy.deepcopy(cfg)
    cfg["application"]["option"]["precision"] = "fp32"
    return cfg


def _sim(cfg):
    import picnix_amd
    return picnix_amd.Simulation(helpers.dumps(cfg))


def _by_id(xu):
    ids = xu[:, 6].view(np.int64)
    order = np.argsort(ids)
    return ids[order], xu[order]


def _particle_errors(o, g, dh):
    dx_err, du_err = 0.0, 0.0
    for c in range(o.nchunks):
        for s in range(o.Ns):
            ia, a = _by_id(o.particles(c, s))
            ib, b = _by_id(g.particles(c, s))
            assert np.array_equal(ia, ib), f"chunk {c} species {s}: particle sets differ"
            if a.shape[0] == 0:
                continue
            dx_err = max(dx_err, float(np.max(np.abs(a[:, 0:3] - b[:, 0:3]))) / dh)
            scale = max(float(np.max(np.abs(a[:, 3:6]))), 1e-30)
            du_err = max(du_err, float(np.max(np.abs(a[:, 3:6] - b[:, 3:6]))) / scale)
    return dx_err, du_err


def _field_errors(o, g, dim):
    worst, scale = np.zeros(10), np.zeros(10)
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        fa = helpers.interior(o.field(c, 0), info, dim)
        fb = helpers.interior(g.field(c, 0), info, dim)
        ja = helpers.interior(o.field(c, 1), info, dim)
        jb = helpers.interior(g.field(c, 1), info, dim)
        for k in range(6):
            worst[k] = max(worst[k], np.max(np.abs(fa[..., k] - fb[..., k])))
            scale[k] = max(scale[k], np.max(np.abs(fa[..., k])))
        for k in range(4):
            worst[6 + k] = max(worst[6 + k], np.max(np.abs(ja[..., k] - jb[..., k])))
            scale[6 + k] = max(scale[6 + k], np.max(np.abs(ja[..., k])))
    return np.where(scale > 0, worst / np.where(scale > 0, scale, 1), worst)


CASES = {
    "thermal3d_o2": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2)),
    "thermal3d_o1": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=1)),
    "thermal3d_o4": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=4, order=4)),
    "twostream2d_o1": (2, lambda: helpers.small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1)),
    # 1D (Nz = Ny = 1): y and z stay absolute coordinates, x is cell-relative
    "twostream1d_o1": (1, lambda: helpers.small_twostream(N=(1, 1, 128), C=(1, 1, 8), np_=64, order=1)),
    "thermal1d_o2": (1, lambda: helpers.small_thermal(N=(1, 1, 64), C=(1, 1, 4), np_=32, order=2)),
    "thermal1d_o3": (1, lambda: helpers.small_thermal(N=(1, 1, 64), C=(1, 1, 4), np_=32, order=3)),
}


@pytest.mark.parametrize("kind", list(CASES))
def test_fp32_against_fp64_oracle(kind):
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    dh = float(cfg["parameter"]["delh"])
    _say(f"[fp32] {kind}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(_fp32(cfg))
    dx, du = _particle_errors(o, g, dh)
    _say(f"[fp32] {kind}: loader position err {dx:.2e} cells, momentum {du:.2e}")
    assert dx <= 1e-6 and du <= 1e-6
    for nstep in (1, 9):
        for _ in range(nstep):
            o.step()
            g.step(dt)
        g.synchronize()
        dx, du = _particle_errors(o, g, dh)
        rel = _field_errors(o, g, dim)
        _say(f"[fp32] {kind}: after {'1' if nstep == 1 else '10'} steps: position {dx:.2e} cells, "
             f"momentum {du:.2e}, fields {np.array2string(rel, precision=2)}")
        assert dx <= 1e-5 and du <= 1e-4
        assert np.all(rel <= 1e-4)
    ho, hg = o.history(), g.history()
    _say(f"[fp32] {kind}: history oracle {ho} fp32 {hg}")
    assert np.allclose(hg[2:], ho[2:], rtol=1e-4, atol=1e-12)
    # charge-conservation residual, against the total charge density scale
    rho_abs = 0.0
    for c in range(o.nchunks):
        rho_abs += np.abs(helpers.interior(o.field(c, 1), o.chunk_info(c), dim)[..., 0]).sum()
    assert abs(hg[0] - ho[0]) <= 1e-4 * max(rho_abs, 1.0)
    g.close()
    o.close()


def test_fp32_graph_and_operators():
    """FP32 steps replayed from hipGraphs and an operator-by-operator step in
    between agree with eager FP32 steps on the same inputs."""
    dim, make = CASES["thermal3d_o2"]
    cfg = _fp32(make())
    dt = float(cfg["parameter"]["delt"])
    a = _sim(cfg)
    b = _sim(cfg)
    b.step_graph(True)
    for it in range(5):
        a.step(dt)
        if it == 2:
            b.push_bfd(0.5 * dt)
            b.push_velocity(dt)
            b.push_position(dt)
            b.deposit_current(dt)
            b.push_bfd(0.5 * dt)
            b.exchange(1)
            b.push_efd(dt)
            b.exchange(0)
            b.exchange(3)
        else:
            b.step(dt)
    a.synchronize()
    b.synchronize()
    ha, hb = a.history(), b.history()
    _say(f"[fp32] graph: eager {ha} graph {hb}")
    assert np.allclose(ha[2:], hb[2:], rtol=1e-5, atol=1e-12)
    for c in range(a.nchunks):
        for s in range(a.Ns):
            ia, pa = _by_id(a.particles(c, s))
            ib, pb = _by_id(b.particles(c, s))
            assert np.array_equal(ia, ib)
            assert np.max(np.abs(pa[:, :6] - pb[:, :6])) <= 1e-5
    a.close()
    b.close()


def test_fp32_rho_carried_by_continuity():
    """The FP32 step carries rho by rho -= dt div J after its first step; after
    20 steps a direct FP32 deposit from the same positions agrees within the
    FP32 bar (1e-4 of max |rho|)."""
    import picnix_amd
    dim, make = CASES["thermal3d_o2"]
    cfg = _fp32(make())
    dt = float(cfg["parameter"]["delt"])
    g = _sim(cfg)
    for _ in range(20):
        g.step(dt)
    g.synchronize()

    def rho():
        return [helpers.interior(g.field(c, 1), g.chunk_info(c), dim)[..., 0].copy()
                for c in range(g.nchunks)]

    carried = rho()
    g.push_position(0.0)
    g.deposit_current(dt)
    g.exchange(picnix_amd.BOUNDARY_CUR)
    direct = rho()
    worst = max(float(np.max(np.abs(a - b))) for a, b in zip(carried, direct))
    scale = max(float(np.max(np.abs(b))) for b in direct)
    _say(f"[fp32] rho continuity: max |carried - direct| {worst:.3e}, max |rho| {scale:.3e}")
    assert worst <= 1e-4 * scale
    g.close()


@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o1", "twostream2d_o1", "shock3d"])
def test_fp32_fused_push(kind, monkeypatch):
    """The FP32 step's row gather runs each particle's position push too
    (PICNIX_FUSED_PUSH=1, the default).  The separate position kernel does the
    same float operations, so after one step from the same state every particle
    (matched by id: slots within a bin are taken in atomic order) has the same
    position and momentum bit for bit; the fields then differ only by the order
    of the current sums, and after 6 steps both runs agree to the FP32 bars."""
    if kind == "shock3d":
        # wall reflection and upstream marking in the position push
        # (which particles are redrawn at x = Lx follows their order in the arrays,
        # which the atomic slots make differ between runs: compare the particles
        # both runs hold, after a step in which the cold plasma has none to redraw)
        cfg = helpers.small_shock(N=(8, 8, 32), C=(2, 2, 2), nppc=4, order=2,
                                  precision="fp32")
    else:
        cfg = _fp32(CASES[kind][1]())
    dt = float(cfg["parameter"]["delt"])
    sims = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("PICNIX_FUSED_PUSH", fused)
        sims[fused] = _sim(cfg)
    a, b = sims["1"], sims["0"]

    def compare(exact):
        dx, du, fe = 0.0, 0.0, 0.0
        for c in range(a.nchunks):
            for s in range(a.Ns):
                ia, pa = _by_id(a.particles(c, s))
                ib, pb = _by_id(b.particles(c, s))
                if kind == "shock3d":
                    common, ka, kb = np.intersect1d(ia, ib, return_indices=True)
                    assert common.size >= 0.99 * max(ia.size, ib.size)
                    pa, pb = pa[ka], pb[kb]
                else:
                    assert np.array_equal(ia, ib), f"{kind}: chunk {c} species {s} particle sets"
                if exact:
                    assert np.array_equal(pa[:, :6], pb[:, :6]), f"{kind}: chunk {c} species {s}"
                elif pa.shape[0]:
                    dx = max(dx, float(np.max(np.abs(pa[:, :3] - pb[:, :3]))))
                    du = max(du, helpers.rel_err(pa[:, 3:6], pb[:, 3:6]))
            for k in range(2):  # uf (E, B) and uj (rho, J)
                fe = max(fe, helpers.rel_err(a.field(c, k), b.field(c, k)))
        return dx, du, fe

    a.step(dt)
    b.step(dt)
    _, _, fe1 = compare(exact=True)
    for _ in range(5):
        a.step(dt)
        b.step(dt)
    dx, du, fe = compare(exact=False)
    _say(f"[fp32] fused push {kind}: step 1 identical particles, fields {fe1:.1e}; "
         f"step 6 dx {dx:.1e} du {du:.1e} fields {fe:.1e}")
    assert fe1 <= 1e-5
    assert dx <= 1e-5 * float(cfg["parameter"]["delh"]) and du <= 1e-4 and fe <= 1e-4
    a.close()
    b.close()


def test_fp32_keeps_particles_at_the_upper_chunk_face():
    """A particle just below a chunk's upper face (x = xmax - 1e-12 in FP64) has
    an offset u - floor(u) that rounds to 1.0f in FP32; it stays in its FP64
    cell (the reference's bin) with the largest offset below 1, and is neither
    dropped nor moved to the next chunk by the conversion, the sort or a step."""
    import picnix_amd
    cfg = _fp32(helpers.small_thermal(N=(8, 8, 16), C=(1, 1, 2), np_=1, order=2))
    sim = _sim(cfg)
    try:
        for c in range(sim.nchunks):
            for s in range(sim.Ns):
                sim.set_particles(c, s, np.zeros((0, 7)))
        info = sim.chunk_info(0)
        x1 = float(info[5]) + 8.0  # upper x face of chunk 0 (8 cells, dx = 1)
        a = np.zeros((4, 7))
        a[:, 0] = [x1 - 1e-12, x1 - 1e-9, np.nextafter(x1, 0.0), x1 - 0.5]
        a[:, 1] = [0.5, 7.999999999999, 3.0, 4.0]
        a[:, 2] = [np.nextafter(8.0, 0.0), 0.25, 5.0, 1.0]
        a[:, 6] = np.arange(4, dtype=np.int64).view(np.float64)
        sim.set_particles(0, 0, a)
        sim.sort_particle()
        assert sim.np(0, 0) == 4, sim.np(0, 0)
        b = sim.particles(0, 0)
        assert np.all(b[:, 0] < x1) and np.all(b[:, 0] >= x1 - 1.0)
        assert np.all(b[:, 1] < 8.0) and np.all(b[:, 2] < 8.0)
        assert np.max(np.abs(np.sort(b[:, 0]) - np.sort(a[:, 0]))) <= 1e-6
        sim.step(float(cfg["parameter"]["delt"]))
        sim.synchronize()
        n = sum(sim.np(c, s) for c in range(sim.nchunks) for s in range(sim.Ns))
        assert n == 4, n
    finally:
        sim.close()


@pytest.mark.parametrize("tile", ["stage1", "stage3", "stage4", "stage16", "comp", "plain"])
def test_fp32_deposit_tile_forms(tile, monkeypatch):
    """The FP32 march deposit's tile forms against the default (an FP64 tile):
    staged compensated pairs with 1, 3 (a row of 8 cells in phases of 3, 3 and
    2), 4 and 16 cells per phase, the pairs updated by LDS atomics, and a plain
    FP32 tile.  From the same state, one step gives
    particles identical by id (the current does not act on them within the step)
    and currents equal to FP32 summation round-off: within 1e-5 relative (the
    plain tile, with no compensation, within 1e-4), and every form within 1e-4 of
    the FP64 oracle."""
    base = CASES["thermal3d_o2"][1]()
    base["parameter"].update(Nx=32, Ny=8, Nz=8, Cx=4, Cy=1, Cz=1)  # chunks of 8 x 8 x 8 cells
    cfg = _fp32(base)
    dt = float(cfg["parameter"]["delt"])
    if tile.startswith("stage"):
        monkeypatch.setenv("PICNIX_FP32_TILE", "staged")
        monkeypatch.setenv("PICNIX_STAGE_CELLS", tile[5:])
    else:
        monkeypatch.setenv("PICNIX_FP32_TILE", tile)
    g = _sim(cfg)
    monkeypatch.delenv("PICNIX_STAGE_CELLS", raising=False)
    monkeypatch.delenv("PICNIX_FP32_TILE", raising=False)
    d = _sim(cfg)
    o = oracle_lib.Oracle(base)
    for sim in (g, d):
        sim.step(dt)
        sim.synchronize()
    o.step()
    worst = 0.0
    for c in range(g.nchunks):
        for s in range(g.Ns):
            ia, a = _by_id(g.particles(c, s))
            ib, b = _by_id(d.particles(c, s))
            assert np.array_equal(ia, ib) and np.array_equal(a[:, :6], b[:, :6]), (tile, c, s)
        info = o.chunk_info(c)
        ja = helpers.interior(g.field(c, 1), info, 3)
        jb = helpers.interior(d.field(c, 1), info, 3)
        for k in range(1, 4):
            sc = max(float(np.max(np.abs(jb[..., k]))), 1e-30)
            worst = max(worst, float(np.max(np.abs(ja[..., k] - jb[..., k]))) / sc)
    rel = _field_errors(o, g, 3)
    _say(f"[fp32] tile {tile}: J vs default {worst:.2e}, fields vs oracle {rel.max():.2e}")
    assert worst <= (1e-4 if tile == "plain" else 1e-5)
    assert rel.max() <= 1e-4
    g.close()
    d.close()
    o.close()
