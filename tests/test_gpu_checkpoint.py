"""Checkpoint / restart (StateHandler::save / load, nix/statehandler.hpp:36-147;
PicChunk::pack / unpack, pic/pic_chunk.cpp:59-104), `-m gpu`.

A run saved after k steps and restarted in a fresh handle from the same
configuration continues as the uninterrupted run does: after m more steps the
particle order (ids) is the same and particle values, E, B and J agree to 1e-12
relative (FP64; the current sums run with atomics, so later bits may differ),
or to 1e-5 in the FP32 form (whose order inside a cell follows the atomics of
its count, so particles are matched by id there).  A handle from another configuration refuses the
file (PN_ERR_CONFIG), as Application::from_json refuses an inconsistent state.
"""
import sys

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(cfg, load=True):
    import picnix_amd
    return picnix_amd.Simulation(helpers.dumps(cfg), load=load)


def _compare(a, b, dim, by_id=False):
    worst_p, worst_f = 0.0, 0.0
    for c in range(a.nchunks):
        for s in range(a.Ns):
            pa, pb = a.particles(c, s), b.particles(c, s)
            if by_id:
                pa = pa[np.argsort(pa[:, 6].view(np.int64))]
                pb = pb[np.argsort(pb[:, 6].view(np.int64))]
            assert np.array_equal(pa[:, 6].view(np.int64), pb[:, 6].view(np.int64))
            worst_p = max(worst_p, helpers.rel_err(pb[:, :6], pa[:, :6]))
        info = a.chunk_info(c)
        for which in (0, 1):
            fa = helpers.interior(a.field(c, which), info, dim)
            fb = helpers.interior(b.field(c, which), info, dim)
            for k in range(fa.shape[-1]):
                worst_f = max(worst_f, helpers.rel_err(fb[..., k], fa[..., k]))
    return worst_p, worst_f


CASES = {
    "thermal3d_o2": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2), 1e-12),
    "friedman2d_o1": (2, lambda: helpers.small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1,
                                                        friedman=0.2), 1e-12),
    "thermal3d_o2_fp32": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2,
                                                          precision="fp32"), 1e-5),
}


@pytest.mark.parametrize("kind", list(CASES))
def test_restart_continues_the_run(kind, tmp_path):
    dim, make, tol = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    a = _sim(cfg)
    for _ in range(4):
        a.step(dt)
    a.synchronize()
    prefix = tmp_path / "state"
    a.save_state(prefix, 4, 4 * dt)
    b = _sim(cfg, load=False)
    step, time = b.load_state(prefix)
    assert step == 4 and time == pytest.approx(4 * dt)
    wp, wf = _compare(a, b, dim)
    _say(f"[ckpt] {kind}: after load particles {wp:.2e} fields {wf:.2e}")
    assert wp == 0.0 and wf == 0.0
    for _ in range(5):
        a.step(dt)
        b.step(dt)
    a.synchronize()
    b.synchronize()
    wp, wf = _compare(a, b, dim, by_id="fp32" in kind)
    ha, hb = a.history(), b.history()
    _say(f"[ckpt] {kind}: after 5 more steps particles {wp:.2e} fields {wf:.2e}; history {ha} {hb}")
    assert wp <= tol and wf <= tol
    assert np.allclose(hb[2:], ha[2:], rtol=tol, atol=1e-14)
    a.close()
    b.close()


def test_restart_shock_draws_made_ahead(tmp_path):
    """The shock's injection draws for the next step are made ahead of it on a host
    thread, and a step uses them only when the chunks' random_seed and lastid are
    the ones they were drawn for.  A run restarted from a checkpoint in a fresh
    handle (no draws ahead) and one loaded back into a handle that had run on (its
    draws ahead are for a later step) both continue exactly as the uninterrupted
    run, whose draws were made ahead: the same particles in the same order, the
    same fields and boundary states after 5 more steps."""
    cfg = helpers.small_shock(N=(8, 8, 32), C=(2, 2, 2), nppc=2, order=2, emf_order=1)
    dt = float(cfg["parameter"]["delt"])
    a = _sim(cfg)
    for _ in range(4):
        a.step(dt)
    a.synchronize()
    prefix = tmp_path / "state"
    a.save_state(prefix, 4, 4 * dt)
    b = _sim(cfg, load=False)
    b.load_state(prefix)
    c = _sim(cfg)
    for _ in range(6):
        c.step(dt)
    c.synchronize()
    c.load_state(prefix)
    for _ in range(5):
        for sim in (a, b, c):
            sim.step(dt)
    for sim in (a, b, c):
        sim.synchronize()
    for name, other in (("fresh handle", b), ("handle loaded back", c)):
        wp, wf = _compare(a, other, 3)
        _say(f"[ckpt] shock3d_o2 restarted in a {name}: particles {wp:.2e} fields {wf:.2e}")
        assert wp <= 1e-12 and wf <= 1e-12
        for k in range(a.nchunks):
            assert other.boundary_state(k) == a.boundary_state(k), (name, k)
    for sim in (a, b, c):
        sim.close()


def test_restart_refuses_other_configuration(tmp_path):
    import picnix_amd
    cfg = helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2)
    a = _sim(cfg)
    a.save_state(tmp_path / "state", 0, 0.0)
    other = helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2)
    other["parameter"]["delt"] = 0.25
    b = _sim(other, load=False)
    with pytest.raises(picnix_amd.PicnixError) as err:
        b.load_state(tmp_path / "state")
    assert err.value.code == -2  # PN_ERR_CONFIG
    a.close()
    b.close()


def test_driver_save_and_load(tmp_path):
    """picnix_run -s saves at the end of a run and -l continues from it
    (Application::finalize / setup_chunks): a run split at step 4 writes the
    same history.txt as one uninterrupted run."""
    import os
    import subprocess
    exe = os.path.join(helpers.ROOT, "pic-nix_amd", "picnix_run")
    cfg = helpers.small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1)
    dt = float(cfg["parameter"]["delt"])

    def run(base, *args):
        c = dict(cfg)
        c["application"] = dict(cfg["application"], basedir=str(tmp_path / base))
        conf = tmp_path / f"{base}.toml"
        conf.write_text(helpers.dumps(c) + "[[diagnostic]]\n  name = 'history'\n  interval = 2\n")
        r = subprocess.run([exe, "-c", str(conf), *args], capture_output=True, text=True, timeout=120)
        _say(f"[ckpt] driver {base} {args}: rc {r.returncode} {r.stderr[-200:]}")
        assert r.returncode == 0
        return (tmp_path / base / "history.txt").read_text().splitlines()

    whole = run("whole", "-t", str(7 * dt))
    run("split", "-t", str(3 * dt), "-s", "ckpt")
    split = run("split", "-t", str(7 * dt), "-l", "ckpt")
    assert len(whole) == len(split) == 5
    a = np.array([[float(v) for v in ln.split()] for ln in whole[1:]])
    b = np.array([[float(v) for v in ln.split()] for ln in split[1:]])
    assert np.allclose(a, b, rtol=1e-6, atol=1e-12)
