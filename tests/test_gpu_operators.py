"""GPU parity of the individual PicChunk operators against the oracle.

The sequence is the one PicApplication::push_openmp runs for a step
(pic/pic_application.cpp:219-292), issued operator by operator through the C
ABI, with a comparison after each operator:
  * push_bfd / push_efd: interior fields bit-exact (one expression per cell)
    while the inputs are; once the current has entered E (it is summed in
    another order), within 1e-12 relative;
  * push_velocity / push_position: particles bit-exact on bit-exact fields
    (within 1e-12 afterwards), cell indices (XtensorParticle::gindex) bit-exact;
  * deposit_current + Cur exchange: uj interior <= 1e-12 relative (the sum
    over particles runs in another order);
  * Particle exchange (migration + sort): particle order (ids) and pindex
    bit-exact.
The second group places particles by hand (pn_set_particles): unsorted input
takes the generic one-thread-per-particle kernels, empty segments, particles on
cell and chunk edges, and particles outside their chunk (dropped by the sort,
as the reference's bin Ng is).
"""
import sys

import numpy as np
import pytest

import helpers
import oracle_lib

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(cfg):
    import picnix_amd
    return picnix_amd.Simulation(helpers.dumps(cfg))


def _fields_rel(o, g, which, dim, ncomp):
    worst = np.zeros(ncomp)
    scale = np.zeros(ncomp)
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        a = helpers.interior(o.field(c, which), info, dim)
        b = helpers.interior(g.field(c, which), info, dim)
        for k in range(ncomp):
            worst[k] = max(worst[k], np.max(np.abs(a[..., k] - b[..., k])))
            scale[k] = max(scale[k], np.max(np.abs(a[..., k])))
    return np.where(scale > 0, worst / np.where(scale > 0, scale, 1), worst)


def _particles_equal(o, g, what, exact=True):
    for c in range(o.nchunks):
        for s in range(o.Ns):
            a = o.particles(c, s)
            b = g.particles(c, s)
            assert a.shape == b.shape, f"{what}: chunk {c} species {s}: np differs"
            assert np.array_equal(a[:, 6].view(np.int64), b[:, 6].view(np.int64)), \
                f"{what}: chunk {c} species {s}: particle order differs"
            if exact:
                assert np.array_equal(a[:, :6], b[:, :6]), \
                    f"{what}: chunk {c} species {s}: values differ"
            else:
                assert helpers.rel_err(b[:, :6], a[:, :6]) <= TOL, \
                    f"{what}: chunk {c} species {s}: values differ beyond {TOL}"


def _gindex_equal(o, g, what):
    for c in range(o.nchunks):
        for s in range(o.Ns):
            assert np.array_equal(o.gindex(c, s), g.gindex(c, s)), \
                f"{what}: chunk {c} species {s}: gindex differs"


def _pindex_equal(o, g, what):
    for c in range(o.nchunks):
        for s in range(o.Ns):
            assert np.array_equal(o.pindex(c, s), g.pindex(c, s)), \
                f"{what}: chunk {c} species {s}: pindex differs"


CASES = {
    "thermal3d_o2": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=2)),
    "thermal3d_o1": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=8, order=1)),
    "thermal3d_o3": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=4, order=3)),
    "thermal3d_o4": (3, lambda: helpers.small_thermal(N=(16, 16, 16), C=(2, 2, 2), np_=4, order=4)),
    "twostream2d_o1": (2, lambda: helpers.small_twostream(N=(1, 16, 64), C=(1, 1, 4), np_=16, order=1)),
}

EMF, CUR, MOM, PARTICLE = 0, 1, 2, 3


def _run_sequence(o, g, dim, dt, tag, exact):
    """One step, operator by operator; `exact`: the fields still match bit for bit."""
    o.push_bfd(0.5 * dt)
    g.push_bfd(0.5 * dt)
    rel = _fields_rel(o, g, 0, dim, 6)
    _say(f"[ops] {tag}: push_bfd(dt/2) rel {rel}")
    assert np.all(rel == 0.0) if exact else np.all(rel <= TOL)

    o.push_velocity(dt)
    g.push_velocity(dt)
    _particles_equal(o, g, f"{tag} push_velocity", exact)
    _say(f"[ops] {tag}: push_velocity ok")

    o.push_position(dt)
    g.push_position(dt)
    _particles_equal(o, g, f"{tag} push_position", exact)
    _gindex_equal(o, g, f"{tag} count")
    _say(f"[ops] {tag}: push_position + count ok")

    o.deposit_current(dt)
    g.deposit_current(dt)
    o.push_bfd(0.5 * dt)
    g.push_bfd(0.5 * dt)
    o.exchange(CUR)
    g.exchange(CUR)
    rel = _fields_rel(o, g, 1, dim, 4)
    _say(f"[ops] {tag}: deposit_current + Cur exchange rel {rel}")
    assert np.all(rel <= TOL)

    o.push_efd(dt)
    g.push_efd(dt)
    rel = _fields_rel(o, g, 0, dim, 6)
    _say(f"[ops] {tag}: push_efd rel {rel}")
    assert np.all(rel <= TOL)
    o.exchange(EMF)
    g.exchange(EMF)

    o.exchange(PARTICLE)
    g.exchange(PARTICLE)
    _particles_equal(o, g, f"{tag} particle exchange", exact)
    _pindex_equal(o, g, f"{tag} particle exchange")
    _say(f"[ops] {tag}: particle exchange + sort ok")


@pytest.mark.parametrize("kind", list(CASES))
def test_operator_sequence(kind):
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[ops] {kind}: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    for it in range(2):
        _run_sequence(o, g, dim, dt, f"{kind} step {it + 1}", exact=(it == 0))
    g.close()
    o.close()


def _placed_particles(o, c, s, n, rng, outside=0, edges=0):
    """n particles inside chunk c; `edges` of them exactly on cell/chunk edges,
    `outside` extra ones beyond the chunk (never within reach of its stencil)."""
    info = o.chunk_info(c)
    nb = int(info[9])
    off = np.array(info[3:6], dtype=np.float64)          # z, y, x cells
    ext = np.array(info[6:9], dtype=np.float64) - 2 * nb  # z, y, x cells
    ext = np.maximum(ext, 1)
    lo = off
    hi = off + ext
    xu = np.zeros((n + outside, 7))
    # columns: x, y, z, ux, uy, uz, id (x <-> index 2 of the z, y, x tables)
    for a, col in ((2, 0), (1, 1), (0, 2)):
        xu[:n, col] = rng.uniform(lo[a], hi[a], n)
    if edges:
        k = min(edges, n)
        for a, col in ((2, 0), (1, 1), (0, 2)):
            xu[:k, col] = np.floor(xu[:k, col])         # on a cell edge
        xu[0, 0] = lo[2]                                # on the chunk's low edge
    if outside:
        for a, col in ((2, 0), (1, 1), (0, 2)):
            xu[n:, col] = rng.uniform(lo[a], hi[a], outside)
        xu[n:, 0] = hi[2] + 1.5                         # beyond the chunk in x
    xu[:, 3:6] = rng.normal(0.0, 0.05, (n + outside, 3))
    ids = np.arange(n + outside, dtype=np.int64) + 1000 * (c * o.Ns + s)
    xu[:, 6] = ids.view(np.float64)
    return xu


def _scaled(o, xu, cfg):
    d = float(cfg["parameter"]["delh"])
    out = xu.copy()
    out[:, 0:3] *= d
    return out


@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o1", "twostream2d_o1"])
def test_placed_particles(kind):
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[ops] {kind} placed: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    rng = np.random.default_rng(7)
    for c in range(o.nchunks):
        for s in range(o.Ns):
            if (c + s) % 5 == 3:
                xu = np.zeros((0, 7))                       # empty segment
            else:
                n = int(rng.integers(1, 200))
                xu = _scaled(o, _placed_particles(o, c, s, n, rng, edges=8), cfg)
            o.set_particles(c, s, xu)
            g.set_particles(c, s, xu)
    _say(f"[ops] {kind} placed: unsorted operators")
    # unsorted input: generic kernels for gather and deposit
    o.push_velocity(dt)
    g.push_velocity(dt)
    _particles_equal(o, g, f"{kind} placed push_velocity")
    o.push_position(dt)
    g.push_position(dt)
    _particles_equal(o, g, f"{kind} placed push_position")
    _gindex_equal(o, g, f"{kind} placed count")
    o.deposit_current(dt)
    g.deposit_current(dt)
    o.exchange(CUR)
    g.exchange(CUR)
    rel = _fields_rel(o, g, 1, dim, 4)
    _say(f"[ops] {kind} placed: deposit rel {rel}")
    assert np.all(rel <= TOL)
    o.exchange(PARTICLE)
    g.exchange(PARTICLE)
    _particles_equal(o, g, f"{kind} placed exchange")
    _pindex_equal(o, g, f"{kind} placed exchange")
    # sorted now: one more sequence on the LDS kernels
    _run_sequence(o, g, dim, dt, f"{kind} placed sorted", exact=False)
    g.close()
    o.close()


@pytest.mark.parametrize("kind", ["thermal3d_o2", "twostream2d_o1"])
def test_sort_drops_outside(kind):
    dim, make = CASES[kind]
    cfg = make()
    _say(f"[ops] {kind} outside: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    rng = np.random.default_rng(11)
    for c in range(o.nchunks):
        for s in range(o.Ns):
            xu = _scaled(o, _placed_particles(o, c, s, 50, rng, outside=5), cfg)
            o.set_particles(c, s, xu)
            g.set_particles(c, s, xu)
    o.sort_particle()
    g.sort_particle()
    _particles_equal(o, g, f"{kind} outside sort")
    _pindex_equal(o, g, f"{kind} outside sort")
    for c in range(o.nchunks):
        for s in range(o.Ns):
            assert g.np(c, s) == 50
    g.close()
    o.close()


def _moments_rel(o, g, dim):
    worst, scale = 0.0, 0.0
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        a = helpers.interior(o.field(c, 3), info, dim)
        b = helpers.interior(g.field(c, 3), info, dim)
        worst = max(worst, float(np.max(np.abs(a - b))))
        scale = max(scale, float(np.max(np.abs(a))))
    return worst / scale if scale > 0 else worst


@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o1", "thermal3d_o3", "thermal3d_o4",
                                  "twostream2d_o1"])
def test_moment_deposit(kind):
    """PicChunk::deposit_moment + Mom exchange on sorted particles (LDS row
    kernels: per-bin register sums up to order 3, LDS atomics per particle for
    order 4 in 3D) and on unsorted ones (after push_position: one thread per
    particle)."""
    import picnix_amd
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[ops] {kind} moment: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    for _ in range(2):
        o.step()
        g.step(dt)
    o.deposit_moment()
    g.deposit_moment()
    g.exchange(picnix_amd.BOUNDARY_MOM)
    rel = _moments_rel(o, g, dim)
    _say(f"[ops] {kind} moment (sorted): rel {rel:.3e}")
    assert rel <= TOL
    # unsorted: after a position push the particles are out of bin order
    o.push_velocity(dt)
    g.push_velocity(dt)
    o.push_position(dt)
    g.push_position(dt)
    o.deposit_moment()
    g.deposit_moment()
    g.exchange(picnix_amd.BOUNDARY_MOM)
    rel = _moments_rel(o, g, dim)
    _say(f"[ops] {kind} moment (unsorted): rel {rel:.3e}")
    assert rel <= TOL
    g.close()
    o.close()


def _rho_interior(g, dim):
    return [helpers.interior(g.field(c, 1), g.chunk_info(c), dim)[..., 0].copy()
            for c in range(g.nchunks)]


@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o1", "twostream2d_o1"])
def test_rho_carried_by_continuity(kind):
    """After its first step pn_step carries rho by rho -= dt div J (the
    continuity equation the Esirkepov current satisfies node by node) instead of
    depositing it again.  After 20 steps a direct deposit from the same
    positions (position push with dt = 0, then deposit_current + Cur exchange)
    gives the same rho to round-off."""
    import picnix_amd
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[ops] {kind} rho continuity: create")
    g = _sim(cfg)
    for _ in range(20):
        g.step(dt)
    g.synchronize()
    carried = _rho_interior(g, dim)
    g.push_position(0.0)
    g.deposit_current(dt)
    g.exchange(picnix_amd.BOUNDARY_CUR)
    direct = _rho_interior(g, dim)
    worst = max(float(np.max(np.abs(a - b))) for a, b in zip(carried, direct))
    scale = max(float(np.max(np.abs(b))) for b in direct)
    _say(f"[ops] {kind} rho continuity: max |carried - direct| {worst:.3e}, max |rho| {scale:.3e}")
    assert worst <= TOL * scale
    g.close()


@pytest.mark.parametrize("kind", ["thermal3d_o2", "thermal3d_o1", "twostream2d_o1"])
def test_operator_sequence_per_chunk(kind):
    """The step driven one chunk at a time, as the reference's PicChunk methods
    are called inside push_openmp's omp-for loops (pic/pic_application.cpp:
    219-292), against the oracle's whole-domain operators: field pushes on each
    chunk's interior, the particle operators on each chunk's segments (the
    others untouched), the deposits of all chunks summed once the Cur exchange
    folds them, and a per-chunk sort.  Two steps; the bars of the whole-domain
    sequence."""
    dim, make = CASES[kind]
    cfg = make()
    dt = float(cfg["parameter"]["delt"])
    _say(f"[ops] {kind} per chunk: create")
    o = oracle_lib.Oracle(cfg)
    g = _sim(cfg)
    chunks = list(range(o.nchunks))
    # a chunk order that differs from the chunk ids
    order = chunks[1::2] + chunks[0::2]
    for it in range(2):
        tag = f"{kind} per chunk step {it + 1}"
        exact = it == 0
        o.push_bfd(0.5 * dt)
        for c in order:
            g.push_bfd(0.5 * dt, c)
        g.exchange(EMF)
        rel = _fields_rel(o, g, 0, dim, 6)
        assert np.all(rel == 0.0) if exact else np.all(rel <= TOL), rel
        o.push_velocity(dt)
        for c in order:
            g.push_velocity(dt, c)
        _particles_equal(o, g, f"{tag} push_velocity", exact)
        o.push_position(dt)
        for c in order:
            g.push_position(dt, c)
        _particles_equal(o, g, f"{tag} push_position", exact)
        _gindex_equal(o, g, f"{tag} count")
        o.deposit_current(dt)
        for c in order:
            g.deposit_current(dt, c)
        o.push_bfd(0.5 * dt)
        for c in order:
            g.push_bfd(0.5 * dt, c)
        o.exchange(CUR)
        g.exchange(CUR)
        rel = _fields_rel(o, g, 1, dim, 4)
        _say(f"[ops] {tag}: per-chunk deposits + Cur exchange rel {rel}")
        assert np.all(rel <= TOL), rel
        o.push_efd(dt)
        for c in order:
            g.push_efd(dt, c)
        o.exchange(EMF)
        g.exchange(EMF)
        rel = _fields_rel(o, g, 0, dim, 6)
        assert np.all(rel <= TOL), rel
        o.exchange(PARTICLE)
        g.exchange(PARTICLE)
        _particles_equal(o, g, f"{tag} particle exchange", exact)
        _pindex_equal(o, g, f"{tag} particle exchange")
    # a per-chunk sort: each chunk's particles end in the reference order, the
    # other chunks untouched; particles placed outside the chunk are dropped
    rng = np.random.default_rng(3)
    for c in order:
        for s in range(o.Ns):
            if c in order[:2]:
                xu = _placed_particles(o, c, s, 50, rng, outside=3, edges=2)
            else:
                xu = o.particles(c, s)  # the same records on both sides
            o.set_particles(c, s, xu)
            g.set_particles(c, s, xu)
    o.sort_particle()
    before = {(c, s): g.particles(c, s) for c in order[1:] for s in range(o.Ns)}
    g.sort_particle(order[0])
    for (c, s), a in before.items():
        assert np.array_equal(a, g.particles(c, s)), (c, s)
    for c in order[1:]:
        g.sort_particle(c)
    _particles_equal(o, g, f"{kind} per-chunk sort", True)
    _pindex_equal(o, g, f"{kind} per-chunk sort")
    g.close()
    o.close()


def test_per_chunk_errors():
    """Chunk indices out of range, the FP32 form and the moment deposit reject a
    per-chunk call with a clear error."""
    import picnix_amd
    dim, make = CASES["thermal3d_o2"]
    cfg = make()
    g = _sim(cfg)
    try:
        with pytest.raises(picnix_amd.PicnixError) as e:
            g.push_bfd(0.25, g.nchunks)
        assert e.value.code == -1
        with pytest.raises(picnix_amd.PicnixError) as e:
            g.push_velocity(0.25, -2)
        assert e.value.code == -1
        with pytest.raises(picnix_amd.PicnixError) as e:
            g._chk(g.L.pn_deposit_moment(g.h, 0))
        assert e.value.code == -5
    finally:
        g.close()
    cfg32 = make()
    cfg32["application"]["option"]["precision"] = "fp32"
    g = _sim(cfg32)
    try:
        with pytest.raises(picnix_amd.PicnixError) as e:
            g.push_velocity(0.25, 0)
        assert e.value.code == -5
    finally:
        g.close()
