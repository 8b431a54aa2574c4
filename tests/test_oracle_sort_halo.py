"""Pin the oracle's counting sort, halo exchanges and particle migration (CPU only).

  * sort == numpy stable sort on the key (cell, ip mod 8, ip) with out-of-chunk
    particles dropped (XtensorParticle::sort, nix/xtensor_particle.hpp:260-321;
    membership check of nix/unittest/test_xtensor_particle.cpp:144-205);
  * the chunked domain (halo copies, current fold, migration with periodic wrap)
    evolves the same particles and fields as a single-chunk run of the same
    domain to round-off.
"""
import numpy as np
import pytest

import helpers
import oracle_lib as ol


def brute_sort(xu, lim, dh, order, dims, has):
    """Reference-order sort of one chunk's records with numpy."""
    is_odd = order % 2
    n = xu.shape[0]
    ix = np.zeros(n, dtype=np.int64)
    iy = np.zeros(n, dtype=np.int64)
    iz = np.zeros(n, dtype=np.int64)
    off = [lim[a][0] - 0.5 * dh * is_odd for a in range(3)]
    if has[2]:
        ix = np.floor((xu[:, 0] - off[2]) * (1 / dh)).astype(np.int64)
    if has[1]:
        iy = np.floor((xu[:, 1] - off[1]) * (1 / dh)).astype(np.int64)
    if has[0]:
        iz = np.floor((xu[:, 2] - off[0]) * (1 / dh)).astype(np.int64)
    sy = (dims[2] + 1) if has[2] else 2
    sz = sy * ((dims[1] + 1) if has[1] else 2)
    cell = iz * sz + iy * sy + ix
    oob = np.zeros(n, dtype=bool)
    for a, col in ((2, 0), (1, 1), (0, 2)):
        if has[a]:
            oob |= (xu[:, col] < lim[a][0]) | (xu[:, col] >= lim[a][1])
    ip = np.arange(n)
    order_ = np.lexsort((ip, ip % 8, cell))
    order_ = order_[~oob[order_]]
    return xu[order_]


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_sort_matches_brute_force(order, dim):
    N = {3: (8, 8, 8), 2: (1, 8, 8), 1: (1, 1, 16)}[dim]
    C = {3: (2, 2, 2), 2: (1, 2, 2), 1: (1, 1, 2)}[dim]
    cfg = helpers.small_thermal(N=N, C=C, np_=1, order=order)
    o = ol.Oracle(cfg, load=False)
    rng = np.random.default_rng(order * 10 + dim)
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        off = info[3:6]
        dims = [N[a] // C[a] for a in range(3)]
        lim = [(off[a] * 1.0, off[a] * 1.0 + dims[a] * 1.0) for a in range(3)]
        n = 2000
        xu = np.zeros((n, 7))
        # mostly inside, some outside (dropped), a few exactly on cell faces
        for col, a in ((0, 2), (1, 1), (2, 0)):
            xu[:, col] = rng.uniform(lim[a][0] - 0.6, lim[a][1] + 0.6, size=n)
        xu[:50, 0] = np.floor(xu[:50, 0])
        if dim <= 2:
            xu[:, 2] = rng.uniform(-3, 3, size=n)  # z is ignored in 2D
        if dim == 1:
            xu[:, 1] = rng.uniform(-3, 3, size=n)  # and y in 1D
        xu[:, 3:6] = rng.normal(size=(n, 3))
        xu[:, 6] = np.arange(n).astype(np.int64).view(np.float64)
        o.set_particles(c, 0, xu)
        o.sort_particle()
        got = o.particles(c, 0)
        has = [dim == 3, dim >= 2, True]
        ref = brute_sort(xu, lim, 1.0, order, dims, has)
        assert np.array_equal(got.view(np.int64), ref.view(np.int64))
        # pindex: every bin holds exactly the particles that digitize to it
        pidx = o.pindex(c, 0)
        assert pidx[-1] == got.shape[0]
        o.set_particles(c, 0, np.zeros((0, 7)))


def _all_particles(o):
    recs = [o.particles(c, s) for c in range(o.nchunks) for s in range(o.Ns)]
    sp = [s for c in range(o.nchunks) for s in range(o.Ns)]
    return recs, sp


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_chunked_equals_single_chunk(dim):
    if dim == 3:
        N, C = (8, 8, 8), (2, 2, 2)
    elif dim == 2:
        N, C = (1, 8, 16), (1, 2, 4)
    else:
        N, C = (1, 1, 32), (1, 1, 4)
    cfg = helpers.small_thermal(N=N, C=C, np_=4, order=2)
    for sp in cfg["parameter"]["particle"]:
        sp["vt"] = 0.3  # many chunk crossings
    a = ol.Oracle(cfg)
    cfg1 = helpers.small_thermal(N=N, C=(1, 1, 1), np_=4, order=2)
    for sp in cfg1["parameter"]["particle"]:
        sp["vt"] = 0.3
    cfg1["application"]["option"]["buffer_ratio"] = 0.5
    b = ol.Oracle(cfg1, load=False)
    for s in range(a.Ns):
        xs = np.concatenate([a.particles(c, s) for c in range(a.nchunks)])
        b.set_particles(0, s, xs)
    b.sort_particle()
    a.step(6)
    b.step(6)
    # particles by id
    for s in range(a.Ns):
        pa = np.concatenate([a.particles(c, s) for c in range(a.nchunks)])
        pb = b.particles(0, s)
        ia = np.argsort(pa[:, 6].view(np.int64))
        ib = np.argsort(pb[:, 6].view(np.int64))
        assert pa.shape == pb.shape
        assert helpers.rel_err(pa[ia, :6], pb[ib, :6]) < 1e-12
    # fields: assemble the chunked interiors into the global grid
    ib = b.chunk_info(0)
    gb = helpers.interior(b.field(0, 0), ib, dim)
    jb = helpers.interior(b.field(0, 1), ib, dim)
    for c in range(a.nchunks):
        info = a.chunk_info(c)
        oz, oy, ox = (int(v) for v in info[3:6])
        fa = helpers.interior(a.field(c, 0), info, dim)
        ja = helpers.interior(a.field(c, 1), info, dim)
        sl = (slice(oz, oz + fa.shape[0]) if dim == 3 else slice(0, 1),
              slice(oy, oy + fa.shape[1]), slice(ox, ox + fa.shape[2]))
        assert np.max(np.abs(fa - gb[sl])) <= 1e-12 * max(1e-300, np.max(np.abs(gb)))
        assert np.max(np.abs(ja - jb[sl])) <= 1e-12 * max(1e-300, np.max(np.abs(jb)))
    ha, hb = a.history(), b.history()
    assert np.allclose(ha[2:], hb[2:], rtol=1e-11)


def test_migration_conserves_particles_and_wraps():
    cfg = helpers.small_thermal(N=(8, 8, 8), C=(2, 2, 2), np_=4, order=1)
    for sp in cfg["parameter"]["particle"]:
        sp["vt"] = 0.4
    o = ol.Oracle(cfg)
    n0 = sum(o.np(c, s) for c in range(o.nchunks) for s in range(o.Ns))
    o.step(10)
    n1 = sum(o.np(c, s) for c in range(o.nchunks) for s in range(o.Ns))
    assert n0 == n1
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        off = info[3:6]
        for s in range(o.Ns):
            x = o.particles(c, s)
            for col, a in ((0, 2), (1, 1), (2, 0)):
                assert np.all(x[:, col] >= off[a]) and np.all(x[:, col] < off[a] + 4)
