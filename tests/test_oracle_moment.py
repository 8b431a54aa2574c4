"""Pin the oracle's moment deposit + Mom exchange (CPU only).

PicChunk::deposit_moment (pic/engine/moment.hpp:74-98, local2d/local3d
:271-394) spreads 14 per-particle quantities over the Order+1 cell-centred
nodes around each particle with the shape_mc weights; the Mom exchange adds the
ghosts into the neighbours (nix/xtensor_halo3d.hpp Moment3D).  Two checks that a
dropped term, a swapped component or a wrong node fails:
  * summed over the interior, each component equals the particle sum of its
    defining quantity (partition of unity + exchange), per species;
  * one particle alone: the mass moment on each node is m times the product of
    centred cardinal B-splines of the node-particle distance.
"""
import numpy as np
import pytest

import helpers
import oracle_lib as ol
from test_oracle_primitives import bspline

# moment.hpp:19-33 component order
def _quantities(u, m, c):
    gm = np.sqrt(1 + (u * u).sum(axis=1) / (c * c))
    ux, uy, uz = u[:, 0], u[:, 1], u[:, 2]
    q = np.zeros((u.shape[0], 14))
    q[:, 0] = m
    q[:, 1] = m * ux / gm
    q[:, 2] = m * uy / gm
    q[:, 3] = m * uz / gm
    q[:, 4] = m * gm * c
    q[:, 5] = m * ux * ux / gm
    q[:, 6] = m * uy * uy / gm
    q[:, 7] = m * uz * uz / gm
    q[:, 8] = m * ux
    q[:, 9] = m * uy
    q[:, 10] = m * uz
    q[:, 11] = m * ux * uy / gm
    q[:, 12] = m * uy * uz / gm
    q[:, 13] = m * uz * ux / gm
    return q


def _species_mass(cfg, s):
    sp = cfg["parameter"]["particle"][s]
    return float(sp["ro"]) / float(sp["np"])


def _chunk_box(o, c, dh):
    info = o.chunk_info(c)
    nb = int(info[9])
    off = np.array(info[3:6], dtype=np.float64)
    ext = np.maximum(np.array(info[6:9], dtype=np.float64) - 2 * nb, 1)
    return off * dh, (off + ext) * dh


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_moment_sums(order, dim):
    N = {3: (8, 8, 8), 2: (1, 8, 16), 1: (1, 1, 16)}[dim]
    C = {3: (2, 2, 2), 2: (1, 2, 2), 1: (1, 1, 2)}[dim]
    cfg = helpers.small_thermal(N=N, C=C, np_=1, order=order)
    dh = float(cfg["parameter"]["delh"])
    cc = float(cfg["parameter"]["cc"])
    o = ol.Oracle(cfg, load=False)
    rng = np.random.default_rng(100 * order + dim)
    want = np.zeros((o.Ns, 14))
    for c in range(o.nchunks):
        lo, hi = _chunk_box(o, c, dh)
        for s in range(o.Ns):
            n = 300
            xu = np.zeros((n, 7))
            for a, col in ((2, 0), (1, 1), (0, 2)):
                xu[:, col] = rng.uniform(lo[a], hi[a], n)
            xu[:, 3:6] = rng.normal(0.0, 0.3, (n, 3))
            xu[:, 6] = np.arange(n, dtype=np.int64).view(np.float64)
            o.set_particles(c, s, xu)
            want[s] += _quantities(xu[:, 3:6], _species_mass(cfg, s), cc).sum(axis=0)
    o.deposit_moment()
    got = np.zeros((o.Ns, 14))
    for c in range(o.nchunks):
        um = helpers.interior(o.field(c, 3), o.chunk_info(c), dim)
        got += um.reshape(-1, o.Ns, 14).sum(axis=0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_moment_single_particle_weights(order):
    cfg = helpers.small_thermal(N=(8, 8, 8), C=(1, 1, 1), np_=1, order=order)
    dh = float(cfg["parameter"]["delh"])
    o = ol.Oracle(cfg, load=False)
    pos = np.array([3.3, 4.71, 2.05]) * dh  # x, y, z
    xu = np.zeros((1, 7))
    xu[0, 0:3] = pos
    xu[0, 3:6] = [0.1, -0.2, 0.05]
    o.set_particles(0, 0, xu)
    o.set_particles(0, 1, np.zeros((0, 7)))
    o.deposit_moment()
    info = o.chunk_info(0)
    nb = int(info[9])
    um = o.field(0, 3)[..., 0]  # species 0, mass moment, ghost-inclusive
    m = _species_mass(cfg, 0)
    # cell centre of local index i along an axis: (i - nb + 0.5) dh; periodic images
    # of the 8-cell box land in the interior after the Mom exchange
    inner = helpers.interior(um[..., None], info, 3)[..., 0]
    ref = np.zeros_like(inner)
    for iz in range(8):
        for iy in range(8):
            for ix in range(8):
                w = 1.0
                for a, i in ((0, ix), (1, iy), (2, iz)):
                    d = (pos[a] / dh - (i + 0.5))
                    d = (d + 4.0) % 8.0 - 4.0  # nearest periodic image
                    w *= bspline(order, d)
                ref[iz, iy, ix] = m * w
    np.testing.assert_allclose(inner, ref, rtol=0, atol=1e-14)
    assert abs(inner.sum() - m) < 1e-14
