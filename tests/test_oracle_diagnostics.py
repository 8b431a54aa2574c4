"""Pin the oracle's history diagnostics on their own (CPU only).

PicChunk::get_energy (pic/pic_chunk.cpp:407-439) sums E^2/2 and B^2/2 over the
interior cells of every chunk; BaseMaxwell::get_diverror_{1,2,3}d
(pic/engine/maxwell.hpp:63-83,203-229,350-381) sums, over the interior,
div(E) - rho with the forward difference of E and div(B) with the backward
difference of B.  Each check below computes the expected value another way:

  * energies: closed forms (a sinusoid over whole periods sums to N/2, a
    constant field to N), with large values in every ghost cell, so including a
    ghost layer or dropping a component fails;
  * div errors: along each line the forward-difference sum telescopes to
    (E at the first ghost above the interior) - (E at the first interior cell),
    and the backward one to (B at the last interior cell) - (B at the last ghost
    below).  The ghost layers hold values unrelated to the interior, so a stencil
    shifted by one cell, a backward/forward swap, a wrong component or a sign
    error on rho changes the sum by far more than round-off;
  * a deliberately shifted stencil (computed here) does not match, so the checks
    can tell the difference.
"""
import math

import numpy as np
import pytest

import helpers
import oracle_lib as ol

GEOMS = {
    "3d": ((8, 8, 16), (2, 2, 2)),
    "2d": ((1, 8, 16), (1, 2, 2)),
    "1d": ((1, 1, 32), (1, 1, 4)),
}


def _oracle(kind, order=2):
    N, C = GEOMS[kind]
    cfg = helpers.small_thermal(N=N, C=C, np_=1, order=order)
    return ol.Oracle(cfg, load=False), cfg


def _interior_slices(info, dim):
    nb = int(info[9])
    nz, ny, nx = (int(v) for v in info[6:9])
    zs = slice(nb, nz - nb) if dim == 3 else slice(nb, nb + 1)
    ys = slice(nb, ny - nb) if dim >= 2 else slice(nb, nb + 1)
    return zs, ys, slice(nb, nx - nb)


@pytest.mark.parametrize("kind", list(GEOMS))
def test_field_energy_closed_form(kind):
    o, cfg = _oracle(kind)
    dim = {"3d": 3, "2d": 2, "1d": 1}[kind]
    par = cfg["parameter"]
    Lx = float(par["Nx"])
    ncell = int(par["Nx"]) * int(par["Ny"]) * int(par["Nz"])
    A, B0, m = 0.75, 0.375, 3
    rng = np.random.default_rng(1)
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        shape = tuple(int(v) for v in info[6:9])
        uf = 1e3 * (1.0 + rng.random(shape + (6,)))  # ghosts: large, unrelated values
        zs, ys, xs = _interior_slices(info, dim)
        nb = int(info[9])
        xi = np.arange(shape[2]) - nb + int(info[5])
        xc = (xi + 0.5)[xs]
        uf[zs, ys, xs, :] = 0.0
        uf[zs, ys, xs, 1] = A * np.sin(2 * math.pi * m * xc / Lx)  # Ey
        uf[zs, ys, xs, 5] = B0                                      # Bz
        o.set_field(c, 0, uf)
        uj = np.zeros(shape + (4,))
        o.set_field(c, 1, uj)
    h = o.history()
    # sum of sin^2 over whole periods = N/2; constant: N
    assert h[2] == pytest.approx(0.5 * A * A * ncell / 2, rel=1e-13)
    assert h[3] == pytest.approx(0.5 * B0 * B0 * ncell, rel=1e-13)
    o.close()


def _telescoped(o, ufs, ujs, dim, shift=0):
    """Expected (div(E) - rho, div(B)) from the telescoped line sums.  shift=1
    evaluates the same sums one cell off (a deliberate stencil mistake)."""
    de = db = 0.0
    for c, (uf, uj) in enumerate(zip(ufs, ujs)):
        info = o.chunk_info(c)
        nb = int(info[9])
        nz, ny, nx = (int(v) for v in info[6:9])
        zs, ys, xs = _interior_slices(info, dim)
        lbx, ubx = nb, nx - nb - 1
        lby, uby = nb, ny - nb - 1
        lbz, ubz = nb, nz - nb - 1
        s = shift
        # x lines (Ex forward, Bx backward)
        de += np.sum(uf[zs, ys, ubx + 1 - s, 0] - uf[zs, ys, lbx - s, 0])
        db += np.sum(uf[zs, ys, ubx - s, 3] - uf[zs, ys, lbx - 1 - s, 3])
        if dim >= 2:
            de += np.sum(uf[zs, uby + 1, xs, 1] - uf[zs, lby, xs, 1])
            db += np.sum(uf[zs, uby, xs, 4] - uf[zs, lby - 1, xs, 4])
        if dim == 3:
            de += np.sum(uf[ubz + 1, ys, xs, 2] - uf[lbz, ys, xs, 2])
            db += np.sum(uf[ubz, ys, xs, 5] - uf[lbz - 1, ys, xs, 5])
        de -= np.sum(uj[zs, ys, xs, 0])
    return de, db


@pytest.mark.parametrize("kind", list(GEOMS))
def test_diverror_telescoped(kind):
    o, cfg = _oracle(kind)
    dim = {"3d": 3, "2d": 2, "1d": 1}[kind]
    rng = np.random.default_rng(2)
    ufs, ujs = [], []
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        shape = tuple(int(v) for v in info[6:9])
        uf = rng.standard_normal(shape + (6,))
        uj = rng.standard_normal(shape + (4,))
        o.set_field(c, 0, uf)
        o.set_field(c, 1, uj)
        ufs.append(uf)
        ujs.append(uj)
    h = o.history()
    de, db = _telescoped(o, ufs, ujs, dim)
    scale = sum(float(np.sum(np.abs(a))) for a in ufs) + sum(float(np.sum(np.abs(b))) for b in ujs)
    assert abs(h[0] - de) <= 1e-13 * scale
    assert abs(h[1] - db) <= 1e-13 * scale
    # a stencil one cell off gives a different sum: the check discriminates
    de1, db1 = _telescoped(o, ufs, ujs, dim, shift=1)
    assert abs(h[0] - de1) > 1e-3 and abs(h[1] - db1) > 1e-3
    o.close()


def test_diverror_rho_sign_and_interior():
    """E = 0 everywhere: div(E) - rho = -(sum of rho over the interior cells);
    rho in the ghosts must not count."""
    o, cfg = _oracle("3d")
    tot = 0.0
    for c in range(o.nchunks):
        info = o.chunk_info(c)
        shape = tuple(int(v) for v in info[6:9])
        o.set_field(c, 0, np.zeros(shape + (6,)))
        uj = np.full(shape + (4,), 100.0)
        zs, ys, xs = _interior_slices(info, 3)
        uj[zs, ys, xs, 0] = 0.25
        tot += 0.25 * uj[zs, ys, xs, 0].size
        o.set_field(c, 1, uj)
    h = o.history()
    assert h[0] == pytest.approx(-tot, rel=1e-15)
    assert h[1] == 0.0
    o.close()
