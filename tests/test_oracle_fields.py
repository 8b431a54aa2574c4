"""Pin the oracle's Yee solver, EMF halo and field gather (CPU only).

  * vacuum plane wave: an exact eigenmode of the discrete Yee scheme advances
    with the numerical dispersion sin(w dt/2)/(c dt) = sin(k dx/2)/dx; a sign
    or half-step error in push_bfd / push_efd or a wrong ghost copy breaks it;
  * div B of arbitrary fields is preserved cell by cell;
  * the MC gather reproduces linear fields exactly at every staggered location
    (a half-cell error in any component's grid shows up as an offset).
"""
import math

import numpy as np
import pytest

import helpers
import oracle_lib as ol
from test_oracle_primitives import boris_textbook


def vacuum_cfg(N, C, order=2):
    cfg = helpers.small_thermal(N=N, C=C, np_=1, order=order)
    return cfg


@pytest.mark.parametrize("N,C,axis", [((4, 4, 32), (1, 1, 2), "x"), ((32, 4, 4), (2, 1, 1), "z"),
                                      ((1, 32, 4), (1, 4, 1), "y"), ((1, 1, 32), (1, 1, 2), "x")])
def test_plane_wave_dispersion(N, C, axis):
    cfg = vacuum_cfg(N, C)
    o = ol.Oracle(cfg, load=False)
    dt, dh, c = 0.5, 1.0, 1.0
    ax = {"x": 2, "y": 1, "z": 0}[axis]
    L = N[ax] * dh
    k = 2 * math.pi * 3 / L
    w = 2 / dt * math.asin(c * dt / dh * math.sin(k * dh / 2))
    # field pair transverse to the propagation axis: E_a, B_b with b = a x n
    if axis == "x":
        ecomp, bcomp, sgn = 1, 5, +1.0   # Ey, Bz
    elif axis == "y":
        ecomp, bcomp, sgn = 2, 3, +1.0   # Ez, Bx
    else:
        ecomp, bcomp, sgn = 0, 4, +1.0   # Ex, By

    # staggering along the propagation axis, for (E, B) component pairs:
    # E_a is cell-centred along the axis, B_b face-centred
    def setup(c_id, t_e, t_b):
        info = o.chunk_info(c_id)
        nb = int(info[9])
        off = info[3:6]
        uf = np.zeros(tuple(int(v) for v in info[6:9]) + (6,))
        n_ax = int(info[6 + ax])
        idx = np.arange(n_ax) - nb + off[ax]
        xf = idx * dh
        xc = xf + 0.5 * dh
        e = np.cos(k * xc - w * t_e)
        b = sgn * np.cos(k * xf - w * t_b)
        shape = [1, 1, 1]
        shape[ax] = n_ax
        uf[..., ecomp] = e.reshape(shape)
        uf[..., bcomp] = b.reshape(shape)
        return uf

    for c_id in range(o.nchunks):
        o.set_field(c_id, 0, setup(c_id, 0.0, -0.5 * dt))
    nstep = 40
    o.step(nstep)
    for c_id in range(o.nchunks):
        info = o.chunk_info(c_id)
        got = helpers.interior(o.field(c_id, 0), info, 3 if N[0] > 1 else 2)
        ref = helpers.interior(setup(c_id, nstep * dt, (nstep - 0.5) * dt), info,
                               3 if N[0] > 1 else 2)
        assert np.max(np.abs(got - ref)) < 1e-11, np.max(np.abs(got - ref))


def _divb(uf_full):
    # single-chunk periodic domain: use np.roll on the interior
    bx, by, bz = uf_full[..., 3], uf_full[..., 4], uf_full[..., 5]
    return (bx - np.roll(bx, 1, axis=2)) + (by - np.roll(by, 1, axis=1)) + \
           (bz - np.roll(bz, 1, axis=0))


def test_divb_preserved():
    cfg = vacuum_cfg((8, 8, 8), (1, 1, 1))
    o = ol.Oracle(cfg, load=False)
    info = o.chunk_info(0)
    rng = np.random.default_rng(5)
    uf = np.zeros(tuple(int(v) for v in info[6:9]) + (6,))
    uf[...] = rng.normal(size=uf.shape)
    o.set_field(0, 0, uf)
    o.exchange(0)
    d0 = _divb(helpers.interior(o.field(0, 0), info, 3))
    o.step(5)
    d1 = _divb(helpers.interior(o.field(0, 0), info, 3))
    assert np.max(np.abs(d1 - d0)) < 1e-12


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("field", ["E", "B"])
@pytest.mark.parametrize("dim", [3, 1])
def test_gather_linear_fields(order, field, dim):
    """dim 1: fields linear in x only (the 1D gather reads the x line,
    pic/engine/velocity.hpp:362-388)."""
    cfg = vacuum_cfg((8, 8, 8) if dim == 3 else (1, 1, 16), (1, 1, 1), order=order)
    o = ol.Oracle(cfg, load=False)
    info = o.chunk_info(0)
    nb = int(info[9])
    n = [int(v) for v in info[6:9]]
    rng = np.random.default_rng(order)
    coef = rng.normal(size=(6, 4)) * 0.1
    if dim == 1:
        coef[:, 2:] = 0.0
    # staggered locations (z, y, x) per component: 0 face, 0.5 centre
    stag = {0: (0.5, 0.5, 0.0), 1: (0.5, 0.0, 0.5), 2: (0.0, 0.5, 0.5),
            3: (0.0, 0.0, 0.5), 4: (0.0, 0.5, 0.0), 5: (0.5, 0.0, 0.0)}
    uf = np.zeros(tuple(n) + (6,))
    zz, yy, xx = np.meshgrid(*(np.arange(m) - nb for m in n), indexing="ij")
    comps = range(3) if field == "E" else range(3, 6)
    for kk in comps:
        sz, sy, sx = stag[kk]
        uf[..., kk] = coef[kk, 0] + coef[kk, 1] * (xx + sx) + coef[kk, 2] * (yy + sy) + \
                      coef[kk, 3] * (zz + sz)
    o.set_field(0, 0, uf)
    pos = rng.uniform(1.0, 7.0, size=(64, 3))
    if dim == 1:
        pos[:, 0] = rng.uniform(1.0, 15.0, size=64)
        pos[:, 1:] = rng.uniform(0.0, 1.0, size=(64, 2))
    u0 = rng.normal(size=(64, 3)) * 0.1 if field == "B" else np.zeros((64, 3))
    xu = np.zeros((64, 7))
    xu[:, :3] = pos
    xu[:, 3:6] = u0
    xu[:, 6] = np.arange(64).astype(np.int64).view(np.float64)
    o.set_particles(0, 0, xu)
    dt = 0.5
    o.push_velocity(dt)
    got = o.particles(0, 0)[:, 3:6]
    q = 1.0 * cfg["parameter"]["particle"][0]["qm"]
    qmdt = 0.5 * cfg["parameter"]["particle"][0]["qm"] * dt
    c = float(cfg["parameter"]["cc"])

    def lin(kk, p):
        return coef[kk, 0] + coef[kk, 1] * p[0] + coef[kk, 2] * p[1] + coef[kk, 3] * p[2]

    for i in range(64):
        e = np.array([lin(k, pos[i]) for k in range(3)]) if field == "E" else np.zeros(3)
        b = np.array([lin(k, pos[i]) for k in range(3, 6)]) if field == "B" else np.zeros(3)
        ref = boris_textbook(u0[i], qmdt * e, qmdt * b, c)
        assert np.max(np.abs(got[i] - ref)) < 1e-13, (i, got[i], ref)
    del q


# --- Friedman time filter (pic/engine/maxwell.hpp:85-146 history, :148-201 combine) ---

def _yee_friedman_numpy(E, B, theta, nstep, c, dt, dh):
    """Periodic Yee scheme with the Friedman filter written in the published
    form (Greenwood et al., J. Comput. Phys. 201, 665 (2004), Eqs. 9-10):
        E_f^n = (1 + th/2) E^n - (1 - th/2) th E^{n-1} + 1/2 (1 - th)^2 th Ebar^{n-2},
        Ebar^{n-2} = E^{n-2} + th Ebar^{n-3},
    started as PicChunk::init_friedman does (E^{-1} = Ebar^{-2} = E^0).  B is
    advanced in two half steps with the same E_f, as push_openmp does.
    Arrays are [3][z][y][x]; E_k sits on the k faces, B_k on the edges, so the
    curl of E uses backward and the curl of B forward differences."""
    A = 1 + 0.5 * theta
    Bc = -(1 - 0.5 * theta) * theta
    C = 0.5 * (1 - theta) ** 2 * theta
    E = E.copy()
    B = B.copy()
    Eprev = E.copy()
    Ebar = E.copy()
    axis = {0: 2, 1: 1, 2: 0}  # component x, y, z -> array axis

    def dback(f, k):
        return f - np.roll(f, 1, axis=axis[k])

    def dfwd(f, k):
        return np.roll(f, -1, axis=axis[k]) - f

    for _ in range(nstep):
        Ef = A * E + Bc * Eprev + C * Ebar
        for _half in range(2):
            cfl = c * 0.5 * dt / dh
            B[0] -= cfl * (dback(Ef[2], 1) - dback(Ef[1], 2))
            B[1] -= cfl * (dback(Ef[0], 2) - dback(Ef[2], 0))
            B[2] -= cfl * (dback(Ef[1], 0) - dback(Ef[0], 1))
        cfl = c * dt / dh
        Enew = E.copy()
        Enew[0] += cfl * (dfwd(B[2], 1) - dfwd(B[1], 2))
        Enew[1] += cfl * (dfwd(B[0], 2) - dfwd(B[2], 0))
        Enew[2] += cfl * (dfwd(B[1], 0) - dfwd(B[0], 1))
        Ebar = Eprev + theta * Ebar
        Eprev = E
        E = Enew
    return E, B


@pytest.mark.parametrize("theta", [0.0, 0.3])
@pytest.mark.parametrize("N,C", [((8, 8, 8), (1, 1, 1)), ((8, 8, 8), (2, 2, 2)),
                                 ((1, 1, 16), (1, 1, 2))])
def test_friedman_filter_matches_published_recursion(theta, N, C):
    cfg = helpers.small_thermal(N=N, C=C, np_=1, order=2, friedman=theta)
    o = ol.Oracle(cfg, load=False)
    dt, dh, c = 0.5, 1.0, 1.0
    rng = np.random.default_rng(21)
    E0 = rng.normal(size=(3,) + N)
    B0 = rng.normal(size=(3,) + N)
    for c_id in range(o.nchunks):
        info = o.chunk_info(c_id)
        nb = int(info[9])
        off = [int(v) for v in info[3:6]]
        ext = [int(v) for v in info[6:9]]
        idx = [(np.arange(ext[a]) - nb + off[a]) % N[a] for a in range(3)]
        uf = np.zeros(tuple(ext) + (6,))
        sel = np.ix_(idx[0], idx[1], idx[2])
        for k in range(3):
            uf[..., k] = E0[k][sel]
            uf[..., 3 + k] = B0[k][sel]
        o.set_field(c_id, 0, uf)
    o.init_friedman()
    nstep = 12
    for _ in range(nstep):
        o.step()
    E, B = _yee_friedman_numpy(E0, B0, theta, nstep, c, dt, dh)
    for c_id in range(o.nchunks):
        info = o.chunk_info(c_id)
        nb = int(info[9])
        off = [int(v) for v in info[3:6]]
        n = [N[a] // C[a] for a in range(3)]
        got = helpers.interior(o.field(c_id, 0), info, 3)
        sl = tuple(slice(off[a], off[a] + n[a]) for a in range(3))
        for k in range(3):
            np.testing.assert_allclose(got[..., k], E[k][sl], rtol=0, atol=1e-11)
            np.testing.assert_allclose(got[..., 3 + k], B[k][sl], rtol=0, atol=1e-11)


def test_friedman_damps_the_shortest_waves():
    """theta > 0 removes energy from grid-scale noise while theta = 0 keeps it
    (to the Yee scheme's bounded oscillation).  theta = 0.1 at c dt / dx = 0.5:
    larger theta at this 3D Courant number grows instead (the filtered scheme's
    stability bound is lower), which the recursion test above does not reach in
    its 12 steps."""
    N = (8, 8, 8)
    rng = np.random.default_rng(3)
    E0 = rng.normal(size=(3,) + N)
    B0 = np.zeros((3,) + N)
    e0 = 0.5 * (E0 ** 2).sum()
    E_a, B_a = _yee_friedman_numpy(E0, B0, 0.0, 200, 1.0, 0.5, 1.0)
    E_b, B_b = _yee_friedman_numpy(E0, B0, 0.1, 200, 1.0, 0.5, 1.0)
    w_a = 0.5 * ((E_a ** 2).sum() + (B_a ** 2).sum())
    w_b = 0.5 * ((E_b ** 2).sum() + (B_b ** 2).sum())
    assert w_a > 0.5 * e0
    assert w_b < 0.5 * w_a
