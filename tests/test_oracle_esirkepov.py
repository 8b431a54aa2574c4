"""Pin the oracle's Esirkepov deposit (CPU only).

Checks, per the reference's own test design (nix/unittest/test_esirkepov.cpp):
  * the new-position charge of one particle sums to q;
  * discrete continuity  rho_new - rho_old + (dt/dh) sum_d (J_d[j+1] - J_d[j]) = 0
    node by node (test_esirkepov.cpp:480-501, 724-754);
and two more that the continuity check alone cannot see:
  * the total current of a particle equals q (x_new - x_old) / dt in every
    direction (fixes the sign and the dh/dt scale of each component);
  * for 2D, Jz = q vz times the time-averaged area weight;
  * for 1D, Jy and Jz on every node = q v times the time-averaged closed-form
    B-spline weight (S_old + S_new) / 2 (nix/esirkepov.hpp:48-67).
Particles move by up to one cell per step in any direction, as in
set_random_particle (test_esirkepov.cpp:388-402).
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def tile3d(order, xv, xu, q, dt, dh):
    L = ol.lib()
    N = order + 3
    t = np.zeros((N, N, N, 4))
    o = np.zeros(3, dtype=np.int32)
    L.pno_esirkepov3d(order, _dp(np.array(xv, float)), _dp(np.array(xu, float)), q, dt, dh, dh, dh,
                      _dp(t), _ip(o))
    return t, o


def tile2d(order, xv, xu, q, dt, dh):
    L = ol.lib()
    N = order + 3
    t = np.zeros((N, N, 4))
    o = np.zeros(3, dtype=np.int32)
    L.pno_esirkepov2d(order, _dp(np.array(xv, float)), _dp(np.array(xu, float)), q, dt, dh, dh,
                      _dp(t), _ip(o))
    return t, o


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("dt,dh", [(0.5, 1.0), (1.0, 0.5), (2.0, 2.0)])
def test_continuity_3d(order, dt, dh):
    rng = np.random.default_rng(100 * order + int(10 * dt))
    N = order + 3
    for _ in range(300):
        x0 = rng.uniform(4, 5, size=3) * dh
        x1 = x0 + rng.uniform(-1, 1, size=3) * dh
        q = rng.uniform(0.1, 2.0)
        t, o = tile3d(order, x0, x1, q, dt, dh)
        s, o2 = tile3d(order, x0, x0, q, dt, dh)
        assert np.array_equal(o, o2)
        rho1, rho0 = t[..., 0], s[..., 0]
        assert abs(rho1.sum() - q) < 1e-13 * q
        J = np.zeros((N + 1, N + 1, N + 1, 4))
        J[:N, :N, :N] = t
        div = (J[:N, :N, 1:, 1] - J[:N, :N, :N, 1]) + (J[:N, 1:, :N, 2] - J[:N, :N, :N, 2]) + \
              (J[1:, :N, :N, 3] - J[:N, :N, :N, 3])
        res = (rho1 - rho0) + (dt / dh) * div
        assert np.sum(np.abs(res)) < 1e-13 * np.sum(np.abs(rho1))
        # total current = q v in every direction: x (comp 1) <-> axis 2 of the tile
        v = (x1 - x0) / dt
        for comp, ax in ((1, 0), (2, 1), (3, 2)):
            tot = t[..., comp].sum() * dh
            assert abs(tot - q * v[ax] * dh) < 1e-12 * q * max(1.0, dh / dt)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_continuity_2d(order):
    rng = np.random.default_rng(order)
    dt, dh = 0.5, 1.0
    N = order + 3
    for _ in range(300):
        x0 = rng.uniform(4, 5, size=3)
        x1 = x0 + rng.uniform(-1, 1, size=3)
        q = rng.uniform(0.1, 2.0)
        t, o = tile2d(order, x0, x1, q, dt, dh)
        s, _ = tile2d(order, x0, x0, q, dt, dh)
        rho1, rho0 = t[..., 0], s[..., 0]
        assert abs(rho1.sum() - q) < 1e-13 * q
        J = np.zeros((N + 1, N + 1, 4))
        J[:N, :N] = t
        div = (J[:N, 1:, 1] - J[:N, :N, 1]) + (J[1:, :N, 2] - J[:N, :N, 2])
        res = (rho1 - rho0) + (dt / dh) * div
        assert np.sum(np.abs(res)) < 1e-13 * np.sum(np.abs(rho1))
        v = (x1 - x0) / dt
        assert abs(t[..., 1].sum() - q * v[0]) < 1e-12 * q
        assert abs(t[..., 2].sum() - q * v[1]) < 1e-12 * q
        # Jz: q vz times weights summing to one
        assert abs(t[..., 3].sum() - q * v[2]) < 1e-12 * q


def tile1d(order, xv, xu, q, dt, dh):
    L = ol.lib()
    N = order + 3
    t = np.zeros((N, 4))
    o = np.zeros(3, dtype=np.int32)
    L.pno_esirkepov1d(order, _dp(np.array(xv, float)), _dp(np.array(xu, float)), q, dt, dh,
                      _dp(t), _ip(o))
    return t, o


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("dt,dh", [(0.5, 1.0), (2.0, 2.0)])
def test_continuity_1d(order, dt, dh):
    from test_oracle_primitives import bspline
    rng = np.random.default_rng(7 * order + int(dt))
    N = order + 3
    is_odd = order % 2
    for _ in range(300):
        x0 = rng.uniform(4, 5, size=3) * dh
        x1 = x0 + rng.uniform(-1, 1, size=3) * dh
        q = rng.uniform(0.1, 2.0)
        t, o = tile1d(order, x0, x1, q, dt, dh)
        s, o2 = tile1d(order, x0, x0, q, dt, dh)
        assert np.array_equal(o, o2)
        rho1, rho0 = t[:, 0], s[:, 0]
        assert abs(rho1.sum() - q) < 1e-13 * q
        J = np.zeros(N + 1)
        J[:N] = t[:, 1]
        res = (rho1 - rho0) + (dt / dh) * (J[1:] - J[:N])
        assert np.sum(np.abs(res)) < 1e-13 * np.sum(np.abs(rho1))
        v = (x1 - x0) / dt
        assert abs(t[:, 1].sum() * dh - q * v[0] * dh) < 1e-12 * q * max(1.0, dh / dt)
        # node j of the tile sits at cell index o_x + j - Nb (Nb = 0 here): centre (i + 1/2) dh
        cell0 = int(np.floor(x0[0] / dh - 0.5 * is_odd))
        assert o[2] == cell0 - order // 2 - 1
        for j in range(N):
            xc = (o[2] + j + 0.5) * dh
            w = 0.5 * (bspline(order, x0[0] / dh - xc / dh) + bspline(order, x1[0] / dh - xc / dh))
            assert abs(t[j, 2] - q * v[1] * w) < 1e-13 * q * max(1.0, abs(v[1]))
            assert abs(t[j, 3] - q * v[2] * w) < 1e-13 * q * max(1.0, abs(v[2]))


def test_origin_tracks_old_cell():
    # the tile starts at the old cell - Order/2 - 1 (pic/engine/current.hpp:390-392)
    for order in (1, 2, 3, 4):
        is_odd = order % 2
        x0 = np.array([3.3, 4.6, 5.1])
        _, o = tile3d(order, x0, x0, 1.0, 0.5, 1.0)
        cell = np.floor(x0 - 0.5 * is_odd).astype(int)[::-1]  # (z, y, x)
        assert np.array_equal(o, cell - order // 2 - 1)
