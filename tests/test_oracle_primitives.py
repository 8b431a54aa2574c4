"""Pin the oracle's primitives against closed forms and invariants (CPU only).

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, an off-by-half-cell grid) fails it:
  * shape_mc1..4 against the analytic B-splines used by the reference's own
    tests (nix/unittest/test_primitives.cpp:271-530) and partition of unity;
  * digitize against floor((x - xmin) / dx) cases (test_primitives.cpp:73-117);
  * lorentz_factor against sqrt(1 + u^2 / c^2);
  * push_boris against an independently written textbook Boris rotation, plus
    |u| conservation for E = 0, the exact rotation angle 2 atan(|t|) and the
    E x B drift.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as ol


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def bspline(order, x):
    """Centred cardinal B-spline of the given degree (support (order+1)/2)."""
    ax = abs(x)
    if order == 1:
        return max(0.0, 1 - ax)
    if order == 2:
        if ax < 0.5:
            return 0.75 - x * x
        if ax < 1.5:
            return (1.5 - ax) ** 2 / 2
        return 0.0
    if order == 3:
        if ax < 1:
            return 2.0 / 3 - ax * ax + ax ** 3 / 2
        if ax < 2:
            return (2 - ax) ** 3 / 6
        return 0.0
    if ax < 0.5:
        return 115.0 / 192 - 5 * ax ** 2 / 8 + ax ** 4 / 4
    if ax < 1.5:
        return (55 + 20 * ax - 120 * ax ** 2 + 80 * ax ** 3 - 16 * ax ** 4) / 96
    if ax < 2.5:
        return (5 - 2 * ax) ** 4 / 384
    return 0.0


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_shape_mc_closed_form(order):
    L = ol.lib()
    rng = np.random.default_rng(order)
    # grid point offsets of the weights, from the shape_mc documentation
    # (nix/primitives.hpp:482-512): odd orders X <= x < X + dx, even X - dx/2 <= x < X + dx/2
    offs = {1: [0, 1], 2: [-1, 0, 1], 3: [-1, 0, 1, 2], 4: [-2, -1, 0, 1, 2]}[order]
    lo, hi = (0.0, 1.0) if order % 2 == 1 else (-0.5, 0.5)
    for _ in range(200):
        d = rng.uniform(lo, hi)
        s = np.zeros(order + 1)
        X = 3.0
        L.pno_shape_mc(order, X + d, X, 1.0, _dp(s))
        ref = np.array([bspline(order, o - d) for o in offs])
        assert np.max(np.abs(s - ref)) < 1e-14
        assert abs(s.sum() - 1.0) < 1e-14


def test_digitize():
    L = ol.lib()
    for x, xmin, dx in [(0.0, 0.0, 1.0), (0.999999, 0.0, 1.0), (1.0, 0.0, 1.0), (-0.1, 0.0, 1.0),
                        (5.25, 0.5, 0.25), (-3.7, -1.0, 0.5)]:
        assert L.pno_digitize(x, xmin, 1 / dx) == math.floor((x - xmin) * (1 / dx))


def test_lorentz_factor():
    L = ol.lib()
    rng = np.random.default_rng(1)
    for _ in range(100):
        u = rng.normal(size=3) * 3
        c = rng.uniform(0.5, 5)
        ref = math.sqrt(1 + (u @ u) / c ** 2)
        assert abs(L.pno_lorentz_factor(u[0], u[1], u[2], 1 / c) - ref) < 1e-14 * ref


def boris_textbook(u, e, b, c):
    """Birdsall & Langdon Boris step with the reference's normalisation:
    e, b are already multiplied by (q/m) dt / 2; the rotation uses b / (gamma c)."""
    um = u + e
    gamma_c = math.sqrt(c * c + um @ um)
    t = b / gamma_c
    s = 2.0 / (1.0 + t @ t)
    up = um + np.cross(um, t)
    upp = um + s * np.cross(up, t)
    return upp + e


def _boris(u, eb, c):
    L = ol.lib()
    uu = np.array(u, dtype=np.float64)
    ebb = np.array(eb, dtype=np.float64)
    L.pno_push_boris(_dp(uu), _dp(ebb), c)
    return uu


def test_boris_matches_textbook():
    rng = np.random.default_rng(7)
    for _ in range(500):
        u = rng.normal(size=3)
        e = rng.normal(size=3) * 0.1
        b = rng.normal(size=3) * 0.5
        c = rng.uniform(0.5, 3)
        got = _boris(u, np.concatenate([e, b]), c)
        ref = boris_textbook(u, e, b, c)
        assert np.max(np.abs(got - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_boris_rotation_invariants():
    rng = np.random.default_rng(3)
    for _ in range(200):
        u = rng.normal(size=3)
        b = rng.normal(size=3)
        c = 1.3
        got = _boris(u, np.concatenate([np.zeros(3), b]), c)
        # |u| conserved for E = 0
        assert abs(np.linalg.norm(got) - np.linalg.norm(u)) < 1e-14 * np.linalg.norm(u)
        # rotation angle about b is 2 atan(|t|), t = b / (gamma c)
        t = np.linalg.norm(b) / math.sqrt(c * c + u @ u)
        bh = b / np.linalg.norm(b)
        p0 = u - (u @ bh) * bh
        p1 = got - (got @ bh) * bh
        ang = math.atan2(np.linalg.norm(np.cross(p0, p1)), p0 @ p1)
        assert abs(ang - 2 * math.atan(t)) < 1e-12


def test_boris_exb_drift():
    # non-relativistic E x B drift: <v> -> c E x B / B^2 (Gaussian units), with
    # e, b scaled by (q/m) dt / 2 as the pusher expects
    c = 100.0
    qm_dt_half = 0.05
    E = np.array([0.0, 1e-2, 0.0]) * c
    B = np.array([0.0, 0.0, 1.0]) * c
    u = np.zeros(3)
    vs = []
    for _ in range(20000):
        u = _boris(u, np.concatenate([qm_dt_half * E, qm_dt_half * B]), c)
        vs.append(u / math.sqrt(1 + (u @ u) / c ** 2))
    vd = np.mean(vs, axis=0)
    ref = c * np.cross(E, B) / (B @ B)
    assert np.max(np.abs(vd - ref)) < 2e-3 * np.linalg.norm(ref)


# --- Vay (2008) and Higuera-Cary (2017) pushers (nix/primitives.hpp:191-253) ---
#
# The textbook forms below are written from the papers (c = 1 variables: U = u/c,
# t-vector tau = b/c with b = (q/m)(dt/2) B, eps = e/c):
#   Vay, Phys. Plasmas 15, 056701 (2008), Eqs. (11)-(14);
#   Higuera & Cary, Phys. Plasmas 24, 052104 (2017), Eqs. (17)-(22);
# the rotation is applied with Rodrigues' formula, not the Boris cross products.
# The reference's expressions reduce to these at c = 1 (its gamma solve is not
# scale-free in c, so other c values follow the reference as written).

def _gamma_new(uprime, tau):
    """Root of the quadratic for the new Lorentz factor (Vay Eq. 11 / HC Eq. 20)."""
    sigma = 1.0 + uprime @ uprime - tau @ tau
    ustar = uprime @ tau
    return math.sqrt(0.5 * (sigma + math.sqrt(sigma * sigma + 4 * (tau @ tau + ustar * ustar))))


def _rotate(u, axis, angle):
    n = np.linalg.norm(axis)
    if n == 0:
        return u.copy()
    k = axis / n
    return u * math.cos(angle) + np.cross(k, u) * math.sin(angle) + k * (k @ u) * (1 - math.cos(angle))


def vay_textbook(u, e, b):
    gamma = math.sqrt(1 + u @ u)
    uprime = u + 2 * e + np.cross(u / gamma, b)
    g1 = _gamma_new(uprime, b)
    t = b / g1
    s = 1.0 / (1.0 + t @ t)
    return s * (uprime + (uprime @ t) * t + np.cross(uprime, t))


def hc_textbook(u, e, b):
    uminus = u + e
    g1 = _gamma_new(uminus, b)
    t = b / g1
    # u+ is u- rotated about t by -2 atan|t| (the magnetic half of the step)
    uplus = _rotate(uminus, t, -2 * math.atan(np.linalg.norm(t)))
    return uplus + e


def _push(name, u, eb, c):
    L = ol.lib()
    uu = np.array(u, dtype=np.float64)
    ebb = np.array(eb, dtype=np.float64)
    getattr(L, name)(_dp(uu), _dp(ebb), c)
    return uu


@pytest.mark.parametrize("name,ref", [("pno_push_vay", vay_textbook),
                                      ("pno_push_higuera_cary", hc_textbook)])
def test_vay_hc_match_papers(name, ref):
    rng = np.random.default_rng(11)
    for _ in range(500):
        u = rng.normal(size=3) * rng.choice([0.1, 1.0, 10.0])
        e = rng.normal(size=3) * 0.1
        b = rng.normal(size=3) * 0.7
        got = _push(name, u, np.concatenate([e, b]), 1.0)
        want = ref(u, e, b)
        assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("name", ["pno_push_vay", "pno_push_higuera_cary"])
def test_vay_hc_magnetic_invariants(name):
    rng = np.random.default_rng(5)
    for _ in range(200):
        u = rng.normal(size=3) * 3
        b = rng.normal(size=3)
        got = _push(name, u, np.concatenate([np.zeros(3), b]), 1.0)
        # E = 0: pure rotation, |u| conserved
        assert abs(np.linalg.norm(got) - np.linalg.norm(u)) < 1e-13 * np.linalg.norm(u)


@pytest.mark.parametrize("name", ["pno_push_vay", "pno_push_higuera_cary"])
def test_vay_hc_relativistic_exb_is_exact(name):
    """A particle at the E x B drift velocity feels no force (E + v x B / c = 0):
    Vay's and Higuera-Cary's schemes keep it unchanged for any drift speed (the
    property they were built for); Boris does not at gamma >> 1."""
    c = 1.0
    for ratio in (0.3, 0.9, 0.99):
        b0 = 0.4
        e = np.array([ratio * b0, 0.0, 0.0])  # (q/m)(dt/2) E, E along x
        b = np.array([0.0, 0.0, b0])          # (q/m)(dt/2) B, B along z
        vd = -c * ratio * np.array([0.0, 1.0, 0.0])  # c E x B / B^2
        gam = 1 / math.sqrt(1 - vd @ vd / c ** 2)
        u = gam * vd
        got = _push(name, u, np.concatenate([e, b]), c)
        assert np.max(np.abs(got - u)) < 1e-12 * np.linalg.norm(u), (ratio, got, u)
    # Boris drifts off at gamma ~ 7
    got = _boris(u, np.concatenate([e, b]), c)
    assert np.max(np.abs(got - u)) > 1e-6 * np.linalg.norm(u)


# --- WT shape functions (nix/primitives.hpp:331-566) ---

def _box(x):
    return 1.0 if -0.5 <= x < 0.5 else 0.0


def _wt_reference(order, delta, dt):
    """Weights of the WT shape of the given order at its nodes, by direct
    quadrature: the degree-(order-1) centred B-spline smeared by a top hat of
    half width dt (Lu et al., J. Comput. Phys. 413, 109388 (2020))."""
    from scipy.integrate import quad
    nodes = {1: [0, 1], 2: [-1, 0, 1], 3: [-1, 0, 1, 2], 4: [-2, -1, 0, 1, 2]}[order]
    deg = order - 1
    out = []
    for n in nodes:
        f = (lambda u, n=n: _box(delta + u - n)) if deg == 0 else \
            (lambda u, n=n: bspline(deg, delta + u - n))
        # breakpoints of the piecewise polynomial inside [-dt, dt]
        bps = sorted({b for k in range(-4, 6) for b in (k - 0.5 - delta + n, k - delta + n)
                      if -dt < b < dt})
        val, _ = quad(f, -dt, dt, points=bps or None, epsabs=1e-15, epsrel=1e-13, limit=200)
        out.append(val / (2 * dt))
    return np.array(out)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_shape_wt_is_smeared_bspline(order):
    L = ol.lib()
    L.pno_shape_wt.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double,
                               ctypes.POINTER(ctypes.c_double)]
    rng = np.random.default_rng(order)
    lo, hi = (0.0, 1.0) if order % 2 == 1 else (-0.5, 0.5)
    for dt in (0.1, 0.3, 0.45, 0.5):
        for delta in np.concatenate([rng.uniform(lo, hi, 40), [lo, 0.5 - dt, 0.5 + dt - 1e-9,
                                                                 -dt, dt - 1e-9]]):
            if not (lo <= delta < hi):
                continue
            s = np.zeros(order + 1)
            # X = 0, dx = 1: delta = x
            L.pno_shape_wt(order, float(delta), 0.0, 1.0, dt, 1 / dt, _dp(s))
            ref = _wt_reference(order, float(delta), dt)
            assert np.max(np.abs(s - ref)) < 1e-12, (order, dt, delta, s, ref)
            assert abs(s.sum() - 1) < 1e-13
