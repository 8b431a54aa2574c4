"""The CPU tangent-linear reference of the adjoint form (pno_adj_tangent) against
central finite differences of its own loss, on runs short and small enough for
the difference quotient to converge (order-1 shapes make the loss piecewise
smooth), and its linearity in the direction."""
import numpy as np

import oracle_lib


def _state(nx, nppc, seed):
    rng = np.random.default_rng(seed)
    n = nx * nx * nppc
    xu = np.zeros((n, 6))
    xu[:, 0] = rng.uniform(0, nx, n)
    xu[:, 1] = rng.uniform(0, nx, n)
    xu[:, 3] = 0.3 * rng.standard_normal(n)
    xu[:, 4:6] = 0.05 * rng.standard_normal((n, 2))
    return xu


def test_tangent_matches_finite_differences():
    nx, nppc, steps = 8, 4, 12
    xu = _state(nx, nppc, 1)
    rng = np.random.default_rng(2)
    args = (nx, nx, 1.0, 0.5, 1.0, -1.0 / nppc, 1.0 / nppc)
    for trial in range(3):
        d = np.zeros_like(xu)
        d[:, :2] = rng.standard_normal((xu.shape[0], 2))
        d[:, 3:] = 0.1 * rng.standard_normal((xu.shape[0], 3))
        L0, dL = oracle_lib.adj_tangent(*args, xu, d, steps)
        h = 1e-7
        Lp, _ = oracle_lib.adj_tangent(*args, xu + h * d, np.zeros_like(d), steps)
        Lm, _ = oracle_lib.adj_tangent(*args, xu - h * d, np.zeros_like(d), steps)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - dL) <= 1e-5 * abs(dL), (trial, dL, fd)


def test_tangent_is_linear_in_the_direction():
    nx, nppc, steps = 8, 4, 10
    xu = _state(nx, nppc, 3)
    rng = np.random.default_rng(4)
    args = (nx, nx, 1.0, 0.5, 1.0, -1.0 / nppc, 1.0 / nppc)
    a = rng.standard_normal(xu.shape)
    b = rng.standard_normal(xu.shape)
    _, da = oracle_lib.adj_tangent(*args, xu, a, steps)
    _, db = oracle_lib.adj_tangent(*args, xu, b, steps)
    _, dab = oracle_lib.adj_tangent(*args, xu, 2.0 * a - 3.0 * b, steps)
    assert abs(dab - (2.0 * da - 3.0 * db)) <= 1e-9 * (abs(da) + abs(db))
