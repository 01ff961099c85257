"""The particle pass's float cell lookup for power-of-two voxel sizes.

The CPU path computes the three cell indices of an axis from the float
coordinate a as int(a / h), int((a - h/2) / h) and int((a + h/2) / h) in
double (psfsolver.cpp:80-100), then clamps them to [0, n - 1].  When h = 2^-k
the HIP tracer computes them in float instead: q = a * 2^k, int(q),
int(q - 0.5f) and (q < 0.5f ? 0 : int(q + 0.5f)).  This checks that the two
agree after the clamp on coordinates chosen where a mistake would show: on
and a few ulps either side of every cell and half-cell plane, around q = 1/2
at the finest float spacing, subnormal, huge, infinite and NaN values.
Conversions follow the device's: truncation, saturating to int32, NaN to 0.
"""
import numpy as np
import pytest


def _cvt(x):
    x = np.asarray(x, np.float64)
    out = np.where(np.isnan(x), 0.0, np.clip(np.trunc(x), -2.0**31, 2.0**31 - 1))
    return out.astype(np.int64)


def _double_path(a, k, n):
    a = a.astype(np.float64)
    h = 2.0**-k
    with np.errstate(invalid="ignore", over="ignore"):
        idx = [_cvt(a / h), _cvt((a - h / 2) / h), _cvt((a + h / 2) / h)]
    return [np.clip(i, 0, n - 1) for i in idx]


def _float_path(a, k, n):
    a = a.astype(np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        q = a * np.float32(2.0**k)
        O = _cvt(q)
        M = _cvt(q - np.float32(0.5))
        P = np.where(q < np.float32(0.5), 0, _cvt(q + np.float32(0.5)))
    return [np.clip(i, 0, n - 1) for i in (O, M, P)]


def _inputs(k, n, rng):
    h = 2.0**-k
    planes = np.concatenate([np.arange(-2, n + 3) * h, (np.arange(-2, n + 2) + 0.5) * h]).astype(np.float32)
    vals = [planes]
    for steps in range(1, 9):
        vals.append(np.nextafter(planes, np.float32(np.inf)))
        vals.append(np.nextafter(planes, np.float32(-np.inf)))
        planes = vals[-2]
        vals.append(planes)
    # the finest floats just below q = 1/2 (a = h/2 minus a few ulps) and near 0
    half = np.float32(h / 2)
    below = [half]
    for _ in range(64):
        below.append(np.nextafter(below[-1], np.float32(0)))
    vals.append(np.array(below, np.float32))
    tiny = np.array([0.0, -0.0, 1e-45, -1e-45, 1e-38, -1e-38, 2.0**-126, 2.0**-140], np.float32)
    huge = np.array([2.0**23, 2.0**23 - 0.5, 2.0**24 + 2, 2.0**31, 2.0**32, 1e30, -1e30, 3.4e38, -3.4e38,
                     np.inf, -np.inf, np.nan], np.float32)
    vals += [tiny, huge, huge * np.float32(h)]
    vals.append(rng.uniform(-0.25 * n * h, 1.25 * n * h, 20000).astype(np.float32))
    return np.concatenate(vals).astype(np.float32)


@pytest.mark.parametrize("k,n", [(0, 1), (0, 2), (1, 3), (3, 8), (5, 32), (6, 64), (7, 128), (10, 700),
                                 (11, 2048), (4, 2048)])
def test_float_cell_indices_match_double(k, n):
    rng = np.random.default_rng(k * 4099 + n)
    a = _inputs(k, n, rng)
    ref = _double_path(a, k, n)
    got = _float_path(a, k, n)
    for name, r, g in zip("OMP", ref, got):
        bad = np.flatnonzero(r != g)
        assert bad.size == 0, f"{name}: a={a[bad[:5]]!r} double {r[bad[:5]]} float {g[bad[:5]]}"


def test_unguarded_plus_half_would_differ():
    """Without the q < 1/2 select, q + 0.5f rounds up to 1 for q just below 1/2
    (a tie at 1 - 2^-25 goes to the even 1.0): the select is needed."""
    q = np.float32(0.5) - np.float32(2.0**-25)
    assert q < np.float32(0.5)
    assert np.float32(q + np.float32(0.5)) == np.float32(1.0)
    assert int(np.trunc(np.float64(q) + 0.5)) == 0
