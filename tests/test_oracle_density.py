"""The oracle's particle density (histogram -> r16f volume of
psfsolvergpu.cpp:876-913 and Resources/Volume/volume.comp:1-17), pinned by
closed forms, an independent scalar restatement and IEEE half conversion by
the standard library."""
import math
import struct

import numpy as np
import pytest

import oracle as O


def _f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def _cell_scalar(p, amin, amax, n):
    # float32 arithmetic emulated with doubles rounded to float32: exact for
    # one +, -, *, / of float32 operands (53 >= 2 * 24 + 2 bits)
    ext = _f32(_f32(amax) - _f32(amin))
    t = _f32(_f32(_f32(p) - _f32(amin)) / ext) if ext != 0 else math.nan
    t = _f32(t * float(n))
    if t >= n:
        return n - 1
    if t > 0:
        return int(t)
    return 0


def test_half_conversion_matches_ieee_binary16():
    """float(count) / 1e6 to binary16 equals the standard library's IEEE
    round-to-nearest-even packing, over every count up to 2^17 and large ones."""
    counts = np.concatenate([np.arange(0, 1 << 17), np.random.default_rng(0).integers(0, 1 << 26, 20000)])
    for c in counts[::97]:
        q = _f32(float(int(c)))
        x = _f32(q / 1.0e6)
        ref = struct.unpack("<H", struct.pack("<e", x))[0]
        got = (np.asarray([c], np.uint32).astype(np.float32) / np.float32(1e6)).astype(np.float16).view(np.uint16)[0]
        assert int(got) == ref, c
    # through the oracle entry point: a cell that holds c particles
    for c in (1, 2, 17, 1000, 65535, 100000):
        p = np.zeros((c, 4), np.float32)
        counts_o, vol = O.density(p, (-0.5,) * 3, (0.5,) * 3, (2, 2, 2))
        assert counts_o.sum() == c
        v = vol[1, 1, 1]
        assert struct.pack("<e", _f32(c / 1.0e6)) == v.tobytes()


def test_cell_centres_land_in_their_cell():
    """A particle at the centre of cell (i, j, k) of a non-cubic grid is
    counted in that cell: catches swapped axes or a transposed layout."""
    dims = (7, 5, 3)
    amin, amax = np.array([-0.5, -0.25, -1.0]), np.array([0.5, 1.25, 2.0])
    rows = []
    for k in range(dims[2]):
        for j in range(dims[1]):
            for i in range(dims[0]):
                c = amin + (np.array([i, j, k]) + 0.5) / np.array(dims) * (amax - amin)
                rows.append([c[0], c[1], c[2], 1.0])
    p = np.array(rows, np.float32)
    counts, vol = O.density(p, amin, amax, dims)
    assert counts.shape == (3, 5, 7)
    assert np.array_equal(counts, np.ones_like(counts))
    assert np.all(vol == np.float16(1e-6))
    # one particle in cell (6, 0, 2) only
    counts, _ = O.density(p[[2 * 35 + 0 * 7 + 6]], amin, amax, dims)
    assert counts[2, 0, 6] == 1 and counts.sum() == 1


def test_edges_clamp_and_nan():
    dims = (4, 4, 4)
    p = np.array([[-0.5, -0.5, -0.5, 1],      # aabb_min -> cell 0
                  [0.5, 0.5, 0.5, 1],         # aabb_max -> last cell
                  [-9.0, 0.0, 9.0, 1],        # outside: clamped
                  [np.nan, 0.1, 0.3, 1],      # NaN: cell 0 on x
                  [np.inf, -np.inf, 0.0, 1]], np.float32)
    counts, _ = O.density(p, (-0.5,) * 3, (0.5,) * 3, dims)
    assert counts[0, 0, 0] == 1
    assert counts[3, 3, 3] == 1
    assert counts[3, 2, 0] == 2                # the clamped one and the NaN one
    assert counts[2, 0, 3] == 1
    assert counts.sum() == 5


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 8, 5), (16, 16, 16)])
def test_against_scalar_restatement(dims):
    """Random particles, partly outside the AABB and on cell boundaries,
    against a per-particle scalar loop with float32-emulated arithmetic."""
    rng = np.random.default_rng(sum(dims))
    amin, amax = (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)
    p = np.empty((3000, 4), np.float32)
    p[:, :3] = rng.uniform(-0.6, 0.6, (3000, 3)).astype(np.float32)
    p[:500, :3] = (rng.integers(0, 17, (500, 3)) / 16.0 - 0.5).astype(np.float32)   # on boundaries
    p[:, 3] = 1.0
    counts, vol = O.density(p, amin, amax, dims)
    ref = np.zeros((dims[2], dims[1], dims[0]), np.uint32)
    for row in p:
        c = [_cell_scalar(float(row[a]), amin[a], amax[a], dims[a]) for a in range(3)]
        ref[c[2], c[1], c[0]] += 1
    assert np.array_equal(counts, ref)
    assert counts.sum() == p.shape[0]
    assert np.array_equal(vol, (ref.astype(np.float32) / np.float32(1e6)).astype(np.float16))
