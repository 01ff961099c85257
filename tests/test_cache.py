"""Scene cache files (psf_cache_write / psf_cache_read_header / psf_cache_read):
round trip bit for bit, optional arrays, and refusal of corrupt, truncated
and foreign files.  No GPU is touched."""
import os
import struct

import numpy as np
import pytest

from scenes import make_scene


@pytest.fixture(scope="module")
def hip():
    from parallelspectralfluids3d_amd import hip as H
    return H


def _arrays(sd):
    return dict(lam=sd.lam, C=sd.C, Phi_cm=sd.Phi, gravity=sd.g)


def test_round_trip(tmp_path, hip):
    sd = make_scene(n=6, K=8)
    p = str(tmp_path / "s.psfcache")
    hip.write_cache(p, sd.grid, sd.resolution, sd.lam, sd.C, Phi_cm=sd.Phi, gravity=sd.g,
                    aabb_min=sd.aabb_min, aabb_max=sd.aabb_max)
    d, has_phi, has_g = hip.read_cache_header(p)
    assert (d.nx, d.ny, d.nz, d.K) == (6, 6, 6, 8) and has_phi and has_g
    assert d.resolution == sd.resolution
    r = hip.read_cache(p)
    for got, ref in ((r["lam"], sd.lam), (r["C"], sd.C), (r["Phi"], sd.Phi), (r["gravity"], sd.g)):
        assert np.array_equal(np.asarray(got).view(np.uint64), np.ascontiguousarray(ref, np.float64).view(np.uint64))
    assert not os.path.exists(p + ".tmp")


def test_without_phi_and_gravity(tmp_path, hip):
    sd = make_scene(n=4, K=5)
    p = str(tmp_path / "s.psfcache")
    hip.write_cache(p, sd.grid, sd.resolution, sd.lam, sd.C)
    d, has_phi, has_g = hip.read_cache_header(p)
    assert not has_phi and not has_g
    r = hip.read_cache(p)
    assert r["Phi"] is None and r["gravity"] is None
    assert np.array_equal(r["C"], sd.C)


def test_corrupt_truncated_foreign(tmp_path, hip):
    from parallelspectralfluids3d_amd import _lib as L
    sd = make_scene(n=4, K=6)
    p = str(tmp_path / "s.psfcache")
    hip.write_cache(p, sd.grid, sd.resolution, sd.lam, sd.C, Phi_cm=sd.Phi)
    raw = bytearray(open(p, "rb").read())
    # one flipped payload bit: the checksum refuses it
    bad = bytearray(raw)
    bad[128 + 8 * 3] ^= 0x10
    q = str(tmp_path / "bad.psfcache")
    open(q, "wb").write(bad)
    with pytest.raises(L.PsfError) as e:
        hip.read_cache(q)
    assert e.value.status == L.PSF_ERR_ARG
    # truncated
    open(q, "wb").write(raw[:-9])
    with pytest.raises(L.PsfError):
        hip.read_cache(q)
    # another format version
    other = bytearray(raw)
    other[8:12] = struct.pack("<I", 99)
    open(q, "wb").write(other)
    with pytest.raises(L.PsfError) as e:
        hip.read_cache_header(q)
    assert e.value.status == L.PSF_ERR_UNSUPPORTED
    # not a cache
    open(q, "wb").write(b"x" * 200)
    with pytest.raises(L.PsfError):
        hip.read_cache_header(q)
