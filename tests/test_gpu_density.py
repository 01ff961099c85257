"""Particle density volume (psfsolvergpu.cpp:876-913,
Resources/Volume/volume.comp:1-17): counts and the half-precision volume from
the device, bit for bit against the oracle, from a count pass and from
particle passes that count as they store, across state changes."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu

C_FP64, C_BF16X2 = 0, 4
PHI_FP64, PHI_FP32 = 0, 1


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _scene(ctx, sd, N, window=1, c_prec=C_FP64, phi_prec=PHI_FP64):
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, c_prec, phi_prec, N, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    s.set_params(0.01, 0.0, False, lifetime=2.0)
    s.set_frame_window(window)
    return s


def _check(s, sd, dims):
    c, v = s.density()
    rc, rv = O.density(s.get_particles(), sd.aabb_min, sd.aabb_max, dims)
    assert np.array_equal(c, rc)
    assert np.array_equal(v.view(np.uint16), rv.view(np.uint16))
    assert int(c.sum()) == s.get_particles().shape[0]
    return c


@pytest.mark.parametrize("dims", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (256, 256, 256)])
def test_count_pass_matches_oracle(hip_ctx, dims):
    """Stored particles, some outside the AABB, on cell boundaries and NaN."""
    _say(f"density count pass dims={dims}")
    sd = make_scene(n=8, K=8)
    N = 20000
    rng = np.random.default_rng(5)
    p = np.empty((N, 4), np.float32)
    p[:, :3] = rng.uniform(-0.6, 0.6, (N, 3)).astype(np.float32)
    p[:2000, :3] = (rng.integers(0, 33, (2000, 3)) / 32.0 - 0.5).astype(np.float32)
    p[5, 0] = np.nan
    p[:, 3] = 1.0
    s = _scene(hip_ctx, sd, N)
    s.set_particles(p)
    s.set_density_grid(dims, count_in_trace=False)
    _check(s, sd, dims)
    _check(s, sd, dims)          # again: the volume pass cleared the counters, so it recounts
    s.close()


@pytest.mark.parametrize("window,c_prec,phi_prec", [(1, C_FP64, PHI_FP64), (8, C_BF16X2, PHI_FP32)])
@pytest.mark.parametrize("in_trace", [True, False])
def test_density_after_frames(hip_ctx, window, c_prec, phi_prec, in_trace):
    """Counting inside the particle pass gives the counts of the particles it
    stored, frame after frame, with respawns, and so does the count pass."""
    _say(f"density after frames window={window} in_trace={in_trace}")
    sd = make_scene(n=8, K=16)
    N = 5000
    s = _scene(hip_ctx, sd, N, window, c_prec, phi_prec)
    s.set_particles(particles_uniform(sd, N, seed=11, lifetime=0.5))
    dims = (32, 24, 16)
    s.set_density_grid(dims, count_in_trace=in_trace)
    for frames in (1, 3, 1, 9):
        s.integrate(frames)
        _check(s, sd, dims)
    s.close()


def test_state_changes(hip_ctx):
    """Frames without density calls, new particles, a grid change and
    removal: each density call counts the particles stored at that moment."""
    _say("density state changes")
    sd = make_scene(n=8, K=16)
    N = 3000
    s = _scene(hip_ctx, sd, N)
    s.set_particles(particles_uniform(sd, N, seed=2, lifetime=0.5))
    s.set_density_grid((16, 16, 16), count_in_trace=True)
    s.integrate(4)                                   # counted 4 times without a density call
    _check(s, sd, (16, 16, 16))
    s.integrate(2)
    s.set_particles(particles_uniform(sd, N // 2, seed=3, lifetime=0.5))   # the counted ones are gone
    _check(s, sd, (16, 16, 16))
    s.trace()
    _check(s, sd, (16, 16, 16))
    s.set_density_grid((9, 4, 11), count_in_trace=True)
    s.integrate(1)
    _check(s, sd, (9, 4, 11))
    s.set_density_grid((0, 0, 0))
    s.integrate(1)
    from parallelspectralfluids3d_amd import _lib as L
    with pytest.raises(L.PsfError) as e:
        s.density()
    assert e.value.status == L.PSF_ERR_STATE
    with pytest.raises(L.PsfError) as e:
        s.set_density_grid((0, 4, 4))
    assert e.value.status == L.PSF_ERR_ARG
    with pytest.raises(L.PsfError) as e:
        s.set_density_grid((1025, 4, 4))
    assert e.value.status == L.PSF_ERR_ARG
    s.close()


def test_density_large(hip_ctx):
    """2^20 particles into the reference's 256^3 grid after traces."""
    _say("density 2^20 particles, 256^3")
    sd = make_scene(n=16, K=16)
    N = 1 << 20
    s = _scene(hip_ctx, sd, N, window=4, c_prec=C_BF16X2, phi_prec=PHI_FP32)
    s.set_particles(particles_uniform(sd, N, seed=9, lifetime=1.0))
    s.set_density_grid((256, 256, 256), count_in_trace=True)
    s.integrate(6)
    _check(s, sd, (256, 256, 256))
    s.close()
