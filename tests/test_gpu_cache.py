"""A scene loaded from a cache file steps, reconstructs and traces bit for
bit like the same scene uploaded from the arrays."""
import sys

import numpy as np
import pytest

from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu


def test_cached_scene_equals_uploaded(hip_ctx, tmp_path):
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import hip
    print("[gpu-case] scene cache vs upload", flush=True)
    sys.stdout.flush()
    sd = make_scene(n=8, K=16)
    p = str(tmp_path / "s.psfcache")
    hip.write_cache(p, sd.grid, sd.resolution, sd.lam, sd.C, Phi_cm=sd.Phi, gravity=sd.g,
                    aabb_min=sd.aabb_min, aabb_max=sd.aabb_max)
    parts = particles_uniform(sd, 2000, seed=4, lifetime=20.0)
    out = []
    for via_cache in (False, True):
        d = hip.make_desc(sd.grid, sd.resolution, sd.K, L.C_BF16X2, L.PHI_FP32, 2000, sd.aabb_min, sd.aabb_max)
        s = hip.Scene(hip_ctx, d)
        if via_cache:
            s.load_cache(p)
        else:
            s.upload_basis(sd.Phi, sd.lam, sd.g)
            s.upload_advection(sd.C)
        s.set_omega(sd.omega)
        s.set_params(0.01, 0.0, True, lifetime=20.0)
        s.set_particles(parts)
        s.integrate(25)
        out.append((s.get_omega(), s.get_velocity(), s.get_particles()))
        s.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    # a scene of another K refuses the file
    d = hip.make_desc(sd.grid, sd.resolution, sd.K + 1, L.C_FP64, L.PHI_FP64, 0)
    s = hip.Scene(hip_ctx, d)
    with pytest.raises(L.PsfError):
        s.load_cache(p)
    s.close()
