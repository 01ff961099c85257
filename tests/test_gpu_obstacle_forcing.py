"""The plain drop-in form on the GPU (FP32 C): a voxelised central sphere
obstacle (cells whose centre lies inside radius 0.25 are solid), the
reference's own pruned-M eigenbasis, an upward force on modes 0-7, and
particles seeded in fluid cells -- the 32^3/K=128 configuration at 16^3/K=32,
run frame by frame against the oracle (psfsolver.cpp:21-234)."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu

C_FP32, C_FP64, PHI_FP32 = 1, 0, 1


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _upward_force(sd, modes=8):
    """-g (gravity_field points down) kept on modes 0..modes-1."""
    f = -np.asarray(sd.g, np.float64).copy()
    f[modes:] = 0.0
    return f


@pytest.mark.parametrize("c_prec,window,tol_w", [(C_FP32, 1, 1e-8), (C_FP32, 16, 1e-8), (C_FP64, 1, 1e-12)])
def test_sphere_obstacle_forcing_frames(hip_ctx, c_prec, window, tol_w):
    _say(f"sphere obstacle, forcing on modes 0-7, c={c_prec} window={window}")
    from parallelspectralfluids3d_amd import hip
    sd = make_scene(n=16, K=32, basis_kind="eigen", obstacle=True)
    force = _upward_force(sd)
    N, frames, dt = 1 << 14, 300, 0.005
    p0 = particles_uniform(sd, N, seed=21, fluid_only=True)
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, c_prec, PHI_FP32, N, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(hip_ctx, d)
    s.upload_basis(sd.Phi, sd.lam, force)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    s.set_params(dt, 0.0, True, lifetime=60.0)
    s.set_particles(p0)
    s.set_frame_window(window)
    s.integrate(frames)
    wg, ug, pg = s.get_omega(), s.get_velocity(), s.get_particles()
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    w, p = sd.omega.copy(), p0.copy()
    for f in range(frames):
        w = O.reduced_step(w, sd.C, sd.lam, force, dt, 0.0, True)
        u = O.reconstruct(sd.Phi, w)
        p = O.particles(m, u, p, resolution=sd.resolution, timestep=dt, aabb_min=sd.aabb_min,
                        aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=f)
    rel = np.linalg.norm(wg - w) / np.linalg.norm(w)
    print(f"  omega rel err {rel:.3e}, u max err {np.abs(ug - u).max():.3e}, "
          f"particle max err {np.abs(pg[:, :3] - p[:, :3]).max():.3e}", flush=True)
    assert rel <= tol_w
    assert np.abs(ug - u).max() <= 1e-4 * np.abs(u).max()
    assert np.abs(pg[:, :3] - p[:, :3]).max() <= 1e-4
    assert np.array_equal(pg[:, 3], p[:, 3])
    # the forcing acted: the trajectory differs from the unforced one
    w_free = sd.omega.copy()
    for f in range(frames):
        w_free = O.reduced_step(w_free, sd.C, sd.lam, force, dt, 0.0, False)
    assert np.linalg.norm(w - w_free) > 1e3 * tol_w * np.linalg.norm(w)
