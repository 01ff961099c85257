"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs (reference: Solver/SolverImpl/psfsolver.cpp:21-234)."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _scene(ctx, sd, c_prec=0, phi_prec=0, max_particles=0):
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, c_prec, phi_prec, max_particles, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    return s


@pytest.mark.parametrize("case", [dict(n=8, K=16), dict(n=16, K=32, basis_kind="eigen"),
                                  dict(dims=(6, 5, 7), K=20)])
def test_step_fp64_matches_oracle(hip_ctx, case):
    """1000 reduced steps at dt=0.01; omega every 100 steps within 1e-12 relative
    (a 1-ulp perturbation of omega_0 grows to ~4e-15 over the same run)."""
    _say(f"step fp64 {case}")
    sd = make_scene(**case)
    s = _scene(hip_ctx, sd)
    s.set_params(0.01, 0.0, False)
    w = sd.omega.copy()
    for blk in range(10):
        for _ in range(100):
            w = O.reduced_step(w, sd.C, sd.lam, sd.g, 0.01, 0.0, False)
        s.step(100)
        wg = s.get_omega()
        rel = np.linalg.norm(wg - w) / np.linalg.norm(w)
        assert rel <= 1e-12, (blk, rel)


def test_step_fp64_viscosity_forcing(hip_ctx):
    """viscosity and forcing on: same order of operations as psfsolver.cpp:45-59."""
    _say("step fp64 viscosity+forcing")
    sd = make_scene(n=8, K=16)
    s = _scene(hip_ctx, sd)
    s.set_params(0.005, 0.3, True)
    w = sd.omega.copy()
    for _ in range(200):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, 0.005, 0.3, True)
    s.step(200)
    wg = s.get_omega()
    assert np.linalg.norm(wg - w) / np.linalg.norm(w) <= 1e-12


def test_step_fp32_close_to_fp64(hip_ctx):
    _say("step fp32")
    sd = make_scene(n=8, K=16)
    s = _scene(hip_ctx, sd, c_prec=1)
    s.set_params(0.01, 0.0, False)
    w = sd.omega.copy()
    for _ in range(1000):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, 0.01, 0.0, False)
    s.step(1000)
    rel = np.linalg.norm(s.get_omega() - w) / np.linalg.norm(w)
    assert rel <= 1e-5, rel


@pytest.mark.parametrize("phi_prec", [0, 1])
def test_reconstruct_matches_oracle(hip_ctx, phi_prec):
    """u = float(Phi w): FP64 Phi within 1 float ulp of the oracle; FP32 Phi within 2e-6 of max|u|."""
    _say(f"reconstruct phi_prec={phi_prec}")
    sd = make_scene(n=16, K=32, basis_kind="eigen")
    s = _scene(hip_ctx, sd, phi_prec=phi_prec)
    s.reconstruct()
    u = s.get_velocity()
    ref = O.reconstruct(sd.Phi, sd.omega)
    if phi_prec == 0:
        ulp = np.spacing(np.abs(ref).astype(np.float32))
        assert np.all(np.abs(u - ref) <= ulp)
    else:
        assert np.abs(u - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("case", [dict(n=8, K=16), dict(dims=(6, 5, 7), K=20)])
def test_trace_bit_exact(hip_ctx, case):
    """Given the same u, particle positions and cell lookups match the oracle bit for bit."""
    _say(f"trace {case}")
    sd = make_scene(**case)
    N = 4096
    s = _scene(hip_ctx, sd, max_particles=N)
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    p0 = particles_uniform(sd, N, seed=3)
    p0[::7, 3] = -1.0                     # some particles take the respawn branch
    s.set_particles(p0)
    s.reconstruct()
    u = s.get_velocity()
    cells = s.trace(cells=True)
    pg = s.get_particles()
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    pr, cr = O.particles(m, u, p0, resolution=sd.resolution, timestep=0.01, aabb_min=sd.aabb_min,
                         aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=0, cells=True)
    assert np.array_equal(cells, cr)
    assert np.array_equal(pg.view(np.uint32), pr.view(np.uint32))


def test_integrate_frames_match_oracle(hip_ctx):
    """50 integrate() frames (step, u = Phi w, particles) against the oracle run frame by frame."""
    _say("integrate frames")
    sd = make_scene(n=8, K=16)
    N = 2048
    s = _scene(hip_ctx, sd, max_particles=N)
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    p = particles_uniform(sd, N, seed=5)
    s.set_particles(p)
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    w = sd.omega.copy()
    for f in range(50):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, 0.01, 0.0, False)
        u = O.reconstruct(sd.Phi, w)
        p = O.particles(m, u, p, resolution=sd.resolution, timestep=0.01, aabb_min=sd.aabb_min,
                        aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=f)
    s.integrate(50)
    wg, ug, pg = s.get_omega(), s.get_velocity(), s.get_particles()
    assert np.linalg.norm(wg - w) / np.linalg.norm(w) <= 1e-12
    assert np.abs(ug - u).max() <= 2 * np.spacing(np.float32(np.abs(u).max()))
    assert np.abs(pg[:, :3] - p[:, :3]).max() <= 1e-6


def test_check_signs_and_errors(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    _say("signs and error paths")
    sd = make_scene(n=8, K=16)
    s = _scene(hip_ctx, sd)
    m = O.Mesh((8, 8, 8), sd.resolution, sd.aabb_min)
    s.check_signs(m.sign_words())
    bad = m.sign_words().copy()
    bad[3] ^= 1
    with pytest.raises(L.PsfError) as e:
        s.check_signs(bad)
    assert e.value.status == L.PSF_ERR_ARG
    # a NaN in omega is caught by the FP64 energy check, for this scene only
    w = sd.omega.copy()
    w[2] = np.nan
    s.set_omega(w)
    s.step(1)
    with pytest.raises(L.PsfError) as e:
        s.get_omega()
    assert e.value.status == L.PSF_ERR_NONFINITE
    # the other scene on the context is unaffected
    s2 = _scene(hip_ctx, sd)
    s2.set_params(0.01)
    s2.step(3)
    assert np.all(np.isfinite(s2.get_omega()))


@pytest.mark.parametrize("case", [dict(n=8, K=16), dict(n=12, K=16)])
def test_trace_on_voxel_planes(hip_ctx, case):
    """Particles on voxel planes, half-cell planes and the AABB walls, and a few
    float ulps either side: the cell index math (double quotient truncated,
    psfsolver.cpp:80-100) and the clamped neighbour columns match the oracle
    bit for bit.  n=12 makes h = 1/12 inexact in binary."""
    _say(f"trace on voxel planes {case}")
    sd = make_scene(**case)
    n = sd.grid.nx
    h = sd.resolution
    rng = np.random.default_rng(7)
    planes = np.concatenate([np.arange(n + 1) * h, (np.arange(n) + 0.5) * h]) + sd.aabb_min[0]
    N = 4096
    p0 = np.zeros((N, 4), np.float32)
    for a in range(3):
        c = rng.choice(planes, N).astype(np.float32)
        steps = rng.integers(-3, 4, N)
        for j in range(N):
            for _ in range(abs(int(steps[j]))):
                c[j] = np.nextafter(c[j], np.float32(np.sign(steps[j]) * 10.0))
        p0[:, a] = np.clip(c, sd.aabb_min[a], sd.aabb_max[a])
    p0[:, 3] = 60.0
    s = _scene(hip_ctx, sd, max_particles=N)
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    s.set_particles(p0)
    s.reconstruct()
    u = s.get_velocity()
    cells = s.trace(cells=True)
    pg = s.get_particles()
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    pr, cr = O.particles(m, u, p0, resolution=sd.resolution, timestep=0.01, aabb_min=sd.aabb_min,
                         aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=0, cells=True)
    assert np.array_equal(cells, cr)
    assert np.array_equal(pg.view(np.uint32), pr.view(np.uint32))
