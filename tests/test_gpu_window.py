"""Frame window: W frames' reduced steps in one launch and their
reconstructions in one pass over Phi, then one trace per integrate().  The
reduced dynamics (psfsolver.cpp:34-59) do not depend on the particles, so
omega, u and the particles must be bitwise those of W single frames
(psfsolver.cpp:21-234), whatever the calls in between."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu

C_FP64, C_FP32, C_BF16X2 = 0, 1, 4
PHI_FP64, PHI_FP32, PHI_FP32_ACC32 = 0, 1, 2


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _pair(ctx, sd, c_prec, phi_prec, N, window, lifetime=60.0, seed=3):
    from parallelspectralfluids3d_amd import hip
    out = []
    p = particles_uniform(sd, N, seed=seed, lifetime=10.0)
    for W in (1, window):
        d = hip.make_desc(sd.grid, sd.resolution, sd.K, c_prec, phi_prec, N, sd.aabb_min, sd.aabb_max)
        s = hip.Scene(ctx, d)
        s.upload_basis(sd.Phi, sd.lam, sd.g)
        s.upload_advection(sd.C)
        s.set_omega(sd.omega)
        s.set_params(0.01, 0.0, False, lifetime=lifetime)
        s.set_particles(p)
        s.set_frame_window(W)
        out.append(s)
    return out


def _same(a, b):
    wa, wb = a.get_omega(), b.get_omega()
    ua, ub = a.get_velocity(), b.get_velocity()
    pa, pb = a.get_particles(), b.get_particles()
    assert np.array_equal(wa.view(np.uint64), wb.view(np.uint64))
    assert np.array_equal(ua.view(np.uint32), ub.view(np.uint32))
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))


@pytest.mark.parametrize("window", [2, 8, 32])
@pytest.mark.parametrize("c_prec,phi_prec", [(C_FP64, PHI_FP64), (C_FP64, PHI_FP32), (C_BF16X2, PHI_FP32),
                                             (C_BF16X2, PHI_FP32_ACC32)])
def test_window_bitwise_equal_single_frames(hip_ctx, window, c_prec, phi_prec):
    """Read back mid-window and at window ends: omega, u and particles equal."""
    _say(f"window={window} c={c_prec} phi={phi_prec}")
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, c_prec, phi_prec, 1024, window)
    for n in (1, window - 1 if window > 2 else 1, window + 3, 2 * window):
        one.integrate(n)
        win.integrate(n)
        _same(one, win)


def test_window_state_changes_mid_window(hip_ctx):
    """set_params, psf_step, psf_reconstruct, set_omega and a window change in
    the middle of a window give what single frames give."""
    _say("window state changes")
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, C_FP64, PHI_FP32, 512, 8)
    for s in (one, win):
        s.integrate(3)
        s.set_params(0.02, 0.5, True, lifetime=60.0)
        s.integrate(5)
        s.step(2)
        s.integrate(9)
        s.reconstruct()
        s.trace()
        s.integrate(4)
    _same(one, win)
    w = np.linspace(-1.0, 1.0, sd.K)
    for s in (one, win):
        s.integrate(2)
        s.set_omega(w)
        s.integrate(11)
    _same(one, win)
    win.set_frame_window(32)
    for s in (one, win):
        s.integrate(40)
    _same(one, win)
    win.set_frame_window(1)
    for s in (one, win):
        s.integrate(5)
    _same(one, win)


def test_deferred_particle_passes(hip_ctx):
    """Particle passes of windowed frames run several frames per launch; a
    new particle set, a density read, a context sync and short lifetimes
    (respawn draws follow the frame number) see every frame advanced so far."""
    _say("deferred particle passes")
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, C_FP64, PHI_FP32, 700, 16, lifetime=3.0)
    fresh = particles_uniform(sd, 700, seed=9, lifetime=2.0)
    for s in (one, win):
        s.integrate(3)
        s.set_particles(fresh)          # the 3 deferred passes ran on the old set first
        s.integrate(5)
    _same(one, win)
    for s in (one, win):
        s.set_density_grid((4, 4, 4), count_in_trace=True)
        s.integrate(6)
    c1, v1 = one.density()
    c2, v2 = win.density()
    assert np.array_equal(c1, c2) and np.array_equal(v1.view(np.uint16), v2.view(np.uint16))
    for s in (one, win):
        s.integrate(7)
    hip_ctx.sync()
    _same(one, win)
    for s in (one, win):
        s.integrate(21)
    _same(one, win)


def test_window_frames_match_oracle(hip_ctx):
    """40 frames with a window of 16 against the oracle run frame by frame."""
    _say("window vs oracle")
    from parallelspectralfluids3d_amd import hip
    sd = make_scene(n=8, K=16)
    N = 1024
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, C_FP64, PHI_FP32, N, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(hip_ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    p = particles_uniform(sd, N, seed=9)
    s.set_particles(p)
    s.set_frame_window(16)
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    w = sd.omega.copy()
    for f in range(40):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, 0.01, 0.0, False)
        u = O.reconstruct(sd.Phi, w)
        p = O.particles(m, u, p, resolution=sd.resolution, timestep=0.01, aabb_min=sd.aabb_min,
                        aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=f)
    s.integrate(40)
    wg, ug, pg = s.get_omega(), s.get_velocity(), s.get_particles()
    assert np.linalg.norm(wg - w) / np.linalg.norm(w) <= 1e-12
    assert np.abs(ug - u).max() <= 2 * np.spacing(np.float32(np.abs(u).max()))
    assert np.abs(pg[:, :3] - p[:, :3]).max() <= 1e-6


def test_window_rejects_bad_sizes(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, C_FP64, PHI_FP32, 64, 4)
    for bad in (0, 3, 64):
        with pytest.raises(L.PsfError) as e:
            win.set_frame_window(bad)
        assert e.value.status == L.PSF_ERR_ARG


@pytest.mark.parametrize("window", [1, 8])
def test_new_particles_after_frames_match_oracle(hip_ctx, window):
    """integrate(n), a fresh particle set, integrate(m): the upload lands
    after the passes already queued on the scene's stream, so the fresh set is
    advanced from its own positions (the oracle advances it from frame n)."""
    _say(f"set_particles after frames, window={window}")
    from parallelspectralfluids3d_amd import hip
    sd = make_scene(n=8, K=16)
    N, n1, n2, dt = 1 << 18, 9, 6, 0.01
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, C_FP64, PHI_FP64, N, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(hip_ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    s.set_params(dt, 0.0, False, lifetime=60.0)
    s.set_particles(particles_uniform(sd, N, seed=21))
    s.set_frame_window(window)
    s.integrate(n1)
    fresh = particles_uniform(sd, N, seed=22)
    s.set_particles(fresh)
    s.integrate(n2)
    pg = s.get_particles()
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    w, p = sd.omega.copy(), fresh.copy()
    for f in range(n1 + n2):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, dt, 0.0, False)
        if f >= n1:
            u = O.reconstruct(sd.Phi, w)
            p = O.particles(m, u, p, resolution=sd.resolution, timestep=dt, aabb_min=sd.aabb_min,
                            aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=f)
    assert np.abs(pg[:, :3] - p[:, :3]).max() <= 1e-6


def test_bounded_frame_phase(hip_ctx):
    """The bench's frame phase (psf_time_phase phase 3) starts at a window
    boundary and computes exactly the frames it times: with a window of 32,
    20 timed frames fill 20 frames (none ahead), and the results are the bits
    of 20 single frames; a caller's frames afterwards continue as usual.
    Phase 11 is the same region with each launch timed by events of its own."""
    _say("bounded frame phase")
    from parallelspectralfluids3d_amd import _lib as L
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, C_FP64, PHI_FP32, 1024, 32)
    win.integrate(5)                      # 5 frames consumed, a window and the next computed ahead
    one.integrate(5)
    c0 = win.counters()
    win.time_phase(L.PHASE_FRAME, 20)
    one.integrate(20)
    _same(one, win)
    rk = win.region_kernels()             # phase 3 times no launch of its own
    assert rk["recon_launches"] == rk["particle_launches"] == rk["step_launches"] == 0
    c1 = win.counters()
    assert c1["windows_dropped"] - c0["windows_dropped"] == 1       # the frames ahead of frame 5
    assert c1["window_frames"] - c0["window_frames"] == 20          # exactly the timed frames
    assert c1["frames"] - c0["frames"] == 20
    # the same region with its launches timed by their own dispatch events:
    # one window reconstruction of 20 fields, particle passes of 8 + 8 + 4
    # frames (the plain path's reduced steps are not timed), same bits
    region_ms = win.time_phase(L.PHASE_FRAME_TIMED, 20)
    one.integrate(20)
    _same(one, win)
    rk = win.region_kernels()
    assert (rk["step_launches"], rk["steps"]) == (0, 0)
    assert (rk["recon_launches"], rk["recon_frames"]) == (1, 20)
    assert (rk["particle_launches"], rk["particle_frames"]) == (3, 20)
    assert rk["recon_ms"] > 0 and rk["particle_ms"] > 0
    assert rk["recon_ms"] + rk["particle_ms"] <= region_ms * 1.001
    c1 = win.counters()
    assert c1["window_frames"] - c0["window_frames"] == 40
    assert c1["frames"] - c0["frames"] == 40
    for s in (one, win):
        s.integrate(45)
    _same(one, win)
    c2 = win.counters()
    assert c2["windows_dropped"] == c1["windows_dropped"]
    # ordinary frames: one window of 32 at the first frame and the next one ahead
    assert c2["window_fills"] - c1["window_fills"] == 3
    # on the persistent path the region's one launch of 20 reduced steps is timed too
    from parallelspectralfluids3d_amd import hip
    p = hip.Scene(hip_ctx, hip.make_desc(sd.grid, sd.resolution, 16, C_BF16X2, PHI_FP32, 1024, sd.aabb_min, sd.aabb_max))
    p.upload_basis(sd.Phi, sd.lam, sd.g)
    p.upload_advection(sd.C)
    p.set_omega(sd.omega)
    p.set_params(0.01, 0.0, False, lifetime=60.0)
    p.set_particles(particles_uniform(sd, 1024, seed=9, lifetime=60.0))
    p.set_frame_window(32)
    region_ms = p.time_phase(L.PHASE_FRAME_TIMED, 20)
    rk = p.region_kernels()
    assert (rk["step_launches"], rk["steps"]) == (1, 20)
    assert (rk["recon_launches"], rk["recon_frames"]) == (1, 20)
    assert (rk["particle_launches"], rk["particle_frames"]) == (3, 20)
    assert min(rk["step_ms"], rk["recon_ms"], rk["particle_ms"]) > 0
    assert rk["step_ms"] + rk["recon_ms"] + rk["particle_ms"] <= region_ms * 1.001
    p.close()


def test_unchanged_params_keep_the_window(hip_ctx):
    """setTimestep & co. with the values already set (the reference's
    integrate() reads them every frame) do not drop the frames computed
    ahead; a changed value does."""
    _say("unchanged params keep the window")
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, C_FP64, PHI_FP32, 512, 16)
    for f in range(40):
        for s in (one, win):
            s.set_params(0.01, 0.0, False, lifetime=60.0)
            s.integrate(1)
    _same(one, win)
    c = win.counters()
    assert c["windows_dropped"] == 0 and c["window_fills"] == 4      # frames 0, 16, 32 (+1 ahead)
    for s in (one, win):
        s.set_params(0.02, 0.0, False, lifetime=60.0)
        s.integrate(3)
    _same(one, win)
    assert win.counters()["windows_dropped"] == 1


def test_device_view(hip_ctx):
    """psf_scene_device_view: the addresses of the resident buffers; u follows
    the current frame of a window; packed-precision scenes expose no C."""
    _say("device view")
    sd = make_scene(n=8, K=16)
    one, win = _pair(hip_ctx, sd, C_FP64, PHI_FP32, 256, 8)
    v = one.device_view()
    assert all(v[k] for k in ("omega", "lam", "gravity", "visc", "C", "Phi", "u", "particles", "pid", "status"))
    assert v["n_particles"] == 256 and v["ld_phi"] % 256 == 0 and v["ld_phi"] >= sd.grid.nF
    win.integrate(3)
    u3 = win.device_view()["u"]
    win.integrate(1)
    assert win.device_view()["u"] - u3 == 4 * win.device_view()["ld_phi"]    # the next row of the window
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, C_BF16X2, PHI_FP32, 0, sd.aabb_min, sd.aabb_max)
    b = hip.Scene(hip_ctx, d)
    assert b.device_view()["C"] is None


@pytest.mark.parametrize("window", [1, 16, 32])
def test_fp32_accumulation_within_bound(hip_ctx, window):
    """PSF_PHI_FP32_ACC32 (FP32 Phi and w, FP32 sums on the matrix cores)
    against the oracle's FP64 u = float(Phi w) (psfsolver.cpp:60), every face
    within the a-priori bound of the three roundings -- Phi and w to FP32
    (2^-24 each) and K FP32 additions (K 2^-24) -- times sum_k |Phi_fk w_k|,
    plus the float output's half ulp; particles after 40 frames within 1e-6."""
    _say(f"fp32 accumulation bound, window={window}")
    from parallelspectralfluids3d_amd import hip
    sd = make_scene(n=8, K=16)
    N = 2048
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, C_FP64, PHI_FP32_ACC32, N, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(hip_ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    p = particles_uniform(sd, N, seed=31)
    s.set_particles(p)
    s.set_frame_window(window)
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    w = sd.omega.copy()
    worst = 0.0
    for f in range(40):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, 0.01, 0.0, False)
        u = O.reconstruct(sd.Phi, w)
        p = O.particles(m, u, p, resolution=sd.resolution, timestep=0.01, aabb_min=sd.aabb_min,
                        aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=f)
        s.integrate(1)
        ug = s.get_velocity().astype(np.float64)
        mag = np.abs(sd.Phi * w[:, None]).sum(axis=0)
        bound = (sd.K + 2) * 2.0 ** -24 * mag + 2.0 ** -24 * np.abs(u.astype(np.float64)) + 1e-300
        ratio = np.abs(ug - u) / bound
        worst = max(worst, float(ratio.max()))
        assert np.all(ratio <= 1.0), (f, float(ratio.max()))
    print(f"  worst error / bound {worst:.3f}", flush=True)
    assert np.abs(s.get_particles()[:, :3] - p[:, :3]).max() <= 1e-6


def test_fp32_accumulation_exact_columns(hip_ctx):
    """u = Phi e_k in FP32 accumulation is column k of the FP32 Phi exactly."""
    _say("fp32 accumulation columns")
    from parallelspectralfluids3d_amd import hip
    sd = make_scene(n=8, K=16)
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, C_FP64, PHI_FP32_ACC32, 0, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(hip_ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    for k in (0, 7, sd.K - 1):
        e = np.zeros(sd.K)
        e[k] = 1.0
        s.set_omega(e)
        s.reconstruct()
        assert np.array_equal(s.get_velocity(), sd.Phi[k].astype(np.float32))
