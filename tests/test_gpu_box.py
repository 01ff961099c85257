"""Closed-box basis and C built on the device from the mode table, against the
host constructions (basis.analytic_box_modes, advection.build_advection_box_*),
against the oracle's literal buildAdvection, and through whole frames against
scenes built from host uploads."""
import sys

import numpy as np
import pytest

import oracle as O
from parallelspectralfluids3d_amd import advection, basis
from parallelspectralfluids3d_amd.grid import Grid

pytestmark = pytest.mark.gpu

C_FP64, C_FP32, C_BF16X2 = 0, 1, 4
PHI_FP64, PHI_FP32 = 0, 1


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _desc(g, K, c_prec, phi_prec, N=0):
    from parallelspectralfluids3d_amd import hip
    n = max(g.nx, g.ny, g.nz)
    half = [0.5 * d / n for d in (g.nx, g.ny, g.nz)]
    return hip.make_desc(g, 1.0 / n, K, c_prec, phi_prec, N, tuple(-x for x in half), tuple(half))


@pytest.mark.parametrize("dims,K", [((8, 8, 8), 24), ((6, 5, 7), 20)])
def test_device_basis_equals_host(hip_ctx, dims, K):
    _say(f"device box basis {dims} K={K}")
    from parallelspectralfluids3d_amd import hip
    g = Grid(*dims)
    Phi, lam, tab = basis.analytic_box_modes(g, K)
    for phi_prec in (PHI_FP64, PHI_FP32):
        s = hip.Scene(hip_ctx, _desc(g, K, C_FP64, phi_prec))
        s.generate_box_basis(tab)
        s.upload_advection(advection.build_advection_box_analytic(g, 1.0 / max(dims), Phi, lam, tab))
        for k in (0, K // 2, K - 1):      # u = Phi e_k reads column k back
            e = np.zeros(K); e[k] = 1.0
            s.set_omega(e)
            s.reconstruct()
            u = s.get_velocity()
            ref = Phi[k].astype(np.float32) if phi_prec == PHI_FP32 else Phi[k]
            err = np.abs(u.astype(np.float64) - ref.astype(np.float32).astype(np.float64)).max()
            assert err <= 2 * np.spacing(np.float32(np.abs(ref).max())), (phi_prec, k, err)
        s.close()


@pytest.mark.parametrize("dims,K", [((10, 9, 8), 40), ((8, 8, 8), 32)])
def test_device_box_advection_equals_oracle(hip_ctx, dims, K):
    """FP64 C from psf_scene_build_box_advection (separable 1-D sums) against
    the oracle's literal voxel loop (psfsolver.cpp:431-551) on the same modes."""
    _say(f"device box C vs oracle {dims} K={K}")
    from parallelspectralfluids3d_amd import hip
    g = Grid(*dims)
    n = max(dims)
    h = 1.0 / n
    Phi, lam, tab = basis.analytic_box_modes(g, K)
    s = hip.Scene(hip_ctx, _desc(g, K, C_FP64, PHI_FP64))
    s.generate_box_basis(tab)
    s.build_box_advection(tab)
    C = s.get_advection()
    s.close()
    V = (basis.curl_operator(g, h) @ Phi.T).T
    amin = tuple(-0.5 * d / n for d in dims)
    m = O.Mesh(dims, h, amin, inside=np.ones(g.nV, bool))
    Co = O.build_advection(m, Phi, lam, V)
    assert np.max(np.abs(C - Co)) <= 1e-12 * np.abs(Co).max()


@pytest.mark.parametrize("c_prec,tol", [(C_FP64, 1e-13), (C_FP32, 1e-6), (C_BF16X2, 1e-7)])
def test_device_advection_steps_like_host_upload(hip_ctx, c_prec, tol):
    """200 reduced steps with C built on the device and with C uploaded from
    the host separable construction agree to the precision's rounding."""
    _say(f"device box C, c_prec={c_prec}")
    from parallelspectralfluids3d_amd import hip
    g = Grid(10, 9, 8)
    K = 40
    h = 1.0 / 10
    Phi, lam, tab = basis.analytic_box_modes(g, K)
    C = advection.build_advection_box_separable(g, h, lam, tab)
    w0 = basis.unit_energy_omega(K, 3)
    out = []
    for device in (True, False):
        s = hip.Scene(hip_ctx, _desc(g, K, c_prec, PHI_FP64))
        if device:
            s.generate_box_basis(tab)
            s.build_box_advection(tab)
        else:
            s.upload_basis(Phi, lam)
            s.upload_advection(C)
        s.set_omega(w0)
        s.set_params(0.01)
        s.step(200)
        out.append(s.get_omega())
        s.close()
    assert np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1]) <= tol


def test_device_built_scene_frames(hip_ctx):
    """Whole frames (window 8, BF16 hi+lo, FP32 Phi, particles) of a scene built on
    the device against the same scene built from host uploads."""
    _say("device-built scene frames")
    from parallelspectralfluids3d_amd import hip
    g = Grid(12, 12, 12)
    K, N = 48, 4096
    h = 1.0 / 12
    Phi, lam, tab = basis.analytic_box_modes(g, K)
    C = advection.build_advection_box_separable(g, h, lam, tab)
    rng = np.random.default_rng(5)
    p0 = np.zeros((N, 4), np.float32)
    p0[:, :3] = rng.random((N, 3)) - 0.5
    p0[:, 3] = 60.0
    w0 = basis.power_law_omega(K, 1)
    res = []
    for device in (True, False):
        s = hip.Scene(hip_ctx, _desc(g, K, C_BF16X2, PHI_FP32, N))
        if device:
            s.generate_box_basis(tab)
            s.build_box_advection(tab)
        else:
            s.upload_basis(Phi, lam)
            s.upload_advection(C)
        s.set_omega(w0)
        s.set_params(0.01, 0.0, False, lifetime=60.0)
        s.set_particles(p0)
        s.set_frame_window(8)
        s.integrate(37)
        res.append((s.get_omega(), s.get_velocity(), s.get_particles()))
        s.close()
    (w1, u1, p1), (w2, u2, p2) = res
    assert np.linalg.norm(w1 - w2) / np.linalg.norm(w2) <= 1e-7
    assert np.abs(u1 - u2).max() <= 1e-5 * np.abs(u2).max()
    assert np.abs(p1[:, :3] - p2[:, :3]).max() <= 1e-5
