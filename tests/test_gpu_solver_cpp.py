"""The C++ drop-in (include/psf/solver.hpp: HipSolver, a SolverImpl in the
place of PSFSolverGPU) driven frame by frame like GLCanvas::simulate drives
integrate() (Widgets/glcanvas.cpp:143-156), checked against the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_build", "solver_check")   # built by __graft_entry__.build()


@pytest.mark.parametrize("cprec,phiprec,tol_w", [(0, 0, 1e-12), (4, 1, 2e-6)])
def test_hip_solver_frames(tmp_path, cprec, phiprec, tol_w):
    print(f"[gpu-case] solver_check cprec={cprec} phiprec={phiprec}", flush=True)
    sd = make_scene(n=8, K=16)
    N, frames, dt = 1024, 20, 0.01
    p0 = particles_uniform(sd, N, seed=11)
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as fh:
        fh.write(f"8 8 8 {sd.resolution!r} {sd.K} {N} {frames} {dt!r} 0.0 0 60.0 {cprec} {phiprec}\n")
    np.ascontiguousarray(sd.Phi, np.float64).tofile(os.path.join(d, "phi.bin"))
    sd.lam.astype(np.float64).tofile(os.path.join(d, "lambda.bin"))
    sd.g.astype(np.float64).tofile(os.path.join(d, "gravity.bin"))
    np.ascontiguousarray(sd.C, np.float64).tofile(os.path.join(d, "C.bin"))
    sd.omega.astype(np.float64).tofile(os.path.join(d, "omega.bin"))
    p0.astype(np.float32).tofile(os.path.join(d, "parts.bin"))
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    w_gpu = np.fromfile(os.path.join(d, "omega_out.bin"), np.float64)
    u_gpu = np.fromfile(os.path.join(d, "u_out.bin"), np.float32)
    p_gpu = np.fromfile(os.path.join(d, "parts_out.bin"), np.float32).reshape(N, 4)
    m = O.Mesh((8, 8, 8), sd.resolution, sd.aabb_min, inside=sd.inside)
    w, p = sd.omega.copy(), p0.copy()
    for f in range(frames):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, dt, 0.0, False)
        u = O.reconstruct(sd.Phi, w)
        p = O.particles(m, u, p, resolution=sd.resolution, timestep=dt, aabb_min=sd.aabb_min,
                        aabb_max=sd.aabb_max, lifetime=60.0, seed=0, step=f)
    assert np.linalg.norm(w_gpu - w) / np.linalg.norm(w) <= tol_w
    assert np.abs(u_gpu - u).max() <= 1e-5 * np.abs(u).max()
    assert np.abs(p_gpu[:, :3] - p[:, :3]).max() <= 1e-5


def test_hip_solver_density_and_benchmark(tmp_path):
    """The density volume of HipSolver's particles equals the oracle's; the
    reference's benchmark.csv is written with per-frame GPU times."""
    print("[gpu-case] solver_check extras", flush=True)
    sd = make_scene(n=8, K=16)
    N, frames, dt = 1024, 10, 0.01
    p0 = particles_uniform(sd, N, seed=12)
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as fh:
        fh.write(f"8 8 8 {sd.resolution!r} {sd.K} {N} {frames} {dt!r} 0.0 0 60.0 4 1\n")
    np.ascontiguousarray(sd.Phi, np.float64).tofile(os.path.join(d, "phi.bin"))
    sd.lam.astype(np.float64).tofile(os.path.join(d, "lambda.bin"))
    sd.g.astype(np.float64).tofile(os.path.join(d, "gravity.bin"))
    np.ascontiguousarray(sd.C, np.float64).tofile(os.path.join(d, "C.bin"))
    sd.omega.astype(np.float64).tofile(os.path.join(d, "omega.bin"))
    p0.astype(np.float32).tofile(os.path.join(d, "parts.bin"))
    r = subprocess.run([EXE, d, "extras"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p_gpu = np.fromfile(os.path.join(d, "parts_out.bin"), np.float32).reshape(N, 4)
    vol = np.fromfile(os.path.join(d, "volume_out.bin"), np.uint16).reshape(16, 16, 16)
    _, ref = O.density(p_gpu, sd.aabb_min, sd.aabb_max, (16, 16, 16))
    assert np.array_equal(vol, ref.view(np.uint16))
    rows = open(os.path.join(d, "benchmark.csv")).read().split("\n")
    assert rows[0] == str(N) and rows[1] == str(sd.K) and rows[2] == "8,8,8"
    assert len(rows[3].split(",")) == 3
    times = [r for r in rows[5:9]]
    assert all(len(t.split(",")) == 2 and min(float(x) for x in t.split(",")) > 0.0 for t in times)
    avg = [float(x) for x in rows[-1].split(",")]
    assert avg[0] > 0.0 and avg[1] > 0.0


def test_eigensolve_from_cpp():
    """psf_eigensolve_csr called from C++ (no Python in between) on the
    Dirichlet path Laplacian: eigenvalues to 1e-12, eigenvectors the sines."""
    print("[gpu-case] solver_check eigen", flush=True)
    r = subprocess.run([EXE, "eigen"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "eigen_check:" in r.stdout


def test_hip_solver_frame_window(tmp_path):
    """HipSolver with a frame window of 32, driven like the reference's canvas
    (setTimestep, then integrate(), once per frame), gives the bits of the
    same solver without a window and computes one window per 32 frames (plus
    the one computed ahead), dropping none."""
    print("[gpu-case] solver_check window 32", flush=True)
    sd = make_scene(n=8, K=16)
    N, frames, dt = 2048, 96, 0.01
    p0 = particles_uniform(sd, N, seed=13)
    outs = []
    for W in (1, 32):
        d = str(tmp_path / f"w{W}")
        os.makedirs(d)
        with open(os.path.join(d, "meta.txt"), "w") as fh:
            fh.write(f"8 8 8 {sd.resolution!r} {sd.K} {N} {frames} {dt!r} 0.0 0 60.0 4 1\n")
        np.ascontiguousarray(sd.Phi, np.float64).tofile(os.path.join(d, "phi.bin"))
        sd.lam.astype(np.float64).tofile(os.path.join(d, "lambda.bin"))
        sd.g.astype(np.float64).tofile(os.path.join(d, "gravity.bin"))
        np.ascontiguousarray(sd.C, np.float64).tofile(os.path.join(d, "C.bin"))
        sd.omega.astype(np.float64).tofile(os.path.join(d, "omega.bin"))
        p0.astype(np.float32).tofile(os.path.join(d, "parts.bin"))
        args = [EXE, d] + (["window", str(W)] if W > 1 else [])
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append((d, r.stdout))
    for name in ("omega_out.bin", "u_out.bin", "parts_out.bin"):
        a = open(os.path.join(outs[0][0], name), "rb").read()
        b = open(os.path.join(outs[1][0], name), "rb").read()
        assert a == b, name
    line = [ln for ln in outs[1][1].splitlines() if ln.startswith("counters ")][0]
    c = {k: int(v) for k, v in (kv.split("=") for kv in line.split()[1:])}
    assert c["frames"] == frames
    assert c["windows_dropped"] == 0
    assert c["window_fills"] == frames // 32 + 1
    assert c["step_launches"] == c["window_fills"] and c["recon_launches"] == c["window_fills"]


def test_live_context_at_exit():
    """A C++ caller that returns from main with a context, a scene and a
    running persistent launch still alive: the library's exit handler
    releases them before the HIP runtime's teardown, and the process exits
    with status 0 and nothing on stderr from the runtime's teardown."""
    print("[gpu-case] solver_check leak", flush=True)
    r = subprocess.run([EXE, "leak"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    assert "leak_check:" in r.stdout
    assert "terminate" not in r.stderr and "Segmentation" not in r.stderr, r.stderr
