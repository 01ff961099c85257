#!/usr/bin/env python3
"""Benchmark of the reduced-space fluid step on one MI355X.

Workload: the 64^3 closed box of the precision configuration -- the K = 256
lowest eigenpairs of the box's DEC 1-form Laplacian M as the reference
builds it (psfsolver.cpp:236-342: wall edges weighted, eigenvalues above
1e-10), solved on the GPU, and C built from them on the GPU; omega ~ N(0,1)
from seed 0 at unit energy, dt = 0.01, no viscosity or forcing, 2^20
particles uniform in the box.  (--basis analytic: the analytic sin/cos box
modes instead, eigenvectors of the free-slip Laplacian.)  A "step" is one
AbstractSolver::integrate() frame (psfsolver.cpp:21-234): reduced step
(BF16 hi+lo MFMA contraction, FP64 update), u = Phi w (FP32 Phi, FP32 sums)
and one particle pass.  With a frame window W (default 32) the reduced
steps of W frames run in one persistent launch and their W velocity fields
in one pass over Phi; every frame still gets its own omega, u and particle
pass, bitwise equal to single frames.

Printed on stdout: one JSON line.  value = particle-steps/s of the frame
loop; beside it the same frames one at a time, the reduced steps/s of one
persistent launch of 1000 steps and of the plain FP32 path, omega error
against the FP64 path every 100 steps, frame latency p50/p99, per-kernel
time per frame with the roofline of each kernel (the headline one is the
kernel with the largest share of a frame) and the CPU oracle timed on this
host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "reduced steps/s + particle-steps/s at 64³,K=256; p99 latency; err vs FP64"
HBM_PEAK = 8.0e12
FP64_PEAK = 78.6e12
FP32_PEAK = 157.3e12
# HBM bytes per launch at 64^3/K=256 with 2^20 particles, measured with
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE over `bench.py --counter-pass`
# (fetched KiB x 2, the factor a 1 GiB hipMemcpy reads back, plus written
# KiB); counters cannot be read from inside this process.
# (the reconstructions: FP32 Phi and w with FP32 sums, the PHI_PREC below)
RECON_TRAFFIC_64_256 = 825_237_504          # k_reconstruct_mfma<float, 1, float>: 401,388 KiB read, 3,120 KiB written
WINDOW_TRAFFIC_64_256_W32 = 927_565_824     # k_reconstruct_mfma<float, 2, float>: 402,993 / 99,840 KiB
WINDOW_TRAFFIC_64_256_W20 = 888_993_792     # the same kernel over the 20-frame timed region: 402,879 / 62,400 KiB
TRACE_TRAFFIC_64_1M = 37_605_376            # k_trace: 10,170 / 16,384 KiB
TRACE_FRAMES_TRAFFIC_64_1M = 69_654_528     # k_trace_frames (8 frames per launch): 25,819 / 16,384 KiB


# u = Phi w of the frame loop: FP32 Phi and w with FP32 sums on the FP32 matrix
# cores (FP64 is kept for the renormalisation, Euler update and viscosity)
PHI_PREC = 2


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scene(n, K, N, seed=0, basis_kind="eigen", ctx=None):
    """The closed box's basis, C, forcing, omega and particles.  basis_kind
    "eigen": the K lowest eigenpairs of the reference's M (the device
    eigensolver, penalty-lifted) and C built on the device in an FP64 scene
    and read back; "analytic": the analytic sin/cos modes and their C on the
    host."""
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import advection, basis, hip
    from parallelspectralfluids3d_amd.grid import Grid, box_mask
    g = Grid(n, n, n)
    h = 1.0 / n
    info = None
    t = time.time()
    if basis_kind == "eigen":
        Phi, lam, info = basis.laplacian_eigenbasis_device(g, h, box_mask(g), K, ctx)
        log(f"basis: {K} eigenpairs of M on {g.nF} faces in {time.time() - t:.1f}s ({info['products']} products)")
        t = time.time()
        s64 = hip.Scene(ctx, hip.make_desc(g, h, K, L.C_FP64, L.PHI_FP64, 0))
        advection.build_advection_device(s64, g, h, box_mask(g), Phi, lam)
        C = s64.get_advection()
        s64.close()
        log(f"structure tensor C ({K}^3) on the device in {time.time() - t:.1f}s")
    else:
        Phi, lam, tab = basis.analytic_box_modes(g, K)
        V = (basis.curl_operator(g, h) @ Phi.T).T
        log(f"basis {g.nF} faces x {K} modes in {time.time() - t:.1f}s")
        t = time.time()
        C = advection.build_advection_box_analytic(g, h, Phi, lam, tab, V)
        log(f"structure tensor C ({K}^3) in {time.time() - t:.1f}s")
        del V
    gvec = basis.project(Phi, basis.gravity_field(g, box_mask(g)))
    omega = basis.unit_energy_omega(K, seed)
    rng = np.random.default_rng(seed + 1)
    parts = np.zeros((N, 4), np.float32)
    parts[:, :3] = rng.random((N, 3)) - 0.5
    parts[:, 3] = np.inf
    return dict(grid=g, h=h, Phi=Phi, lam=lam, C=C, g=gvec, omega=omega, parts=parts, basis=basis_kind,
                basis_info=info)


def make_hip_scene(ctx, sc, c_prec, phi_prec, with_phi, N):
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sc["grid"], sc["h"], len(sc["lam"]), c_prec, phi_prec, N)
    s = hip.Scene(ctx, d)
    s.upload_basis(sc["Phi"] if with_phi else None, sc["lam"], sc["g"])
    s.upload_advection(sc["C"])
    s.set_omega(sc["omega"])
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    if N:
        s.set_particles(sc["parts"][:N])
    return s


def oracle_precision_errors(sc, traj):
    """The FP64 oracle's reduced steps (psfsolver.cpp:34-59) from the bench's
    initial omega, every 100 steps for 1000 steps, and the relative omega
    error of each GPU trajectory against it (the checker of the precision
    section; run in the CPU leg)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    nthreads = min(16, os.cpu_count() or 1)
    K = sc["lam"].size
    w = sc["omega"].copy()
    ora = []
    for _ in range(10):
        for _ in range(100):
            w = O.reduced_step(w, sc["C"], sc["lam"], np.zeros(K), 0.01, 0.0, False, nthreads=nthreads)
        ora.append(w.copy())
    return {name: [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(t, ora)]
            for name, t in traj.items()}


def cpu_baseline(sc, budget_s=12.0):
    """The oracle (a C restatement of psfsolver.cpp:21-234, OpenMP as in the
    reference) on this host: whole integrate() frames until budget_s elapse."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    nthreads = min(16, os.cpu_count() or 1)
    g = sc["grid"]
    m = O.Mesh((g.nx, g.ny, g.nz), sc["h"], (-0.5, -0.5, -0.5), inside=np.ones(g.nV, bool))
    Phi = np.ascontiguousarray(sc["Phi"])
    w = sc["omega"].copy()
    p = sc["parts"].copy()
    frames, t_red, t_rec, t_par = 0, 0.0, 0.0, 0.0
    t_start = time.time()
    while frames < 2 or (time.time() - t_start < budget_s and frames < 200):
        t0 = time.perf_counter()
        w = O.reduced_step(w, sc["C"], sc["lam"], sc["g"], 0.01, 0.0, False, nthreads=nthreads)
        t1 = time.perf_counter()
        u = O.reconstruct(Phi, w, nthreads=nthreads)
        t2 = time.perf_counter()
        p = O.particles(m, u, p, resolution=sc["h"], timestep=0.01, aabb_min=(-.5, -.5, -.5),
                        aabb_max=(.5, .5, .5), lifetime=60.0, seed=0, step=frames, nthreads=nthreads)
        t3 = time.perf_counter()
        if frames > 0:                      # first frame is warm-up
            t_red += t1 - t0; t_rec += t2 - t1; t_par += t3 - t2
        frames += 1
    timed = frames - 1
    tot = t_red + t_rec + t_par
    N = p.shape[0]
    ref_state = dict(frames=frames, omega=w.copy(), parts=p.copy(), u=u.copy())   # the oracle after `frames` frames
    # one thread, a few frames of the same workload
    t1_frames, t1_tot = 3, 0.0
    for f in range(t1_frames):
        t0 = time.perf_counter()
        w = O.reduced_step(w, sc["C"], sc["lam"], sc["g"], 0.01, 0.0, False, nthreads=1)
        u = O.reconstruct(Phi, w, nthreads=1)
        p = O.particles(m, u, p, resolution=sc["h"], timestep=0.01, aabb_min=(-.5, -.5, -.5),
                        aabb_max=(.5, .5, .5), lifetime=60.0, seed=0, step=frames + f, nthreads=1)
        t1_tot += time.perf_counter() - t0
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return ref_state, {"value": N * timed / tot, "unit": "particle-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"{timed} integrate() frames of the same {g.nx}^3/K={len(sc['lam'])}/{N}-particle workload "
                      f"after 1 warm-up frame (FP64 C and Phi as in psfsolver.cpp)",
            "ms_per_frame": 1e3 * tot / timed, "ms_reduced": 1e3 * t_red / timed,
            "ms_reconstruct": 1e3 * t_rec / timed, "ms_particles": 1e3 * t_par / timed,
            # AbstractSolver's benchmark.csv averages (abstractsolver.cpp:292-294): t0 = psfsolver.cpp:32-67
            # (step with u = Phi w), t1 = the particle loop (psfsolver.cpp:70-231)
            "benchmark_csv_ms": {"t0": 1e3 * (t_red + t_rec) / timed, "t1": 1e3 * t_par / timed},
            "value_1thread": N * t1_frames / t1_tot, "ms_per_frame_1thread": 1e3 * t1_tot / t1_frames,
            "host_cpus": os.cpu_count(), "omp_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": cpu_model}


SERVE_GRIDS = (16, 24, 32, 48)
SERVE_K0 = 192
# modes per workgroup of the persistent layout, each mode at its best of
# {default, 2, 4, 8} measured on the 512-request stream (0 = per-request default)
SERVE_MPW_BATCHED = 0
SERVE_MPW_ALONE = 2
# velocity fields with FP32 sums (PSF_PHI_FP32_ACC32), as in the frame loop, in
# every serving run
SERVE_U_FP32_SUMS = True
# batched runs: the scenes of a grid read one register-resident copy of its
# C; one at a time: each request's own C (K modes), the fastest lone-scene form
SERVE_SHARED_BATCHED = True


def serving_workload(n_requests, rate, seed=0, particles=1 << 14):
    """The serving stream: per request a closed box from {16,24,32,48}^3, K
    uniform in [16, 192], a frame count uniform in [100, 2000], omega ~ N(0,1)
    at unit energy and `particles` uniform in the box; Poisson arrivals at
    `rate` per second.  All draws from one seeded generator."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    for j in range(n_requests):
        n = int(rng.choice(SERVE_GRIDS))
        K = int(rng.integers(16, SERVE_K0 + 1))
        frames = int(rng.integers(100, 2001))
        w = rng.standard_normal(K)
        w /= np.linalg.norm(w)
        p = np.empty((particles, 4), np.float32)
        p[:, :3] = rng.random((particles, 3)) - 0.5
        p[:, 3] = np.inf
        reqs.append(dict(arrival=float(t[j]), grid=n, K=K, frames=frames, omega=w, parts=p))
    return reqs


def serving_bases():
    """One K0 = 192 basis per grid: analytic box modes, their C and forcing."""
    from parallelspectralfluids3d_amd import advection, basis
    from parallelspectralfluids3d_amd.grid import Grid, box_mask
    out = {}
    for n in SERVE_GRIDS:
        g = Grid(n, n, n)
        h = 1.0 / n
        Phi, lam, tab = basis.analytic_box_modes(g, SERVE_K0)
        V = (basis.curl_operator(g, h) @ Phi.T).T
        C = advection.build_advection_box_analytic(g, h, Phi, lam, tab, V)
        gv = basis.project(Phi, basis.gravity_field(g, box_mask(g)))
        out[n] = dict(grid=g, h=h, Phi=Phi, lam=lam, C=C, g=gv)
    return out


def run_stream(ctx, bases, reqs, max_active, window=32, saturate=False, modes_per_wg=0, u_fp32_sums=False,
               shared_basis=False):
    """Feeds the requests to a server at their arrival times (all at once when
    saturate) and runs rounds until every request has finished.  Returns the
    throughput over the makespan and the latency of each request from its
    arrival."""
    from parallelspectralfluids3d_amd import hip
    sv = hip.Server(ctx, window=window, max_active=max_active, modes_per_wg=modes_per_wg, u_fp32_sums=u_fp32_sums,
                    shared_basis=shared_basis)
    bid = {n: sv.add_basis(b["grid"], b["h"], b["Phi"], b["lam"], b["g"], b["C"]) for n, b in bases.items()}
    pending = sorted(range(len(reqs)), key=lambda j: reqs[j]["arrival"])
    outstanding = {}
    lat = np.zeros(len(reqs))
    frames_done = 0
    nxt = 0
    enqueue_s = None
    if saturate:
        # capacity: the clock starts with every request in the queue (each
        # submit copies the request and sorts its particles on the host)
        te = time.perf_counter()
        for j in pending:
            r = reqs[j]
            rid = sv.submit(bid[r["grid"]], r["omega"], r["parts"], r["frames"])
            outstanding[rid] = (j, 0.0)
        nxt = len(pending)
        enqueue_s = time.perf_counter() - te
    t0 = time.perf_counter()
    while nxt < len(pending) or outstanding:
        now = time.perf_counter() - t0
        while nxt < len(pending) and (saturate or reqs[pending[nxt]]["arrival"] <= now):
            j = pending[nxt]
            r = reqs[j]
            rid = sv.submit(bid[r["grid"]], r["omega"], r["parts"], r["frames"])
            outstanding[rid] = (j, time.perf_counter() - t0)
            nxt += 1
        st = sv.stats()
        if st["active"] == 0 and st["queued"] == 0:
            if nxt < len(pending):
                time.sleep(max(0.0, reqs[pending[nxt]]["arrival"] - (time.perf_counter() - t0)))
            continue
        sv.round()
        for rid in sv.completed():
            j, t_sub = outstanding.pop(rid)
            _, _, ms = sv.result(rid)
            arrival = 0.0 if saturate else reqs[j]["arrival"]
            lat[j] = t_sub + ms * 1e-3 - arrival
            frames_done += reqs[j]["frames"]
    makespan = time.perf_counter() - t0
    st = sv.stats()
    sv.close()
    first = 0.0 if saturate else min(r["arrival"] for r in reqs)
    span = makespan - first
    return dict(makespan_s=makespan, enqueue_s=enqueue_s, scene_frames=frames_done, scene_frames_per_s=frames_done / span,
                particle_steps_per_s=frames_done * reqs[0]["parts"].shape[0] / span,
                latency_ms_p50=float(np.percentile(lat, 50) * 1e3), latency_ms_p99=float(np.percentile(lat, 99) * 1e3),
                rounds=int(st["rounds"]), busy_s=st["busy_ms"] * 1e-3, admit_s=st["admit_ms"] * 1e-3,
                retire_s=st["retire_ms"] * 1e-3, failed=int(st["failed"]),
                mem_peak_gb=st["mem_peak"] / 1e9)


def dropin_sequential(ctx, bases, reqs):
    """The requests one after another through the single-scene drop-in: a
    scene per request (upload of Phi's first K columns, C[:K,:K,:K], omega and
    particles), integrate() frame by frame with no frame window, read back.
    Returns scene-frames/s over the wall time of all of it."""
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import hip
    frames = 0
    t0 = time.perf_counter()
    for r in reqs:
        b = bases[r["grid"]]
        K = r["K"]
        d = hip.make_desc(b["grid"], b["h"], K, L.C_BF16X2, L.PHI_FP32_ACC32 if SERVE_U_FP32_SUMS else L.PHI_FP32,
                          r["parts"].shape[0])
        s = hip.Scene(ctx, d)
        s.upload_basis(b["Phi"][:K], b["lam"][:K], b["g"][:K])
        s.upload_advection(np.ascontiguousarray(b["C"][:K, :K, :K]))
        s.set_omega(r["omega"])
        s.set_params(0.01, 0.0, False, lifetime=60.0)
        s.set_particles(r["parts"])
        s.integrate(r["frames"])
        s.get_omega()
        s.get_particles()
        s.close()
        frames += r["frames"]
    return frames / (time.perf_counter() - t0)


def serving_bench(ctx, n_requests, rate):
    t = time.time()
    bases = serving_bases()
    reqs = serving_workload(n_requests, rate)
    log(f"serving: {len(bases)} bases (K0={SERVE_K0}) and {n_requests} requests in {time.time() - t:.1f}s")
    # Each mode at its best measured modes-per-workgroup setting: batched, the
    # per-request default (the most modes whose C blocks fit in registers:
    # fewer workgroups per scene, more scenes per persistent launch); one at a
    # time, two (more workgroups share a lone scene's contraction).
    batched = run_stream(ctx, bases, reqs, max_active=64, modes_per_wg=SERVE_MPW_BATCHED,
                         u_fp32_sums=SERVE_U_FP32_SUMS, shared_basis=SERVE_SHARED_BATCHED)
    log(f"serving batched: {batched['scene_frames_per_s']:.0f} scene-frames/s, "
        f"p50 {batched['latency_ms_p50']:.1f} ms, p99 {batched['latency_ms_p99']:.1f} ms")
    alone = run_stream(ctx, bases, reqs, max_active=1, modes_per_wg=SERVE_MPW_ALONE, u_fp32_sums=SERVE_U_FP32_SUMS)
    log(f"serving one at a time: {alone['scene_frames_per_s']:.0f} scene-frames/s, "
        f"p50 {alone['latency_ms_p50']:.1f} ms, p99 {alone['latency_ms_p99']:.1f} ms")
    sat = run_stream(ctx, bases, reqs, max_active=64, saturate=True, modes_per_wg=SERVE_MPW_BATCHED,
                     u_fp32_sums=SERVE_U_FP32_SUMS, shared_basis=SERVE_SHARED_BATCHED)
    log(f"serving batched, all queued at once: {sat['scene_frames_per_s']:.0f} scene-frames/s "
        f"(enqueue {sat['enqueue_s']:.3f} s before the clock)")
    sat_alone = run_stream(ctx, bases, reqs, max_active=1, saturate=True, modes_per_wg=SERVE_MPW_ALONE,
                           u_fp32_sums=SERVE_U_FP32_SUMS)
    log(f"serving one at a time, all queued at once: {sat_alone['scene_frames_per_s']:.0f} scene-frames/s")
    sat_own = run_stream(ctx, bases, reqs, max_active=64, saturate=True, modes_per_wg=SERVE_MPW_BATCHED,
                         u_fp32_sums=SERVE_U_FP32_SUMS, shared_basis=False)
    log(f"serving batched with per-request C, all queued at once: {sat_own['scene_frames_per_s']:.0f} scene-frames/s")
    nd = min(16, n_requests)
    dropin = dropin_sequential(ctx, bases, reqs[:nd])
    log(f"serving: first {nd} requests one after another through the drop-in, frame by frame: {dropin:.0f} "
        f"scene-frames/s")
    return {
        "workload": f"{n_requests} requests, Poisson arrivals at {rate:g}/s, grids {SERVE_GRIDS}^3, K in [16, 192], "
                    f"frames in [100, 2000], 2^14 particles each; one K0={SERVE_K0} basis per grid; BF16 hi+lo C, "
                    f"u with {'FP32' if SERVE_U_FP32_SUMS else 'FP64'} sums",
        "modes_per_wg": {"batched": SERVE_MPW_BATCHED or "per-request default", "one_at_a_time": SERVE_MPW_ALONE},
        "batched": batched,
        "one_at_a_time": alone,
        "batched_all_queued": sat,
        "one_at_a_time_all_queued": sat_alone,
        "batched_all_queued_per_request_c": sat_own,
        "shared_basis": {"batched": SERVE_SHARED_BATCHED, "one_at_a_time": False},
        "throughput_vs_one_at_a_time": batched["scene_frames_per_s"] / alone["scene_frames_per_s"],
        "capacity_vs_one_at_a_time": sat["scene_frames_per_s"] / sat_alone["scene_frames_per_s"],
        "capacity_note": "both with all 512 requests queued before the clock starts; one at a time is each request "
                         "on its own C (K modes) and the frame window, the fastest lone-scene form",
        "dropin_frame_by_frame": {"requests": nd, "scene_frames_per_s": dropin,
                                  "note": "the first requests of the stream, each through psf_integrate with no "
                                          "frame window, uploads included"},
        "capacity_vs_dropin_frame_by_frame": sat["scene_frames_per_s"] / dropin,
    }


def large_bench(args):
    """The large single scene: 128^3 closed box, K=512 analytic modes (Phi
    12.99 GB in FP32), power-law omega (|omega_k| ~ k^-5/6, random signs),
    500 frames at dt=0.005, 2^24 particles.  The oracle cannot run this size,
    so the checks are size-independent ones: omega against the FP64 plain
    path after all frames, u on sampled faces against the host FP64 Phi, one
    particle pass of a particle sample bit for bit against the oracle, and
    unit energy."""
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import basis, hip
    from parallelspectralfluids3d_amd.grid import Grid
    n, K, N, frames, dt = 128, 512, 1 << 24, 500, 0.005
    g = Grid(n, n, n)
    h = 1.0 / n
    tab = basis.box_mode_table_normalised(g, K)
    lam = tab[:, 0].copy()
    omega = basis.power_law_omega(K, 0)
    rng = np.random.default_rng(1)
    parts = np.empty((N, 4), np.float32)
    parts[:, :3] = rng.random((N, 3), dtype=np.float32) - np.float32(0.5)
    parts[:, 3] = np.inf
    ctx = hip.Context(0)
    t = time.time()
    d = hip.make_desc(g, h, K, L.C_BF16X2, L.PHI_FP32, N)
    s = hip.Scene(ctx, d)
    s.generate_box_basis(tab)
    s.build_box_advection(tab)
    s.set_omega(omega)
    s.set_params(dt, 0.0, False, lifetime=60.0)
    s.set_particles(parts)
    s.set_frame_window(32)
    t_build = time.time() - t
    log(f"large: Phi ({g.nF * K * 4 / 1e9:.2f} GB) and C built on the device, scene ready in {t_build:.1f}s")
    s.time_phase(L.PHASE_FRAME, 32)            # warm-up window
    s.set_omega(omega)
    s.set_particles(parts)
    ms = s.time_phase(L.PHASE_FRAME, frames)
    log(f"large: {frames} frames in {ms:.1f} ms")
    ms_tr = s.time_phase(L.PHASE_TRACE, 5) / 5
    ms_w = s.time_phase(L.PHASE_WINDOW_RECON, 2) / 2
    w_gpu = s.get_omega()
    u_gpu = s.get_velocity()
    # omega against the FP64 plain path over the same frames
    ref = hip.Scene(ctx, hip.make_desc(g, h, K, L.C_FP64, L.PHI_FP32, 0))
    ref.upload_basis(None, lam)
    ref.build_box_advection(tab)
    ref.set_omega(omega)
    ref.set_params(dt, 0.0, False, lifetime=60.0)
    ref.step(frames)
    w_ref = ref.get_omega()
    ref.close()
    rel_w = float(np.linalg.norm(w_gpu - w_ref) / np.linalg.norm(w_ref))
    # u on sampled faces against the FP64 sum over the host's FP32-rounded Phi
    fs = rng.choice(g.nF, 4096, replace=False)
    u_host = (basis.box_mode_values(g, tab, fs).astype(np.float32).astype(np.float64).T @ w_gpu).astype(np.float32)
    u_err = float(np.max(np.abs(u_gpu[fs] - u_host)) / np.max(np.abs(u_host)))
    # one particle pass of a sample against the oracle, with the GPU's u
    p_now = s.get_particles()
    sample = p_now[rng.choice(N, 4096, replace=False)].copy()
    s.set_particles(sample)
    s.trace()
    p_gpu = s.get_particles()
    s.close()
    ctx.close()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    m = O.Mesh((n, n, n), h, (-0.5, -0.5, -0.5), inside=np.ones(g.nV, bool))
    p_ora = O.particles(m, u_gpu, sample, resolution=h, timestep=dt, aabb_min=(-.5, -.5, -.5),
                        aabb_max=(.5, .5, .5), lifetime=60.0, seed=0, step=frames + 32)
    trace_exact = bool(np.array_equal(p_gpu.view(np.uint32), p_ora.view(np.uint32)))
    out = {"large": {
        "workload": f"{n}^3 closed box, K={K} analytic modes, power-law omega, {frames} frames at dt={dt}, "
                    f"{N} particles, frame window 32, BF16 hi+lo C, FP32 Phi",
        "phi_gb_fp32": g.nF * K * 4 / 1e9, "device_build_s": t_build,
        "ms_per_frame": ms / frames, "particle_steps_per_s": N * frames / (ms * 1e-3),
        "trace_ms": ms_tr, "window_reconstruct_ms": ms_w,
        "window_reconstruct_tflops": 2.0 * g.nF * K * 32 / (ms_w * 1e-3) / 1e12,
        "omega_rel_err_vs_fp64_path": rel_w, "omega_norm": float(np.linalg.norm(w_gpu)),
        "u_sample_max_rel_err": u_err, "trace_sample_bit_exact_vs_oracle": trace_exact}}
    print(json.dumps(out), flush=True)


def config2_bench(args):
    """The plain drop-in form at full size: 32^3 box with a voxelised central
    sphere (cells whose centre lies within radius 0.25 are solid), the K=128
    lowest eigenpairs of the pruned DEC 1-form Laplacian and C (both on the
    device, or on the host with --config2-host-eig),
    an upward force on modes 0-7, FP32 C on the plain path and FP32 Phi, one
    integrate() per frame (no frame window), 2000 frames at dt=0.005, 2^18
    particles in fluid cells.  Checks: omega after all frames against the
    FP64 oracle, and the particles against the oracle's particle loop fed
    with the oracle's own u (FP64 Phi), frame by frame."""
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import advection, basis, hip
    from parallelspectralfluids3d_amd.grid import Grid, sphere_obstacle_mask
    n, K, N, frames, dt = 32, 128, 1 << 18, 2000, 0.005
    g = Grid(n, n, n)
    h = 1.0 / n
    amin, amax = (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)
    inside = sphere_obstacle_mask(g, h, amin)
    cache = args.config2_cache
    t_eig = t_c = None
    eig_info = None
    ctx = None if (args.config2_build_only and args.config2_host_eig) else hip.Context(0)
    if cache and os.path.exists(cache):
        sc = hip.read_cache(cache)
        assert (sc["desc"].nx, sc["desc"].K) == (n, K)
        Phi, lam, C, force = sc["Phi"], sc["lam"], sc["C"], sc["gravity"]
        log(f"config2: basis, C and forcing read from the scene cache {cache}")
    else:
        t = time.time()
        if args.config2_host_eig:
            Phi, lam = basis.laplacian_eigenbasis(g, h, inside, K)
        else:
            Phi, lam, eig_info = basis.laplacian_eigenbasis_device(g, h, inside, K, ctx)
        t_eig = time.time() - t
        log(f"config2: eigenbasis K={K} in {t_eig:.1f}s ({'host' if args.config2_host_eig else 'device'})")
        t = time.time()
        V = (basis.curl_operator(g, h) @ Phi.T).T
        if args.config2_host_eig:
            C = advection.build_advection(g, h, inside, Phi, lam, V)
        else:
            # on the device, in an FP64 scene whose C is read back for the
            # oracle, the cache and the FP32 scene's upload
            sc64 = hip.Scene(ctx, hip.make_desc(g, h, K, L.C_FP64, L.PHI_FP64, 0, amin, amax))
            advection.build_advection_device(sc64, g, h, inside, Phi, lam, V)
            C = sc64.get_advection()
            sc64.close()
        del V
        t_c = time.time() - t
        log(f"config2: C in {t_c:.1f}s ({'host' if args.config2_host_eig else 'device'})")
        force = -basis.project(Phi, basis.gravity_field(g, inside))
        force[8:] = 0.0
        if cache:
            hip.write_cache(cache, g, h, lam, C, Phi_cm=Phi, gravity=force, aabb_min=amin, aabb_max=amax)
            log(f"config2: scene cache written to {cache}")
        if args.config2_build_only:
            if ctx is not None:
                ctx.close()
            return
    omega = basis.unit_energy_omega(K, 0)
    rng = np.random.default_rng(21)
    x, y, z = g.voxel_coords()
    vox = np.flatnonzero(inside)
    pick = vox[rng.integers(0, vox.size, N)]
    cell = np.stack([x[pick], y[pick], z[pick]], 1).astype(np.float64)
    parts = np.zeros((N, 4), np.float32)
    parts[:, :3] = np.asarray(amin) + (cell + rng.random((N, 3))) * h
    parts[:, 3] = 60.0
    s = hip.Scene(ctx, hip.make_desc(g, h, K, L.C_FP32, L.PHI_FP32, N, amin, amax))
    s.upload_basis(Phi, lam, force)
    s.upload_advection(C)
    s.set_omega(omega)
    s.set_params(dt, 0.0, True, lifetime=60.0)
    s.set_particles(parts)
    ms = s.time_phase(L.PHASE_FRAME, frames)
    w_gpu, p_gpu = s.get_omega(), s.get_particles()
    s.close()
    ctx.close()
    log(f"config2: {frames} frames in {ms:.1f} ms on the GPU")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    nthreads = min(16, os.cpu_count() or 1)
    m = O.Mesh((n, n, n), h, amin, inside=inside)
    w, p = omega.copy(), parts.copy()
    t = time.time()
    for f in range(frames):
        w = O.reduced_step(w, C, lam, force, dt, 0.0, True, nthreads=nthreads)
        u = O.reconstruct(Phi, w, nthreads=nthreads)
        p = O.particles(m, u, p, resolution=h, timestep=dt, aabb_min=amin, aabb_max=amax, lifetime=60.0,
                        seed=0, step=f, nthreads=nthreads)
        if f % 500 == 499:
            log(f"config2: oracle frame {f + 1}")
    t_ora = time.time() - t
    err = np.abs(p_gpu[:, :3] - p[:, :3]).max(axis=1)
    out = {"config2": {
        "workload": f"{n}^3 box, central sphere obstacle r=0.25, K={K} eigenbasis of the pruned M, upward force on "
                    f"modes 0-7, FP32 C (plain path), FP32 Phi, {frames} frames at dt={dt}, one integrate() per "
                    f"frame, {N} particles in fluid cells",
        "eigensolve_s": t_eig, "eigensolve": ("host shift-invert" if args.config2_host_eig else eig_info),
        "C_build_s": t_c, "C_build": "host" if args.config2_host_eig else "device",
        "ms_per_frame": ms / frames, "particle_steps_per_s": N * frames / (ms * 1e-3),
        "oracle_ms_per_frame": 1e3 * t_ora / frames, "oracle_threads": nthreads,
        "omega_rel_err_vs_fp64_oracle": float(np.linalg.norm(w_gpu - w) / np.linalg.norm(w)),
        "particle_err_max": float(err.max()), "particle_err_p99": float(np.percentile(err, 99)),
        "lifetimes_equal": bool(np.array_equal(p_gpu[:, 3], p[:, 3]))}}
    print(json.dumps(out), flush=True)


def obstacle_bench(args):
    """A 64^3 box with the voxelised central sphere of the 32^3 obstacle scene, K=256, built
    entirely on the device: the pruned Laplacian's eigenbasis
    (psf_eigensolve_csr, penalty-lifted), C from that basis
    (psf_scene_build_advection), then the windowed frame loop with 2^20
    particles in fluid cells, BF16 hi+lo C and FP32 Phi.  Checks that need
    no oracle at this size: every mode's residual against M_II and its
    divergence on deep voxels, and omega's norm after the frames (no
    viscosity or forcing: the renormalisation keeps it 1)."""
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import advection, basis, hip
    from parallelspectralfluids3d_amd.grid import Grid, sphere_obstacle_mask
    n, K, N = 64, 256, 1 << 20
    g = Grid(n, n, n)
    h = 1.0 / n
    amin, amax = (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)
    inside = sphere_obstacle_mask(g, h, amin).astype(bool).reshape(-1)
    ctx = hip.Context(0)
    t = time.time()
    Phi, lam, info = basis.laplacian_eigenbasis_device(g, h, inside, K, ctx)
    t_eig = time.time() - t
    log(f"obstacle: eigenbasis K={K} in {t_eig:.1f}s ({info['products']} products)")
    MII, act = basis.pruned_laplacian(g, h, inside)
    Va = Phi[:, act].T
    res = float(np.linalg.norm(MII @ Va - Va * lam, axis=0).max())
    div = float(np.abs(basis.deep_divergence(g, inside, act) @ Va).max())
    del Va, MII
    t = time.time()
    s = hip.Scene(ctx, hip.make_desc(g, h, K, L.C_BF16X2, L.PHI_FP32, N, amin, amax))
    s.upload_basis(Phi, lam, basis.project(Phi, basis.gravity_field(g, inside)))
    advection.build_advection_device(s, g, h, inside, Phi, lam)
    t_c = time.time() - t
    log(f"obstacle: C and upload in {t_c:.1f}s")
    del Phi
    rng = np.random.default_rng(3)
    x, y, z = g.voxel_coords()
    vox = np.flatnonzero(inside)
    pick = vox[rng.integers(0, vox.size, N)]
    parts = np.zeros((N, 4), np.float32)
    parts[:, :3] = np.asarray(amin) + (np.stack([x[pick], y[pick], z[pick]], 1) + rng.random((N, 3))) * h
    parts[:, 3] = 60.0
    s.set_omega(basis.unit_energy_omega(K, 0))
    s.set_params(0.01, 0.0, False, lifetime=60.0)
    s.set_particles(parts)
    s.set_frame_window(32)
    s.time_phase(L.PHASE_FRAME, 64)
    frames = 256
    ms = s.time_phase(L.PHASE_FRAME, frames) / frames
    w = s.get_omega()
    s.close()
    ctx.close()
    out = {"obstacle": {
        "workload": f"{n}^3 box, central sphere obstacle r=0.25, K={K} eigenbasis of the pruned M and C both built on "
                    f"the GPU, BF16 hi+lo C, FP32 Phi, frame window 32, {N} particles in fluid cells, dt=0.01",
        "active_faces": int(act.size), "eigensolve_s": t_eig, "eigensolve": info, "C_build_and_upload_s": t_c,
        "max_mode_residual": res, "max_mode_divergence": div,
        "ms_per_frame": ms, "particle_steps_per_s": N / (ms * 1e-3), "frames": frames + 64,
        "omega_norm_minus_1": float(np.linalg.norm(w) - 1.0)}}
    print(json.dumps(out), flush=True)


def counter_pass_copies(frames, ms):
    """Three device-to-device copies of 1 GiB with the HIP runtime's own copy
    kernel: the reference for the FETCH_SIZE / WRITE_SIZE counters (each copy
    reads 2^30 bytes and writes 2^30 bytes)."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    n = ct.c_size_t(1 << 30)
    src, dst = ct.c_void_p(), ct.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ct.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, n) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, n, 3) != 0:        # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    print(json.dumps({"counter_pass": True, "frames": frames, "frame_ms": ms / frames,
                      "copy_bytes_each_way": 1 << 30, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--modes", type=int, default=256)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--reduced-steps", type=int, default=1000)
    ap.add_argument("--basis", choices=("eigen", "analytic"), default="eigen",
                    help="eigen: the K lowest eigenpairs of the reference's M, solved on the GPU (default); "
                         "analytic: the analytic sin/cos box modes")
    ap.add_argument("--window", type=int, default=32,
                    help="frame window: reduced steps and reconstructions of this many frames are computed "
                         "together (1 = one frame at a time)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-density", action="store_true")
    ap.add_argument("--config2", action="store_true",
                    help="the 32^3 sphere-obstacle K=128 drop-in configuration at full size, against the oracle")
    ap.add_argument("--config2-cache", default=None,
                    help="scene cache file of the configuration: read if present, else built and written")
    ap.add_argument("--config2-build-only", action="store_true", help="build the basis, C and forcing (and write the cache), then stop; with --config2-host-eig no GPU work")
    ap.add_argument("--config2-host-eig", action="store_true",
                    help="the 32^3 sphere scene: basis (scipy shift-invert) and C on the host instead of the device")
    ap.add_argument("--serving-only", action="store_true", help="only the serving section")
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-profile", type=int, default=0, metavar="N",
                    help="only run the first N requests of the serving stream, all queued at once, through the "
                         "batched server (a short run for a profiler); prints one line and exits")
    ap.add_argument("--obstacle", action="store_true",
                    help="only the 64^3 sphere-obstacle scene (K=256) built on the GPU from its mask: eigensolve, "
                         "C, frame loop; prints one line and exits (about a minute)")
    ap.add_argument("--large", action="store_true",
                    help="only the large single scene (128^3, K=512, 2^24 particles, 500 frames): timing and "
                         "size-independent checks; prints one line and exits (several minutes, ~60 GB of host memory)")
    ap.add_argument("--counter-pass", action="store_true",
                    help="short run for hardware-counter collection: the frame loop, one persistent launch of "
                         "--reduced-steps steps, 20 single-frame particle passes and reconstructions, then three "
                         "plain 1 GiB device copies that calibrate the byte counters; prints one line and exits")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import hip

    if args.large:
        large_bench(args)
        return

    if args.config2:
        config2_bench(args)
        return

    if args.obstacle:
        obstacle_bench(args)
        return

    if args.serving_only:
        ctx = hip.Context(0)
        r = serving_bench(ctx, args.serving_requests, args.serving_rate)
        ctx.close()
        print(json.dumps({"serving": r}), flush=True)
        return

    if args.serving_profile:
        ctx = hip.Context(0)
        bases = serving_bases()
        reqs = serving_workload(args.serving_profile, args.serving_rate)
        r = run_stream(ctx, bases, reqs, max_active=64, saturate=True, modes_per_wg=SERVE_MPW_BATCHED,
                       u_fp32_sums=SERVE_U_FP32_SUMS, shared_basis=SERVE_SHARED_BATCHED)
        ctx.close()
        print(json.dumps({"serving_profile": args.serving_profile, **r}), flush=True)
        return

    n, K, N = args.grid, args.modes, args.particles
    ctx = hip.Context(0)
    sc = build_scene(n, K, N, basis_kind=args.basis, ctx=ctx)
    F = sc["grid"].nF

    # ---- frame loop (the value) ---------------------------------------------------
    s = make_hip_scene(ctx, sc, L.C_BF16X2, PHI_PREC, True, N)
    W = args.window
    s.set_frame_window(W)
    log(f"frame scene uploaded (frame window {W})")
    # The timed region (psf_time_phase phase 3) starts at a window boundary
    # with nothing computed ahead, computes no frame past the timed ones and
    # ends after every stream's work: it holds exactly the timed frames' reduced
    # steps, velocity fields and particle passes.  The warm-up runs the same way.
    s.time_phase(L.PHASE_FRAME, args.warmup)
    ctx.sync()
    c0 = s.counters()
    t0 = time.perf_counter()
    ms = s.time_phase(L.PHASE_FRAME, args.steps)
    wall = time.perf_counter() - t0
    c1 = s.counters()
    ms_frame = ms / args.steps
    timed = {k: c1[k] - c0[k] for k in c1}
    timed_region = {
        "frames": timed["frames"], "frames_computed": timed["window_frames"],
        "window_fills": timed["window_fills"], "frames_dropped": timed["frames_dropped"],
        "launches": {"k_persistent": timed["step_launches"], "k_reconstruct_mfma": timed["recon_launches"],
                     "k_trace_frames": timed["particle_launches"]},
        "note": "between two k_timed_region_mark launches: the reduced steps (k_persistent, all of a window's "
                "steps per launch), velocity fields (k_reconstruct_mfma, all of a window's frames per launch) "
                "and particle passes (k_trace_frames, up to 8 frames per launch) of exactly the timed frames; "
                "nothing computed before the start or for frames after the end"}
    # The same region again with each launch timed by HIP events on its own
    # stream (phase 11, psf_scene_region_kernels): the events lengthen the
    # region by about a tenth, so the headline region above runs without them.
    ms_replica = s.time_phase(L.PHASE_FRAME_TIMED, args.steps)
    rk = s.region_kernels()
    timed_region["kernel_ms"] = {"k_persistent": rk["step_ms"], "k_reconstruct_mfma": rk["recon_ms"],
                                 "k_trace_frames": rk["particle_ms"]}
    timed_region["kernel_units"] = {"steps": rk["steps"], "fields": rk["recon_frames"],
                                    "particle_frames": rk["particle_frames"]}
    timed_region["kernel_ms_note"] = (f"per-launch HIP events over an identical second region of {args.steps} "
                                      f"frames ({ms_replica:.3f} ms with the events, the headline region "
                                      f"{ms:.3f} ms without)")
    log(f"{args.steps} frames: {ms:.2f} ms (events), {1e3 * wall:.2f} ms (wall); in the region: "
        f"{timed_region['launches']}, {timed['window_frames']} frames computed")
    if args.counter_pass:
        s.step(args.reduced_steps)     # one persistent launch of that many steps
        s.time_phase(L.PHASE_TRACE, 20)        # single-frame kernels of the one-at-a-time path
        s.time_phase(L.PHASE_RECON, 20)
        ctx.sync()
        s.close()
        ctx.close()
        counter_pass_copies(args.steps, ms)
        return
    # frame latency distribution: one integrate() per call continuing the
    # window, as a caller's frame loop runs them, 8 windows' worth
    lat = np.array([s.time_phase(L.PHASE_FRAME_CONTINUE, 1) for _ in range(max(200, 8 * W))])
    # per-kernel averages, each per frame
    reps = 100
    ms_tr = s.time_phase(L.PHASE_TRACE, reps) / reps
    kern = {"k_trace": ms_tr}
    ms_trf = None
    if W >= L.TRACE_FUSE:
        # the frame loop's particle launches: TRACE_FUSE frames each
        ms_trf = s.time_phase(L.PHASE_TRACE_FUSED, reps // 4) / (reps // 4)
        kern = {"k_trace_frames": ms_trf / L.TRACE_FUSE}
    if W > 1:
        ms_wrec = s.time_phase(L.PHASE_WINDOW_RECON, 10) / 10
        ms_wstep = s.time_phase(L.PHASE_NSTEP, W)
        kern["k_reconstruct_window"] = ms_wrec / W
        kern["k_persistent"] = ms_wstep / W
    ms_rec = s.time_phase(L.PHASE_RECON, reps) / reps
    ms_st = s.time_phase(L.PHASE_STEP, reps) / reps
    if W <= 1:
        kern["k_reconstruct"] = ms_rec
        kern["k_persistent"] = ms_st
    # particle density volume on the reference's 256^3 grid, counted inside
    # the particle pass: a frame followed by psf_density, and each pass alone
    dens = None
    if not args.no_density:
        dims = (256, 256, 256)
        s.set_density_grid(dims, count_in_trace=True)
        s.time_phase(L.PHASE_FRAME_DENSITY, args.warmup)
        ms_fd = s.time_phase(L.PHASE_FRAME_DENSITY, args.steps) / args.steps
        ms_vol = s.time_phase(L.PHASE_DENSITY_VOLUME, reps) / reps
        ms_cnt = s.time_phase(L.PHASE_DENSITY_COUNT, reps) / reps
        s.set_density_grid((0, 0, 0))
        cells = dims[0] * dims[1] * dims[2]
        dens = {"grid": list(dims), "ms_per_frame_with_density": ms_fd, "ms_per_frame_without": ms_frame,
                "volume_pass_ms": ms_vol, "count_pass_ms": ms_cnt,
                "count_pass_particles_per_s": N / (ms_cnt * 1e-3)}
        log(f"density {dims}: frame + volume {1e3 * ms_fd:.1f} us, volume pass {1e3 * ms_vol:.1f} us, "
            f"count pass {1e3 * ms_cnt:.1f} us")
    # the same frames one at a time
    s.set_frame_window(1)
    s.time_phase(L.PHASE_FRAME, args.warmup)
    ms_frame1 = s.time_phase(L.PHASE_FRAME, args.steps) / args.steps
    s.close()
    log(f"per frame: " + ", ".join(f"{k} {1e3 * v:.1f} us" for k, v in kern.items()) +
        f"; one frame at a time {1e3 * ms_frame1:.1f} us")

    # roofline of each kernel, per launch; the dominant one (largest share of a
    # frame) is the headline
    roof = {}
    trace_bytes = N * 32 + F * 4
    roof["k_trace"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9, achieved=trace_bytes / (ms_tr * 1e-3) / 1e9,
                           algorithmic_bytes_per_launch=trace_bytes, traffic=TRACE_TRAFFIC_64_1M if
                           (n, N) == (64, 1 << 20) else None)
    if ms_trf is not None:
        # the per-unit bytes of a particle-step (32 B record in + out, plus 4 F of
        # u per frame) times the units of one launch (N particles x TRACE_FUSE
        # frames); the fused launch itself moves the records once
        fb = L.TRACE_FUSE * (N * 32 + F * 4)
        fb_moved = N * 32 + L.TRACE_FUSE * F * 4
        roof["k_trace_frames"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9,
                                      achieved=fb / (ms_trf * 1e-3) / 1e9, algorithmic_bytes_per_launch=fb,
                                      compulsory_bytes_per_launch=fb_moved, frames_per_launch=L.TRACE_FUSE,
                                      traffic=TRACE_FRAMES_TRAFFIC_64_1M if (n, N) == (64, 1 << 20) else None,
                                      note="algorithmic bytes = particle-steps x 32 B + frames x 4F; the fused "
                                           "launch keeps records in registers across its frames and moves "
                                           "compulsory_bytes_per_launch. Bound by the vector-memory path, not HBM: 204 "
                                           "vector instructions and 12 dependent flux-pair gathers per "
                                           "particle-step, 22.7 L1 accesses per gather (the address unit serves "
                                           "four lanes per cycle); texture-data unit busy 80% (TD_TD_BUSY_sum per "
                                           "CU), address unit 69% (TA_BUSY_avr), 63% of wave-cycles waiting on "
                                           "memory (SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES)")
    rec_bytes = F * K * 4 + F * 4 + K * 8
    # (the calibrated traffic constants were measured for the FP32-sum kernels)
    fp64_sums = PHI_PREC != 2
    roof["k_reconstruct"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9,
                                 achieved=rec_bytes / (ms_rec * 1e-3) / 1e9, algorithmic_bytes_per_launch=rec_bytes,
                                 traffic=RECON_TRAFFIC_64_256 if (n, K) == (64, 256) and not fp64_sums else None)
    if W > 1:
        wflops = 2.0 * F * K * W
        wbytes = F * K * 4 + W * (F * 4 + K * 8)
        if fp64_sums:
            roof["k_reconstruct_window"] = dict(bound="fp64", unit="TFLOP/s", peak=FP64_PEAK / 1e12,
                                                achieved=wflops / (ms_wrec * 1e-3) / 1e12,
                                                algorithmic_flops_per_launch=wflops, algorithmic_bytes_per_launch=wbytes,
                                                traffic=None)
        else:
            # FP32 matrix cores: 2 F K W flops take less time at their peak than
            # the pass over Phi at the HBM peak, so HBM bounds the launch
            roof["k_reconstruct_window"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9,
                                                achieved=wbytes / (ms_wrec * 1e-3) / 1e9,
                                                algorithmic_bytes_per_launch=wbytes, algorithmic_flops_per_launch=wflops,
                                                fp32_mfma_tflops=wflops / (ms_wrec * 1e-3) / 1e12,
                                                fp32_mfma_peak_tflops=FP32_PEAK / 1e12,
                                                traffic=WINDOW_TRAFFIC_64_256_W32 if (n, K, W) == (64, 256, 32) else None)
        cbytes = K * K * (K + 1) // 2 * 4 + 32 * K
        roof["k_persistent"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9,
                                    achieved=cbytes / (kern["k_persistent"] * 1e-3) / 1e9,
                                    algorithmic_bytes_per_step=cbytes, traffic=None,
                                    note="C stays in registers across the steps of a launch: effective bandwidth")
    if dens is not None:
        vbytes = 10 * cells         # counters read and cleared (4 + 4 B), volume written (2 B)
        roof["k_density_volume"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9,
                                        achieved=vbytes / (ms_vol * 1e-3) / 1e9, algorithmic_bytes_per_launch=vbytes,
                                        traffic=None)
        cbytes_d = 16 * N
        roof["k_density_count"] = dict(bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9,
                                       achieved=cbytes_d / (ms_cnt * 1e-3) / 1e9,
                                       algorithmic_bytes_per_launch=cbytes_d, traffic=None,
                                       note="plus one counter increment per particle (L2 atomics)")
    for r in roof.values():
        r["frac"] = r["achieved"] / r["peak"]
    # The headline roofline: the kernel with the largest share of the timed
    # frame region, at its average launch duration there (HIP events around
    # each launch on its stream, in the replica of the headline region).
    region_ms = timed_region["kernel_ms"]
    region_roof = {}
    if rk["particle_launches"] > 0:
        per_launch_ms = rk["particle_ms"] / rk["particle_launches"]
        frames_per_launch = rk["particle_frames"] / rk["particle_launches"]
        pb = frames_per_launch * (N * 32 + F * 4)
        region_roof["k_trace_frames"] = dict(
            bound="hbm", unit="GB/s", peak=HBM_PEAK / 1e9, achieved=pb / (per_launch_ms * 1e-3) / 1e9,
            algorithmic_bytes_per_launch=pb, frames_per_launch=frames_per_launch,
            launches=rk["particle_launches"], avg_launch_ms=per_launch_ms,
            traffic=TRACE_FRAMES_TRAFFIC_64_1M if (n, N) == (64, 1 << 20) else None,
            ms_per_frame=rk["particle_ms"] / max(1, rk["particle_frames"]),
            note="k_trace_frames in the timed frame region (8 + 8 + 4 frames for 20 frames); algorithmic "
                 "bytes = particle-steps x 32 B + frames x 4F per launch. Bound by the vector-memory path "
                 "(the texture address unit serves four lanes per cycle, one cache line each) and 201 vector "
                 "instructions per particle-step, not HBM")
    if rk["recon_launches"] > 0:
        per_launch_ms = rk["recon_ms"] / rk["recon_launches"]
        wr = rk["recon_frames"] / rk["recon_launches"]
        rbytes = F * K * 4 + wr * (F * 4 + K * 8)
        rflops = 2.0 * F * K * wr
        r = dict(unit="GB/s", peak=HBM_PEAK / 1e9, bound="hbm", achieved=rbytes / (per_launch_ms * 1e-3) / 1e9,
                 algorithmic_bytes_per_launch=rbytes, algorithmic_flops_per_launch=rflops, frames_per_launch=wr,
                 launches=rk["recon_launches"], avg_launch_ms=per_launch_ms,
                 ms_per_frame=rk["recon_ms"] / max(1, rk["recon_frames"]),
                 traffic=WINDOW_TRAFFIC_64_256_W20 if (n, K, wr) == (64, 256, 20) and not fp64_sums else None,
                 note="k_reconstruct_mfma in the timed frame region: the 20 frames' velocity fields in one "
                      "pass over Phi; algorithmic bytes = F K 4 (Phi) + frames x (4F + 8K). FP32 matrix cores "
                      "in 16-frame column tiles (two tiles for 20 frames), so the FP32 MFMA work is that of 32 "
                      "frames: HBM and the matrix cores bound it together")
        if fp64_sums:
            r.update(bound="fp64", unit="TFLOP/s", peak=FP64_PEAK / 1e12,
                     achieved=rflops / (per_launch_ms * 1e-3) / 1e12)
        else:
            r.update(fp32_mfma_tflops=rflops / (per_launch_ms * 1e-3) / 1e12, fp32_mfma_peak_tflops=FP32_PEAK / 1e12)
        region_roof["k_reconstruct_mfma"] = r
    for r in region_roof.values():
        r["frac"] = r["achieved"] / r["peak"]
        if rk["recon_launches"] > 1:
            r["overlap"] = ("the region spans several frame windows: the next window's reduced steps and "
                            "velocity fields run on the side stream beside the current window's particle "
                            "passes, so each launch's duration includes the time it shares the GPU")
    dominant = max(region_ms, key=region_ms.get)
    if dominant in region_roof:
        head = dict(region_roof[dominant])
    else:
        dominant = max(kern, key=kern.get)
        head = dict(roof[dominant])
        head["ms_per_frame"] = kern[dominant]
    head["kernel"] = dominant

    # ---- reduced steps/s: one persistent launch vs the plain FP32 path -------------------
    nred = args.reduced_steps
    res = {}
    for name, prec in (("bf16x2", L.C_BF16X2), ("bf16", L.C_BF16), ("fp16", L.C_FP16)):
        r = make_hip_scene(ctx, sc, prec, L.PHI_FP32, False, 0)
        r.time_phase(L.PHASE_NSTEP, 50)
        r.set_omega(sc["omega"])
        res[name] = nred / (1e-3 * r.time_phase(L.PHASE_NSTEP, nred))
        r.close()
    for name, prec in (("plain_fp32", L.C_FP32), ("plain_fp64", L.C_FP64)):
        r = make_hip_scene(ctx, sc, prec, L.PHI_FP32, False, 0)
        r.time_phase(L.PHASE_STEP, 20)
        res[name] = nred / (1e-3 * r.time_phase(L.PHASE_STEP, nred))
        r.close()
    log("reduced steps/s " + ", ".join(f"{k}={v:.0f}" for k, v in res.items()))

    # ---- omega trajectories of each precision, every 100 steps (compared with the
    # FP64 oracle in the CPU leg below, and with the GPU's FP64 path) -----------------------
    errors, traj = {}, {}
    if not args.no_precision:
        for name, prec in (("fp64", L.C_FP64), ("bf16x2", L.C_BF16X2), ("bf16", L.C_BF16), ("fp16", L.C_FP16)):
            r = make_hip_scene(ctx, sc, prec, L.PHI_FP32, False, 0)
            traj[name] = []
            for _ in range(10):
                r.step(100)
                traj[name].append(r.get_omega())
            r.close()
        for name in ("bf16x2", "bf16", "fp16"):
            errors[name] = [float(np.linalg.norm(w - v) / np.linalg.norm(v)) for w, v in zip(traj[name], traj["fp64"])]
        log("omega rel err at 1000 steps vs the GPU FP64 path: " +
            ", ".join(f"{k}={v[-1]:.2e}" for k, v in errors.items()))
    ctx.close()
    serving = None
    if not args.no_serving:
        # a context of its own: no memory pool, streams or scenes left by the sections above
        sctx = hip.Context(0)
        serving = serving_bench(sctx, args.serving_requests, args.serving_rate)
        sctx.close()

    cpu, accuracy, ora_err = None, None, None
    if not args.no_cpu_baseline:
        ref_state, cpu = cpu_baseline(sc)
        if traj:
            ora_err = oracle_precision_errors(sc, traj)
            log("omega rel err at 1000 steps vs the FP64 oracle: " +
                ", ".join(f"{k}={v[-1]:.2e}" for k, v in ora_err.items()))
        log(f"cpu baseline {cpu['value']:.3e} particle-steps/s on {cpu['cores']} threads")
        # the frame loop's arithmetic against the FP64 oracle over the frames the CPU leg ran
        actx = hip.Context(0)
        a = make_hip_scene(actx, sc, L.C_BF16X2, PHI_PREC, True, N)
        a.set_frame_window(W)
        a.integrate(ref_state["frames"])
        wg, ug, pg = a.get_omega(), a.get_velocity(), a.get_particles()
        a.close()
        actx.close()
        accuracy = {"frames": ref_state["frames"],
                    "omega_rel_err": float(np.linalg.norm(wg - ref_state["omega"]) / np.linalg.norm(ref_state["omega"])),
                    "u_max_abs_err_over_max_u": float(np.abs(ug - ref_state["u"]).max() / np.abs(ref_state["u"]).max()),
                    "particle_max_abs_err": float(np.abs(pg[:, :3] - ref_state["parts"][:, :3]).max()),
                    "against": "the oracle (FP64 C, Phi and sums, FP32 particles) from the same initial state"}
        log(f"frame loop vs FP64 oracle after {ref_state['frames']} frames: omega {accuracy['omega_rel_err']:.2e}, "
            f"u {accuracy['u_max_abs_err_over_max_u']:.2e} of max|u|, particles {accuracy['particle_max_abs_err']:.2e}")

    out = {
        "metric": METRIC,
        "value": N / (ms_frame * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_frame,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "dtype_detail": "C in BF16 hi+lo on MFMA with FP32 accumulation; FP64 omega update (energy, Euler, "
                        "renormalisation, viscosity); u = Phi w with FP32 Phi and w and FP32 sums on the FP32 matrix "
                        "cores (one pass over Phi per frame window); FP32 particles",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 closed box, K={K} "
                               f"{'lowest eigenpairs of the DEC 1-form Laplacian M (wall edges weighted)' if sc['basis'] == 'eigen' else 'analytic sin/cos box modes'}, "
                               f"{N} particles, dt=0.01, "
                               f"one integrate() frame per step", "grid": n, "K": K, "particles": N,
                   "frame_window": W},
        "basis": {"kind": sc["basis"], "solve": sc["basis_info"],
                  "note": "the frame loop and the precision sections run on this basis; the CPU oracle gets the "
                          "same Phi, lambda and C"},
        "ms_per_frame_one_at_a_time": ms_frame1,
        "particle_steps_per_s_one_at_a_time": N / (ms_frame1 * 1e-3),
        "reduced_steps_per_s": res.get("bf16x2"),
        "reduced_steps_per_s_by_path": res,
        "reduced_speedup_vs_plain_fp32": res["bf16x2"] / res["plain_fp32"],
        "omega_rel_err_vs_fp64_oracle_every_100_steps": ora_err,
        "omega_rel_err_vs_gpu_fp64_every_100_steps": errors,
        "timed_region": timed_region,
        "frame_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                             "samples": int(lat.size),
                             "note": "one integrate() call continuing the frame window, its particle pass included"},
        "kernel_ms_per_frame": kern,
        "phase_ms_one_at_a_time": {"reduced_step": ms_st, "reconstruct": ms_rec, "trace": ms_tr},
        "roofline": head,
        "roofline_in_timed_region": region_roof,
        "roofline_by_kernel": roof,
        "density": dens,
        "serving": serving,
        "cpu_baseline": cpu,
        "frame_loop_error_vs_fp64_oracle": accuracy,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
