"""Scene server (continuous batching): every request's result is the same
alone and in a mix, matches the oracle run frame by frame
(psfsolver.cpp:21-234) on the first K modes of the shared basis, and
admission respects the memory budget.  Each case runs with per-request C
(shared_basis=False) and with the basis's C shared by its scenes
(shared_basis=True)."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene, particles_uniform

pytestmark = pytest.mark.gpu

K0 = 24


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _server(ctx, sd, **kw):
    from parallelspectralfluids3d_amd import hip
    sv = hip.Server(ctx, **kw)
    bid = sv.add_basis(sd.grid, sd.resolution, sd.Phi, sd.lam, sd.g, sd.C, sd.aabb_min, sd.aabb_max)
    return sv, bid


def _requests(sd, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for j, (K, frames, N) in enumerate([(16, 37, 700), (24, 70, 1000), (8, 5, 300), (20, 64, 0), (12, 33, 257)]):
        w = rng.standard_normal(K)
        w /= np.linalg.norm(w)
        p = particles_uniform(sd, N, seed=100 + j, lifetime=25.0) if N else np.zeros((0, 4), np.float32)
        out.append((K, frames, w, p))
    return out


def _run(sv, bid, reqs, dt=0.01):
    ids = [sv.submit(bid, w, p, frames, dt=dt, lifetime=30.0) for (K, frames, w, p) in reqs]
    guard = 0
    finished = []
    while not all(sv.done(i) for i in ids):
        sv.round()
        finished += sv.completed()
        guard += 1
        assert guard < 10000
    finished += sv.completed()
    assert sorted(finished) == sorted(ids)      # psf_server_completed names each exactly once
    return [sv.result(i) for i in ids]


@pytest.mark.parametrize("shared_basis", [False, True])
@pytest.mark.parametrize("u_fp32_sums", [False, True])
def test_alone_equals_mix_bitwise(hip_ctx, u_fp32_sums, shared_basis):
    _say(f"server alone vs mix, u_fp32_sums={u_fp32_sums}, shared_basis={shared_basis}")
    sd = make_scene(n=8, K=K0)
    reqs = _requests(sd)
    sv, bid = _server(hip_ctx, sd, window=16, u_fp32_sums=u_fp32_sums, shared_basis=shared_basis)
    mix = _run(sv, bid, reqs)
    st = sv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for j, rq in enumerate(reqs):
        alone = _run(sv, bid, [rq])[0]
        assert np.array_equal(alone[0].view(np.uint64), mix[j][0].view(np.uint64)), j
        assert np.array_equal(alone[1].view(np.uint32), mix[j][1].view(np.uint32)), j
    sv.close()


@pytest.mark.parametrize("shared_basis", [False, True])
def test_paired_fields_alone_equals_mix_bitwise(hip_ctx, shared_basis):
    """With FP32 sums the server computes the velocity fields of two scenes on
    one basis in one pass over Phi (paired by K).  Mode counts that are not
    multiples of the MFMA's k-step of 4 and the two-tile window (32) cover the
    partial step and the second scene's zero modes; each result must keep the
    bits it has alone."""
    _say(f"server paired velocity fields, window 32, shared_basis={shared_basis}")
    sd = make_scene(n=8, K=K0)
    rng = np.random.default_rng(7)
    reqs = []
    for j, (K, frames, N) in enumerate([(23, 40, 600), (17, 70, 800), (9, 33, 300), (24, 65, 500), (13, 50, 257)]):
        w = rng.standard_normal(K)
        w /= np.linalg.norm(w)
        reqs.append((K, frames, w, particles_uniform(sd, N, seed=200 + j, lifetime=25.0)))
    sv, bid = _server(hip_ctx, sd, window=32, u_fp32_sums=True, shared_basis=shared_basis)
    mix = _run(sv, bid, reqs)
    for j, rq in enumerate(reqs):
        alone = _run(sv, bid, [rq])[0]
        assert np.array_equal(alone[0].view(np.uint64), mix[j][0].view(np.uint64)), j
        assert np.array_equal(alone[1].view(np.uint32), mix[j][1].view(np.uint32)), j
    sv.close()


@pytest.mark.parametrize("shared_basis", [False, True])
@pytest.mark.parametrize("u_fp32_sums", [False, True])
def test_server_matches_oracle(hip_ctx, u_fp32_sums, shared_basis):
    """omega within the BF16 hi+lo budget of the FP64 oracle; particles within
    1e-5 (they see u from that omega, with FP64 or FP32 sums) after up to 70
    frames with respawns."""
    _say(f"server vs oracle, u_fp32_sums={u_fp32_sums}, shared_basis={shared_basis}")
    sd = make_scene(n=8, K=K0)
    reqs = _requests(sd, seed=1)
    sv, bid = _server(hip_ctx, sd, window=8, u_fp32_sums=u_fp32_sums, shared_basis=shared_basis)
    res = _run(sv, bid, reqs)
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    for (K, frames, w0, p0), (wg, pg, lat) in zip(reqs, res):
        C = np.ascontiguousarray(sd.C[:K, :K, :K])
        Phi = np.ascontiguousarray(sd.Phi[:K])
        w, p = w0.copy(), p0.copy()
        for f in range(frames):
            w = O.reduced_step(w, C, sd.lam[:K], sd.g[:K], 0.01, 0.0, False)
            if p.shape[0]:
                u = O.reconstruct(Phi, w)
                p = O.particles(m, u, p, resolution=sd.resolution, timestep=0.01, aabb_min=sd.aabb_min,
                                aabb_max=sd.aabb_max, lifetime=30.0, seed=0, step=f)
        assert np.linalg.norm(wg - w) / np.linalg.norm(w) <= 2e-6, K
        if p.shape[0]:
            assert np.abs(pg[:, :3] - p[:, :3]).max() <= 1e-5, K
            assert np.array_equal(pg[:, 3], p[:, 3]), K
        assert lat > 0.0
    sv.close()


@pytest.mark.parametrize("shared_basis", [False, True])
def test_admission_by_footprint(hip_ctx, shared_basis):
    """A budget that holds two of the requests keeps at most two on the
    device; all finish with the results they give alone."""
    _say(f"server admission by footprint, shared_basis={shared_basis}")
    sd = make_scene(n=8, K=K0)
    reqs = _requests(sd, seed=2)
    sv, bid = _server(hip_ctx, sd, window=4, shared_basis=shared_basis)
    fps = sorted(sv.footprint(bid, K, p.shape[0]) for (K, _, _, p) in reqs)
    sv.close()
    budget = fps[-1] + fps[-2]
    sv, bid = _server(hip_ctx, sd, window=4, mem_budget=budget, shared_basis=shared_basis)
    ids = [sv.submit(bid, w, p, frames, lifetime=30.0) for (K, frames, w, p) in reqs]
    peak_active = 0
    while not all(sv.done(i) for i in ids):
        sv.round()
        st = sv.stats()
        peak_active = max(peak_active, st["active"])
        assert st["mem_in_use"] <= budget
    assert 1 <= peak_active <= 2
    assert sv.stats()["mem_peak"] <= budget
    got = [sv.result(i) for i in ids]
    sv.close()
    sv, bid = _server(hip_ctx, sd, window=4, shared_basis=shared_basis)
    for rq, g in zip(reqs, got):
        a = _run(sv, bid, [rq])[0]
        assert np.array_equal(a[0].view(np.uint64), g[0].view(np.uint64))
        assert np.array_equal(a[1].view(np.uint32), g[1].view(np.uint32))
    sv.close()


def test_server_errors(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    _say("server errors")
    sd = make_scene(n=8, K=K0)
    sv, bid = _server(hip_ctx, sd)
    w = np.ones(K0 + 1) / np.sqrt(K0 + 1)
    with pytest.raises(L.PsfError) as e:
        sv.submit(bid, w, np.zeros((0, 4), np.float32), 10)
    assert e.value.status == L.PSF_ERR_ARG
    with pytest.raises(L.PsfError) as e:
        sv.submit(bid + 7, w[:4], np.zeros((0, 4), np.float32), 10)
    assert e.value.status == L.PSF_ERR_ARG
    rid = sv.submit(bid, w[:8], np.zeros((0, 4), np.float32), 40)
    with pytest.raises(L.PsfError) as e:
        sv.result(rid)
    assert e.value.status == L.PSF_ERR_STATE
    with pytest.raises(L.PsfError):
        type(sv)(hip_ctx, window=3)
    sv.close()


def test_shared_basis_many_scenes(hip_ctx):
    """Twenty requests on one basis and five on a second (two instances of the
    first basis in each round): each gives the bits it gives alone, and a
    sampled request of each basis agrees with the oracle."""
    _say("server shared basis, 25 requests on two bases")
    sd = make_scene(n=8, K=K0)
    sd2 = make_scene(n=6, K=16, seed=3)
    sv, bid = _server(hip_ctx, sd, window=8, shared_basis=True, max_active=32)
    bid2 = sv.add_basis(sd2.grid, sd2.resolution, sd2.Phi, sd2.lam, sd2.g, sd2.C, sd2.aabb_min, sd2.aabb_max)
    rng = np.random.default_rng(7)
    reqs = []
    for j in range(25):
        b, d = (bid, sd) if j < 20 else (bid2, sd2)
        K = int(rng.integers(1, d.lam.size + 1))
        w = rng.standard_normal(K)
        w /= np.linalg.norm(w)
        frames = int(rng.integers(1, 40))
        p = particles_uniform(d, 64, seed=300 + j, lifetime=25.0)
        reqs.append((b, d, K, frames, w, p))

    def run(sel):
        ids = [sv.submit(reqs[j][0], reqs[j][4], reqs[j][5], reqs[j][3], dt=0.01, lifetime=30.0) for j in sel]
        guard = 0
        while not all(sv.done(i) for i in ids):
            sv.round()
            guard += 1
            assert guard < 10000
        return [sv.result(i) for i in ids]

    mix = run(range(len(reqs)))
    for j in range(len(reqs)):
        alone = run([j])[0]
        assert np.array_equal(alone[0].view(np.uint64), mix[j][0].view(np.uint64)), j
        assert np.array_equal(alone[1].view(np.uint32), mix[j][1].view(np.uint32)), j
    assert sv.stats()["failed"] == 0
    sv.close()
    for j in (3, 22):
        b, d, K, frames, w0, p0 = reqs[j]
        m = O.Mesh((d.grid.nx, d.grid.ny, d.grid.nz), d.resolution, d.aabb_min, inside=d.inside)
        C = np.ascontiguousarray(d.C[:K, :K, :K])
        Phi = np.ascontiguousarray(d.Phi[:K])
        w, p = w0.copy(), p0.copy()
        for f in range(frames):
            w = O.reduced_step(w, C, d.lam[:K], d.g[:K], 0.01, 0.0, False)
            u = O.reconstruct(Phi, w)
            p = O.particles(m, u, p, resolution=d.resolution, timestep=0.01, aabb_min=d.aabb_min,
                            aabb_max=d.aabb_max, lifetime=30.0, seed=0, step=f)
        wg, pg, _ = mix[j]
        assert np.linalg.norm(wg - w) / np.linalg.norm(w) <= 2e-6, j
        assert np.abs(pg[:, :3] - p[:, :3]).max() <= 1e-5, j
        assert np.array_equal(pg[:, 3], p[:, 3]), j


def test_shared_basis_rejects_large_basis(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import hip
    _say("server shared basis rejects K0 > 256")
    sd = make_scene(n=8, K=K0)
    sv = hip.Server(hip_ctx, shared_basis=True)
    K = 257
    Phi = np.zeros((K, sd.grid.nF))
    with pytest.raises(L.PsfError) as e:
        sv.add_basis(sd.grid, sd.resolution, Phi, np.ones(K), np.zeros(K), np.zeros(K ** 3))
    assert e.value.status == L.PSF_ERR_UNSUPPORTED
    sv.close()


@pytest.mark.parametrize("shared_basis", [False, True])
def test_exchange_wait_limit(hip_ctx, shared_basis):
    """A wait limit of 10 ns (one tick of the wait clock) makes the reduced
    steps' exchange give up instead of waiting: the launch ends, and each
    request fails with the timeout (with the basis-shared step, the error
    names the wait that gave up: instance, workgroup, step, mode and scene);
    a server with the default limit then runs the same requests."""
    from parallelspectralfluids3d_amd import _lib as L
    _say(f"server exchange wait limit, shared_basis={shared_basis}")
    sd = make_scene(n=8, K=96)          # 12 workgroups per scene or instance
    rng = np.random.default_rng(11)
    reqs = []
    for K in (96, 80):
        w = rng.standard_normal(K)
        w /= np.linalg.norm(w)
        reqs.append((K, 64, w, particles_uniform(sd, 128, seed=5, lifetime=25.0)))
    sv, bid = _server(hip_ctx, sd, window=32, shared_basis=shared_basis, wait_limit_ns=10)
    ids = [sv.submit(bid, w, p, frames, lifetime=30.0) for (K, frames, w, p) in reqs]
    guard = 0
    while not all(sv.done(i) for i in ids):
        sv.round()
        guard += 1
        assert guard < 1000
    assert sv.stats()["failed"] == len(ids)
    for i in ids:
        with pytest.raises(L.PsfError) as e:
            sv.result(i)
        assert e.value.status == L.PSF_ERR_HIP and "timed out" in str(e.value)
        if shared_basis:
            assert "workgroup" in str(e.value) and "scene" in str(e.value), str(e.value)
    sv.close()
    sv, bid = _server(hip_ctx, sd, window=32, shared_basis=shared_basis)
    res = _run(sv, bid, reqs)
    assert all(np.all(np.isfinite(r[0])) for r in res)
    sv.close()
