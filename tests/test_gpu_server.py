"""Scene server (continuous batching): every request's result is the same
alone and in a mix, matches the oracle run frame by frame
(psfsolver.cpp:21-234) on the first K modes of the shared basis, and
admission respects the memory budget."""
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
    while not all(sv.done(i) for i in ids):
        sv.round()
        guard += 1
        assert guard < 10000
    return [sv.result(i) for i in ids]


@pytest.mark.parametrize("u_fp32_sums", [False, True])
def test_alone_equals_mix_bitwise(hip_ctx, u_fp32_sums):
    _say(f"server alone vs mix, u_fp32_sums={u_fp32_sums}")
    sd = make_scene(n=8, K=K0)
    reqs = _requests(sd)
    sv, bid = _server(hip_ctx, sd, window=16, u_fp32_sums=u_fp32_sums)
    mix = _run(sv, bid, reqs)
    st = sv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for j, rq in enumerate(reqs):
        alone = _run(sv, bid, [rq])[0]
        assert np.array_equal(alone[0].view(np.uint64), mix[j][0].view(np.uint64)), j
        assert np.array_equal(alone[1].view(np.uint32), mix[j][1].view(np.uint32)), j
    sv.close()


@pytest.mark.parametrize("u_fp32_sums", [False, True])
def test_server_matches_oracle(hip_ctx, u_fp32_sums):
    """omega within the BF16 hi+lo budget of the FP64 oracle; particles within
    1e-5 (they see u from that omega, with FP64 or FP32 sums) after up to 70
    frames with respawns."""
    _say(f"server vs oracle, u_fp32_sums={u_fp32_sums}")
    sd = make_scene(n=8, K=K0)
    reqs = _requests(sd, seed=1)
    sv, bid = _server(hip_ctx, sd, window=8, u_fp32_sums=u_fp32_sums)
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


def test_admission_by_footprint(hip_ctx):
    """A budget that holds two of the requests keeps at most two on the
    device; all finish with the results they give alone."""
    _say("server admission by footprint")
    sd = make_scene(n=8, K=K0)
    reqs = _requests(sd, seed=2)
    sv, bid = _server(hip_ctx, sd, window=4)
    fps = sorted(sv.footprint(bid, K, p.shape[0]) for (K, _, _, p) in reqs)
    sv.close()
    budget = fps[-1] + fps[-2]
    sv, bid = _server(hip_ctx, sd, window=4, mem_budget=budget)
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
    sv, bid = _server(hip_ctx, sd, window=4)
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
