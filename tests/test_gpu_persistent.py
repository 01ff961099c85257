"""Persistent MFMA path (BF16, FP16, BF16 hi+lo) against the FP64 oracle
(psfsolver.cpp:34-59).  Tolerances are the precision budget of each mode,
measured as relative omega error after the stated number of steps."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene

pytestmark = pytest.mark.gpu

C_BF16, C_FP16, C_BF16X2 = 2, 3, 4
# relative omega error vs FP64 after 1000 steps at dt = 0.01; BF16 at the
# precision target of 1e-4 (the 64^3/K=256 case is in test_gpu_configs.py)
BUDGET = {C_BF16: 1e-4, C_FP16: 2e-5, C_BF16X2: 2e-6}


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _scene(ctx, sd, c_prec):
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, c_prec, 0, 0, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    return s


def _oracle_traj(sd, steps, dt=0.01, nu=0.0, grav=False):
    w = sd.omega.copy()
    for _ in range(steps):
        w = O.reduced_step(w, sd.C, sd.lam, sd.g, dt, nu, grav)
    return w


@pytest.mark.parametrize("prec", [C_BF16X2, C_BF16, C_FP16])
@pytest.mark.parametrize("case", [dict(n=8, K=16), dict(n=16, K=32, basis_kind="eigen")])
def test_persistent_within_budget(hip_ctx, prec, case):
    _say(f"persistent prec={prec} {case}")
    sd = make_scene(**case)
    s = _scene(hip_ctx, sd, prec)
    s.set_params(0.01)
    s.step(1000)
    w = _oracle_traj(sd, 1000)
    rel = np.linalg.norm(s.get_omega() - w) / np.linalg.norm(w)
    print(f"  rel err {rel:.3e}", flush=True)
    assert rel <= BUDGET[prec], rel


def test_one_launch_equals_many(hip_ctx):
    """n steps in one launch and n one-step launches give the same omega bit for bit."""
    _say("persistent one launch vs many")
    sd = make_scene(n=8, K=16)
    a = _scene(hip_ctx, sd, C_BF16X2)
    b = _scene(hip_ctx, sd, C_BF16X2)
    a.set_params(0.01)
    b.set_params(0.01)
    a.step(64)
    for _ in range(64):
        b.step(1)
    assert np.array_equal(a.get_omega(), b.get_omega())


def test_several_modes_per_workgroup(hip_ctx):
    """K = 300 > 256 workgroups: two modes per workgroup and blocks streamed
    from HBM beyond the register-resident part."""
    _say("persistent K=300")
    sd = make_scene(n=10, K=300)
    s = _scene(hip_ctx, sd, C_BF16X2)
    s.set_params(0.01)
    s.step(200)
    w = _oracle_traj(sd, 200)
    rel = np.linalg.norm(s.get_omega() - w) / np.linalg.norm(w)
    assert rel <= BUDGET[C_BF16X2], rel


def test_forcing_viscosity_and_nonfinite(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    _say("persistent forcing/viscosity/nan")
    sd = make_scene(n=8, K=16)
    s = _scene(hip_ctx, sd, C_BF16X2)
    s.set_params(0.005, 0.2, True)
    s.step(300)
    w = _oracle_traj(sd, 300, dt=0.005, nu=0.2, grav=True)
    assert np.linalg.norm(s.get_omega() - w) / np.linalg.norm(w) <= BUDGET[C_BF16X2]
    bad = sd.omega.copy()
    bad[1] = np.inf
    s.set_omega(bad)
    s.step(5)
    with pytest.raises(L.PsfError) as e:
        s.get_omega()
    assert e.value.status == L.PSF_ERR_NONFINITE


def test_exchange_watchdog_gives_up(hip_ctx):
    """A wait limit of 10 ns (one tick of the wait clock) makes every
    workgroup's first wait for the others' y give up: the launch ends, the
    status reports the timeout instead of a hang, omega keeps the last step
    all workgroups completed, and after the default limit is restored and
    omega is set again the scene steps like the oracle."""
    from parallelspectralfluids3d_amd import _lib as L
    _say("persistent watchdog")
    sd = make_scene(n=8, K=16)
    s = _scene(hip_ctx, sd, C_BF16X2)
    s.set_params(0.01)
    s.step(3)
    w3 = s.get_omega()
    s.set_wait_limit(1e-8)
    s.step(50)
    with pytest.raises(L.PsfError) as e:
        s.get_omega()
    assert e.value.status == L.PSF_ERR_HIP and "timed out" in str(e.value)
    s.set_wait_limit(0.0)
    s.set_omega(sd.omega)
    s.step(100)
    w = _oracle_traj(sd, 100)
    assert np.linalg.norm(s.get_omega() - w) / np.linalg.norm(w) <= BUDGET[C_BF16X2]
    assert np.all(np.isfinite(w3))
    with pytest.raises(L.PsfError) as e:
        s.set_wait_limit(-1.0)
    assert e.value.status == L.PSF_ERR_ARG
    p = _scene(hip_ctx, sd, 0)          # FP64: the plain path has no exchange
    with pytest.raises(L.PsfError) as e:
        p.set_wait_limit(1.0)
    assert e.value.status == L.PSF_ERR_UNSUPPORTED
