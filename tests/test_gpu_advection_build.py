"""C built on the device from any basis (psf_scene_build_advection) against
the oracle's literal buildAdvection (O.build_advection, psfsolver.cpp:431-551)
and the host restatement (advection.build_advection), on obstacle scenes with
eigensolved bases and on the analytic closed box, in FP64, FP32 and BF16
hi+lo; plus the argument errors."""
import sys

import numpy as np
import pytest

import oracle as O
from scenes import make_scene

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _scene(ctx, sd, prec):
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, prec, L.PHI_FP64, 0, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(ctx, d)
    s.upload_basis(sd.Phi, sd.lam, sd.g)
    return s


def _build(s, sd):
    from parallelspectralfluids3d_amd import advection
    advection.build_advection_device(s, sd.grid, sd.resolution, sd.inside, sd.Phi, sd.lam, sd.V)


@pytest.mark.parametrize("case", [dict(n=8, K=16, basis_kind="eigen", obstacle=True),
                                  dict(n=10, K=24, basis_kind="eigen", obstacle=True),
                                  dict(dims=(6, 8, 10), K=12, basis_kind="eigen", obstacle=True),
                                  dict(n=8, K=16, basis_kind="analytic", obstacle=False)])
def test_fp64_matches_host(hip_ctx, case):
    from parallelspectralfluids3d_amd import _lib as L
    _say(f"device C build fp64 {case}")
    sd = make_scene(**case)
    s = _scene(hip_ctx, sd, L.C_FP64)
    _build(s, sd)
    C = s.get_advection()
    s.close()
    scale = np.abs(sd.C).max()
    assert np.max(np.abs(C - sd.C)) <= 1e-12 * scale
    # the oracle's literal construction (voxel loop, wedge slots, V^T W_i / lambda_i)
    m = O.Mesh((sd.grid.nx, sd.grid.ny, sd.grid.nz), sd.resolution, sd.aabb_min, inside=sd.inside)
    Co = O.build_advection(m, sd.Phi, sd.lam, sd.V)
    assert np.max(np.abs(C - Co)) <= 1e-12 * np.abs(Co).max()
    # the reference's own structure: lambda_i C_i(k, i) = 0 (the diagonal slot of each C_i vanishes)
    i = np.arange(sd.K)
    assert np.max(np.abs(C[i, i, :])) <= 1e-12 * scale


def test_fp32_and_packed(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    _say("device C build fp32 and bf16 hi+lo")
    sd = make_scene(n=8, K=16, basis_kind="eigen", obstacle=True)
    s = _scene(hip_ctx, sd, L.C_FP32)
    _build(s, sd)
    C32 = s.get_advection()
    s.close()
    # one float ulp, or the FP64 rounding noise of entries that vanish (C_i(k, i))
    ref = sd.C.astype(np.float32).astype(np.float64)
    ulp = np.spacing(np.abs(sd.C).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(C32 - ref) <= np.maximum(ulp, 1e-12 * np.abs(sd.C).max()))
    # BF16 hi+lo: same steps as the scene given the host C
    om = []
    for device in (True, False):
        s = _scene(hip_ctx, sd, L.C_BF16X2)
        if device:
            _build(s, sd)
            with pytest.raises(L.PsfError) as e:
                s.get_advection()
            assert e.value.status == L.PSF_ERR_UNSUPPORTED
        else:
            s.upload_advection(sd.C)
        s.set_omega(sd.omega)
        s.set_params(0.01, 0.0, False)
        s.step(40)
        om.append(s.get_omega())
        s.close()
    assert np.linalg.norm(om[0] - om[1]) <= 1e-9 * np.linalg.norm(om[1])


def test_argument_errors(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import advection
    _say("device C build argument errors")
    sd = make_scene(n=8, K=16, basis_kind="eigen", obstacle=True)
    s = _scene(hip_ctx, sd, L.C_FP64)
    e, a, b, c = advection.advection_terms(sd.grid, sd.resolution, sd.inside)
    with pytest.raises(L.PsfError) as err:          # V with a row too few
        s.build_advection(sd.Phi, sd.V[:, :-1], sd.lam, (e, a, b, c))
    assert err.value.status == L.PSF_ERR_ARG
    bad = a.copy()
    bad[5] = sd.grid.nF                              # face index out of range
    with pytest.raises(L.PsfError) as err:
        s.build_advection(sd.Phi, sd.V, sd.lam, (e, bad, b, c))
    assert err.value.status == L.PSF_ERR_ARG
    lam = sd.lam.copy()
    lam[3] = 0.0
    with pytest.raises(L.PsfError) as err:
        s.build_advection(sd.Phi, sd.V, lam, (e, a, b, c))
    assert err.value.status == L.PSF_ERR_ARG
    with pytest.raises(L.PsfError) as err:          # nothing built yet
        s.get_advection()
    assert err.value.status == L.PSF_ERR_STATE
    s.close()
    s = _scene(hip_ctx, sd, L.C_FP16)
    with pytest.raises(L.PsfError) as err:
        s.build_advection(sd.Phi, sd.V, sd.lam, (e, a, b, c))
    assert err.value.status == L.PSF_ERR_UNSUPPORTED
    s.close()
