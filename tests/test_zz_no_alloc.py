"""psf_step_device, the allocation-free variant of the plain reduced step:
its steps on a scene's own device buffers (psf_scene_device_view) give the
bits of psf_step on an identical scene, for two calls in a row, the second
with a smaller K; sizes past its scratch array are refused."""
import ctypes as ct
import sys

import numpy as np
import pytest

from scenes import make_scene

pytestmark = pytest.mark.gpu
C_FP64, C_FP32, C_BF16X2 = 0, 1, 4


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _scene(ctx, sd, prec, grav):
    from parallelspectralfluids3d_amd import hip
    d = hip.make_desc(sd.grid, sd.resolution, sd.K, prec, 0, 0, sd.aabb_min, sd.aabb_max)
    s = hip.Scene(ctx, d)
    s.upload_basis(None, sd.lam, sd.g)
    s.upload_advection(sd.C)
    s.set_omega(sd.omega)
    s.set_params(0.01, 0.3, grav)
    return s


@pytest.mark.parametrize("prec", [C_FP64, C_FP32])
def test_no_alloc_steps_equal_psf_step(hip_ctx, prec):
    from parallelspectralfluids3d_amd import _lib as L
    lib = L.load()
    for case, nsteps, grav in ((dict(n=10, K=300), 120, True), (dict(n=8, K=24), 60, False)):
        _say(f"psf_step_device prec={prec} K={case['K']} steps={nsteps}")
        sd = make_scene(**case)
        ref = _scene(hip_ctx, sd, prec, grav)
        ref.step(nsteps)
        w_ref = ref.get_omega()
        dev = _scene(hip_ctx, sd, prec, grav)
        v = dev.device_view()
        hip_ctx.sync()
        rc = lib.psf_step_device(prec, v["C"], v["omega"], v["visc"], v["gravity"], sd.K, 0.01, 1 if grav else 0,
                                 nsteps, v["status"], hip_ctx.stream())
        assert rc == 0
        w = dev.get_omega()                  # synchronises the context stream
        assert np.array_equal(w.view(np.uint64), w_ref.view(np.uint64))
        ref.close()
        dev.close()


def test_no_alloc_refuses_bad_sizes(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    lib = L.load()
    sd = make_scene(n=8, K=16)
    s = _scene(hip_ctx, sd, C_FP64, False)
    v = s.device_view()
    for K, prec, steps in ((1025, C_FP64, 1), (0, C_FP64, 1), (16, C_FP64, -1)):
        assert lib.psf_step_device(prec, v["C"], v["omega"], v["visc"], v["gravity"], K, 0.01, 0, steps,
                                   v["status"], hip_ctx.stream()) == L.PSF_ERR_ARG
    assert lib.psf_step_device(C_BF16X2, v["C"], v["omega"], v["visc"], v["gravity"], 16, 0.01, 0, 1,
                               v["status"], hip_ctx.stream()) == L.PSF_ERR_UNSUPPORTED
    assert lib.psf_step_device(C_FP64, None, v["omega"], v["visc"], v["gravity"], 16, 0.01, 0, 1,
                               v["status"], hip_ctx.stream()) == L.PSF_ERR_ARG
    s.close()
