"""ctypes binding of libpsf_hip.so (include/psf/psf.h).

The library is built in-tree by parallelspectralfluids3d_amd.build; loading
fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpsf_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "psf", "psf.h")

PSF_OK, PSF_ERR_ARG, PSF_ERR_HIP, PSF_ERR_NOMEM, PSF_ERR_STATE, PSF_ERR_UNSUPPORTED, PSF_ERR_NONFINITE = \
    0, -1, -2, -3, -4, -5, -6
C_FP64, C_FP32, C_BF16, C_FP16, C_BF16X2 = 0, 1, 2, 3, 4
PHI_FP64, PHI_FP32, PHI_FP32_ACC32 = 0, 1, 2
PATH_PLAIN, PATH_PERSISTENT = 0, 1
PHASE_STEP, PHASE_RECON, PHASE_TRACE, PHASE_FRAME, PHASE_NSTEP, PHASE_WINDOW_RECON = 0, 1, 2, 3, 4, 5
PHASE_DENSITY_COUNT, PHASE_DENSITY_VOLUME, PHASE_FRAME_DENSITY, PHASE_TRACE_FUSED = 6, 7, 8, 9
PHASE_FRAME_CONTINUE = 10   # frames continuing the current window (a caller's frame loop)
PHASE_FRAME_TIMED = 11      # PHASE_FRAME with each launch timed (psf_scene_region_kernels)
TRACE_FUSE = 8   # frames per fused particle launch of the windowed frame loop

_STATUS = {0: "PSF_OK", -1: "PSF_ERR_ARG", -2: "PSF_ERR_HIP", -3: "PSF_ERR_NOMEM", -4: "PSF_ERR_STATE",
           -5: "PSF_ERR_UNSUPPORTED", -6: "PSF_ERR_NONFINITE"}


class PsfError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class SceneDesc(ct.Structure):
    _fields_ = [("nx", ct.c_uint32), ("ny", ct.c_uint32), ("nz", ct.c_uint32),
                ("resolution", ct.c_double),
                ("aabb_min", ct.c_float * 3), ("aabb_max", ct.c_float * 3),
                ("K", ct.c_uint32), ("c_precision", ct.c_int32), ("phi_precision", ct.c_int32),
                ("max_particles", ct.c_uint32)]


class ServerConfig(ct.Structure):
    _fields_ = [("window", ct.c_uint32), ("max_active", ct.c_uint32), ("mem_budget", ct.c_uint64),
                ("modes_per_wg", ct.c_uint32), ("u_fp32_sums", ct.c_uint32), ("shared_basis", ct.c_uint32),
                ("wait_limit_ns", ct.c_uint32)]


class SceneCounters(ct.Structure):
    _fields_ = [("frames", ct.c_uint64), ("window_fills", ct.c_uint64), ("window_frames", ct.c_uint64),
                ("windows_dropped", ct.c_uint64), ("frames_dropped", ct.c_uint64),
                ("step_launches", ct.c_uint64), ("recon_launches", ct.c_uint64),
                ("particle_launches", ct.c_uint64)]


class RegionKernels(ct.Structure):
    _fields_ = [("step_ms", ct.c_double), ("recon_ms", ct.c_double), ("particle_ms", ct.c_double),
                ("step_launches", ct.c_uint32), ("recon_launches", ct.c_uint32), ("particle_launches", ct.c_uint32),
                ("steps", ct.c_uint32), ("recon_frames", ct.c_uint32), ("particle_frames", ct.c_uint32)]


class SceneView(ct.Structure):
    _fields_ = [("omega", ct.c_void_p), ("lam", ct.c_void_p), ("gravity", ct.c_void_p), ("visc", ct.c_void_p),
                ("C", ct.c_void_p), ("Phi", ct.c_void_p), ("ld_phi", ct.c_uint64), ("u", ct.c_void_p),
                ("particles", ct.c_void_p), ("pid", ct.c_void_p), ("n_particles", ct.c_uint32),
                ("status", ct.c_void_p)]


class ServerStats(ct.Structure):
    _fields_ = [("submitted", ct.c_uint64), ("completed", ct.c_uint64), ("failed", ct.c_uint64),
                ("rounds", ct.c_uint64), ("scene_frames", ct.c_uint64), ("active", ct.c_uint32),
                ("queued", ct.c_uint32), ("mem_in_use", ct.c_uint64), ("mem_peak", ct.c_uint64),
                ("busy_ms", ct.c_double), ("admit_ms", ct.c_double), ("retire_ms", ct.c_double)]


_lib = None


def load():
    """Load libpsf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m parallelspectralfluids3d_amd.build` "
                           "(the HIP path has no fallback)")
    L = ct.CDLL(LIB_PATH)
    vp, i32, u32, dbl, flt = ct.c_void_p, ct.c_int, ct.c_uint32, ct.c_double, ct.c_float
    pp = ct.POINTER(ct.c_void_p)
    sig = {
        "psf_abi_version": (i32, []),
        "psf_ctx_create": (i32, [i32, pp]),
        "psf_ctx_destroy": (i32, [vp]),
        "psf_last_error": (ct.c_char_p, [vp]),
        "psf_ctx_sync": (i32, [vp]),
        "psf_ctx_stream": (vp, [vp]),
        "psf_scene_create": (i32, [vp, ct.POINTER(SceneDesc), pp]),
        "psf_scene_destroy": (i32, [vp]),
        "psf_scene_footprint": (i32, [ct.POINTER(SceneDesc), ct.POINTER(ct.c_uint64)]),
        "psf_scene_upload_basis": (i32, [vp, vp, vp, vp]),
        "psf_scene_upload_advection": (i32, [vp, vp]),
        "psf_scene_upload_advection_matrices": (i32, [vp, vp]),
        "psf_scene_check_signs": (i32, [vp, vp, ct.c_size_t]),
        "psf_scene_generate_box_basis": (i32, [vp, vp, vp, vp, vp]),
        "psf_scene_build_box_advection": (i32, [vp, vp, vp, vp]),
        "psf_set_density_grid": (i32, [vp, u32, u32, u32, i32]),
        "psf_cache_write": (i32, [ct.c_char_p, vp, vp, vp, vp, vp]),
        "psf_cache_read_header": (i32, [ct.c_char_p, vp, vp, vp]),
        "psf_cache_read": (i32, [ct.c_char_p, vp, vp, vp, vp]),
        "psf_scene_load_cache": (i32, [vp, ct.c_char_p]),
        "psf_scene_build_advection": (i32, [vp, vp, vp, u32, vp, u32, vp, vp, vp, vp]),
        "psf_scene_get_advection": (i32, [vp, vp]),
        "psf_eigensolve_csr": (i32, [vp, u32, vp, vp, vp, u32, u32, u32, dbl, u32, ct.c_uint64, vp, vp,
                                     ct.POINTER(ct.c_uint32)]),
        "psf_density": (i32, [vp, vp, vp]),
        "psf_set_params": (i32, [vp, dbl, dbl, i32, flt]),
        "psf_set_step_path": (i32, [vp, i32]),
        "psf_set_wait_limit": (i32, [vp, dbl]),
        "psf_set_omega": (i32, [vp, vp]),
        "psf_get_omega": (i32, [vp, vp]),
        "psf_set_particles": (i32, [vp, vp, u32]),
        "psf_get_particles": (i32, [vp, vp, u32]),
        "psf_step": (i32, [vp, i32]),
        "psf_reconstruct": (i32, [vp]),
        "psf_get_velocity": (i32, [vp, vp]),
        "psf_trace": (i32, [vp, vp]),
        "psf_integrate": (i32, [vp, i32]),
        "psf_set_frame_window": (i32, [vp, ct.c_uint32]),
        "psf_scene_get_counters": (i32, [vp, ct.POINTER(SceneCounters)]),
        "psf_scene_region_kernels": (i32, [vp, ct.POINTER(RegionKernels)]),
        "psf_scene_device_view": (i32, [vp, ct.POINTER(SceneView)]),
        "psf_step_device": (i32, [i32, vp, vp, vp, vp, u32, dbl, i32, i32, vp, vp]),
        "psf_time_phase": (i32, [vp, i32, i32, ct.POINTER(ct.c_double)]),
        "psf_server_create": (i32, [vp, ct.POINTER(ServerConfig), pp]),
        "psf_server_destroy": (i32, [vp]),
        "psf_server_add_basis": (i32, [vp, ct.POINTER(SceneDesc), vp, vp, vp, vp, ct.POINTER(ct.c_uint32)]),
        "psf_server_request_footprint": (i32, [vp, u32, u32, u32, ct.POINTER(ct.c_uint64)]),
        "psf_server_submit": (i32, [vp, u32, u32, u32, dbl, dbl, i32, flt, vp, vp, u32,
                                    ct.POINTER(ct.c_uint64)]),
        "psf_server_round": (i32, [vp, ct.POINTER(ct.c_uint32)]),
        "psf_server_done": (i32, [vp, ct.c_uint64]),
        "psf_server_completed": (i32, [vp, ct.POINTER(ct.c_uint64), ct.c_uint32, ct.POINTER(ct.c_uint32)]),
        "psf_server_result": (i32, [vp, ct.c_uint64, vp, vp, ct.POINTER(ct.c_double)]),
        "psf_server_get_stats": (i32, [vp, ct.POINTER(ServerStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def declared_symbols():
    """Names of every function include/psf/psf.h declares."""
    import re
    with open(HEADER) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"^PSF_API\s+[\w\s\*]*?\b(psf_\w+)\s*\(", txt, flags=re.M)))
