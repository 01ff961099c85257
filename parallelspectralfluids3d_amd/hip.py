"""Python handles over the C ABI: Context (one GPU, one stream) and Scene.

Every call goes through libpsf_hip.so; numpy arrays are passed as host
pointers and copied by the library.
"""
from __future__ import annotations

import atexit
import ctypes as ct
import os
import weakref

import numpy as np

from . import _lib as L
from .grid import Grid


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


_LIVE_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_all():
    """Release device memory before the HIP runtime is torn down at exit."""
    for c in list(_LIVE_CONTEXTS):
        c.close()


class Context:
    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = ct.c_void_p()
        rc = self.lib.psf_ctx_create(device, ct.byref(h))
        if rc != 0:
            raise L.PsfError(rc, f"psf_ctx_create(device={device}) failed")
        self.h = h
        self.scenes = weakref.WeakSet()
        _LIVE_CONTEXTS.add(self)

    def check(self, rc, what=""):
        if rc != 0:
            msg = self.lib.psf_last_error(self.h)
            raise L.PsfError(rc, f"{what}: {msg.decode() if msg else ''}")

    def sync(self):
        self.check(self.lib.psf_ctx_sync(self.h), "sync")

    def stream(self):
        """The context's hipStream_t as an integer (psf_ctx_stream)."""
        return self.lib.psf_ctx_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            for sc in list(self.scenes):
                sc.close()
            self.lib.psf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_desc(grid: Grid, resolution, K, c_precision=L.C_FP64, phi_precision=L.PHI_FP64, max_particles=0,
              aabb_min=(-0.5, -0.5, -0.5), aabb_max=(0.5, 0.5, 0.5)):
    d = L.SceneDesc()
    d.nx, d.ny, d.nz = grid.nx, grid.ny, grid.nz
    d.resolution = float(resolution)
    d.aabb_min = (ct.c_float * 3)(*aabb_min)
    d.aabb_max = (ct.c_float * 3)(*aabb_max)
    d.K = int(K)
    d.c_precision = int(c_precision)
    d.phi_precision = int(phi_precision)
    d.max_particles = int(max_particles)
    return d


def footprint(desc) -> int:
    lib = L.load()
    out = ct.c_uint64()
    rc = lib.psf_scene_footprint(ct.byref(desc), ct.byref(out))
    if rc != 0:
        raise L.PsfError(rc, "psf_scene_footprint")
    return int(out.value)


def write_cache(path, grid: Grid, resolution, lam, C, Phi_cm=None, gravity=None, aabb_min=(-0.5, -0.5, -0.5),
                aabb_max=(0.5, 0.5, 0.5)):
    """Scene cache file (psf_cache_write): lambda (K,), C (K, K, K) with
    C[i, b, a] = C_i(a, b), and optionally Phi (K, F) and g (K,), all FP64."""
    lib = L.load()
    lam = np.ascontiguousarray(lam, np.float64)
    K = lam.size
    d = make_desc(grid, resolution, K, aabb_min=aabb_min, aabb_max=aabb_max)
    C = np.ascontiguousarray(C, np.float64)
    assert C.shape == (K, K, K)
    P = None if Phi_cm is None else np.ascontiguousarray(Phi_cm, np.float64)
    if P is not None:
        assert P.shape == (K, grid.nF)
    g = None if gravity is None else np.ascontiguousarray(gravity, np.float64)
    rc = lib.psf_cache_write(os.fsencode(path), ct.byref(d), None if P is None else _ptr(P), _ptr(lam),
                             None if g is None else _ptr(g), _ptr(C))
    if rc != 0:
        raise L.PsfError(rc, f"psf_cache_write({path})")


def read_cache_header(path):
    """(desc, has_phi, has_gravity) of a scene cache file."""
    lib = L.load()
    d = L.SceneDesc()
    hp, hg = ct.c_uint32(), ct.c_uint32()
    rc = lib.psf_cache_read_header(os.fsencode(path), ct.byref(d), ct.byref(hp), ct.byref(hg))
    if rc != 0:
        raise L.PsfError(rc, f"psf_cache_read_header({path})")
    return d, bool(hp.value), bool(hg.value)


def read_cache(path):
    """dict(lam, C, Phi or None, gravity or None, desc) of a scene cache file,
    checked against its checksum."""
    d, has_phi, has_g = read_cache_header(path)
    K = int(d.K)
    nF = Grid(d.nx, d.ny, d.nz).nF
    lam = np.empty(K)
    C = np.empty((K, K, K))
    P = np.empty((K, nF)) if has_phi else None
    g = np.empty(K) if has_g else None
    rc = L.load().psf_cache_read(os.fsencode(path), None if P is None else _ptr(P), _ptr(lam),
                                 None if g is None else _ptr(g), _ptr(C))
    if rc != 0:
        raise L.PsfError(rc, f"psf_cache_read({path})")
    return dict(desc=d, lam=lam, C=C, Phi=P, gravity=g)


def eigensolve_csr(ctx: Context, A, K: int, guard: int = 0, degree: int = 20, tol: float = 1e-11,
                   max_products: int = 20000, seed: int = 0):
    """The K smallest nonzero eigenpairs of the symmetric positive
    semi-definite sparse A (anything scipy.sparse.csr_matrix accepts) on the
    context's GPU (psf_eigensolve_csr: Chebyshev-filtered subspace iteration,
    replacing the shift-invert eigensolve of psfsolver.cpp:267-341).
    Returns (lam[K] ascending, X[n, K] orthonormal columns, products with A).
    guard = 0 picks max(16, K // 2) extra block vectors."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("A must be square")
    if guard <= 0:
        guard = max(16, K // 2)
    rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(A.indices, dtype=np.int32)
    av = np.ascontiguousarray(A.data, dtype=np.float64)
    lam = np.empty(K)
    X = np.empty((K, n))                 # column-major n x K
    used = ct.c_uint32(0)
    rc = L.load().psf_eigensolve_csr(ctx.h, n, _ptr(rp), _ptr(ci), _ptr(av), K, guard, degree, tol, max_products,
                                     seed, _ptr(lam), _ptr(X), ct.byref(used))
    ctx.check(rc, "psf_eigensolve_csr")
    return lam, X.T, int(used.value)


class Scene:
    """One scene resident on the device (C, Phi, omega, u, particles)."""

    def __init__(self, ctx: Context, desc):
        self.ctx = ctx
        self.desc = desc
        self.K = int(desc.K)
        self.grid = Grid(desc.nx, desc.ny, desc.nz)
        h = ct.c_void_p()
        if not ctx.h:
            raise L.PsfError(L.PSF_ERR_STATE, "context is closed")
        ctx.check(ctx.lib.psf_scene_create(ctx.h, ct.byref(desc), ct.byref(h)), "psf_scene_create")
        self.h = h
        self.n_particles = 0
        ctx.scenes.add(self)

    def _c(self, rc, what):
        self.ctx.check(rc, what)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.ctx.lib.psf_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # uploads
    def upload_basis(self, Phi_cm, lam, gravity=None):
        """Phi_cm (K, F) float64, or None for a reduced-step-only scene."""
        Phi = None
        if Phi_cm is not None:
            Phi = np.ascontiguousarray(Phi_cm, dtype=np.float64)
            assert Phi.shape == (self.K, self.grid.nF)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = None if gravity is None else np.ascontiguousarray(gravity, dtype=np.float64)
        self._c(self.ctx.lib.psf_scene_upload_basis(self.h, None if Phi is None else _ptr(Phi), _ptr(lam),
                                                    None if g is None else _ptr(g)), "upload_basis")

    def upload_advection(self, C):
        C = np.ascontiguousarray(C, dtype=np.float64)
        assert C.shape == (self.K, self.K, self.K)
        self._c(self.ctx.lib.psf_scene_upload_advection(self.h, _ptr(C)), "upload_advection")

    def build_advection(self, Phi_cm, V_cm, lam, terms):
        """C from any basis on the device (psf_scene_build_advection).  Phi_cm
        (K, F) and V_cm = curl Phi (K, E) as basis.py returns them; terms =
        (edge, face a, face b, coefficient) arrays, one entry per (inside
        voxel, edge slot), as advection.wedge_terms and scale_sh give them."""
        Phi = np.ascontiguousarray(Phi_cm, dtype=np.float64)     # (K, F): column-major F x K
        V = np.ascontiguousarray(V_cm, dtype=np.float64)         # (K, E): column-major E x K
        assert Phi.shape == (self.K, self.grid.nF) and V.shape[0] == self.K
        e, a, b, c = terms
        e = np.ascontiguousarray(e, dtype=np.uint32)
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        c = np.ascontiguousarray(c, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._c(self.ctx.lib.psf_scene_build_advection(self.h, _ptr(Phi), _ptr(V), V.shape[1], _ptr(lam), e.size,
                                                       _ptr(e), _ptr(a), _ptr(b), _ptr(c)), "build_advection")

    def get_advection(self):
        """C (K, K, K) in the upload layout; FP64 and FP32 scenes only."""
        C = np.empty((self.K, self.K, self.K))
        self._c(self.ctx.lib.psf_scene_get_advection(self.h, _ptr(C)), "get_advection")
        return C

    def generate_box_basis(self, table, gravity=None):
        """Phi, lambda (and gravity) of a closed box's analytic modes, written on
        the device from the normalised mode table (columns lambda, a, b, c and
        the x-, y-, z-face amplitudes, as basis.analytic_box_modes returns it)."""
        tab = np.asarray(table, np.float64)
        freqs = np.ascontiguousarray(tab[:, 1:4].astype(np.int32))
        amps = np.ascontiguousarray(tab[:, 4:7])
        lam = np.ascontiguousarray(tab[:, 0])
        g = None if gravity is None else np.ascontiguousarray(gravity, np.float64)
        self._c(self.ctx.lib.psf_scene_generate_box_basis(self.h, _ptr(freqs), _ptr(amps), _ptr(lam),
                                                          None if g is None else _ptr(g)), "generate_box_basis")

    def build_box_advection(self, table):
        """C of the same modes, built on the device in the scene's c_precision."""
        tab = np.asarray(table, np.float64)
        freqs = np.ascontiguousarray(tab[:, 1:4].astype(np.int32))
        amps = np.ascontiguousarray(tab[:, 4:7])
        lam = np.ascontiguousarray(tab[:, 0])
        self._c(self.ctx.lib.psf_scene_build_box_advection(self.h, _ptr(freqs), _ptr(amps), _ptr(lam)),
                "build_box_advection")

    def load_cache(self, path):
        """Phi, lambda, g and C from a scene cache file (psf_scene_load_cache)."""
        self._c(self.ctx.lib.psf_scene_load_cache(self.h, os.fsencode(path)), f"load_cache({path})")

    def check_signs(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        self._c(self.ctx.lib.psf_scene_check_signs(self.h, _ptr(w), w.size), "check_signs")

    def set_params(self, dt, viscosity=0.0, gravity_active=False, lifetime=60.0):
        self._c(self.ctx.lib.psf_set_params(self.h, float(dt), float(viscosity), int(bool(gravity_active)),
                                            float(lifetime)), "set_params")

    def set_wait_limit(self, seconds):
        """Watchdog of the persistent exchange (psf_set_wait_limit; 0 = default 2 s)."""
        self._c(self.ctx.lib.psf_set_wait_limit(self.h, float(seconds)), "set_wait_limit")

    def set_path(self, path):
        self._c(self.ctx.lib.psf_set_step_path(self.h, int(path)), "set_step_path")

    def set_omega(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == (self.K,)
        self._c(self.ctx.lib.psf_set_omega(self.h, _ptr(w)), "set_omega")

    def get_omega(self):
        w = np.zeros(self.K)
        self._c(self.ctx.lib.psf_get_omega(self.h, _ptr(w)), "get_omega")
        return w

    def set_particles(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.ndim == 2 and p.shape[1] == 4
        self._c(self.ctx.lib.psf_set_particles(self.h, _ptr(p), p.shape[0]), "set_particles")
        self.n_particles = p.shape[0]

    def get_particles(self):
        p = np.zeros((self.n_particles, 4), np.float32)
        self._c(self.ctx.lib.psf_get_particles(self.h, _ptr(p), self.n_particles), "get_particles")
        return p

    # hot path
    def step(self, n=1):
        self._c(self.ctx.lib.psf_step(self.h, int(n)), "step")

    def reconstruct(self):
        self._c(self.ctx.lib.psf_reconstruct(self.h), "reconstruct")

    def get_velocity(self):
        u = np.zeros(self.grid.nF, np.float32)
        self._c(self.ctx.lib.psf_get_velocity(self.h, _ptr(u)), "get_velocity")
        return u

    def trace(self, cells=False):
        if cells:
            c = np.zeros((self.n_particles, 12), np.int32)
            self._c(self.ctx.lib.psf_trace(self.h, _ptr(c)), "trace")
            return c
        self._c(self.ctx.lib.psf_trace(self.h, None), "trace")
        return None

    def integrate(self, nframes=1):
        self._c(self.ctx.lib.psf_integrate(self.h, int(nframes)), "integrate")

    def set_frame_window(self, frames):
        """Frames whose reduced steps and reconstructions are computed together
        (1 = one frame at a time)."""
        self._c(self.ctx.lib.psf_set_frame_window(self.h, int(frames)), "set_frame_window")

    def counters(self) -> dict:
        """What the scene's calls have queued (psf_scene_get_counters)."""
        c = L.SceneCounters()
        self._c(self.ctx.lib.psf_scene_get_counters(self.h, ct.byref(c)), "get_counters")
        return {name: int(getattr(c, name)) for name, _ in L.SceneCounters._fields_}

    def region_kernels(self) -> dict:
        """Kernel time by kind inside the last PHASE_FRAME_TIMED region, each
        launch bracketed by events on its own stream (psf_scene_region_kernels)."""
        r = L.RegionKernels()
        self._c(self.ctx.lib.psf_scene_region_kernels(self.h, ct.byref(r)), "region_kernels")
        return {name: (float(getattr(r, name)) if name.endswith("_ms") else int(getattr(r, name)))
                for name, _ in L.RegionKernels._fields_}

    def device_view(self) -> dict:
        """Device addresses of the scene's buffers (psf_scene_device_view)."""
        v = L.SceneView()
        self._c(self.ctx.lib.psf_scene_device_view(self.h, ct.byref(v)), "device_view")
        return {name: getattr(v, name) for name, _ in L.SceneView._fields_}

    def set_density_grid(self, dims, count_in_trace=True):
        """Particle density grid of dims = (rx, ry, rz) cells over the AABB
        ((0, 0, 0) removes it); count_in_trace makes every particle pass count
        the particles' new positions."""
        rx, ry, rz = (int(v) for v in dims)
        self._c(self.ctx.lib.psf_set_density_grid(self.h, rx, ry, rz, 1 if count_in_trace else 0),
                "set_density_grid")
        self._dens = (rx, ry, rz)

    def density(self, counts=True, volume=True):
        """Density of the stored particles: (counts uint32, volume float16), each
        shaped (rz, ry, rx), or None where not asked for."""
        rx, ry, rz = self._dens
        c = np.empty((rz, ry, rx), np.uint32) if counts else None
        v = np.empty((rz, ry, rx), np.uint16) if volume else None
        self._c(self.ctx.lib.psf_density(self.h, None if c is None else _ptr(c), None if v is None else _ptr(v)),
                "density")
        return c, (None if v is None else v.view(np.float16))

    def time_phase(self, phase, count):
        ms = ct.c_double()
        self._c(self.ctx.lib.psf_time_phase(self.h, int(phase), int(count), ct.byref(ms)), "time_phase")
        return float(ms.value)


class Server:
    """Continuous batching of independent scenes on one context (psf_server_*).

    add_basis() registers one grid with K0 modes (Phi, lambda, gravity, C);
    submit() queues a request on the first K modes of a basis; round()
    admits queued requests and advances every active scene by up to `window`
    frames; result() returns omega, particles and the latency of a finished
    request."""

    def __init__(self, ctx: Context, window=32, max_active=32, mem_budget=0, modes_per_wg=0, u_fp32_sums=False,
                 shared_basis=False, wait_limit_ns=0):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = L.ServerConfig()
        cfg.window, cfg.max_active, cfg.mem_budget, cfg.modes_per_wg = window, max_active, mem_budget, modes_per_wg
        cfg.u_fp32_sums = 1 if u_fp32_sums else 0
        cfg.shared_basis = 1 if shared_basis else 0
        cfg.wait_limit_ns = int(wait_limit_ns)
        h = ct.c_void_p()
        ctx.check(self.lib.psf_server_create(ctx.h, ct.byref(cfg), ct.byref(h)), "psf_server_create")
        self.h = h
        self.window = window
        self._n = {}
        self._K = {}
        ctx.scenes.add(self)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.psf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_basis(self, grid: Grid, resolution, Phi_cm, lam, gravity, C, aabb_min=(-0.5, -0.5, -0.5),
                  aabb_max=(0.5, 0.5, 0.5)) -> int:
        K0 = int(np.asarray(lam).size)
        d = make_desc(grid, resolution, K0, L.C_BF16X2, L.PHI_FP32, 0, aabb_min, aabb_max)
        Phi = np.ascontiguousarray(Phi_cm, np.float64)
        lam = np.ascontiguousarray(lam, np.float64)
        g = np.ascontiguousarray(gravity if gravity is not None else np.zeros(K0), np.float64)
        C = np.ascontiguousarray(C, np.float64)
        assert Phi.shape == (K0, grid.nF) and C.size == K0 ** 3
        bid = ct.c_uint32()
        self.ctx.check(self.lib.psf_server_add_basis(self.h, ct.byref(d), _ptr(Phi), _ptr(lam), _ptr(g), _ptr(C),
                                                     ct.byref(bid)), "psf_server_add_basis")
        return int(bid.value)

    def footprint(self, basis, K, n_particles) -> int:
        out = ct.c_uint64()
        self.ctx.check(self.lib.psf_server_request_footprint(self.h, basis, K, n_particles, ct.byref(out)),
                       "psf_server_request_footprint")
        return int(out.value)

    def submit(self, basis, omega, particles, nframes, dt=0.01, viscosity=0.0, gravity_active=False,
               lifetime=60.0) -> int:
        w = np.ascontiguousarray(omega, np.float64)
        p = np.ascontiguousarray(particles, np.float32).reshape(-1, 4)
        rid = ct.c_uint64()
        self.ctx.check(self.lib.psf_server_submit(self.h, basis, w.size, int(nframes), float(dt), float(viscosity),
                                                  1 if gravity_active else 0, float(lifetime), _ptr(w), _ptr(p),
                                                  p.shape[0], ct.byref(rid)), "psf_server_submit")
        self._n[rid.value] = p.shape[0]
        self._K[rid.value] = w.size
        return int(rid.value)

    def round(self) -> int:
        c = ct.c_uint32()
        self.ctx.check(self.lib.psf_server_round(self.h, ct.byref(c)), "psf_server_round")
        return int(c.value)

    def done(self, rid) -> bool:
        rc = self.lib.psf_server_done(self.h, ct.c_uint64(rid))
        if rc < 0:
            self.ctx.check(rc, "psf_server_done")
        return rc == 1

    def completed(self, capacity=1024) -> list:
        """Ids of the requests that finished since the last call (done or
        failed), in the order they finished."""
        ids = (ct.c_uint64 * capacity)()
        n = ct.c_uint32()
        out = []
        while True:
            self.ctx.check(self.lib.psf_server_completed(self.h, ids, capacity, ct.byref(n)), "psf_server_completed")
            out.extend(int(ids[j]) for j in range(n.value))
            if n.value < capacity:
                return out

    def result(self, rid):
        """(omega, particles, latency_ms) of a finished request."""
        w = np.empty(self._K[rid], np.float64)
        p = np.empty((self._n[rid], 4), np.float32)
        lat = ct.c_double()
        self.ctx.check(self.lib.psf_server_result(self.h, ct.c_uint64(rid), _ptr(w), _ptr(p), ct.byref(lat)),
                       "psf_server_result")
        del self._n[rid], self._K[rid]
        return w, p, float(lat.value)

    def stats(self) -> dict:
        st = L.ServerStats()
        self.ctx.check(self.lib.psf_server_get_stats(self.h, ct.byref(st)), "psf_server_get_stats")
        return {name: getattr(st, name) for name, _ in L.ServerStats._fields_}
