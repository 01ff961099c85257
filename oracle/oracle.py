"""ctypes front end of the CPU oracle (oracle/psf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.  The product package never imports this module.

Array conventions follow the reference's Eigen storage:
  Phi   F x K column-major  (AbstractSolver::velBasisField, abstractsolver.h:123)
  V     E x K column-major  (vortBasisField = curl * Phi, psfsolver.cpp:342)
  C     K matrices K x K, each column-major, matrix i at C[i]  (advection,
        abstractsolver.h:118); as a numpy array C[i, b, a] == C_i(a, b).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpsf_oracle.so")
_lib = None

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class ParticleParams(ct.Structure):
    _fields_ = [("aabb_min", ct.c_float * 3), ("aabb_max", ct.c_float * 3),
                ("resolution", ct.c_double), ("timestep", ct.c_double),
                ("lifetime", ct.c_float), ("seed", ct.c_uint32), ("step", ct.c_uint32)]


def build() -> str:
    """Compile the oracle (make -C oracle); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        vp = ct.c_void_p
        L.ora_mesh_create.restype = vp
        L.ora_mesh_create.argtypes = [ct.c_int] * 3 + [ct.c_float] * 4
        L.ora_mesh_destroy.argtypes = [vp]
        L.ora_mesh_set_voxel_inside.argtypes = [vp, ct.c_uint]
        L.ora_mesh_counts.argtypes = [vp, _u32p]
        L.ora_mesh_sign_words.argtypes = [vp, _u32p]
        L.ora_mesh_voxels.argtypes = [vp, _i32p, _i32p, _f64p]
        L.ora_mesh_faces.argtypes = [vp, _i32p, _f64p, _f64p, _i32p]
        L.ora_mesh_edges.argtypes = [vp, _i32p, _i32p, _u32p, _f64p, _i32p]
        L.ora_mesh_points.argtypes = [vp, _i32p, _u32p, _f64p, _i32p]
        for n in ("ora_voxel_index", "ora_zface_index", "ora_yface_index", "ora_xface_index"):
            getattr(L, n).restype = ct.c_uint
            getattr(L, n).argtypes = [vp, ct.c_uint, ct.c_uint, ct.c_uint]
        L.ora_face_signum.restype = ct.c_int
        L.ora_face_signum.argtypes = [vp, ct.c_uint, ct.c_uint]
        L.ora_hodge2_primal.argtypes = [vp, _f64p]
        L.ora_hodge2_dual.argtypes = [vp, _f64p]
        L.ora_derivative1.argtypes = [vp, _i32p, _i32p, _f64p]
        L.ora_derivative2.argtypes = [vp, _i32p, _i32p, _f64p]
        L.ora_gravity_field.argtypes = [vp, _f64p]
        L.ora_build_advection.restype = ct.c_int
        L.ora_build_advection.argtypes = [vp, ct.c_int, _f64p, _f64p, _f64p, _f64p, ct.c_int]
        L.ora_reduced_step.argtypes = [ct.c_int, _f64p, _f64p, _f64p, _f64p, ct.c_double,
                                       ct.c_double, ct.c_int, vp, ct.c_int]
        L.ora_reconstruct.argtypes = [ct.c_int, ct.c_int, _f64p, _f64p, _f32p, ct.c_int]
        L.ora_respawn_draw.restype = ct.c_uint32
        L.ora_respawn_draw.argtypes = [ct.c_uint32] * 4
        L.ora_particles.argtypes = [vp, ct.POINTER(ParticleParams), _f32p, _f32p, ct.c_int, vp, ct.c_int]
        _lib = L
    return _lib


class Mesh:
    """Literal DECMesh3D (decmesh3d.cpp:62-106) plus setVoxelInside for a mask."""

    def __init__(self, dims, resolution, aabb_min, inside=None):
        self.dims = tuple(int(d) for d in dims)
        self.resolution = float(np.float32(resolution))
        self.aabb_min = np.asarray(aabb_min, dtype=np.float32)
        L = lib()
        self.h = L.ora_mesh_create(*self.dims, ct.c_float(resolution),
                                   *[ct.c_float(v) for v in self.aabb_min])
        if not self.h:
            raise ValueError("bad mesh dims")
        c = np.zeros(11, dtype=np.uint32)
        L.ora_mesh_counts(self.h, c)
        self.nV, self.nF, self.nE, self.nP, self.nSign = (int(v) for v in c[:5])
        self.nZF, self.nYF, self.nXF, self.nZE, self.nYE, self.nXE = (int(v) for v in c[5:])
        if inside is not None:
            for v in np.flatnonzero(np.asarray(inside).reshape(-1)):
                L.ora_mesh_set_voxel_inside(self.h, int(v))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_mesh_destroy(self.h)
            self.h = None

    # --- tables -------------------------------------------------------------
    def sign_words(self):
        out = np.zeros(self.nSign, dtype=np.uint32)
        lib().ora_mesh_sign_words(self.h, out)
        return out

    def voxels(self):
        f6 = np.zeros((self.nV, 6), np.int32); ins = np.zeros(self.nV, np.int32)
        c = np.zeros((self.nV, 3), np.float64)
        lib().ora_mesh_voxels(self.h, f6, ins, c)
        return f6, ins, c

    def faces(self):
        e4 = np.zeros((self.nF, 4), np.int32); n = np.zeros((self.nF, 3)); c = np.zeros((self.nF, 3))
        ins = np.zeros(self.nF, np.int32)
        lib().ora_mesh_faces(self.h, e4, n, c, ins)
        return e4, n, c, ins

    def edges(self):
        v2 = np.zeros((self.nE, 2), np.int32); f4 = np.zeros((self.nE, 4), np.int32)
        dc = np.zeros(self.nE, np.uint32); c = np.zeros((self.nE, 3)); ins = np.zeros(self.nE, np.int32)
        lib().ora_mesh_edges(self.h, v2, f4, dc, c, ins)
        return v2, f4, dc, c, ins

    def points(self):
        e6 = np.zeros((self.nP, 6), np.int32); nd = np.zeros(self.nP, np.uint32)
        c = np.zeros((self.nP, 3)); ins = np.zeros(self.nP, np.int32)
        lib().ora_mesh_points(self.h, e6, nd, c, ins)
        return e6, nd, c, ins

    def face_signum(self, vid, fidx):
        return lib().ora_face_signum(self.h, vid, fidx)

    # --- operators (dec.cpp) -------------------------------------------------
    def hodge2_primal(self):
        d = np.zeros(self.nF); lib().ora_hodge2_primal(self.h, d); return d

    def hodge2_dual(self):
        d = np.zeros(self.nE); lib().ora_hodge2_dual(self.h, d); return d

    def derivative1(self):
        """scipy CSR E x F (dec.cpp:381-455, derivative1(mesh, true))."""
        import scipy.sparse as sp
        n = 4 * self.nF
        r = np.zeros(n, np.int32); c = np.zeros(n, np.int32); v = np.zeros(n)
        lib().ora_derivative1(self.h, r, c, v)
        return sp.csr_matrix((v, (r, c)), shape=(self.nE, self.nF))

    def derivative2(self):
        """scipy CSC V x F with the reference's explicit entries kept (dec.cpp:351-378)."""
        import scipy.sparse as sp
        n = 6 * self.nV
        r = np.zeros(n, np.int32); c = np.zeros(n, np.int32); v = np.zeros(n)
        lib().ora_derivative2(self.h, r, c, v)
        m = sp.coo_matrix((v, (r, c)), shape=(self.nV, self.nF)).tocsc()
        m.sum_duplicates()
        return m

    def laplacian(self):
        """M = d1^T He d1 Hf with rows of faces that do not have exactly two
        inside voxels removed (psfsolver.cpp:242-265); also returns curl = d1 Hf."""
        import scipy.sparse as sp
        d1 = self.derivative1()
        He = sp.diags(self.hodge2_dual())
        Hf = sp.diags(self.hodge2_primal())
        M = (d1.T @ He @ d1 @ Hf).tocsr()
        d2 = self.derivative2()
        counts = np.diff(d2.indptr)          # stored entries per face column
        keep = (counts == 2).astype(np.float64)
        M = (sp.diags(keep) @ M).tocsr()
        M.eliminate_zeros()
        curl = (d1 @ Hf).tocsr()
        return M, curl, counts

    def gravity_field(self):
        f = np.zeros(self.nF); lib().ora_gravity_field(self.h, f); return f


def build_advection(mesh: Mesh, Phi_cm: np.ndarray, lam: np.ndarray, V_cm: np.ndarray, nthreads=0):
    """Literal buildAdvection (psfsolver.cpp:431-551).

    Phi_cm: (K, F) array = Phi column-major; V_cm: (K, E); returns C (K, K, K)
    with C[i, b, a] = C_i(a, b)."""
    K = Phi_cm.shape[0]
    C = np.zeros((K, K, K))
    rc = lib().ora_build_advection(mesh.h, K, np.ascontiguousarray(Phi_cm, np.float64),
                                   np.ascontiguousarray(lam, np.float64),
                                   np.ascontiguousarray(V_cm, np.float64), C, nthreads)
    if rc != 0:
        raise MemoryError("ora_build_advection: allocation failed")
    return C


def reduced_step(omega, C, lam, g, dt, nu, gravity, nthreads=0, return_y=False):
    """One step of psfsolver.cpp:34-59, in place on a copy of omega."""
    w = np.array(omega, dtype=np.float64, copy=True)
    K = w.shape[0]
    y = np.zeros(K) if return_y else None
    lib().ora_reduced_step(K, w, np.ascontiguousarray(C, np.float64), np.ascontiguousarray(lam, np.float64),
                           np.ascontiguousarray(g, np.float64), dt, nu, int(bool(gravity)),
                           y.ctypes.data if return_y else None, nthreads)
    return (w, y) if return_y else w


def reconstruct(Phi_cm, omega, nthreads=0):
    """u = float(Phi omega)  (psfsolver.cpp:60)."""
    K, F = Phi_cm.shape
    u = np.zeros(F, np.float32)
    lib().ora_reconstruct(F, K, np.ascontiguousarray(Phi_cm, np.float64),
                          np.ascontiguousarray(omega, np.float64), u, nthreads)
    return u


def particles(mesh: Mesh, u, parts, *, resolution, timestep, aabb_min, aabb_max,
              lifetime=60.0, seed=0, step=0, nthreads=0, cells=False):
    """One pass of the particle loop (psfsolver.cpp:70-232) on a copy of parts (N,4)."""
    p = np.array(parts, dtype=np.float32, copy=True, order="C")
    N = p.shape[0]
    pp = ParticleParams((ct.c_float * 3)(*np.asarray(aabb_min, np.float32)),
                        (ct.c_float * 3)(*np.asarray(aabb_max, np.float32)),
                        float(resolution), float(timestep), float(lifetime), int(seed), int(step))
    co = np.zeros((N, 12), np.int32) if cells else None
    lib().ora_particles(mesh.h, ct.byref(pp), np.ascontiguousarray(u, np.float32), p, N,
                        co.ctypes.data if cells else None, nthreads)
    return (p, co) if cells else p


def respawn_draw(seed, step, particle, k):
    return int(lib().ora_respawn_draw(seed, step, particle, k))


def density(parts, aabb_min, aabb_max, dims):
    """Particle density of the GPU solver's particle drawing
    (Solver/SolverImpl/psfsolvergpu.cpp:876-913): counts per cell of a
    dims = (nx, ny, nz) grid over the AABB, shaped (nz, ny, nx), and the r16f
    volume float(count) / 1e6 (Resources/Volume/volume.comp:14) as float16.

    The counting shader (Res/Volume/history.comp) is absent from the
    reference, so the cell rule is the library's: per axis
    trunc((p - aabb_min) / extent * n) in float32, clamped to [0, n-1], NaN to
    cell 0, with extent = aabb_max - aabb_min in float32 (AABB::getExtent)."""
    p = np.asarray(parts, np.float32).reshape(-1, 4)
    amin = np.asarray(aabb_min, np.float32)
    ext = (np.asarray(aabb_max, np.float32) - amin).astype(np.float32)
    cell = []
    for a in range(3):
        n = int(dims[a])
        r = np.float32(n)
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            t = ((p[:, a] - amin[a]) / ext[a]) * r
        c = np.zeros(p.shape[0], np.int64)
        hi = t >= r
        mid = (t > np.float32(0.0)) & ~hi
        c[hi] = n - 1
        c[mid] = np.trunc(t[mid]).astype(np.int64)
        cell.append(c)
    nx, ny, nz = (int(v) for v in dims)
    lin = (cell[2] * ny + cell[1]) * nx + cell[0]
    counts = np.bincount(lin, minlength=nx * ny * nz).astype(np.uint32).reshape(nz, ny, nx)
    volume = (counts.astype(np.float32) / np.float32(1.0e6)).astype(np.float16)
    return counts, volume
