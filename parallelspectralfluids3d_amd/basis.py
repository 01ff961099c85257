"""Velocity bases (Phi, lambda) and the derived inputs of the step.

Two sources of Phi, both F x K, stored here as (K, F) arrays so that
Phi_cm[k] is column k of the reference's column-major Eigen matrix
(AbstractSolver::velBasisField, abstractsolver.h:123):

* analytic_box_modes -- discrete sin/cos modes of a closed box.  They are
  exact eigenvectors of the DEC 1-form Laplacian M = d1^T He d1 Hf
  (psfsolver.cpp:242-265) with zero Hodge weight on wall edges (free slip);
  the reference weights wall edges by h and h/2 (dec.cpp:109-128), so for it
  they are close but not exact eigenvectors.  No eigensolve, any size.
* laplacian_eigenbasis -- the K smallest positive eigenpairs of the
  reference's own pruned M, by a sparse eigensolve (scipy), for masked scenes
  of moderate size.

Fluxes are measured along each face's cycle orientation (+x, +y, +z for
interior faces), the orientation in which derivative2 (dec.cpp:351-378) is
the divergence.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid


# ---------------------------------------------------------------------------
def _box_mode_list(g: Grid, K=None):
    """(lambda, a, b, c, pol) of divergence-free box modes, sorted; with K,
    at least the K lowest (all modes whose indices stay below a bound M that
    is raised until every mode beyond it has a larger eigenvalue)."""
    na, nb, nc = g.nx, g.ny, g.nz
    sa = np.sin(np.arange(na) * np.pi / (2 * na))
    sb = np.sin(np.arange(nb) * np.pi / (2 * nb))
    sc = np.sin(np.arange(nc) * np.pi / (2 * nc))
    M = 8 if K is not None else max(na, nb, nc)
    while True:
        ma, mb, mc = min(M, na), min(M, nb), min(M, nc)
        modes = []
        for a in range(ma):
            for b in range(mb):
                for c in range(mc):
                    zeros = (a == 0) + (b == 0) + (c == 0)
                    if zeros >= 2:
                        continue
                    lam = 4.0 * (sa[a] ** 2 + sb[b] ** 2 + sc[c] ** 2)
                    npol = 2 if zeros == 0 else 1
                    for pol in range(npol):
                        modes.append((lam, a, b, c, pol))
        modes.sort(key=lambda m: (round(m[0], 12), m[1], m[2], m[3], m[4]))
        complete = ma == na and mb == nb and mc == nc
        if complete or K is None:
            return modes
        # a mode with some index >= M has lambda >= 4 sin^2(M pi / 2n) on that axis
        bound = min(4.0 * np.sin(min(M, n) * np.pi / (2 * n)) ** 2 for n in (na, nb, nc) if M < n)
        if len(modes) >= K and modes[K - 1][0] < bound:
            return modes
        M *= 2


def _polarisation(kappa, pol, zero_axis):
    kap = np.asarray(kappa, dtype=np.float64)
    if zero_axis is not None:
        t = np.zeros(3); t[zero_axis] = 1.0
        e = np.cross(kap, t)
        return e / np.linalg.norm(e)
    t = np.zeros(3); t[int(np.argmin(np.abs(kap)))] = 1.0
    e1 = np.cross(kap, t); e1 /= np.linalg.norm(e1)
    if pol == 0:
        return e1
    e2 = np.cross(kap, e1)
    return e2 / np.linalg.norm(e2)


def box_mode_table(g: Grid, K: int):
    """(K, 7) array: lambda, a, b, c, A, B, C (unnormalised polarisation)."""
    modes = _box_mode_list(g, K)
    if K > len(modes):
        raise ValueError(f"box {g.nx}x{g.ny}x{g.nz} has only {len(modes)} divergence-free modes")
    out = np.zeros((K, 7))
    for m, (lam, a, b, c, pol) in enumerate(modes[:K]):
        kap = (np.sin(a * np.pi / (2 * g.nx)), np.sin(b * np.pi / (2 * g.ny)), np.sin(c * np.pi / (2 * g.nz)))
        zero_axis = 0 if a == 0 else (1 if b == 0 else (2 if c == 0 else None))
        e = _polarisation(kap, pol, zero_axis)
        out[m] = (lam, a, b, c, e[0], e[1], e[2])
    return out


def mode_profiles(n: int, freq: int):
    """1-D factors: node sine (n+1 values, exact zeros at both walls) and cell cosine (n)."""
    i = np.arange(n + 1)
    s = np.sin(freq * np.pi * i / n)
    s[0] = 0.0
    s[n] = 0.0
    c = np.cos(freq * np.pi * (np.arange(n) + 0.5) / n)
    return s, c


def box_mode_table_normalised(g: Grid, K: int):
    """The mode table of analytic_box_modes (amplitudes of the unit-norm
    modes) without forming Phi: a mode's squared norm is a sum over the three
    face families of products of 1-D sums."""
    tab = box_mode_table(g, K)
    for m in range(K):
        a, b, c = int(tab[m, 1]), int(tab[m, 2]), int(tab[m, 3])
        sx, cx = mode_profiles(g.nx, a)
        sy, cy = mode_profiles(g.ny, b)
        sz, cz = mode_profiles(g.nz, c)
        A, B, C = tab[m, 4:7]
        n2 = (C * C * (sz @ sz) * (cy @ cy) * (cx @ cx) + B * B * (cz @ cz) * (sy @ sy) * (cx @ cx) +
              A * A * (cz @ cz) * (cy @ cy) * (sx @ sx))
        tab[m, 4:7] /= np.sqrt(n2)
    return tab


def box_mode_values(g: Grid, table, faces):
    """Phi[:, faces] (K, len(faces)) of the modes of a normalised mode table,
    evaluated at the given faces only."""
    faces = np.asarray(faces, dtype=np.int64)
    tab = np.asarray(table)
    out = np.zeros((tab.shape[0], faces.size))
    nzf, nyf = g.nZF, g.nYF
    for m in range(tab.shape[0]):
        a, b, c = int(tab[m, 1]), int(tab[m, 2]), int(tab[m, 3])
        A, B, C = tab[m, 4:7]
        sx, cx = mode_profiles(g.nx, a)
        sy, cy = mode_profiles(g.ny, b)
        sz, cz = mode_profiles(g.nz, c)
        for j, f in enumerate(faces):
            if f < nzf:
                z, r = divmod(int(f), g.nx * g.ny); y, x = divmod(r, g.nx)
                out[m, j] = C * ((sz[z] * cy[y]) * cx[x])
            elif f < nzf + nyf:
                z, r = divmod(int(f) - nzf, (g.ny + 1) * g.nx); y, x = divmod(r, g.nx)
                out[m, j] = B * ((cz[z] * sy[y]) * cx[x])
            else:
                z, r = divmod(int(f) - nzf - nyf, g.ny * (g.nx + 1)); y, x = divmod(r, g.nx + 1)
                out[m, j] = A * ((cz[z] * cy[y]) * sx[x])
    return out


def analytic_box_modes(g: Grid, K: int, dtype=np.float64):
    """Phi_cm (K, F), lambda (K,), mode table.  Each column has unit 2-norm."""
    tab = box_mode_table(g, K)
    Phi = np.zeros((K, g.nF), dtype=dtype)
    nzf, nyf = g.nZF, g.nYF
    for m in range(K):
        lam, a, b, c, A, B, C = tab[m]
        a, b, c = int(a), int(b), int(c)
        sx, cx = mode_profiles(g.nx, a)
        sy, cy = mode_profiles(g.ny, b)
        sz, cz = mode_profiles(g.nz, c)
        # Z faces (z node, y cell, x cell) stored [z][y][x]
        vz = C * np.multiply.outer(np.multiply.outer(sz, cy), cx)
        # Y faces (z cell, y node, x cell)
        vy = B * np.multiply.outer(np.multiply.outer(cz, sy), cx)
        # X faces (z cell, y cell, x node)
        vx = A * np.multiply.outer(np.multiply.outer(cz, cy), sx)
        col = np.concatenate([vz.reshape(-1), vy.reshape(-1), vx.reshape(-1)])
        nrm = np.sqrt(np.dot(col, col))
        Phi[m] = col / nrm
        tab[m, 4:7] /= nrm
    return Phi, tab[:, 0].copy(), tab


# ---------------------------------------------------------------------------
def hodge_primal(g: Grid, resolution: float):
    """Hf = 1/(2h) on every face (dec.cpp:143-193 evaluates to this), h the
    mesh's float resolution."""
    h = float(np.float32(resolution))
    return np.full(g.nF, 1.0 / (2.0 * h))


def hodge_dual(g: Grid, resolution: float, inside):
    """He on inside edges: {h/2, h, 2h} for 2, 3, 4 adjacent faces (dec.cpp:109-128)."""
    h = float(np.float32(resolution))
    _, cnt = g.edge_faces
    val = np.select([cnt == 2, cnt == 3, cnt == 4], [1.0, 2.0, 4.0], 0.0) * (h / 2.0) * (h / 2.0) / (h / 2.0)
    return np.where(g.edge_inside(inside), val, 0.0)


def derivative1(g: Grid):
    """d1: E x F, entry sign(edge k of face f) (dec.cpp:381-455)."""
    import scipy.sparse as sp
    rows = g.face_edges.reshape(-1)
    cols = np.repeat(np.arange(g.nF), 4)
    vals = g.face_edge_signs.reshape(-1).astype(np.float64)
    return sp.csr_matrix((vals, (rows, cols)), shape=(g.nE, g.nF))


def curl_operator(g: Grid, resolution: float):
    """curl = d1 * Hf  (psfsolver.cpp:252)."""
    import scipy.sparse as sp
    return (derivative1(g) @ sp.diags(hodge_primal(g, resolution))).tocsr()


def laplacian(g: Grid, resolution: float, inside):
    """M = d1^T He d1 Hf with the rows of faces that do not have exactly two
    inside voxels removed (psfsolver.cpp:247-265)."""
    import scipy.sparse as sp
    d1 = derivative1(g)
    M = (d1.T @ sp.diags(hodge_dual(g, resolution, inside)) @ d1 @ sp.diags(hodge_primal(g, resolution))).tocsr()
    keep = (g.face_inside_count(inside) == 2).astype(np.float64)
    M = (sp.diags(keep) @ M).tocsr()
    M.eliminate_zeros()
    return M


def laplacian_eigenbasis(g: Grid, resolution: float, inside, K: int, penalty=None, tol=1e-12, extra=8):
    """The K smallest eigenpairs of the pruned M with eigenvalue above 1e-10
    (the acceptance filter of psfsolver.cpp:289-290), restricted to the faces
    between two inside voxels (on all other faces eigenvectors of nonzero
    eigenvalue vanish).

    M's exact null space contains the gradient D^T e_c of every inside voxel c
    whose six faces all lie between inside voxels; those directions are lifted
    by adding penalty * D_c^T D_c, which leaves every eigenvector of nonzero
    eigenvalue unchanged (they are orthogonal to the null space).  With the
    reference's wall Hodge weights (dec.cpp:109-128) the remaining spectrum is
    not exactly divergence-free, and the basis keeps that property."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    inside = np.asarray(inside, dtype=bool).reshape(-1)
    MII, act = pruned_laplacian(g, resolution, inside)
    MII = MII.tocsc()
    D = deep_divergence(g, inside, act)
    if penalty is None:
        penalty = 4.0 * max(1.0, float(np.abs(MII).sum(1).max()))
    A = (MII + penalty * (D.T @ D)).tocsc()
    vals, vecs = sla.eigsh(A, k=min(K + extra, act.size - 1), sigma=0.0, which="LM", tol=tol)
    order = np.argsort(vals)
    vals, vecs = vals[order], vecs[:, order]
    keep = vals > 1e-10
    vals, vecs = vals[keep], vecs[:, keep]
    if vals.size < K:
        raise RuntimeError("eigensolve returned fewer than K accepted eigenpairs")
    vals, vecs = vals[:K], vecs[:, :K]
    res = np.linalg.norm(MII @ vecs - vecs * vals, axis=0)
    if np.any(res > 1e-8 * max(1.0, float(vals.max()))):
        raise RuntimeError(f"eigenpair residual {res.max():.3e}: the penalty left a null direction in range")
    return _basis_from_active(g, act, vecs, vals)


def pruned_laplacian(g: Grid, resolution: float, inside):
    """(M_II, act): M (psfsolver.cpp:247) restricted to the faces between two
    inside voxels (the rows psfsolver.cpp:251-265 keep), symmetrised, as CSR,
    and the indices of those faces."""
    inside = np.asarray(inside, dtype=bool).reshape(-1)
    M = laplacian(g, resolution, inside)
    act = np.flatnonzero(g.face_inside_count(inside) == 2)
    MII = M[act][:, act]
    return ((MII + MII.T) * 0.5).tocsr(), act


def deep_divergence(g: Grid, inside, act):
    """D: the signed face sums of every inside voxel whose six faces all lie
    between inside voxels, on the active faces act.  M_II D^T = 0: its rows'
    transposes (voxel gradients) span M_II's null space apart from harmonic
    fields, and every eigenvector of nonzero eigenvalue has D v = 0."""
    import scipy.sparse as sp
    inside = np.asarray(inside, dtype=bool).reshape(-1)
    cnt = g.face_inside_count(inside)
    vf_all = g.voxel_faces
    deep = inside & (cnt[vf_all] == 2).all(1)
    vf = vf_all[deep]
    vs = g.voxel_face_signs[deep].astype(np.float64)
    return sp.csr_matrix((vs.reshape(-1), (np.repeat(np.arange(vf.shape[0]), 6), vf.reshape(-1))),
                         shape=(vf.shape[0], g.nF))[:, act]


def _basis_from_active(g: Grid, act, vecs, vals):
    """Phi (K x F) from eigenvectors on the active faces, zero elsewhere, each
    mode signed so that its largest-magnitude entry is positive."""
    K = vals.size
    Phi = np.zeros((K, g.nF))
    Phi[:, act] = vecs.T
    idx = np.argmax(np.abs(Phi), axis=1)
    Phi *= np.sign(Phi[np.arange(K), idx])[:, None]
    return Phi, vals


def _sparse_times_dense(S, V, workers=None):
    """S @ V with V's columns split over host threads (scipy's sparse times
    dense product runs on one core and releases the GIL)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    w = max(1, min(16, os.cpu_count() or 1, V.shape[1])) if workers is None else workers
    if w == 1:
        return S @ V
    cols = np.array_split(np.arange(V.shape[1]), w)
    with ThreadPoolExecutor(w) as ex:
        parts = list(ex.map(lambda c: S @ np.ascontiguousarray(V[:, c]), cols))
    return np.concatenate(parts, axis=1)


def laplacian_eigenbasis_device(g: Grid, resolution: float, inside, K: int, ctx, tol=1e-11, degree=80,
                                guard=0, seed=0, penalty=16.0):
    """laplacian_eigenbasis on the GPU: hip.eigensolve_csr on
    A = M_II + penalty * D^T D, which lifts M_II's gradient null space and
    leaves every eigenpair of nonzero eigenvalue (D v = 0) unchanged.  A
    modest penalty keeps A's spectrum short (the filter's cost grows with its
    square root).  A returned mode with D v != 0 is a lifted gradient: then the
    penalty is raised fourfold and the solve repeated.  Same acceptance
    (eigenvalues above 1e-10, psfsolver.cpp:289-290), residual check against
    M_II and sign convention as the host path.
    Returns (Phi, lam, dict(products=..., penalty=..., attempts=..., history=[per solve]))."""
    from . import hip
    MII, act = pruned_laplacian(g, resolution, inside)
    D = deep_divergence(g, inside, act)
    DtD = (D.T @ D).tocsr()
    # null vectors the penalty does not lift (gradients of voxels next to the
    # pruned faces, harmonic fields) come back with eigenvalue ~0 and are
    # dropped, as the reference's acceptance test drops them: ask for a few
    # more than K, and for more again if that was not enough
    extra = max(8, K // 8)
    # rows in spatial order (z, y, x of the face), so that a row's neighbours
    # sit near it in the block and the products reuse cached rows; the Z|Y|X
    # face numbering would put them a third of the faces apart
    fc = g.face_coords[act]
    perm = np.lexsort((fc[:, 0], fc[:, 1], fc[:, 2]))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    products = 0
    attempt = 0
    history = []
    while True:
        attempt += 1
        if attempt > 6:
            raise RuntimeError("no K accepted divergence-free eigenpairs after raising the penalty and the count")
        A = (MII + penalty * DtD).tocsr()[perm][:, perm]
        vals, vp, used = hip.eigensolve_csr(ctx, A, K + extra, guard=guard, degree=degree, tol=tol, seed=seed)
        vecs = np.take(vp, inv, axis=0)   # vecs[perm] = vp, as one gather
        products += used
        keep = vals > 1e-10
        step = dict(modes=K + extra, penalty=penalty, products=used, accepted=int(keep.sum()))
        history.append(step)
        if keep.sum() < K:
            extra += K - int(keep.sum()) + 8
            continue
        vals, vecs = vals[keep][:K], vecs[:, keep][:, :K]
        step["max_div"] = float(np.abs(_sparse_times_dense(D, vecs)).max())
        if step["max_div"] <= 1e-6:
            break
        penalty *= 4.0
    res = np.linalg.norm(_sparse_times_dense(MII, vecs) - vecs * vals, axis=0)
    if np.any(res > 1e-8 * max(1.0, float(vals.max()))):
        raise RuntimeError(f"eigenpair residual {res.max():.3e}")
    Phi, lam = _basis_from_active(g, act, vecs, vals)
    return Phi, lam, dict(products=products, penalty=penalty, attempts=attempt, history=history)


# ---------------------------------------------------------------------------
def gravity_field(g: Grid, inside):
    """face field of psfsolver.cpp:344-357: +-0.00981 on inside faces with a y normal."""
    n = g.face_normals
    f = np.zeros(g.nF)
    fin = g.face_inside(inside)
    ny = n[:, 1]
    sel = fin & (np.abs(ny) > np.finfo(np.float64).eps)
    f[sel] = -np.where(ny[sel] > 0, -1.0, 1.0) * 0.00981
    return f


def project(Phi_cm, field):
    """Phi^T field (K,)."""
    return Phi_cm @ field


def unit_energy_omega(K: int, seed: int):
    """omega ~ N(0,1) from numpy's PCG64 with the given seed, scaled to |omega| = 1."""
    w = np.random.default_rng(seed).standard_normal(K)
    return w / np.linalg.norm(w)


def power_law_omega(K: int, seed: int, exponent=-5.0 / 6.0):
    """|omega_k| ~ (k+1)^exponent with random signs, unit energy."""
    rng = np.random.default_rng(seed)
    w = (np.arange(1, K + 1, dtype=np.float64) ** exponent) * rng.choice([-1.0, 1.0], size=K)
    return w / np.linalg.norm(w)
