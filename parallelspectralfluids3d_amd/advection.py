"""Structure tensor C of the reduced advection (buildAdvection, psfsolver.cpp:431-551).

For every inside voxel and each of its 12 edge slots s the reference adds

    W_i[e_s][j] += s_h (Phi[a_s][i] Phi[b_s][j] - Phi[a_s][j] Phi[b_s][i]) c(a_s, b_s)

with s_h = (h/2)^2 in float, c(a,b) = dot(cross(n_a, n_b), (1,1,1)) of the
stored face normals, and then C_i = (1/lambda_i) V^T W_i with V = curl Phi.
The reference materialises all K wedge matrices W_i (E x K each, 2*K*E*K
doubles).  Here W_i = A_i Phi with a sparse E x F matrix A_i whose entries
are the per-voxel terms, built one i at a time, so memory stays O(E K).

Output layout: C[i, b, a] = C_i(a, b), i.e. the K column-major K x K
matrices of std::vector<Eigen::MatrixXd> advection.
"""
from __future__ import annotations

import numpy as np

from .basis import _sparse_times_dense, curl_operator
from .grid import Grid

# (face slot holding the edge, edge slot in that face, face a, face b), faces 0-based f1..f6
SLOTS = [(0, 0, 0, 3), (0, 1, 5, 0), (0, 2, 2, 0), (0, 3, 0, 4),
         (1, 0, 3, 1), (1, 1, 4, 1), (1, 2, 1, 2), (1, 3, 1, 5),
         (4, 0, 4, 3), (4, 2, 2, 4), (5, 0, 3, 5), (5, 2, 5, 2)]


def wedge_terms(g: Grid, inside):
    """Per (inside voxel, slot) term: edge index, faces a and b, coefficient c."""
    inside = np.asarray(inside, dtype=bool).reshape(-1)
    vf = g.voxel_faces[inside]                      # (Vi, 6)
    fe = g.face_edges                               # (F, 4)
    n = g.face_normals
    one = np.ones(3)
    E, A, B, Cc = [], [], [], []
    for (fs, es, a, b) in SLOTS:
        E.append(fe[vf[:, fs], es])
        A.append(vf[:, a])
        B.append(vf[:, b])
        Cc.append(np.cross(n[vf[:, a]], n[vf[:, b]]) @ one)
    return (np.concatenate(E), np.concatenate(A), np.concatenate(B), np.concatenate(Cc))


def scale_sh(resolution: float):
    """(decMesh.resolution/2)^2 evaluated in float (psfsolver.cpp:444)."""
    hr = np.float32(np.float32(resolution) / np.float32(2.0))
    return float(np.float32(hr * hr))


def build_advection(g: Grid, resolution: float, inside, Phi_cm, lam, V_cm=None, progress=None):
    """C (K, K, K) for any mask; cost O(K (T K + E K^2)) with T = 12 * inside voxels."""
    import scipy.sparse as sp
    K, F = Phi_cm.shape
    if V_cm is None:
        V_cm = _sparse_times_dense(curl_operator(g, resolution), Phi_cm.T).T
    PhiT = np.ascontiguousarray(Phi_cm.T)           # F x K
    VT = np.ascontiguousarray(V_cm.T)               # E x K
    e, a, b, c = wedge_terms(g, inside)
    s = scale_sh(resolution) * c
    rows = np.concatenate([e, e])
    cols = np.concatenate([b, a])
    # coalesce duplicate (row, col) pairs once; per i only the values change
    key = rows.astype(np.int64) * F + cols
    uniq, pos = np.unique(key, return_inverse=True)
    ur, uc = uniq // F, uniq % F
    indptr = np.concatenate([[0], np.cumsum(np.bincount(ur, minlength=g.nE))])
    C = np.zeros((K, K, K))
    for i in range(K):
        vals = np.concatenate([s * PhiT[a, i], -s * PhiT[b, i]])
        data = np.bincount(pos, weights=vals, minlength=uniq.size)
        Ai = sp.csr_matrix((data, uc, indptr), shape=(g.nE, F))
        Wi = Ai @ PhiT                               # E x K
        C[i] = (Wi.T @ VT) / lam[i]                  # C[i, j, k] = C_i(k, j)
        if progress:
            progress(i)
    return C


def advection_terms(g: Grid, resolution: float, inside):
    """(edge, a, b, s_h * c) per (inside voxel, edge slot): the inputs of
    hip.Scene.build_advection, which forms C on the device."""
    e, a, b, c = wedge_terms(g, inside)
    return e, a, b, scale_sh(resolution) * c


def build_advection_device(scene, g: Grid, resolution: float, inside, Phi_cm, lam, V_cm=None):
    """build_advection on the GPU, into the scene's C (any precision but FP16)."""
    if V_cm is None:
        V_cm = _sparse_times_dense(curl_operator(g, resolution), Phi_cm.T).T
    scene.build_advection(Phi_cm, V_cm, lam, advection_terms(g, resolution, inside))


# ---------------------------------------------------------------------------
# Closed boxes with the analytic sin/cos basis.
#
# Every voxel is inside and every mode vanishes exactly on wall faces, so a
# slot term survives only when both of its faces are interior; interior faces
# are created BACK/TOP/RIGHT, which fixes each slot's normals (hence c_s) and
# its edge for all surviving terms.  Each mode is a product of 1-D profiles
# (node sine / cell cosine) along x, y and z, so the voxel sum of a slot,
#     T_s[i, j, k] = sum_v V[e_s(v), k] Phi[a_s(v), i] Phi[b_s(v), j],
# factorises into three 1-D contractions over the distinct frequency pairs of
# each axis, with V = curl Phi kept as computed (its wall edges included).

# per face slot f1..f6 of voxel (x,y,z): (face axis, offset along that axis)
_FACE_OF_SLOT = [("z", 0), ("z", 1), ("y", 0), ("y", 1), ("x", 0), ("x", 1)]
# edge of each slot for interior creators (BACK/TOP/RIGHT edge lists):
# (edge axis, offset dx, dy, dz) relative to the voxel
_SLOT_EDGE = [("x", 0, 1, 0), ("y", 0, 0, 0), ("x", 0, 0, 0), ("y", 1, 0, 0),     # f1 (BACK at z)
              ("x", 0, 1, 1), ("y", 0, 0, 1), ("x", 0, 0, 1), ("y", 1, 0, 1),     # f2 (BACK at z+1)
              ("z", 0, 1, 0), ("z", 0, 0, 0),                                     # f5 (RIGHT at x) e1, e3
              ("z", 1, 1, 0), ("z", 1, 0, 0)]                                     # f6 (RIGHT at x+1) e1, e3
_INTERIOR_NORMAL = {"z": np.array([0.0, 0.0, 1.0]), "y": np.array([0.0, -1.0, 0.0]), "x": np.array([-1.0, 0.0, 0.0])}


def _profile_table(n, freqs, kind, off):
    """values over voxel coordinate v = 0..n-1: node sine at v+off or cell cosine at v."""
    v = np.arange(n)
    out = np.empty((len(freqs), n))
    for r, f in enumerate(freqs):
        if kind == "node":
            s = np.sin(f * np.pi * (v + off) / n)
            s[(v + off) == 0] = 0.0
            s[(v + off) == n] = 0.0
            out[r] = s
        else:
            out[r] = np.cos(f * np.pi * (v + 0.5) / n)
    return out


def build_advection_box_analytic(g: Grid, resolution: float, Phi_cm, lam, table, V_cm=None, kchunk=64,
                                 progress=None):
    """C (K, K, K) for a closed box whose basis is analytic_box_modes(g, K)
    (table = its mode table, amplitudes normalised).  progress(slot, nslots),
    if given, is called after each of the 12 edge slots."""
    K = Phi_cm.shape[0]
    if V_cm is None:
        V_cm = _sparse_times_dense(curl_operator(g, resolution), Phi_cm.T).T        # (K, E)
    dims = {"x": g.nx, "y": g.ny, "z": g.nz}
    freq = {"x": table[:, 1].astype(int), "y": table[:, 2].astype(int), "z": table[:, 3].astype(int)}
    amp = {"x": table[:, 4], "y": table[:, 5], "z": table[:, 6]}
    uniq = {ax: np.unique(freq[ax]) for ax in "xyz"}
    pos = {ax: np.searchsorted(uniq[ax], freq[ax]) for ax in "xyz"}
    x, y, z = g.voxel_coords()
    one = np.ones(3)
    sh = scale_sh(resolution)
    C = np.zeros((K, K, K))
    for s, (fs, es, a, b) in enumerate(SLOTS):
        fa, fb = _FACE_OF_SLOT[a], _FACE_OF_SLOT[b]
        c_s = float(np.cross(_INTERIOR_NORMAL[fa[0]], _INTERIOR_NORMAL[fb[0]]) @ one)
        if c_s == 0.0:
            continue
        ax_e, dx, dy, dz = _SLOT_EDGE[s]
        e_idx = g.edge_index(ax_e, x + dx, y + dy, z + dz)
        # per axis: profile tables of face a and face b over voxel coordinates
        G = {}
        for ax in "xyz":
            n = dims[ax]
            ka = ("node", fa[1]) if fa[0] == ax else ("cell", 0)
            kb = ("node", fb[1]) if fb[0] == ax else ("cell", 0)
            Pa = _profile_table(n, uniq[ax], *ka)
            Pb = _profile_table(n, uniq[ax], *kb)
            G[ax] = (Pa[:, None, :] * Pb[None, :, :]).reshape(-1, n)     # (D*D, n)
        Dz, Dy, Dx = G["z"].shape[0], G["y"].shape[0], G["x"].shape[0]
        ia = amp[fa[0]]
        ib = amp[fb[0]]
        # pair index of (i, j) per axis
        du = {ax: len(uniq[ax]) for ax in "xyz"}
        pz = pos["z"][:, None] * du["z"] + pos["z"][None, :]
        py = pos["y"][:, None] * du["y"] + pos["y"][None, :]
        px = pos["x"][:, None] * du["x"] + pos["x"][None, :]
        for k0 in range(0, K, kchunk):
            k1 = min(K, k0 + kchunk)
            Vs = V_cm[k0:k1, e_idx].reshape(k1 - k0, g.nz, g.ny, g.nx)         # [k, z, y, x]
            T1 = np.tensordot(Vs, G["z"], axes=([1], [1]))                     # [k, y, x, Dz]
            T2 = np.tensordot(T1, G["y"], axes=([1], [1]))                     # [k, x, Dz, Dy]
            T3 = np.tensordot(T2, G["x"], axes=([1], [1]))                     # [k, Dz, Dy, Dx]
            T3 = T3.reshape(k1 - k0, -1)
            flat = (pz * Dy + py) * Dx + px                                    # [i, j]
            Tij = T3[:, flat]                                                  # [k, i, j]
            Tij = Tij * (ia[None, :, None] * ib[None, None, :])
            anti = Tij - np.transpose(Tij, (0, 2, 1))                          # T[i,j] - T[j,i]
            C[:, :, k0:k1] += np.transpose(anti, (1, 2, 0)) * (sh * c_s)
        if progress is not None:
            progress(s + 1, len(SLOTS))
    C /= np.asarray(lam)[:, None, None]
    return C


# ---------------------------------------------------------------------------
# Closed boxes, fully separable: C from the mode table alone.
#
# V = curl Phi is separable too.  Around an edge of axis A the curl takes four
# face fluxes with fixed offsets and signs (CURL_FACES below).  Edges on a
# low wall plane (a node coordinate 0) are oriented the other way, which
# flips their sign; edges on two low walls touch only wall faces, where every
# mode vanishes.  With every factor a product of 1-D profiles, each slot's
# voxel sum becomes, per curl term, a product over the three axes of 1-D
# triple sums S[f_k, f_i, f_j] over the distinct frequencies of that axis.
# Nothing of size F or E is formed; the cost is O(K^3) per (slot, term).

# faces around an edge of each axis: (face axis, face coordinates - edge coordinates, sign)
CURL_FACES = {"x": [("y", (0, 0, -1), -1.0), ("y", (0, 0, 0), 1.0), ("z", (0, -1, 0), 1.0), ("z", (0, 0, 0), -1.0)],
              "y": [("x", (0, 0, -1), -1.0), ("x", (0, 0, 0), 1.0), ("z", (-1, 0, 0), 1.0), ("z", (0, 0, 0), -1.0)],
              "z": [("x", (0, -1, 0), -1.0), ("x", (0, 0, 0), 1.0), ("y", (-1, 0, 0), 1.0), ("y", (0, 0, 0), -1.0)]}


def _profile_at(n, f, kind, idx):
    """Node sine at node idx (zero at and beyond the walls) or cell cosine at
    cell idx (zero outside 0..n-1)."""
    idx = np.asarray(idx)
    if kind == "node":
        v = np.sin(f * np.pi * idx / n)
        return np.where((idx <= 0) | (idx >= n), 0.0, v)
    v = np.cos(f * np.pi * (idx + 0.5) / n)
    return np.where((idx < 0) | (idx > n - 1), 0.0, v)


def box_separable_terms(g: Grid, table):
    """The (slot, curl term) list of the separable form: for each, the slot
    constant c_s, the curl sign, the face axes of the curl term and of faces
    a and b, and per axis the table S[f_k, f_i, f_j] over that axis's
    distinct frequencies (ascending).  Returns (terms, uniq) with uniq[axis]
    the distinct frequencies."""
    dims = {"x": g.nx, "y": g.ny, "z": g.nz}
    col = {"x": 1, "y": 2, "z": 3}
    uniq = {ax: np.unique(np.asarray(table)[:, col[ax]].astype(int)) for ax in "xyz"}
    one = np.ones(3)
    terms = []
    for s, (fs, es, a, b) in enumerate(SLOTS):
        fa, fb = _FACE_OF_SLOT[a], _FACE_OF_SLOT[b]
        c_s = float(np.cross(_INTERIOR_NORMAL[fa[0]], _INTERIOR_NORMAL[fb[0]]) @ one)
        if c_s == 0.0:
            continue
        ea, dx, dy, dz = _SLOT_EDGE[s]
        off = {"x": dx, "y": dy, "z": dz}
        for (ft, d, sg) in CURL_FACES[ea]:
            dd = {"x": d[0], "y": d[1], "z": d[2]}
            S = {}
            for al in "xyz":
                n = dims[al]
                v = np.arange(n)
                ec = v + off[al]
                Q = np.stack([_profile_at(n, f, "node" if al == ft else "cell", ec + dd[al]) for f in uniq[al]])
                if al != ea:
                    Q = Q * np.where(ec == 0, -1.0, 1.0)[None, :]
                Pa = np.stack([_profile_at(n, f, "node", v + fa[1]) if al == fa[0] else _profile_at(n, f, "cell", v)
                               for f in uniq[al]])
                Pb = np.stack([_profile_at(n, f, "node", v + fb[1]) if al == fb[0] else _profile_at(n, f, "cell", v)
                               for f in uniq[al]])
                S[al] = np.einsum("kv,iv,jv->kij", Q, Pa, Pb)
            terms.append(dict(c=c_s, sign=sg, face=ft, fa=fa[0], fb=fb[0], S=S))
    return terms, uniq


def build_advection_box_separable(g: Grid, resolution: float, lam, table, kchunk=32):
    """C (K, K, K) for a closed box whose basis is analytic_box_modes(g, K),
    from its mode table alone (no Phi, no curl)."""
    from .basis import hodge_primal
    table = np.asarray(table)
    K = table.shape[0]
    Hf = float(hodge_primal(Grid(1, 1, 1), resolution)[0])
    sh = scale_sh(resolution)
    terms, uniq = box_separable_terms(g, table)
    amp = {"x": table[:, 4], "y": table[:, 5], "z": table[:, 6]}
    pos = {ax: np.searchsorted(uniq[ax], table[:, c].astype(int)) for ax, c in (("x", 1), ("y", 2), ("z", 3))}
    C = np.zeros((K, K, K))
    for k0 in range(0, K, kchunk):
        k1 = min(K, k0 + kchunk)
        kk = np.arange(k0, k1)
        acc = np.zeros((K, K, k1 - k0))          # [i, j, k]
        for t in terms:
            prod = np.ones((k1 - k0, K, K))       # [k, i, j]
            for al in "xyz":
                p = pos[al]
                prod *= t["S"][al][p[kk][:, None, None], p[None, :, None], p[None, None, :]]
            Tk = (t["sign"] * Hf) * amp[t["face"]][kk][:, None, None] * amp[t["fa"]][None, :, None] * \
                amp[t["fb"]][None, None, :] * prod
            acc += (np.transpose(Tk, (1, 2, 0)) - np.transpose(Tk, (2, 1, 0))) * (sh * t["c"])
        C[:, :, k0:k1] = acc
    C /= np.asarray(lam)[:, None, None]
    return C
