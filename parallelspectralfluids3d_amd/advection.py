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

from .basis import curl_operator
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
        V_cm = (curl_operator(g, resolution) @ Phi_cm.T).T
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
