"""The device eigensolver (psf_eigensolve_csr) that replaces the shift-invert
eigensolve of psfsolver.cpp:267-341.

- Path Laplacians: Dirichlet, eigenpairs 2 - 2 cos(pi k / (n + 1)) and
  sin(pi k j / (n + 1)) in closed form; Neumann, whose null space (the
  constants) comes back first with eigenvalue 0.
- The pruned 1-form Laplacian of a voxelised sphere in a box (its null space
  holds one gradient per deep inside voxel): eigenvalues, residuals and, at a
  spectral gap, the spanned subspace against scipy's shift-invert with the
  null space lifted (basis.laplacian_eigenbasis).
- Argument errors come back as status codes with a message.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _path(n, neumann):
    main = np.full(n, 2.0)
    if neumann:
        main[0] = main[-1] = 1.0
    return sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, 1, -1], format="csr")


# block widths K + guard: 40 and 150 (even: two columns per 16-byte load; 150
# leaves the second 128-column chunk partly empty) and 39 (odd: one per load)
@pytest.mark.parametrize("K,guard", [(24, 16), (24, 15), (40, 110)])
def test_path_laplacian_closed_form(hip_ctx, K, guard):
    from parallelspectralfluids3d_amd import hip
    _say(f"eigensolve: Dirichlet path Laplacian n=2000 K={K} guard={guard}")
    n = 2000
    lam, X, used = hip.eigensolve_csr(hip_ctx, _path(n, False), K, guard=guard, degree=30, tol=1e-12)
    k = np.arange(1, K + 1)
    exact = 2.0 - 2.0 * np.cos(np.pi * k / (n + 1))
    assert np.max(np.abs(lam - exact)) < 1e-12
    assert np.max(np.abs(X.T @ X - np.eye(K))) < 1e-10
    # eigenvectors sin(pi k j / (n + 1)), j = 1..n, up to sign (simple eigenvalues)
    j = np.arange(1, n + 1)
    for kk in (1, 5, K):
        v = np.sin(np.pi * kk * j / (n + 1))
        v /= np.linalg.norm(v)
        assert abs(abs(v @ X[:, kk - 1]) - 1.0) < 1e-9
    assert 0 < used < 5000


def test_null_space_is_returned(hip_ctx):
    """A null space is an eigenspace like the others: the Neumann path's
    constants come first with eigenvalue 0, then 2 - 2 cos(pi k / n)."""
    from parallelspectralfluids3d_amd import hip
    _say("eigensolve: Neumann path Laplacian n=600 K=12")
    n, K = 600, 12
    lam, X, used = hip.eigensolve_csr(hip_ctx, _path(n, True), K, guard=12, degree=30, tol=1e-12)
    exact = 2.0 - 2.0 * np.cos(np.pi * np.arange(K) / n)
    assert np.max(np.abs(lam - exact)) < 1e-12
    assert abs(abs(X[:, 0].sum()) - np.sqrt(n)) < 1e-9


@pytest.mark.parametrize("n,K", [(12, 24), (16, 40)])
def test_sphere_scene_against_shift_invert(hip_ctx, n, K):
    from parallelspectralfluids3d_amd import basis
    from parallelspectralfluids3d_amd.grid import Grid, sphere_obstacle_mask
    _say(f"eigensolve: sphere obstacle n={n} K={K}")
    g = Grid(n, n, n)
    h = 1.0 / n
    inside = sphere_obstacle_mask(g, h, (-0.5, -0.5, -0.5)).astype(bool).reshape(-1)
    Kx = K + 12
    Phi_ref, lam_ref = basis.laplacian_eigenbasis(g, h, inside, Kx)
    Phi, lam, info = basis.laplacian_eigenbasis_device(g, h, inside, K, hip_ctx)
    assert np.max(np.abs(lam - lam_ref[:K]) / lam_ref[:K]) < 1e-9
    MII, act = basis.pruned_laplacian(g, h, inside)
    V = Phi[:, act].T
    res = np.linalg.norm(MII @ V - V * lam, axis=0)
    assert res.max() < 1e-9
    assert np.max(np.abs(V.T @ V - np.eye(K))) < 1e-10
    assert np.all(Phi[:, np.setdiff1d(np.arange(g.nF), act)] == 0.0)
    # the spanned space below the largest gap within the first K: each basis
    # is within (block residual) / gap of the exact eigenspace (Davis-Kahan),
    # so the sine of their largest principal angle is bounded by the sum
    gaps = np.diff(lam_ref[:K + 1])
    k = int(np.argmax(gaps / lam_ref[1:K + 1])) + 1
    assert gaps[k - 1] > 1e-3 * lam_ref[k]
    Vr = Phi_ref[:k, act].T
    Vd = V[:, :k]
    res_ref = np.linalg.norm(MII @ Vr - Vr * lam_ref[:k], axis=0)
    bound = (np.linalg.norm(res[:k]) + np.linalg.norm(res_ref)) / gaps[k - 1]
    # sine of the largest principal angle, as the norm of Vd's part outside
    # span(Vr) (sqrt(1 - cos^2) would lose half the digits)
    sin_max = np.linalg.norm(Vd - Vr @ (Vr.T @ Vd), 2)
    assert sin_max <= 2.0 * bound + 1e-12


def test_argument_errors(hip_ctx):
    from parallelspectralfluids3d_amd import _lib as L
    from parallelspectralfluids3d_amd import hip
    _say("eigensolve: argument errors")
    A = sp.identity(50, format="csr") * 2.0
    with pytest.raises(L.PsfError) as e:
        hip.eigensolve_csr(hip_ctx, A, 1000, guard=100)   # block above 1024
    assert e.value.status == L.PSF_ERR_ARG
    with pytest.raises(L.PsfError) as e:
        hip.eigensolve_csr(hip_ctx, A, 40, guard=16)  # block as large as n
    assert e.value.status == L.PSF_ERR_ARG
    B = sp.csr_matrix(A)
    B.indices[3] = 77                                # column out of range
    with pytest.raises(L.PsfError) as e:
        hip.eigensolve_csr(hip_ctx, B, 4, guard=4)
    assert e.value.status == L.PSF_ERR_ARG
    # too few products allowed: PSF_ERR_STATE with the residual reached
    C = _path(3000, False)
    with pytest.raises(L.PsfError) as e:
        hip.eigensolve_csr(hip_ctx, C, 16, guard=16, degree=10, tol=1e-14, max_products=30)
    assert e.value.status == L.PSF_ERR_STATE
