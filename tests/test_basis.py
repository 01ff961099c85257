"""Bases: analytic box modes (exact eigenvectors of the free-slip DEC
Laplacian) and the reference-M eigensolve (psfsolver.cpp:268-341 acceptance)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from parallelspectralfluids3d_amd import basis
from parallelspectralfluids3d_amd.grid import Grid, box_mask


@pytest.mark.parametrize("dims,K", [((8, 8, 8), 30), ((4, 8, 2), 12)])
def test_analytic_modes(dims, K):
    g = Grid(*dims)
    h = 1.0 / max(dims)
    Phi, lam, _ = basis.analytic_box_modes(g, K)
    assert np.abs(Phi @ Phi.T - np.eye(K)).max() < 1e-13
    assert np.all(np.diff(lam) >= -1e-12)
    m = O.Mesh(dims, h, (-0.5, -0.5, -0.5), inside=box_mask(g))
    assert np.abs(m.derivative2() @ Phi.T).max() < 1e-13          # divergence-free
    wall = g.face_inside_count(box_mask(g)) != 2
    assert np.abs(Phi[:, wall]).max() == 0.0                       # no flux through walls
    # free-slip M: the reference operator with zero Hodge weight on wall edges
    d1 = m.derivative1()
    he = m.hodge2_dual()
    _, cnt = g.edge_faces
    he_fs = np.where(cnt == 4, he, 0.0)
    M = d1.T @ sp.diags(he_fs) @ d1 @ sp.diags(m.hodge2_primal())
    M = sp.diags((~wall).astype(float)) @ M
    res = np.linalg.norm(M @ Phi.T - Phi.T * lam, axis=0)
    assert res.max() < 1e-12


def test_laplacian_eigenbasis_box():
    g = Grid(8, 8, 8)
    h = 1.0 / 8
    ins = box_mask(g)
    Phi, lam = basis.laplacian_eigenbasis(g, h, ins, 24)
    m = O.Mesh((8, 8, 8), h, (-0.5, -0.5, -0.5), inside=ins)
    M, _, _ = m.laplacian()
    assert np.linalg.norm(M @ Phi.T - Phi.T * lam, axis=0).max() < 1e-11
    assert np.abs(Phi @ Phi.T - np.eye(24)).max() < 1e-12
    assert np.all(lam > 1e-10) and np.all(np.diff(lam) >= -1e-12)
    # nothing smaller was skipped: a dense eigensolve of the same operator agrees
    act = np.flatnonzero(g.face_inside_count(ins) == 2)
    ev = np.linalg.eigvalsh(M[act][:, act].toarray())
    pos = ev[ev > 1e-10]
    assert np.allclose(lam, pos[:24], rtol=1e-10)


def test_omega_generators():
    w = basis.unit_energy_omega(50, 0)
    assert abs(np.linalg.norm(w) - 1) < 1e-15
    assert np.array_equal(w, basis.unit_energy_omega(50, 0))
    p = basis.power_law_omega(64, 1)
    assert abs(np.linalg.norm(p) - 1) < 1e-15
    r = np.abs(p[1:]) / np.abs(p[:-1])
    assert np.allclose(r, (np.arange(2, 65) / np.arange(1, 64)) ** (-5 / 6))


def test_bounded_mode_enumeration_equals_full():
    """The K lowest box modes from the bounded enumeration are the full
    enumeration's first K, in the same order."""
    from parallelspectralfluids3d_amd import basis as B
    for dims, K in (((8, 8, 8), 64), ((6, 5, 7), 40), ((16, 16, 16), 200), ((12, 12, 12), 300), ((20, 9, 14), 150)):
        g = Grid(*dims)
        assert B._box_mode_list(g, K)[:K] == B._box_mode_list(g)[:K], dims


@pytest.mark.parametrize("dims", [(6, 5, 7), (8, 8, 8)])
def test_mode_table_without_phi(dims):
    """The normalised mode table and sampled Phi columns, computed from 1-D
    profile sums without forming Phi, equal the explicit unit-norm modes."""
    g = Grid(*dims)
    Phi, lam, tab = basis.analytic_box_modes(g, 20)
    t2 = basis.box_mode_table_normalised(g, 20)
    assert np.allclose(t2, tab, rtol=0, atol=1e-14)
    faces = np.random.default_rng(0).choice(g.nF, 200, replace=False)
    assert np.allclose(basis.box_mode_values(g, t2, faces), Phi[:, faces], rtol=0, atol=1e-14)


def test_pruned_laplacian_null_space_and_penalty():
    """The device eigensolve's inputs: M_II (pruned, symmetrised) annihilates
    every deep voxel's gradient D^T e_c, exact eigenvectors of nonzero
    eigenvalue are divergence-free on deep voxels, and M_II + p D^T D has
    the same nonzero eigenpairs (so the penalty lifts only gradients)."""
    from parallelspectralfluids3d_amd.grid import sphere_obstacle_mask
    import scipy.sparse.linalg as sla
    g = Grid(8, 8, 8)
    h = 1.0 / 8
    inside = sphere_obstacle_mask(g, h, (-0.5, -0.5, -0.5)).astype(bool).reshape(-1)
    MII, act = basis.pruned_laplacian(g, h, inside)
    D = basis.deep_divergence(g, inside, act)
    assert D.shape == (int((inside & (g.face_inside_count(inside)[g.voxel_faces] == 2).all(1)).sum()), act.size)
    assert abs(MII - MII.T).max() == 0.0
    scale = abs(MII).max()
    assert abs(MII @ D.T).max() < 1e-13 * scale                 # gradients are null vectors
    w, V = np.linalg.eigh(MII.toarray())
    nz = w > 1e-10 * w.max()
    assert np.abs(D @ V[:, nz]).max() < 1e-10                    # nonzero modes are divergence-free
    for p in (1.0, 16.0):
        A = (MII + p * (D.T @ D)).tocsc()
        wk = sla.eigsh(A, k=6, sigma=w[nz][0] * 0.5, which="LM", return_eigenvectors=False)
        got = np.sort(wk)
        # the six eigenvalues of A nearest half the first nonzero one include
        # the first nonzero eigenvalues of M_II unchanged
        assert np.min(np.abs(got - w[nz][0])) < 1e-12 * w.max()


def test_device_eigenbasis_helpers_reorder():
    """laplacian_eigenbasis_device puts the active faces in spatial order and
    maps the eigenvectors back: the permutation is a bijection of the rows."""
    from parallelspectralfluids3d_amd.grid import sphere_obstacle_mask
    g = Grid(6, 6, 6)
    inside = sphere_obstacle_mask(g, 1 / 6, (-0.5, -0.5, -0.5)).astype(bool).reshape(-1)
    _, act = basis.pruned_laplacian(g, 1 / 6, inside)
    fc = g.face_coords[act]
    perm = np.lexsort((fc[:, 0], fc[:, 1], fc[:, 2]))
    assert np.array_equal(np.sort(perm), np.arange(act.size))
    z = fc[perm, 2]
    assert np.all(np.diff(z) >= 0)


def test_sparse_times_dense_threads():
    """the column-split product equals scipy's one-thread product, for more
    workers than columns and for Fortran-ordered operands too"""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    S = sp.random(300, 200, density=0.05, random_state=4, format="csr")
    for cols in (1, 5, 37):
        V = rng.standard_normal((200, cols))
        ref = S @ V
        for w in (None, 1, 3, 16):
            assert np.array_equal(basis._sparse_times_dense(S, V, w), ref)
        assert np.array_equal(basis._sparse_times_dense(S, np.asfortranarray(V)), ref)
