"""Pins the oracle's literal buildAdvection (oracle/psf_oracle.c,
ora_build_advection, restating psfsolver.cpp:431-551) by geometry, not by a
second transcription of its slot table.

With V = identity (E x E, so K = E) and lambda_i = 1 + i, the oracle returns
C_i(k, j) = W_i[k][j] / lambda_i: the wedge matrices themselves.  The basis has
three non-zero modes, a uniform flux of 1 through every X face (mode 0),
every Y face (mode 1) and every Z face (mode 2).  For a pair of modes the
wedge term s_h (phi_a^i phi_b^j - phi_a^j phi_b^i) c(a, b) of a voxel slot
(a, b) is non-zero only when a and b are faces of the two flux directions,
and then its edge must run along the third axis: cross(n_a, n_b) of an X and
a Y face is parallel to z, and so on.  The values below were worked out by
hand, per voxel, from the face normals and edge lists DECMesh3D stores
(decmesh3d.cpp:1202-1366: a face created at the low wall keeps the normal
pointing into the grid, an interior one the opposite; the Z face's edge list
order flips between z = 0 and z > 0), so a slot with the wrong pair of faces,
a slot on the wrong edge of its face, a dropped scale or 1/lambda factor, or
an index mix-up between axes changes them (checked by planting each in the
oracle).  Swapping a slot's (a, b) is a symmetry of the term -- the bracket
and cross(n_a, n_b) both change sign -- and changes nothing."""
import numpy as np
import pytest

import oracle as O


def _zf(d, x, y, z):
    nx, ny, nz = d
    return z * nx * ny + y * nx + x


def _yf(d, x, y, z):
    nx, ny, nz = d
    return z * (ny + 1) * nx + y * nx + x + (nz + 1) * nx * ny


def _xf(d, x, y, z):
    nx, ny, nz = d
    return z * ny * (nx + 1) + y * (nx + 1) + x + (nz + 1) * nx * ny + (ny + 1) * nx * nz


# edge indices (decmesh3d.cpp:32-45): Z edges, then Y, then X
def _ze(d, x, y, z):
    nx, ny, nz = d
    return z * (nx + 1) * (ny + 1) + y * (nx + 1) + x


def _ye(d, x, y, z):
    nx, ny, nz = d
    return z * (nx + 1) * ny + y * (nx + 1) + x + (nx + 1) * (ny + 1) * nz


def _xe(d, x, y, z):
    nx, ny, nz = d
    return z * nx * (ny + 1) + y * nx + x + (nx + 1) * (ny + 1) * nz + (nx + 1) * ny * (nz + 1)


def _expected(d, inside):
    """Hand-derived wedge sums (without s_h) for the mode pairs (X, Y),
    (X, Z) and (Y, Z).  Per inside voxel (x, y, z):
      s5 = +1 if x == 0 else -1   (normal of its low X face, +x at the wall)
      s3 = +1 if y == 0 else -1   (low Y face)
      t1 = -1 if z == 0 else +1   (low Z face: -z at the wall)
    (X, Y), on Z edges: Ze(x, y+1) -s5, Ze(x, y) +s3 s5, Ze(x+1, y+1) +1, Ze(x+1, y) -s3
    (X, Z), on Y edges: z == 0: Ye(x+1, y, z) -1, Ye(x, y, z) +s5;
                        z > 0:  Ye(x, y, z) +1,   Ye(x+1, y, z) -s5;
                        and Ye(x, y, z+1) -s5, Ye(x+1, y, z+1) +1
    (Y, Z), on X edges: Xe(x, y+1, z) -t1, Xe(x, y, z) +s3 t1, Xe(x, y+1, z+1) -1, Xe(x, y, z+1) +s3"""
    nx, ny, nz = d
    E = (nx + 1) * (ny + 1) * nz + (nx + 1) * ny * (nz + 1) + nx * (ny + 1) * (nz + 1)
    W = {p: np.zeros(E) for p in ((0, 1), (0, 2), (1, 2))}
    v = 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if inside[v]:
                    s5 = 1 if x == 0 else -1
                    s3 = 1 if y == 0 else -1
                    t1 = -1 if z == 0 else 1
                    w = W[(0, 1)]
                    w[_ze(d, x, y + 1, z)] += -s5
                    w[_ze(d, x, y, z)] += s3 * s5
                    w[_ze(d, x + 1, y + 1, z)] += 1
                    w[_ze(d, x + 1, y, z)] += -s3
                    w = W[(0, 2)]
                    if z == 0:
                        w[_ye(d, x + 1, y, z)] += -1
                        w[_ye(d, x, y, z)] += s5
                    else:
                        w[_ye(d, x, y, z)] += 1
                        w[_ye(d, x + 1, y, z)] += -s5
                    w[_ye(d, x, y, z + 1)] += -s5
                    w[_ye(d, x + 1, y, z + 1)] += 1
                    w = W[(1, 2)]
                    w[_xe(d, x, y + 1, z)] += -t1
                    w[_xe(d, x, y, z)] += s3 * t1
                    w[_xe(d, x, y + 1, z + 1)] += -1
                    w[_xe(d, x, y, z + 1)] += s3
                v += 1
    return W, E


@pytest.mark.parametrize("dims,hole", [((1, 1, 1), None), ((2, 2, 2), None), ((3, 2, 2), (1, 0, 1)),
                                       ((2, 3, 1), (1, 2, 0))])
def test_wedges_of_uniform_fluxes(dims, hole):
    nx, ny, nz = dims
    h = 0.25
    inside = np.ones(nx * ny * nz, bool)
    if hole is not None:
        x, y, z = hole
        inside[(z * ny + y) * nx + x] = False
    m = O.Mesh(dims, h, (-0.5 * nx * h, -0.5 * ny * h, -0.5 * nz * h), inside=inside)
    Wx, E = _expected(dims, inside)
    assert m.nE == E
    K = E
    Phi = np.zeros((K, m.nF))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx + 1):
                Phi[0, _xf(dims, x, y, z)] = 1.0
    for z in range(nz):
        for y in range(ny + 1):
            for x in range(nx):
                Phi[1, _yf(dims, x, y, z)] = 1.0
    for z in range(nz + 1):
        for y in range(ny):
            for x in range(nx):
                Phi[2, _zf(dims, x, y, z)] = 1.0
    assert Phi[:3].sum() == m.nF
    lam = 1.0 + np.arange(K, dtype=np.float64)
    V = np.eye(E)                           # V_cm (K, E): column k of V is edge k
    C = O.build_advection(m, Phi, lam, V)   # C[i, j, k] = C_i(k, j) = W_i[k][j] / lambda_i
    hr = np.float32(h) / np.float32(2)
    sh = float(hr * hr)                     # s_h = (h/2)^2 in float (psfsolver.cpp:444)
    W = C * lam[:, None, None]              # W[i, j, e]
    for (i, j), w in Wx.items():
        assert np.array_equal(W[i, j], sh * w), (i, j)
        assert np.array_equal(W[j, i], -sh * w), (j, i)
    # nothing else: the other mode pairs, a mode with itself, the zero modes
    mask = np.ones((K, K), bool)
    for i, j in Wx:
        mask[i, j] = mask[j, i] = False
    assert not np.any(W[mask])
    # each pair lands on edges of the third axis only
    nZE = (nx + 1) * (ny + 1) * nz
    nYE = (nx + 1) * ny * (nz + 1)
    assert not np.any(W[0, 1, nZE:])
    assert not np.any(W[0, 2, :nZE]) and not np.any(W[0, 2, nZE + nYE:])
    assert not np.any(W[1, 2, :nZE + nYE])
    assert np.any(Wx[(0, 1)]) and np.any(Wx[(0, 2)]) and np.any(Wx[(1, 2)])


def test_wedge_edges_follow_the_normals():
    """The edge of each non-zero wedge entry runs along cross(n_a, n_b) of its
    two face normals as the oracle's mesh stores them (decmesh3d.cpp:1202-1366):
    the edge's end points (edges(), signed point ids) differ along that axis
    only."""
    dims = (2, 2, 2)
    m = O.Mesh(dims, 0.25, (-0.25, -0.25, -0.25), inside=np.ones(8, bool))
    _, n, _, _ = m.faces()
    v2, _, _, _, _ = m.edges()
    _, _, pc, _ = m.points()
    Wx, E = _expected(dims, np.ones(8, bool))
    face_of_flux = {0: _xf(dims, 0, 0, 0), 1: _yf(dims, 0, 0, 0), 2: _zf(dims, 0, 0, 0)}
    for (i, j), w in Wx.items():
        axis = np.abs(np.cross(n[face_of_flux[i]], n[face_of_flux[j]]))
        for e in np.flatnonzero(w):
            a, b = (abs(int(v)) - 1 for v in v2[e])     # signed point ids (decmesh3d.cpp:52-60)
            d = np.abs(pc[b] - pc[a])
            assert np.count_nonzero(d) == 1 and np.argmax(d) == np.argmax(axis), (i, j, e)
