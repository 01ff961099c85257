"""Pins the oracle's literal DECMesh3D restatement (DEC/decmesh3d.cpp) and the
product's closed-form tables (parallelspectralfluids3d_amd/grid.py) against
geometry, the reference's own invariants and exact DEC identities."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from parallelspectralfluids3d_amd import basis
from parallelspectralfluids3d_amd.grid import Grid, box_mask, sphere_obstacle_mask

GRIDS = [(1, 1, 1), (2, 3, 4), (4, 4, 4), (5, 2, 3)]


def _mesh(dims, h=0.125, inside=None):
    return O.Mesh(dims, h, (-0.5, -0.5, -0.5), inside=inside)


@pytest.mark.parametrize("dims", GRIDS)
def test_indices_match_geometry(dims):
    """Face/voxel index closed forms (decmesh3d.cpp:12-30) address the element at
    the right place: the stored centre of element index(x,y,z) is at (x,y,z)."""
    h = 0.125
    m = _mesh(dims, h)
    g = Grid(*dims)
    _, _, fc, _ = m.faces()
    _, _, vc = m.voxels()
    mn = -0.5
    for z in range(dims[2] + 1):
        for y in range(dims[1] + 1):
            for x in range(dims[0] + 1):
                if x < dims[0] and y < dims[1]:
                    f = O.lib().ora_zface_index(m.h, x, y, z)
                    assert f == g.zface_index(x, y, z)
                    assert np.allclose(fc[f], [mn + (x + .5) * h, mn + (y + .5) * h, mn + z * h])
                if x < dims[0] and z < dims[2]:
                    f = O.lib().ora_yface_index(m.h, x, y, z)
                    assert f == g.yface_index(x, y, z)
                    assert np.allclose(fc[f], [mn + (x + .5) * h, mn + y * h, mn + (z + .5) * h])
                if y < dims[1] and z < dims[2]:
                    f = O.lib().ora_xface_index(m.h, x, y, z)
                    assert f == g.xface_index(x, y, z)
                    assert np.allclose(fc[f], [mn + x * h, mn + (y + .5) * h, mn + (z + .5) * h])
                if x < dims[0] and y < dims[1] and z < dims[2]:
                    v = O.lib().ora_voxel_index(m.h, x, y, z)
                    assert np.allclose(vc[v], [mn + (x + .5) * h, mn + (y + .5) * h, mn + (z + .5) * h])


@pytest.mark.parametrize("dims", GRIDS)
def test_sign_bits_closed_form_and_reference_invariants(dims):
    """Sign bits {z==0,1,y==0,1,x==0,1} (decmesh3d.cpp:1383-1399) and the
    reference's checkInternalState (decmesh3d.cpp:1409-1435): the face-id sign
    equals the sign bit, opposite signed normals are antiparallel."""
    m = _mesh(dims)
    g = Grid(*dims)
    assert np.array_equal(m.sign_words(), g.sign_words())
    f6, _, vc = m.voxels()
    _, n, fc, _ = m.faces()
    x, y, z = g.voxel_coords()
    for v in range(g.nV):
        expect = [z[v] == 0, True, y[v] == 0, True, x[v] == 0, True]
        for j in range(6):
            s = m.face_signum(v, j)
            assert s == (1 if expect[j] else -1)
            assert s == (1 if f6[v, j] > 0 else -1)
        for j in (0, 2, 4):
            s1 = 1 if f6[v, j] > 0 else -1
            s2 = 1 if f6[v, j + 1] > 0 else -1
            assert np.dot(s1 * n[abs(f6[v, j]) - 1], s2 * n[abs(f6[v, j + 1]) - 1]) < 0


@pytest.mark.parametrize("dims", GRIDS)
def test_dec_exactness(dims):
    """d2 d1^T = 0 (boundary of a boundary) and d1^T d0 = 0 (curl grad): a wrong
    edge sign, edge index or face sign breaks one of them."""
    m = _mesh(dims, inside=np.ones(np.prod(dims), bool))
    d1 = m.derivative1()
    d2 = m.derivative2()
    v2, _, _, _, _ = m.edges()
    rows = np.repeat(np.arange(m.nE), 2)
    d0 = sp.csr_matrix((np.tile([-1.0, 1.0], m.nE), (rows, (np.abs(v2) - 1).reshape(-1))), shape=(m.nE, m.nP))
    assert abs(d2 @ d1.T).max() == 0
    assert abs(d1.T @ d0).max() == 0


@pytest.mark.parametrize("dims", GRIDS)
def test_face_cycles_and_normals(dims):
    """Each face's signed edge list is a closed loop; its orientation is +axis for
    faces created BACK/TOP/RIGHT and -axis for FRONT/BOTTOM/LEFT; the stored normal
    equals that orientation for Z faces and its negative for X and Y faces."""
    m = _mesh(dims)
    e4, n, _, _ = m.faces()
    v2, _, _, _, _ = m.edges()
    _, _, pc, _ = m.points()
    for f in range(m.nF):
        area = np.zeros(3)
        path = []
        for k in range(4):
            e = abs(e4[f, k]) - 1
            a, b = pc[abs(v2[e, 0]) - 1], pc[abs(v2[e, 1]) - 1]
            if e4[f, k] < 0:
                a, b = b, a
            path.append((a, b))
        for k in range(4):
            assert np.allclose(path[k][1], path[(k + 1) % 4][0])     # closed loop
            area += np.cross(path[k][0], path[k][1])
        axis = int(np.argmax(np.abs(area)))
        orient = np.sign(area)
        if axis == 2:
            assert np.array_equal(orient, n[f])
        else:
            assert np.array_equal(orient, -n[f])


@pytest.mark.parametrize("dims", GRIDS)
def test_hodge_closed_forms(dims):
    """Hf = 1/(2h); He = {h/2, h, 2h} for dualCount {2,3,4} on inside edges
    (the reference's asserts, dec.cpp:109-122 and 164-183)."""
    h = 0.125
    ins = np.ones(np.prod(dims), bool)
    m = _mesh(dims, h, ins)
    assert np.allclose(m.hodge2_primal(), 1 / (2 * h), rtol=0, atol=1e-12)
    _, _, dc, _, _ = m.edges()
    he = m.hodge2_dual()
    assert np.allclose(he, np.select([dc == 2, dc == 3, dc == 4], [h / 2, h, 2 * h]), rtol=0, atol=1e-12)
    g = Grid(*dims)
    assert np.array_equal(he, basis.hodge_dual(g, h, ins))


@pytest.mark.parametrize("dims", GRIDS + [(7, 7, 7), (3, 6, 2)])
def test_product_tables_equal_oracle(dims):
    """The product's closed-form tables equal the literal construction bit for bit."""
    m = _mesh(dims)
    g = Grid(*dims)
    f6, _, _ = m.voxels()
    e4, n, _, _ = m.faces()
    _, ef4, dc, _, _ = m.edges()
    assert np.array_equal(f6, (g.voxel_faces + 1) * g.voxel_face_signs)
    assert np.array_equal(e4, (g.face_edges + 1) * g.face_edge_signs)
    assert np.array_equal(n, g.face_normals)
    ef, cnt = g.edge_faces
    assert np.array_equal(ef4, ef) and np.array_equal(dc, cnt)


def test_inside_marking_and_laplacian_sphere():
    """setVoxelInside marks faces, edges and points (decmesh3d.cpp:1668-1704);
    the product's inside edges, face counts and M equal the oracle's (h dyadic,
    so the oracle's float-built Hodge values are exact)."""
    n = 8
    g = Grid(n, n, n)
    h = 1.0 / n
    ins = sphere_obstacle_mask(g, h, (-0.5, -0.5, -0.5), radius=0.3)
    assert 0 < (~ins).sum() < ins.size
    m = O.Mesh((n, n, n), h, (-0.5, -0.5, -0.5), inside=ins)
    _, _, _, _, ei = m.edges()
    assert np.array_equal(ei == 3, g.edge_inside(ins))
    _, _, _, fi = m.faces()
    assert np.array_equal(fi == 3, g.face_inside(ins))
    M, _, counts = m.laplacian()
    assert np.array_equal(counts == 2, g.face_inside_count(ins) == 2)
    assert abs(M - basis.laplacian(g, h, ins)).max() == 0


def test_gravity_field_matches_oracle():
    n = 6
    g = Grid(n, n, n)
    ins = box_mask(g)
    m = O.Mesh((n, n, n), 1.0 / n, (-0.5, -0.5, -0.5), inside=ins)
    assert np.array_equal(m.gravity_field(), basis.gravity_field(g, ins))
    # interior y faces carry -0.00981 (TOP normal -y), wall y faces at y=0 +0.00981
    f = m.gravity_field()
    assert f[g.yface_index(2, 3, 2)] == -0.00981 and f[g.yface_index(2, 0, 2)] == 0.00981
