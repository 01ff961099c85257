"""Voxel-grid DEC tables in closed form (vectorised numpy).

The reference builds its mesh element by element (DEC/decmesh3d.cpp:62-106,
1026-1403): voxels are visited z-major, each voxel creates the faces it is
the first to touch, each face creates the edges it is the first to touch, and
an edge's stored direction is the direction its creating face walks it.
Everything that creation order decides (face normals, face edge lists, edge
signs, the per-voxel sign bits) depends only on the grid dimensions, never on
which voxels are inside.  This module derives all of it without walking the
mesh: the first voxel to touch a face is the one below it along the face's
axis (or voxel 0 of that axis at the min wall), and the first face to touch
an edge is the face with the smallest (creating voxel, face slot, edge slot).

Index layouts (decmesh3d.cpp:12-50): faces [Z | Y | X], edges [Z | Y | X],
voxel v = z*nx*ny + y*nx + x.
"""
from __future__ import annotations

import dataclasses
from functools import cached_property

import numpy as np

# face creation kinds (DECMesh3D::FaceDirection) and their stored normals
FRONT, BACK, TOP, BOTTOM, LEFT, RIGHT = range(6)
NORMALS = np.array([[0, 0, -1], [0, 0, 1], [0, -1, 0], [0, 1, 0], [1, 0, 0], [-1, 0, 0]], dtype=np.float64)
# walk direction of an edge: +1 = POS (towards +axis), -1 = NEG
# per face kind: (edge axis, offsets (dx,dy,dz) of the edge, walk) for e1..e4
# (decmesh3d.cpp:1209-1358)
_EDGE_LISTS = {
    FRONT: [("x", (0, 1, 0), +1), ("y", (1, 0, 0), -1), ("x", (0, 0, 0), -1), ("y", (0, 0, 0), +1)],
    BACK: [("x", (0, 1, 0), -1), ("y", (0, 0, 0), -1), ("x", (0, 0, 0), +1), ("y", (1, 0, 0), +1)],
    TOP: [("x", (0, 0, 1), +1), ("z", (1, 0, 0), -1), ("x", (0, 0, 0), -1), ("z", (0, 0, 0), +1)],
    BOTTOM: [("x", (0, 0, 1), -1), ("z", (0, 0, 0), -1), ("x", (0, 0, 0), +1), ("z", (1, 0, 0), +1)],
    LEFT: [("z", (0, 1, 0), -1), ("y", (0, 0, 0), -1), ("z", (0, 0, 0), +1), ("y", (0, 0, 1), +1)],
    RIGHT: [("z", (0, 1, 0), +1), ("y", (0, 0, 1), -1), ("z", (0, 0, 0), -1), ("y", (0, 0, 0), +1)],
}


@dataclasses.dataclass(frozen=True)
class Grid:
    nx: int
    ny: int
    nz: int

    # ---- counts (decmesh3d.cpp:64-70) ----------------------------------------
    @property
    def nV(self):
        return self.nx * self.ny * self.nz

    @property
    def nZF(self):
        return (self.nz + 1) * self.nx * self.ny

    @property
    def nYF(self):
        return (self.ny + 1) * self.nx * self.nz

    @property
    def nXF(self):
        return (self.nx + 1) * self.ny * self.nz

    @property
    def nF(self):
        return self.nZF + self.nYF + self.nXF

    @property
    def nZE(self):
        return (self.nx + 1) * (self.ny + 1) * self.nz

    @property
    def nYE(self):
        return (self.nx + 1) * self.ny * (self.nz + 1)

    @property
    def nXE(self):
        return self.nx * (self.ny + 1) * (self.nz + 1)

    @property
    def nE(self):
        return self.nZE + self.nYE + self.nXE

    @property
    def nP(self):
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    # ---- index closed forms (decmesh3d.cpp:12-50) -----------------------------
    def voxel_index(self, x, y, z):
        return z * (self.nx * self.ny) + y * self.nx + x

    def zface_index(self, x, y, z):
        return z * (self.nx * self.ny) + y * self.nx + x

    def yface_index(self, x, y, z):
        return z * ((self.ny + 1) * self.nx) + y * self.nx + x + self.nZF

    def xface_index(self, x, y, z):
        return z * (self.ny * (self.nx + 1)) + y * (self.nx + 1) + x + self.nZF + self.nYF

    def zedge_index(self, x, y, z):
        return z * (self.nx + 1) * (self.ny + 1) + y * (self.nx + 1) + x

    def yedge_index(self, x, y, z):
        return z * (self.nx + 1) * self.ny + y * (self.nx + 1) + x + self.nZE

    def xedge_index(self, x, y, z):
        return z * self.nx * (self.ny + 1) + y * self.nx + x + self.nZE + self.nYE

    def edge_index(self, axis, x, y, z):
        return {"x": self.xedge_index, "y": self.yedge_index, "z": self.zedge_index}[axis](x, y, z)

    # ---- sign bits (decmesh3d.cpp:1383-1399, closed form) --------------------
    def voxel_sign_bits(self):
        """6 bits per voxel: bit j set iff face j+1 was created by the voxel,
        i.e. {z==0, 1, y==0, 1, x==0, 1}."""
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        bits = ((z == 0).astype(np.uint32) | 2 | ((y == 0).astype(np.uint32) << 2) | 8 |
                ((x == 0).astype(np.uint32) << 4) | 32)
        return bits.reshape(-1).astype(np.uint32)

    def sign_words(self):
        bits = self.voxel_sign_bits()
        nw = -(-self.nV // 5)
        pad = np.zeros(nw * 5, dtype=np.uint64)
        pad[: self.nV] = bits
        pad = pad.reshape(nw, 5)
        words = np.zeros(nw, dtype=np.uint64)
        for o in range(5):
            words |= pad[:, o] << np.uint64(6 * o)
        return words.astype(np.uint32)

    # ---- voxels ---------------------------------------------------------------
    def voxel_coords(self):
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        return x.reshape(-1), y.reshape(-1), z.reshape(-1)

    @cached_property
    def voxel_faces(self):
        """(V, 6) unsigned face indices f1..f6 = Z(z), Z(z+1), Y(y), Y(y+1), X(x), X(x+1)
        (decmesh3d.cpp:1377-1382)."""
        x, y, z = self.voxel_coords()
        return np.stack([self.zface_index(x, y, z), self.zface_index(x, y, z + 1),
                         self.yface_index(x, y, z), self.yface_index(x, y + 1, z),
                         self.xface_index(x, y, z), self.xface_index(x + 1, y, z)], axis=1).astype(np.int64)

    @cached_property
    def voxel_face_signs(self):
        """(V, 6) +1 where the voxel created the face (the sign of Voxel3D::f[j])."""
        bits = self.voxel_sign_bits()
        return np.stack([np.where((bits >> j) & 1, 1, -1) for j in range(6)], axis=1).astype(np.int64)

    # ---- faces ----------------------------------------------------------------
    @cached_property
    def face_kind(self):
        """creation kind of every face (FRONT/BACK/...); interior faces are
        BACK/TOP/RIGHT, faces on the min walls FRONT/BOTTOM/LEFT."""
        kind = np.empty(self.nF, dtype=np.int64)
        zc = np.arange(self.nZF) // (self.nx * self.ny)
        kind[: self.nZF] = np.where(zc == 0, FRONT, BACK)
        yy = (np.arange(self.nYF) % ((self.ny + 1) * self.nx)) // self.nx
        kind[self.nZF: self.nZF + self.nYF] = np.where(yy == 0, BOTTOM, TOP)
        xx = np.arange(self.nXF) % (self.nx + 1)
        kind[self.nZF + self.nYF:] = np.where(xx == 0, LEFT, RIGHT)
        return kind

    @cached_property
    def face_coords(self):
        """(F, 3) integer coordinates (x, y, z) of each face as addFace receives them."""
        c = np.empty((self.nF, 3), dtype=np.int64)
        i = np.arange(self.nZF)
        c[: self.nZF] = np.stack([i % self.nx, (i // self.nx) % self.ny, i // (self.nx * self.ny)], 1)
        i = np.arange(self.nYF)
        c[self.nZF: self.nZF + self.nYF] = np.stack(
            [i % self.nx, (i // self.nx) % (self.ny + 1), i // ((self.ny + 1) * self.nx)], 1)
        i = np.arange(self.nXF)
        c[self.nZF + self.nYF:] = np.stack(
            [i % (self.nx + 1), (i // (self.nx + 1)) % self.ny, i // (self.ny * (self.nx + 1))], 1)
        return c

    @cached_property
    def face_normals(self):
        return NORMALS[self.face_kind]

    @cached_property
    def face_creation_rank(self):
        """6*creating voxel + slot of the face in that voxel (addVoxel order)."""
        c = self.face_coords
        x, y, z = c[:, 0], c[:, 1], c[:, 2]
        kind = self.face_kind
        rank = np.empty(self.nF, dtype=np.int64)
        vx = np.where(kind == RIGHT, x - 1, x)
        vy = np.where(kind == TOP, y - 1, y)
        vz = np.where(kind == BACK, z - 1, z)
        slot = np.select([kind == FRONT, kind == BACK, kind == BOTTOM, kind == TOP, kind == LEFT], [0, 1, 2, 3, 4], 5)
        rank[:] = self.voxel_index(vx, vy, vz) * 6 + slot
        return rank

    @cached_property
    def _face_edges_walk(self):
        """(F, 4) unsigned edge indices and (F, 4) walk directions (+1/-1)."""
        c = self.face_coords
        kind = self.face_kind
        e = np.empty((self.nF, 4), dtype=np.int64)
        w = np.empty((self.nF, 4), dtype=np.int64)
        for k, lst in _EDGE_LISTS.items():
            m = kind == k
            cx, cy, cz = c[m, 0], c[m, 1], c[m, 2]
            for j, (ax, (ox, oy, oz), walk) in enumerate(lst):
                e[m, j] = self.edge_index(ax, cx + ox, cy + oy, cz + oz)
                w[m, j] = walk
        return e, w

    @property
    def face_edges(self):
        return self._face_edges_walk[0]

    @cached_property
    def edge_walk_stored(self):
        """stored direction of every edge = the walk of its creating face
        (addEdge, decmesh3d.cpp:1043-1199)."""
        e, w = self._face_edges_walk
        key = self.face_creation_rank[:, None] * 4 + np.arange(4)[None, :]
        order = np.argsort(key.reshape(-1), kind="stable")
        ee = e.reshape(-1)[order]
        ww = w.reshape(-1)[order]
        first = np.full(self.nE, -1, dtype=np.int64)
        # the first occurrence of each edge in creation order
        _, idx = np.unique(ee, return_index=True)
        first[ee[idx]] = ww[idx]
        return first

    @cached_property
    def face_edge_signs(self):
        """(F, 4): +1 where the face walks the edge in its stored direction."""
        e, w = self._face_edges_walk
        return np.where(w == self.edge_walk_stored[e], 1, -1).astype(np.int64)

    @cached_property
    def edge_faces(self):
        """(E, 4) signed face ids (+(f+1)) in registration order, 0 padded, and dualCount."""
        e = self.face_edges
        key = self.face_creation_rank[:, None] * 4 + np.arange(4)[None, :]
        order = np.argsort(key.reshape(-1), kind="stable")
        ee = e.reshape(-1)[order]
        ff = np.repeat(np.arange(self.nF), 4)[order]
        o2 = np.argsort(ee, kind="stable")
        ee, ff = ee[o2], ff[o2]
        cnt = np.bincount(ee, minlength=self.nE)
        start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        slot = np.arange(ee.size) - start[ee]
        out = np.zeros((self.nE, 4), dtype=np.int64)
        out[ee, slot] = ff + 1
        return out, cnt

    # ---- masks ----------------------------------------------------------------
    def face_inside_count(self, inside):
        """number of inside voxels adjacent to each face."""
        inside = np.asarray(inside, dtype=bool).reshape(-1)
        vf = self.voxel_faces[inside]
        return np.bincount(vf.reshape(-1), minlength=self.nF)

    def edge_inside(self, inside):
        """edges that belong to a face of an inside voxel (setVoxelInside)."""
        inside = np.asarray(inside, dtype=bool).reshape(-1)
        f = np.unique(self.voxel_faces[inside].reshape(-1))
        out = np.zeros(self.nE, dtype=bool)
        out[self.face_edges[f].reshape(-1)] = True
        return out

    def face_inside(self, inside):
        return self.face_inside_count(inside) > 0


def box_mask(g: Grid):
    return np.ones(g.nV, dtype=bool)


def voxel_centers(g: Grid, resolution: float, aabb_min):
    """voxel centres as Voxel3D::center (decmesh3d.cpp:1375): float products
    x*res, summed in double with 0.5*res."""
    r = np.float32(resolution)
    x, y, z = g.voxel_coords()
    mn = np.asarray(aabb_min, dtype=np.float32).astype(np.float64)
    cx = mn[0] + (x.astype(np.float32) * r).astype(np.float64) + 0.5 * np.float64(r)
    cy = mn[1] + (y.astype(np.float32) * r).astype(np.float64) + 0.5 * np.float64(r)
    cz = mn[2] + (z.astype(np.float32) * r).astype(np.float64) + 0.5 * np.float64(r)
    return np.stack([cx, cy, cz], 1)


def sphere_obstacle_mask(g: Grid, resolution: float, aabb_min, radius=0.25, center=(0.0, 0.0, 0.0)):
    """fluid voxels of a box with a solid sphere: a voxel whose centre lies
    strictly inside the sphere is solid."""
    c = voxel_centers(g, resolution, aabb_min) - np.asarray(center, dtype=np.float64)
    return (c * c).sum(1) >= radius * radius
