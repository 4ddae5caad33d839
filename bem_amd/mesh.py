"""Synthetic surface meshes for the benchmark and test geometries.

Closed, manifold triangle meshes with outward winding, vertices shared
between faces (the impedance assembly treats two triangles as neighbours
when they share a vertex index, src/Triangle.h:23-30, so duplicated corner
vertices would break the singular-pair treatment).  Coordinates are in the
reference's x3d units; the physical scale `a` (metres per unit) is applied
only through the dimensionless frequency.

  cube(n)        axis-aligned cube, each face an n x n grid of squares, each
                 square split on one fixed diagonal into 2 triangles
  cuboid(...)    box with its own grid count per axis
  icosphere(l)   icosahedron subdivided l times, projected onto the sphere
  cone_tip(...)  rounded cone (spherical apex cap) closed by a flat base
"""
from __future__ import annotations

import numpy as np

__all__ = ["cube", "cuboid", "icosphere", "cone_tip", "translate", "Mesh"]


class Mesh:
    """Vertex coordinates (nv, 3) float64 and triangles (nt, 3) int32."""

    def __init__(self, xyz, tri, name=""):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32)
        self.name = name

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nt(self):
        return self.tri.shape[0]

    @property
    def ne(self):
        return 3 * self.nt // 2

    def translated(self, d):
        return Mesh(self.xyz + np.asarray(d, dtype=np.float64)[None, :], self.tri.copy(), self.name)

    def __repr__(self):
        return f"Mesh({self.name!r}, nv={self.nv}, nt={self.nt})"


def translate(m: Mesh, d) -> Mesh:
    return m.translated(d)


def cuboid(lx=1.0, ly=1.0, lz=1.0, nx=1, ny=1, nz=1, origin=(0.0, 0.0, 0.0), name="cuboid") -> Mesh:
    """Box [0,lx]x[0,ly]x[0,lz] (+origin) with nx, ny, nz squares along the axes."""
    counts = (nx, ny, nz)
    lens = (lx, ly, lz)
    index = {}
    pts = []

    def vid(ijk):
        key = tuple(ijk)
        v = index.get(key)
        if v is None:
            v = len(pts)
            index[key] = v
            pts.append([origin[d] + lens[d] * ijk[d] / counts[d] for d in range(3)])
        return v

    tris = []
    for d in range(3):
        u, w = (d + 1) % 3, (d + 2) % 3
        for side in (0, 1):
            for a in range(counts[u]):
                for b in range(counts[w]):
                    def p(da, db):
                        ijk = [0, 0, 0]
                        ijk[d] = side * counts[d]
                        ijk[u] = a + da
                        ijk[w] = b + db
                        return vid(ijk)
                    p00, p10, p11, p01 = p(0, 0), p(1, 0), p(1, 1), p(0, 1)
                    if side == 1:   # outward normal +e_d = e_u x e_w
                        tris.append((p00, p10, p11))
                        tris.append((p00, p11, p01))
                    else:
                        tris.append((p00, p11, p10))
                        tris.append((p00, p01, p11))
    return Mesh(np.array(pts), np.array(tris), name)


def cube(n=2, edge=1.0, origin=(0.0, 0.0, 0.0)) -> Mesh:
    return cuboid(edge, edge, edge, n, n, n, origin, name=f"cube{n}")


def icosphere(level=3, radius=1.0, center=(0.0, 0.0, 0.0)) -> Mesh:
    """Icosahedron subdivided `level` times: 20*4**level triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
             (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            v = cache.get(key)
            if v is None:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                v = len(verts) - 1
                cache[key] = v
            return v

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    xyz = np.array(verts) * radius + np.asarray(center, dtype=np.float64)[None, :]
    return Mesh(xyz, np.array(faces), name=f"icosphere{level}")


def cone_tip(height=2.0, base_radius=1.0, apex_radius=0.05, n_around=40, n_along=30, n_cap=6, n_base=8,
             name="cone_tip") -> Mesh:
    """Cone along -z with a spherical apex cap, apex at the origin, flat base at z=height.

    The apex cap is tangent to the cone flank; rings are spaced so the apex is
    resolved finely (useful for near-field tip geometries)."""
    half = np.arctan2(base_radius - apex_radius, height)   # approximate half-angle
    half = np.arctan2(base_radius, height)
    # sphere of radius r0 tangent to the flank, bottom point at z=0
    r0 = apex_radius
    zc = r0 / np.sin(half) if np.sin(half) > 0 else r0
    zc = r0  # cap centre so the cap's lowest point is the apex at z=0
    rings = []
    # cap rings: polar angle from 0 (apex) to (pi/2 - half)
    theta_max = np.pi / 2 - half
    for k in range(1, n_cap + 1):
        th = theta_max * k / n_cap
        rings.append((r0 * np.sin(th), zc - r0 * np.cos(th)))
    z0 = rings[-1][1]
    r_start = rings[-1][0]
    # flank rings, geometric spacing towards the base
    for k in range(1, n_along + 1):
        s = (k / n_along) ** 1.5
        z = z0 + s * (height - z0)
        r = r_start + (base_radius - r_start) * (z - z0) / (height - z0)
        rings.append((r, z))
    verts = [(0.0, 0.0, 0.0)]
    ring_ids = []
    for r, z in rings:
        ids = []
        for j in range(n_around):
            ph = 2 * np.pi * j / n_around
            ids.append(len(verts))
            verts.append((r * np.cos(ph), r * np.sin(ph), z))
        ring_ids.append(ids)
    tris = []
    first = ring_ids[0]
    for j in range(n_around):
        tris.append((0, first[(j + 1) % n_around], first[j]))
    for a, b in zip(ring_ids[:-1], ring_ids[1:]):
        for j in range(n_around):
            j1 = (j + 1) % n_around
            tris.append((a[j], a[j1], b[j1]))
            tris.append((a[j], b[j1], b[j]))
    # flat base: concentric rings down to a centre vertex
    last = ring_ids[-1]
    prev = last
    for k in range(1, n_base):
        rr = base_radius * (1 - k / n_base)
        ids = []
        for j in range(n_around):
            ph = 2 * np.pi * j / n_around
            ids.append(len(verts))
            verts.append((rr * np.cos(ph), rr * np.sin(ph), height))
        for j in range(n_around):
            j1 = (j + 1) % n_around
            tris.append((prev[j], prev[j1], ids[j1]))
            tris.append((prev[j], ids[j1], ids[j]))
        prev = ids
    c = len(verts)
    verts.append((0.0, 0.0, height))
    for j in range(n_around):
        tris.append((prev[j], prev[(j + 1) % n_around], c))
    return Mesh(np.array(verts), np.array(tris), name)
