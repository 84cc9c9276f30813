"""Deterministic domain generators and an OBJ reader/writer.

The OBJ dialect is the one the reference loader reads (src/mesh.cpp:76-270):
`v x y [z]` records, 2D closed polylines as `l` records (one record per
boundary loop, its index being the segment's boundary id, mesh.cpp:156-212),
3D triangles as `f` records with 1-based indices.

Every generator returns (verts float64 [n, dim], prims int32 [m, dim],
boundary_ids int32 [m] or None).
"""
from __future__ import annotations

import math

import numpy as np


# ----------------------------------------------------------------- 2D loops
def polygon_loop(n: int, radius_fn, ccw: bool = True):
    """Closed polyline through r(theta_k) at theta_k = 2 pi k / n."""
    k = np.arange(n, dtype=np.float64)
    theta = 2.0 * math.pi * k / n
    r = radius_fn(theta)
    pts = np.stack([r * np.cos(theta), r * np.sin(theta)], axis=1)
    if not ccw:
        pts = pts[::-1].copy()
    return pts


def loops_to_mesh(loops):
    """Concatenate closed loops; the loop index is the boundary id."""
    verts, prims, bids = [], [], []
    base = 0
    for b, pts in enumerate(loops):
        n = len(pts)
        verts.append(pts)
        i = np.arange(n, dtype=np.int32)
        prims.append(np.stack([base + i, base + (i + 1) % n], axis=1))
        bids.append(np.full(n, b, dtype=np.int32))
        base += n
    return (np.concatenate(verts).astype(np.float64), np.concatenate(prims).astype(np.int32),
            np.concatenate(bids).astype(np.int32))


def annulus(n: int = 128, r_in: float = 0.25, r_out: float = 1.0):
    """Two n-gons.  The inner loop comes first (boundary id 0) and runs
    clockwise so the hole has winding number 0 (src/inside.cpp:122-137, :199)."""
    inner = polygon_loop(n, lambda t: np.full_like(t, r_in), ccw=False)
    outer = polygon_loop(n, lambda t: np.full_like(t, r_out), ccw=True)
    return loops_to_mesh([inner, outer])


def circle(n: int = 256, r: float = 1.0):
    return loops_to_mesh([polygon_loop(n, lambda t: np.full_like(t, r))])


def rectangle(xmin: float, xmax: float, ymin: float, ymax: float):
    """The four counter-clockwise edges of a rectangle, boundary id 0
    (make_rectangular_plate, apps/welding/include/welding/rectangular_plate.hpp:11-33)."""
    if not (xmax > xmin and ymax > ymin):
        raise ValueError("rectangle needs increasing coordinate bounds")
    verts = np.array([[xmin, ymin], [xmax, ymin], [xmax, ymax], [xmin, ymax]], np.float64)
    prims = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    return verts, prims, np.zeros(4, np.int32)


def star(n: int, a: float, k: int):
    """r(theta) = 1 + a cos(k theta), counter-clockwise."""
    return loops_to_mesh([polygon_loop(n, lambda t: 1.0 + a * np.cos(k * t))])


def star_family(count: int = 16, seed: int = 2024, n_min: int = 512, n_max: int = 16384):
    """The batched star domains: a ~ U(0.1, 0.4), k ~ U{3..9}, segment count
    log-uniform in [n_min, n_max], from a fixed seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        a = float(rng.uniform(0.1, 0.4))
        k = int(rng.integers(3, 10))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        out.append(dict(a=a, k=k, n=n, mesh=star(n, a, k)))
    return out


# ----------------------------------------------------------------- 3D surfaces
def _snap_unit(v):
    """Make axis points exactly +-1/0 so the bounding box is exactly [-1, 1]^3."""
    v = v.copy()
    near_one = np.abs(np.abs(v) - 1.0) < 1e-12
    v[near_one] = np.sign(v[near_one])
    v[np.abs(v) < 1e-15] = 0.0
    return v


def icosphere(subdivisions: int = 4):
    """Unit icosphere: icosahedron (0, +-1, +-phi) cyclic, each triangle split
    in four with midpoints pushed to the sphere, outward orientation.
    Subdivision 4 gives 2562 vertices and 5120 triangles."""
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0),
         (0, -1, phi), (0, 1, phi), (0, -1, -phi), (0, 1, -phi),
         (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / math.sqrt(1 + phi * phi) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = _snap_unit(np.array(verts))
    F = np.array(faces, dtype=np.int32)
    # outward orientation: normal . centroid > 0
    e1 = V[F[:, 1]] - V[F[:, 0]]
    e2 = V[F[:, 2]] - V[F[:, 0]]
    flip = np.einsum("ij,ij->i", np.cross(e1, e2), V[F].mean(axis=1)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F, None


def torus(R: float = 1.0, r: float = 0.35, nu: int = 200, nv: int = 100):
    """Torus from a periodic nu x nv (u, v) grid, two triangles per quad
    (2 nu nv triangles), outward orientation."""
    u = 2.0 * math.pi * np.arange(nu) / nu
    v = 2.0 * math.pi * np.arange(nv) / nv
    U, Vv = np.meshgrid(u, v, indexing="ij")
    X = (R + r * np.cos(Vv)) * np.cos(U)
    Y = (R + r * np.cos(Vv)) * np.sin(U)
    Z = r * np.sin(Vv)
    verts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    idx = lambda i, j: (i % nu) * nv + (j % nv)
    tris = []
    for i in range(nu):
        for j in range(nv):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    F = np.array(tris, dtype=np.int32)
    # outward: normal points away from the tube centre line
    cen = verts[F].mean(axis=1)
    ring = np.stack([np.cos(np.arctan2(cen[:, 1], cen[:, 0])),
                     np.sin(np.arctan2(cen[:, 1], cen[:, 0])), np.zeros(len(F))], axis=1) * R
    e1 = verts[F[:, 1]] - verts[F[:, 0]]
    e2 = verts[F[:, 2]] - verts[F[:, 0]]
    flip = np.einsum("ij,ij->i", np.cross(e1, e2), cen - ring) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return verts, F, None


def box(lx: float, ly: float, lz: float, nx: int, ny: int, nz: int):
    """Closed box [0,lx]x[0,ly]x[0,lz], each face a uniform grid split into
    triangles, shared edge vertices, outward orientation.
    Triangles: 4 (nx ny + nx nz + ny nz)."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    index = {}
    verts = []

    def vid(i, j, k):
        key = (i, j, k)
        if key not in index:
            index[key] = len(verts)
            verts.append((xs[i], ys[j], zs[k]))
        return index[key]

    tris = []

    def face(fixed_axis, fixed_idx, n1, n2, outward_sign):
        for a in range(n1):
            for b in range(n2):
                q = []
                for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    ijk = [0, 0, 0]
                    free = [ax for ax in range(3) if ax != fixed_axis]
                    ijk[fixed_axis] = fixed_idx
                    ijk[free[0]] = a + da
                    ijk[free[1]] = b + db
                    q.append(vid(*ijk))
                t1, t2 = (q[0], q[1], q[2]), (q[0], q[2], q[3])
                # (free0, free1, fixed) right-handed when fixed_axis is 1 flips
                sign = outward_sign * (1 if fixed_axis != 1 else -1)
                if sign < 0:
                    t1, t2 = (t1[0], t1[2], t1[1]), (t2[0], t2[2], t2[1])
                tris.extend([t1, t2])

    face(2, 0, nx, ny, -1)
    face(2, nz, nx, ny, +1)
    face(1, 0, nx, nz, -1)
    face(1, ny, nx, nz, +1)
    face(0, 0, ny, nz, -1)
    face(0, nx, ny, nz, +1)
    return np.array(verts, dtype=np.float64), np.array(tris, dtype=np.int32), None


def plate_200k():
    """120 x 60 x 10 mm plate at about 0.42 mm spacing: 200316 triangles."""
    return box(120.0, 60.0, 10.0, 283, 141, 24)


# ----------------------------------------------------------------- OBJ
def write_obj(path: str, verts, prims, boundary_ids=None):
    """Write the reference OBJ dialect.  2D loops are written one `l` record
    per boundary id, closed by repeating the first vertex."""
    verts = np.asarray(verts)
    prims = np.asarray(prims)
    dim = prims.shape[1]
    with open(path, "w") as f:
        for p in verts:
            if dim == 2:
                f.write("v %.17g %.17g 0\n" % (p[0], p[1]))
            else:
                f.write("v %.17g %.17g %.17g\n" % (p[0], p[1], p[2]))
        if dim == 2:
            bids = np.zeros(len(prims), np.int32) if boundary_ids is None else np.asarray(boundary_ids)
            for b in np.unique(bids):
                seg = prims[bids == b]
                chain = [int(seg[0, 0])] + [int(s[1]) for s in seg]
                f.write("l " + " ".join(str(i + 1) for i in chain) + "\n")
        else:
            for t in prims:
                f.write("f %d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))


def signed_area_2d(verts, prims):
    v = np.asarray(verts)
    s0, s1 = v[prims[:, 0]], v[prims[:, 1]]
    return 0.5 * np.sum(s0[:, 0] * s1[:, 1] - s1[:, 0] * s0[:, 1])


def enclosed_volume_3d(verts, prims):
    v = np.asarray(verts)
    a, b, c = v[prims[:, 0]], v[prims[:, 1]], v[prims[:, 2]]
    return np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0
