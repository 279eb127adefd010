"""Triangle surface meshes used as assembly inputs.

A mesh is the pair (vertices, faces) that CompScienceMeshes' ``Mesh`` holds:
``vertices`` is (nv, 3) float64, ``faces`` is (nt, 3) int32 with 0-based
vertex ids, oriented so that (v1 - v3) x (v2 - v3) points out of the body.
The cell chart is x(b1, b2) = v3 + b1 (v1 - v3) + b2 (v2 - v3), the convention
of the reference's element packing (load_parameters.jl:25-46).

Generators:
  * :func:`icosphere` -- geodesic sphere, level L has 20*4^L faces.
  * :func:`perturb_radial` -- seeded sum of real spherical harmonics applied
    to the radius.
  * :func:`meshcuboid` -- structured cuboid surface with the face-diagonal
    layout whose touching-pair counts match the ones the reference recorded
    for ``meshcuboid(1,1,1,0.5/d)`` (data/GPU/Saut-Swab.txt:50-55).
  * :func:`ellipsoid` -- triaxial ellipsoid from a cuboid grid projected
    radially, with a seeded tangential jitter.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Mesh:
    vertices: np.ndarray  # (nv, 3) float64
    faces: np.ndarray  # (nt, 3) int32

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64)
        self.faces = np.ascontiguousarray(self.faces, dtype=np.int32)
        if self.vertices.ndim != 2 or self.vertices.shape[1] != 3:
            raise ValueError("vertices must be (nv, 3)")
        if self.faces.ndim != 2 or self.faces.shape[1] != 3:
            raise ValueError("faces must be (nt, 3)")

    @property
    def nv(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def nt(self) -> int:
        return int(self.faces.shape[0])

    def areas(self) -> np.ndarray:
        v = self.vertices[self.faces]
        n = np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2])
        return 0.5 * np.linalg.norm(n, axis=1)

    def max_edge(self) -> float:
        v = self.vertices[self.faces]
        e = np.concatenate([v[:, 0] - v[:, 1], v[:, 1] - v[:, 2], v[:, 2] - v[:, 0]])
        return float(np.linalg.norm(e, axis=1).max())

    def vertex_degrees(self) -> np.ndarray:
        return np.bincount(self.faces.ravel(), minlength=self.nv)


def _orient_outward(vertices: np.ndarray, faces: np.ndarray, centre=(0.0, 0.0, 0.0)) -> np.ndarray:
    v = vertices[faces]
    n = np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2])
    c = v.mean(axis=1) - np.asarray(centre)
    flip = np.einsum("ij,ij->i", n, c) < 0
    faces = faces.copy()
    faces[flip] = faces[flip][:, [1, 0, 2]]
    return faces


def icosphere(level: int, radius: float = 1.0) -> Mesh:
    """Geodesic icosphere: icosahedron, ``level`` midpoint subdivisions, vertices
    projected to the sphere after each subdivision.  Children of a face are
    stored consecutively, so face order is spatially coherent."""
    if level < 0:
        raise ValueError("level must be >= 0")
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [
        (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
        (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
        (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
    ]
    faces = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    V = [np.asarray(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    F = list(faces)
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf.extend([(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)])
        F = nf
    vertices = np.array(V) * radius
    faces = _orient_outward(vertices, np.array(F, dtype=np.int32))
    return Mesh(vertices, faces)


def real_sph_harm(l: int, m: int, polar: np.ndarray, azimuth: np.ndarray) -> np.ndarray:
    """Orthonormal real spherical harmonic Y_l^m."""
    from scipy.special import sph_harm_y

    if m == 0:
        return np.real(sph_harm_y(l, 0, polar, azimuth))
    y = sph_harm_y(l, abs(m), polar, azimuth)
    return np.sqrt(2.0) * (np.real(y) if m > 0 else np.imag(y))


def perturb_radial(mesh: Mesh, seed: int, amplitude: float, n_terms: int = 8, lmax: int = 6,
                   cache: dict | None = None) -> Mesh:
    """r -> r * (1 + delta(theta, phi)), delta = sum of ``n_terms`` real spherical
    harmonics with random degree 1..lmax, random order and N(0,1) weights,
    rescaled so that max |delta| over the vertices equals ``amplitude``.
    ``cache`` (a dict kept by the caller for one base mesh) holds the
    harmonics already evaluated on its vertices."""
    v = mesh.vertices
    r = np.linalg.norm(v, axis=1)
    if amplitude == 0.0:
        return Mesh(v.copy(), mesh.faces.copy())
    rng = np.random.default_rng(seed)
    polar = np.arccos(np.clip(v[:, 2] / r, -1.0, 1.0))
    azimuth = np.arctan2(v[:, 1], v[:, 0])
    delta = np.zeros(mesh.nv)
    for _ in range(n_terms):
        l = int(rng.integers(1, lmax + 1))
        m = int(rng.integers(-l, l + 1))
        a = rng.standard_normal()
        if cache is None:
            y = real_sph_harm(l, m, polar, azimuth)
        else:
            y = cache.get((l, m))
            if y is None:
                y = cache[(l, m)] = real_sph_harm(l, m, polar, azimuth)
        delta += a * y
    delta *= amplitude / np.abs(delta).max()
    return Mesh(v * (1.0 + delta)[:, None], mesh.faces.copy())


def _grid_face(origin, du, dv, nu, nv, vid, verts, avoid):
    """Structured nu x nv quad grid on the cuboid face spanned by du, dv from
    origin, every quad split along the diagonal that does not pass through the
    cuboid corner ``avoid``."""
    def key(p):
        return tuple(np.round(np.asarray(p) * 1e9).astype(np.int64))

    def vert(i, j):
        p = origin + du * (i / nu) + dv * (j / nv)
        k = key(p)
        if k not in vid:
            vid[k] = len(verts)
            verts.append(p)
        return vid[k]

    # corners (0,0) and (nu,nv) lie on the main diagonal direction
    c00 = origin
    cnn = origin + du + dv
    main_hits_avoid = np.allclose(c00, avoid) or np.allclose(cnn, avoid)
    tris = []
    for j in range(nv):
        for i in range(nu):
            a, b, c, d = vert(i, j), vert(i + 1, j), vert(i + 1, j + 1), vert(i, j + 1)
            if not main_hits_avoid:
                tris.extend([(a, b, c), (a, c, d)])
            else:
                tris.extend([(a, b, d), (b, c, d)])
    return tris


def meshcuboid(width: float, height: float, length: float, delta: float) -> Mesh:
    """Surface of the cuboid [0,width] x [0,length] x [0,height] with square
    cells of side ``delta`` (two triangles per cell).  In every face the cells
    are split along the face diagonal that avoids the corner (0,0,0) or
    (width,length,height) lying on that face, so those two corners have degree 3,
    the other six degree 5 and every other vertex degree 6 -- the degree pattern
    behind the touching-pair counts recorded for meshcuboid(1,1,1,0.5/d) in the
    reference (CV 248,784 at d=24, 691,152 at d=40; data/GPU/Saut-Swab.txt:50-55)."""
    nx, ny, nz = (int(round(s / delta)) for s in (width, length, height))
    if min(nx, ny, nz) < 1:
        raise ValueError("delta larger than the cuboid")
    W = np.array([width, 0.0, 0.0])
    L = np.array([0.0, length, 0.0])
    H = np.array([0.0, 0.0, height])
    O = np.zeros(3)
    top = W + L + H
    vid: dict = {}
    verts: list = []
    tris: list = []
    tris += _grid_face(O, W, L, nx, ny, vid, verts, O)        # z = 0
    tris += _grid_face(H, W, L, nx, ny, vid, verts, top)      # z = h
    tris += _grid_face(O, W, H, nx, nz, vid, verts, O)        # y = 0
    tris += _grid_face(L, W, H, nx, nz, vid, verts, top)      # y = l
    tris += _grid_face(O, L, H, ny, nz, vid, verts, O)        # x = 0
    tris += _grid_face(W, L, H, ny, nz, vid, verts, top)      # x = w
    vertices = np.array(verts)
    faces = _orient_outward(vertices, np.array(tris, dtype=np.int32), centre=(W + L + H) / 2)
    return Mesh(vertices, faces)


def ellipsoid(axes=(1.0, 0.7, 0.5), cells=(95, 60, 60), seed: int = 3, jitter: float = 0.15) -> Mesh:
    """Triaxial ellipsoid: a cuboid surface grid with ``cells`` = (nx, ny, nz)
    quads along each axis (4(nx ny + ny nz + nz nx) triangles) is projected
    radially onto the ellipsoid with semi-axes ``axes``.  Interior grid vertices
    get a seeded jitter of ``jitter`` cell widths before projection.  The
    default gives 60,000 triangles and 90,000 edges."""
    nx, ny, nz = cells
    W = np.array([2.0, 0.0, 0.0])
    L = np.array([0.0, 2.0, 0.0])
    H = np.array([0.0, 0.0, 2.0])
    O = np.array([-1.0, -1.0, -1.0])
    top = O + W + L + H
    vid: dict = {}
    verts: list = []
    tris: list = []
    tris += _grid_face(O, W, L, nx, ny, vid, verts, O)
    tris += _grid_face(O + H, W, L, nx, ny, vid, verts, top)
    tris += _grid_face(O, W, H, nx, nz, vid, verts, O)
    tris += _grid_face(O + L, W, H, nx, nz, vid, verts, top)
    tris += _grid_face(O, L, H, ny, nz, vid, verts, O)
    tris += _grid_face(O + W, L, H, ny, nz, vid, verts, top)
    p = np.array(verts)
    # equal-angle parametrisation of each cube coordinate
    q = np.tan(p * (np.pi / 4.0))
    rng = np.random.default_rng(seed)
    onface = np.isclose(np.abs(p), 1.0)
    nedge = onface.sum(axis=1)
    interior = nedge == 1
    h = np.array([2.0 / nx, 2.0 / ny, 2.0 / nz])
    jit = rng.uniform(-1.0, 1.0, size=p.shape) * (jitter * h)
    jit[onface] = 0.0
    q = np.where(interior[:, None], np.tan((p + jit) * (np.pi / 4.0)), q)
    d = q / np.linalg.norm(q, axis=1, keepdims=True)
    a = np.asarray(axes, dtype=np.float64)
    r = 1.0 / np.sqrt(((d / a) ** 2).sum(axis=1))
    vertices = d * r[:, None]
    faces = _orient_outward(vertices, np.array(tris, dtype=np.int32))
    return Mesh(vertices, faces)
