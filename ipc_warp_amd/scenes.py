"""Procedural geometry for the contact workloads (host side, numpy).

The reference loads meshes from files that do not ship with it; these
generators rebuild the same kinds of input procedurally and deterministically
(seeded):

* tet-grid cubes split 6 tets per cell with the reference's split table
  (fem/geometry.py:36-49) and their boundary surface;
* subdivided icospheres (the wrecking-ball bodies);
* the reference's ground quad (assets/plane.obj: 40 x 40, y = 0);
* a soup assembly that concatenates bodies the way SimComplexBase does
  (fem/geometry.py:174-229): vertices stacked, edges from boundary faces
  (igl.edges of F), one body id per vertex;
* synthetic contact pair streams with a controlled feature mix and
  distance distribution, used by the kernel benchmarks.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Corner order of a grid cell is J = i*4 + j*2 + k (fem/geometry.py:23-25,60-62);
# the 6-tet split table is fem/geometry.py:44-49.
_CELL_TETS = np.array([[4, 1, 5, 3], [5, 4, 3, 6], [3, 5, 6, 7], [0, 1, 4, 2], [1, 4, 2, 3], [4, 2, 3, 6]])


@dataclass
class Soup:
    """Mesh soup handed to every contact kernel (fem/geometry.py:165-171)."""

    x: np.ndarray  # float32 [V,3] current positions (x_transformed / states.x)
    xcs: np.ndarray  # float32 [V,3] rest positions
    edges: np.ndarray  # int32 [E,2]
    triangles: np.ndarray  # int32 [F,3]
    body: np.ndarray  # int32 [V]
    tets: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))  # int32 [T,4]
    n_bodies: int = 1
    body_vertex_offsets: np.ndarray | None = None

    @property
    def n_vertices(self) -> int:
        return int(self.x.shape[0])


def tet_grid_cube(n: int = 8, size: float = 1.0, origin=(0.0, 0.0, 0.0)):
    """(n+1)^3 nodes, 6 n^3 tets. Node index i*(n+1)^2 + j*(n+1) + k."""
    h = size / n
    g = np.arange(n + 1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    V = np.stack([I, J, K], -1).reshape(-1, 3).astype(np.float64) * h + np.asarray(origin, np.float64)
    c = np.arange(n)
    ci, cj, ck = (a.reshape(-1) for a in np.meshgrid(c, c, c, indexing="ij"))
    corners = np.zeros((ci.size, 8), np.int64)
    for J_ in range(8):
        di, dj, dk = J_ // 4, (J_ % 4) // 2, J_ % 2
        corners[:, J_] = (ci + di) * (n + 1) ** 2 + (cj + dj) * (n + 1) + (ck + dk)
    T = corners[:, _CELL_TETS].reshape(-1, 4)
    return V, T.astype(np.int32)


def boundary_faces(V: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Faces of a tet mesh used by exactly one tet, oriented outward."""
    faces = np.concatenate([T[:, [1, 2, 3]], T[:, [0, 3, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 1]]])
    opposite = np.concatenate([T[:, 0], T[:, 1], T[:, 2], T[:, 3]])
    key = np.sort(faces, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    once = cnt[inv.reshape(-1)] == 1
    F, opp = faces[once], opposite[once]
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    inward = np.einsum("ij,ij->i", n, V[opp] - V[F[:, 0]]) > 0
    F[inward] = F[inward][:, [0, 2, 1]]
    return F.astype(np.int32)


def unique_edges(F: np.ndarray) -> np.ndarray:
    """Undirected edges of a triangle list (igl.edges), sorted (min, max)."""
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e = np.sort(e, axis=1)
    return np.unique(e, axis=0).astype(np.int32)


def icosphere(subdivisions: int = 3, radius: float = 0.5):
    """Loop-free midpoint subdivision of the icosahedron, projected to the sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    for _ in range(subdivisions):
        verts = list(V)
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        V, F = np.array(verts), np.array(nf, np.int64)
    return V * radius, F.astype(np.int32)


def ground_quad(y: float = 0.0, half: float = 20.0):
    """assets/plane.obj: a 40 x 40 quad at y = 0 as two triangles."""
    V = np.array([[-half, y, half], [half, y, half], [-half, y, -half], [half, y, -half]], np.float64)
    F = np.array([[1, 2, 0], [1, 3, 2]], np.int32)
    return V, F


def assemble(bodies) -> Soup:
    """bodies: iterable of (V [n,3], F [m,3], T [k,4] or None)."""
    Vs, Fs, Es, Ts, B, offsets = [], [], [], [], [], [0]
    base = 0
    for b, (V, F, T) in enumerate(bodies):
        Vs.append(np.asarray(V, np.float64))
        Fs.append(np.asarray(F, np.int64) + base)
        Es.append(unique_edges(np.asarray(F)) + base)
        if T is not None and len(T):
            Ts.append(np.asarray(T, np.int64) + base)
        B.append(np.full(len(V), b, np.int32))
        base += len(V)
        offsets.append(base)
    x = np.concatenate(Vs).astype(np.float32)
    return Soup(
        x=x,
        xcs=x.copy(),
        edges=np.concatenate(Es).astype(np.int32),
        triangles=np.concatenate(Fs).astype(np.int32),
        body=np.concatenate(B),
        tets=np.concatenate(Ts).astype(np.int32) if Ts else np.zeros((0, 4), np.int32),
        n_bodies=len(B),
        body_vertex_offsets=np.asarray(offsets, np.int64),
    )


# ------------------------------------------------------------- workloads ----
def two_cubes_on_ground(seed: int = 0, n: int = 8, gap: float = 2.0e-3, jitter: float = 1.0e-4) -> Soup:
    """Two n^3-cell tet cubes (edge 1.0) stacked `gap` apart above the ground quad.

    The bottom cube sits `gap` above y = 0, the top cube `gap` above the
    bottom one (origins 1 + gap apart); vertices get U(-jitter, jitter).
    """
    rng = np.random.default_rng(seed)
    V1, T1 = tet_grid_cube(n, 1.0, (0.0, gap, 0.0))
    V2, T2 = tet_grid_cube(n, 1.0, (0.0, 1.0 + 2.0 * gap, 0.0))
    bodies = []
    for V, T in ((V1, T1), (V2, T2)):
        V = V + rng.uniform(-jitter, jitter, V.shape)
        bodies.append((V, boundary_faces(V, T), T))
    Vg, Fg = ground_quad(0.0)
    bodies.append((Vg, Fg, None))
    return assemble(bodies)


def cube_drop(seed: int = 0, n: int = 8):
    """Seeded variants of the two-cube stack for the warm start's trajectories.

    The bottom cube starts 2e-3 above the ground at rest; the top cube is
    shifted U(-0.25, 0.25) in x and z, starts a gap U(2e-3, 3e-2) above the
    bottom cube and moves with (N(0, 0.2), -U(0, 1.5), N(0, 0.2)) m/s.
    Returns (soup, per-vertex velocity float64 [V, 3], fixed mask [V]).
    """
    rng = np.random.default_rng(seed)
    gap = rng.uniform(2.0e-3, 3.0e-2)
    shift = rng.uniform(-0.25, 0.25, 2)
    vel = np.array([rng.normal(0, 0.2), -rng.uniform(0.0, 1.5), rng.normal(0, 0.2)])
    V1, T1 = tet_grid_cube(n, 1.0, (0.0, 2.0e-3, 0.0))
    V2, T2 = tet_grid_cube(n, 1.0, (shift[0], 1.0 + 2.0e-3 + gap, shift[1]))
    bodies = []
    for V, T in ((V1, T1), (V2, T2)):
        V = V + rng.uniform(-1.0e-4, 1.0e-4, V.shape)
        bodies.append((V, boundary_faces(V, T), T))
    Vg, Fg = ground_quad(0.0)
    bodies.append((Vg, Fg, None))
    soup = assemble(bodies)
    v = np.zeros((soup.x.shape[0], 3))
    v[soup.body == 1] = vel
    return soup, v, soup.body == 2


def icosphere_grid(seed: int = 0, per_axis: int = 4, spacing: float = 1.001, subdivisions: int = 3,
                   radius: float = 0.5):
    """per_axis^3 icospheres on a grid `spacing` apart and seeded N(0,1) body velocities.

    Returns (soup, velocity [n_bodies,3] float64).
    """
    rng = np.random.default_rng(seed)
    Vs, Fs = icosphere(subdivisions, radius)
    bodies = []
    for i in range(per_axis):
        for j in range(per_axis):
            for k in range(per_axis):
                bodies.append((Vs + spacing * np.array([i, j, k], np.float64), Fs, None))
    soup = assemble(bodies)
    vel = rng.standard_normal((soup.n_bodies, 3))
    return soup, vel


def _random_rotations(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def affine_icospheres(seed: int = 0, per_axis: int = 3, spacing: float = 0.999, subdivisions: int = 2,
                      radius: float = 0.5, shear: float = 0.0):
    """Icospheres as affine bodies on a per_axis^3 grid `spacing` apart.

    Body b has A_b = (I + shear N(0,1)) R_b with a seeded rotation R_b and
    p_b at its grid point; xcs holds the body-local rest coordinates (the
    sphere centred at the origin) and x = A_b xcs + p_b in float32.
    Returns (soup, q float32 [n_bodies * 4, 3] as (p, A[:,0], A[:,1], A[:,2])).
    """
    rng = np.random.default_rng(seed)
    Vs, Fs = icosphere(subdivisions, radius)
    nb = per_axis ** 3
    bodies = [(Vs, Fs, None)] * nb
    soup = assemble(bodies)
    R = _random_rotations(rng, nb)
    A = (np.eye(3)[None] + shear * rng.standard_normal((nb, 3, 3))) @ R
    grid = np.stack(np.meshgrid(*[np.arange(per_axis)] * 3, indexing="ij"), -1).reshape(-1, 3)
    p = spacing * grid.astype(np.float64)
    q = np.concatenate([p[:, None, :], A.transpose(0, 2, 1)], axis=1).astype(np.float32)  # rows: p, A[:,0..2]
    rest = soup.xcs.astype(np.float64)
    b = soup.body
    qd = q.astype(np.float64)
    x = qd[b, 0] + np.einsum("vij,vj->vi", qd[b, 1:].transpose(0, 2, 1), rest)
    soup.x = x.astype(np.float32)
    return soup, q.reshape(-1, 3)


def affine_step(q: np.ndarray, rest: np.ndarray, body: np.ndarray, seed: int = 0, move: float = 0.02,
                turn: float = 0.05):
    """A seeded affine step for the bodies of q: p moves N(0, move), A turns
    by a rotation of angle ~ turn. Returns (dq, dx) with q1 = q - dq and
    x1 = x - dx, x1 = A1 rest + p1 evaluated like affine_icospheres."""
    rng = np.random.default_rng(seed)
    nb = q.shape[0] // 4
    qd = q.reshape(nb, 4, 3).astype(np.float64)
    axis = rng.standard_normal((nb, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = turn * rng.standard_normal(nb)
    K = np.zeros((nb, 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -axis[:, 2], axis[:, 1], -axis[:, 0]
    K = K - K.transpose(0, 2, 1)
    Rs = np.eye(3)[None] + np.sin(ang)[:, None, None] * K + (1 - np.cos(ang))[:, None, None] * (K @ K)
    A0 = qd[:, 1:].transpose(0, 2, 1)
    A1 = Rs @ A0
    p1 = qd[:, 0] + move * rng.standard_normal((nb, 3))
    q1 = np.concatenate([p1[:, None], A1.transpose(0, 2, 1)], axis=1).astype(np.float32)
    x0 = qd[body, 0] + np.einsum("vij,vj->vi", A0[body], rest)
    x1 = p1[body] + np.einsum("vij,vj->vi", A1[body], rest)
    dq = (q.reshape(nb, 4, 3) - q1).reshape(-1, 3).astype(np.float32)
    dx = (x0.astype(np.float32) - x1.astype(np.float32)).astype(np.float32)
    return dq, dx


# --------------------------------------------------------- pair streams ----
def _orthonormal_frames(rng, n):
    a = rng.standard_normal((n, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    t = np.where(np.abs(a[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    b = np.cross(a, t)
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    c = np.cross(a, b)
    return a, b, c


def pair_stream(kind: str, n: int, seed: int = 0, dhat: float = 1.0e-3, scale: float = 0.05,
                center_spread: float = 1.0, dtype=np.float32):
    """Synthetic active contacts with 4 private vertices each.

    Distances d/dhat are log-uniform on [1e-3, 1) with 15% forced below
    1e-2 (the smallest distance, 1e-6 at dhat = 1e-3, stays well above the
    fp32 spacing of coordinates in the unit cube). PT feature mix: 50% face, 30% edge, 20% vertex. EE mix: 70%
    interior, 20% endpoint, 10% near-parallel. Primitives have size
    ~`scale`, positions spread over a cube of side `center_spread`.
    Returns (x [4n,3] in `dtype` (float32, or float64 for the fp64-scalar
    kernels), contacts int32 [n,4]) with contacts[i] = 4i .. 4i+3 in the
    reference's slot order (PT: p, t1, t0, t2).
    """
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    ratio = np.exp(np.log(1e-3) + (np.log(1.0) - np.log(1e-3)) * rng.random(n))
    small = rng.random(n) < 0.15
    ratio[small] = np.exp(np.log(1e-3) + (np.log(1e-2) - np.log(1e-3)) * rng.random(small.sum()))
    d = ratio * dhat * 0.999
    nrm, t1, t2 = _orthonormal_frames(rng, n)
    centre = rng.uniform(0, center_spread, (n, 3))
    L = scale * (0.5 + rng.random((n, 1)))
    x = np.zeros((n, 4, 3))
    if kind == "pt":
        # triangle in the plane spanned by t1, t2 through `centre`
        ang = rng.uniform(0, 2 * np.pi, (n, 1)) + np.array([[0.0, 2.1, 4.2]]) + rng.uniform(-0.4, 0.4, (n, 3))
        rad = L * (0.6 + 0.4 * rng.random((n, 3)))
        tri = centre[:, None, :] + rad[..., None] * (np.cos(ang)[..., None] * t1[:, None, :] +
                                                      np.sin(ang)[..., None] * t2[:, None, :])
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        face = u < 0.5
        edge = (u >= 0.5) & (u < 0.8)
        p = np.zeros((n, 3))
        # face: above a random interior point along the face normal
        w = rng.dirichlet([2, 2, 2], n)
        fn = np.cross(b - a, c - a)
        fn /= np.linalg.norm(fn, axis=1, keepdims=True)
        p[face] = (w[:, :1] * a + w[:, 1:2] * b + w[:, 2:] * c)[face] + d[face, None] * fn[face]
        # edge: outside edge ab, in the plane spanned by the outward edge normal and the face normal
        s = rng.uniform(0.1, 0.9, n)[:, None]
        q = a + s * (b - a)
        eo = np.cross(b - a, fn)
        eo /= np.linalg.norm(eo, axis=1, keepdims=True)
        eo *= np.sign(np.einsum("ij,ij->i", eo, q - c))[:, None]
        phi = rng.uniform(0.2, 1.3, n)[:, None]
        p[edge] = (q + d[:, None] * (np.cos(phi) * eo + np.sin(phi) * fn))[edge]
        # vertex: outside vertex a, along the outward bisector-ish direction
        vo = (a - (b + c) / 2)
        vo /= np.linalg.norm(vo, axis=1, keepdims=True)
        vdir = vo + 0.3 * fn
        vdir /= np.linalg.norm(vdir, axis=1, keepdims=True)
        vert = u >= 0.8
        p[vert] = (a + d[:, None] * vdir)[vert]
        # record order (p, t1, t0, t2): slots 1..3 are the triangle (a, b, c)
        x[:, 0], x[:, 1], x[:, 2], x[:, 3] = p, a, b, c
    else:
        # edge a along t1 through centre, edge b along a direction rotated about nrm
        interior = u < 0.7
        endpoint = (u >= 0.7) & (u < 0.9)
        ang = rng.uniform(0.3, np.pi - 0.3, n)[:, None]
        ang[~(interior | endpoint)] = rng.uniform(1e-4, 5e-3, (~(interior | endpoint)).sum())[:, None]
        db = np.cos(ang) * t1 + np.sin(ang) * t2
        sa = rng.uniform(0.2, 0.8, n)[:, None]
        sb = rng.uniform(0.2, 0.8, n)[:, None]
        a0 = centre - sa * L * t1
        a1 = centre + (1 - sa) * L * t1
        pb = centre + d[:, None] * nrm
        b0 = pb - sb * L * db
        b1 = pb + (1 - sb) * L * db
        # endpoint: edge b ends just above edge a's interior and rises away
        # from it, so its end point b1 is the closest feature
        e = endpoint
        rise = 0.5 * db + nrm
        rise /= np.linalg.norm(rise, axis=1, keepdims=True)
        b1[e] = pb[e]
        b0[e] = pb[e] + L[e] * rise[e]
        x[:, 0], x[:, 1], x[:, 2], x[:, 3] = a0, a1, b0, b1
    contacts = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    return x.reshape(-1, 3).astype(dtype), contacts


def ccd_stream(kind: str, n: int, seed: int = 0, hit_fraction: float = 0.3, coord: float = 10.0,
               scale: float = 0.05, dtype=np.float32):
    """Synthetic CCD candidates with 4 private vertices each.

    Primitives of size ~`scale` centred uniformly in [-coord, coord]^3. A
    fraction `hit_fraction` of the candidates move into contact at a time
    t ~ U(0.05, 0.95) of the step (PT: the point crosses the triangle's
    plane inside the moving triangle; EE: edge b sweeps through an interior
    point of edge a); the rest move apart (the point leaves along the face
    normal, edge b leaves along the common normal) and report no impact.
    Returns (x0, x1 `dtype` [4n,3], quads int32 [n,4], t float64 [n] with
    t = 1 for the misses); quads[i] = 4i .. 4i+3 in the TOI argument order
    (PT: p, t0, t1, t2; EE: a0, a1, b0, b1). With dtype float64 the
    positions are the unrounded construction, so t is exact up to double
    rounding.
    """
    rng = np.random.default_rng(seed)
    hit = rng.random(n) < hit_fraction
    t = np.where(hit, rng.uniform(0.05, 0.95, n), 1.0)
    centre = rng.uniform(-coord, coord, (n, 3))
    if kind == "pt":
        tri0 = centre[:, None, :] + scale * rng.standard_normal((n, 3, 3))
        vel = 0.5 * scale * rng.standard_normal((n, 3))
        w = rng.dirichlet([3, 3, 3], n)
        nrm = np.cross(tri0[:, 1] - tri0[:, 0], tri0[:, 2] - tri0[:, 0])
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        speed = scale * rng.uniform(0.5, 2.0, (n, 1))
        th = np.where(hit, t, 0.5)[:, None]
        meet = np.einsum("ij,ijk->ik", w, tri0) + th * vel
        # hits approach the plane against the normal; misses start on the
        # other side and leave, so the point never reaches the plane
        vp = np.where(hit[:, None], -nrm * speed, nrm * speed) + vel
        p0 = np.where(hit[:, None], meet - th * vp, meet + nrm * speed)
        X0 = np.concatenate([p0[:, None], tri0], axis=1)
        X1 = np.concatenate([(p0 + vp)[:, None], tri0 + vel[:, None, :]], axis=1)
    else:
        a0 = centre + scale * rng.standard_normal((n, 3))
        a1 = centre + scale * rng.standard_normal((n, 3))
        b_dir = np.cross(a1 - a0, rng.standard_normal((n, 3)))
        b_dir *= scale / np.linalg.norm(b_dir, axis=1, keepdims=True)
        s, u = rng.uniform(0.2, 0.8, (2, n, 1))
        meet = a0 + s * (a1 - a0)
        common = np.cross(a1 - a0, b_dir)
        common /= np.linalg.norm(common, axis=1, keepdims=True)
        vb = scale * rng.standard_normal((n, 3))
        th = t[:, None]
        # hits pass through `meet` at time t; misses start one edge length off
        # along the common normal and move further away
        away = common * scale * (1.0 + rng.random((n, 1)))
        vb = np.where(hit[:, None], vb, common * np.abs(vb).sum(axis=1, keepdims=True))
        b0 = np.where(hit[:, None], meet - u * b_dir - th * vb, meet - u * b_dir + away)
        b1 = b0 + b_dir
        X0 = np.stack([a0, a1, b0, b1], axis=1)
        X1 = np.stack([a0, a1, b0 + vb, b1 + vb], axis=1)
    x0 = X0.reshape(-1, 3).astype(dtype)
    x1 = X1.reshape(-1, 3).astype(dtype)
    return x0, x1, np.arange(4 * n, dtype=np.int32).reshape(n, 4), t


def borderline_pairs(offset: float, thickness: float, seed: int = 0, n_case: int = 256) -> Soup:
    """Crossing edge pairs and points over triangles at distances within a few
    ulps of 2 thickness (and a spread around it), placed at coordinates
    ~offset: the cases where a DCD box inflated by less than thickness +
    buffer could drop a pair the fp32 distance test accepts. n_case cases
    of 8 vertices each: two edges crossing at distance d, and a triangle
    with a point at distance d over its interior or beyond an edge; one case
    in four is axis-aligned, so that the boxes' gap equals the distance."""
    rng = np.random.default_rng(seed)
    L = 4 * thickness
    ulp = float(np.spacing(np.float32(max(abs(offset), 1.0))))
    xs, edges, tris = [], [], []
    for k in range(n_case):
        base = np.array([offset + (k % 16) * 20 * L, offset + (k // 16) * 20 * L, offset], np.float64)
        step = rng.integers(-24, 25) * ulp if k % 2 else rng.uniform(-0.05, 0.05) * 2 * thickness
        d = 2 * thickness + step
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        a = rng.normal(size=3)
        a -= a.dot(u) * u
        a /= np.linalg.norm(a)
        b = np.cross(u, a)
        if k % 8 >= 6:  # axis-aligned: the boxes' gap along u is the distance
            u, a, b = np.eye(3)[2], np.eye(3)[0], np.eye(3)[1]
        v0 = len(xs)
        # edge (a0, a1) and edge (b0, b1) crossing at distance d along u
        xs += [base - L * a, base + L * a, base + d * u - L * b, base + d * u + L * b]
        edges += [(v0, v0 + 1), (v0 + 2, v0 + 3)]
        # triangle in the plane through base + 6L u normal to u, and a point at
        # distance d from it: over its interior (k mod 4 < 2) or beyond an edge
        c = base + 6 * L * u
        t0, t1, t2 = c - L * a - L * b, c + L * a - L * b, c + L * b
        p = c + d * u if k % 4 < 2 else c - L * b - (d / np.sqrt(2)) * (b - u)
        xs += [t0, t1, t2, p]
        tris.append((v0 + 4, v0 + 5, v0 + 6))
    x = np.asarray(xs, np.float32)
    return Soup(x=x, xcs=x.copy(), edges=np.asarray(edges, np.int32), triangles=np.asarray(tris, np.int32),
                body=np.zeros(len(x), np.int32))


# ------------------------------------------------------- rod bundle (twist) ----
def hex_disk(rings: int, radius: float):
    """Triangulated regular hexagon of circumradius `radius`: the triangular
    lattice points with |i|, |j|, |i + j| <= rings (spacing radius / rings),
    corners at angles 0, 60, ..., 300 degrees. Returns (P float64 [n, 2],
    triangles int32 [6 rings^2, 3]), counter-clockwise."""
    pts, index = [], {}
    for i in range(-rings, rings + 1):
        for j in range(-rings, rings + 1):
            if abs(i + j) <= rings:
                index[(i, j)] = len(pts)
                pts.append((i + 0.5 * j, (3.0 ** 0.5 / 2.0) * j))
    tris = []
    for i in range(-rings - 1, rings + 1):
        for j in range(-rings - 1, rings + 1):
            a, b = index.get((i, j)), index.get((i + 1, j))
            c, d = index.get((i, j + 1)), index.get((i + 1, j + 1))
            if a is not None and b is not None and c is not None:
                tris.append((a, b, c))
            if b is not None and c is not None and d is not None:
                tris.append((b, d, c))
    return np.asarray(pts, np.float64) * (radius / rings), np.asarray(tris, np.int32)


def extruded_prism_tets(section_tris: np.ndarray, n_section: int, segments: int) -> np.ndarray:
    """Tets of a triangle section extruded over `segments` layers, 3 per prism.

    Layer s holds vertices s * n_section + k. A prism over triangle (a, b, c)
    with a < b < c (section indices) splits into (a0 b0 c0 c1), (a0 b0 b1 c1),
    (a0 a1 b1 c1): the diagonal of each side quad runs from the smaller
    index's lower vertex, so neighbouring prisms agree on it (conforming)."""
    tri = np.sort(section_tris, axis=1).astype(np.int64)
    s = np.arange(segments, dtype=np.int64)[:, None, None]
    lo = tri[None] + s * n_section
    hi = lo + n_section
    a0, b0, c0 = lo[..., 0], lo[..., 1], lo[..., 2]
    a1, b1, c1 = hi[..., 0], hi[..., 1], hi[..., 2]
    T = np.stack([np.stack([a0, b0, c0, c1], -1), np.stack([a0, b0, b1, c1], -1),
                  np.stack([a0, a1, b1, c1], -1)], axis=2)
    return T.reshape(-1, 4).astype(np.int32)


def _smoothstep(t):
    t = np.clip(t, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


ROD_TWIST_RATE = 2.0  # rad/s at each end cap (the config); rods_twist.py:36 uses 0.4 pi
ROD_TWIST_DT = 0.025  # RodsTwist's default time step (rods_twist.py:85)
ROD_THICKNESS = 1.0e-3  # rods_twist.py:38,91-92 (thickness and buffer, unit-scaled scene)


def rod_bundle(seed: int = 0, per_side: int = 4, length: float = 10.0, radius: float = 0.1,
               segments: int = 400, spacing: float = 0.25, rings: int = 2, steps: int = 200,
               twist_rate: float = ROD_TWIST_RATE, dt: float = ROD_TWIST_DT, thickness: float = ROD_THICKNESS,
               gap: float = 0.5, gap_jitter: float = 0.15, press: float = 0.6, vertex_noise: float = 0.05):
    """Twisted bundle of hexagonal-prism tet rods at the peak of a rod-twist run.

    per_side^2 rods (length `length`, hexagon circumradius `radius`, `rings`
    lattice rings per section, `segments` layers) along x on a square grid
    `spacing` apart, scaled to unit extent like rods_twist.py:108-113 (so
    `thickness` is in unit-scene coordinates, as the reference's contact
    thickness is). Hexagon corners face the y neighbours, flats the z
    neighbours.

    The state after `steps` steps of `dt` with both end caps turned at
    `twist_rate` in opposite senses (set_twist_displacement,
    rods_twist.py:46-75) is generated rather than simulated: over the middle
    of the bundle the rods are pressed together, each y (z) neighbour pair
    `gap` * l0 apart across corners (flats) with a smooth seeded per-rod
    offset of up to `gap_jitter` * l0, easing back to the rest grid over the
    outer 10 % at each end; then every cross-section is turned about the
    bundle axis by theta(x) = theta_end * (2 x / L - 1). Each cross-section
    is placed rigidly and x is kept, so the rods stay disjoint (the map is
    one-to-one) while neighbours sit within l0 = 2 thickness.

    The step's motion (x1 = x - dx) turns the ends one step further and
    presses the middle by another `press` * l0 (with new per-rod offsets)
    plus U(-1, 1) * `vertex_noise` * l0 per vertex, so the closest
    neighbours meet within the step.

    Returns (soup, dx float32 [V, 3], fixed mask [V] of end-cap vertices).
    """
    rng = np.random.default_rng(seed)
    P2, tri2 = hex_disk(rings, radius)
    nsec = P2.shape[0]
    T1 = extruded_prism_tets(tri2, nsec, segments)
    xs = np.linspace(0.0, length, segments + 1)
    rod_rest = np.zeros(((segments + 1) * nsec, 3))
    rod_rest[:, 0] = np.repeat(xs, nsec)
    rod_rest[:, 1:] = np.tile(P2, (segments + 1, 1))
    # orient every tet positively and find the outward boundary once
    e1 = rod_rest[T1[:, 1]] - rod_rest[T1[:, 0]]
    e2 = rod_rest[T1[:, 2]] - rod_rest[T1[:, 0]]
    e3 = rod_rest[T1[:, 3]] - rod_rest[T1[:, 0]]
    neg = np.einsum("ij,ij->i", np.cross(e1, e2), e3) < 0
    T1[neg] = T1[neg][:, [0, 2, 1, 3]]
    F1 = boundary_faces(rod_rest, T1)

    n_rods = per_side * per_side
    grid = np.array([(a - (per_side - 1) / 2.0, b - (per_side - 1) / 2.0)
                     for a in range(per_side) for b in range(per_side)])
    scale = 1.0 / max(length, spacing * (per_side - 1) + 2.0 * radius)
    L, r, l0 = length * scale, radius * scale, 2.0 * thickness
    rest_c = grid * spacing * scale
    pitch_pressed = np.array([2.0 * r + gap * l0, 3.0 ** 0.5 * r + gap * l0])
    pressed_c = grid * pitch_pressed

    def offsets(rng_):
        # smooth per-rod in-plane offsets, |offset| <= gap_jitter * l0 / 2 per axis
        m = np.arange(1, 4)
        amp = rng_.uniform(-1.0, 1.0, (n_rods, 2, 3)) / m
        phase = rng_.uniform(0, 2 * np.pi, (n_rods, 2, 3))
        norm = 0.5 * gap_jitter * l0 / np.abs(1.0 / m).sum()
        u = (xs * scale / L)[:, None, None, None]
        return norm * (amp[None] * np.sin(2 * np.pi * m * u + phase[None])).sum(-1)  # [S+1, rods, 2]

    w = _smoothstep((xs * scale / L) / 0.1) * _smoothstep((1.0 - xs * scale / L) / 0.1)  # [S+1]
    theta_end = twist_rate * steps * dt

    def place(theta_e, press_by, offs):
        c = (1.0 - w)[:, None, None] * rest_c[None] + w[:, None, None] * (
            pressed_c[None] - press_by * l0 * grid[None] + offs)
        th = theta_e * (2.0 * xs * scale / L - 1.0)
        cs, sn = np.cos(th)[:, None], np.sin(th)[:, None]
        out = np.zeros((n_rods, segments + 1, nsec, 3))
        for k in range(n_rods):
            yz = P2[None] * scale + c[:, k][:, None, :]
            y, z = yz[..., 0], yz[..., 1]
            out[k, ..., 0] = (xs * scale)[:, None]
            out[k, ..., 1] = cs * y - sn * z
            out[k, ..., 2] = sn * y + cs * z
        return out.reshape(n_rods, -1, 3)

    x0 = place(theta_end, 0.0, offsets(rng))
    x1 = place(theta_end + twist_rate * dt, press, offsets(rng))
    x1 = x1 + vertex_noise * l0 * rng.uniform(-1.0, 1.0, x1.shape)
    rest = [rod_rest * scale + np.array([0.0, rest_c[k, 0], rest_c[k, 1]]) for k in range(n_rods)]
    soup = assemble([(rest[k], F1, T1) for k in range(n_rods)])
    soup.x = x0.reshape(-1, 3).astype(np.float32)
    x1 = x1.reshape(-1, 3)
    xr = soup.xcs.astype(np.float64)
    fixed = (xr[:, 0] < 1e-5 * scale) | (xr[:, 0] > L - 1e-5 * scale)
    x1[fixed] = x0.reshape(-1, 3)[fixed]
    dx = (soup.x - x1.astype(np.float32)).astype(np.float32)
    return soup, dx, fixed


# ------------------------------------------------------- screw and nut ----
SCREW_PITCH = 0.2
SCREW_RADIUS = 1.0
# thread profile over one pitch, in sixteenths: crest 2, flank 5, root 4, flank 5
# (a 60-degree ISO-like profile: depth = (5/16) pitch / tan 30 deg)
_PROFILE_16 = np.array([0, 2, 7, 11, 16])


def thread_radius(phase: np.ndarray, major: float = SCREW_RADIUS, pitch: float = SCREW_PITCH) -> np.ndarray:
    """Radius of the screw's thread surface at helical phase `phase` in [0, 1)
    of a pitch: crest (major radius), falling flank, root, rising flank."""
    depth = (5.0 / 16.0) * pitch / np.tan(np.pi / 6.0)
    knots = _PROFILE_16 / 16.0
    level = np.array([0.0, 0.0, 1.0, 1.0, 0.0])
    return major - depth * np.interp(np.mod(phase, 1.0), knots, level)


def _helical_band(n_theta: int, rows: int, turns: int, u0: float, pitch: float, radius_of_u):
    """A surface band wound `turns` times about the y axis.

    Vertex (i, j), i in [0, turns * n_theta], j in [0, rows]: angle
    2 pi i / n_theta, y = u + pitch * i / n_theta with the helical
    coordinate u = u0 + pitch * j / rows, radius radius_of_u(u). (i, rows) is the same point as (i + n_theta, 0)
    and is stored once. Rows follow the helix, so kinks of a thread profile
    at multiples of 1 / rows lie on mesh rows. Returns (V [n, 3], vid(i, j)
    index function over arrays, quads [m, 4] (i, j), (i+1, j), (i+1, j+1),
    (i, j+1))."""
    imax = turns * n_theta
    # stored: j in [0, rows) for every i, plus (i, rows) for i > imax - n_theta
    n_low = (imax + 1) * rows
    top_i = np.arange(imax - n_theta + 1, imax + 1)

    def vid(i, j):
        i = np.asarray(i)
        j = np.asarray(j)
        ii = np.where(j == rows, i + n_theta, i)
        jj = np.where(j == rows, 0, j)
        low = ii * rows + jj
        top = n_low + (i - (imax - n_theta + 1))
        return np.where((j == rows) & (ii > imax), top, low)

    I = np.concatenate([np.repeat(np.arange(imax + 1), rows), top_i])
    J = np.concatenate([np.tile(np.arange(rows), imax + 1), np.full(top_i.size, rows)])
    ang = 2.0 * np.pi * I / n_theta
    rad = radius_of_u(u0 + pitch * (J / rows))
    V = np.stack([rad * np.cos(ang), u0 + pitch * (J / rows + I / n_theta), rad * np.sin(ang)], -1)
    qi, qj = np.meshgrid(np.arange(imax), np.arange(rows), indexing="ij")
    qi, qj = qi.ravel(), qj.ravel()
    quads = np.stack([vid(qi, qj), vid(qi + 1, qj), vid(qi + 1, qj + 1), vid(qi, qj + 1)], -1)
    return V, vid, quads


def _quads_to_tris(q: np.ndarray) -> np.ndarray:
    return np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]])


def _stitch(a: np.ndarray, b: np.ndarray, ta: np.ndarray, tb: np.ndarray) -> np.ndarray:
    """Triangle strip between polylines a and b (vertex ids) with increasing
    parameters ta, tb over the same interval (equal end parameters). Where
    both next parameters are equal the quad a_i a_i+1 b_j+1 b_j is split
    along a_i - b_j+1, so matching points are joined by an edge."""
    tris, i, j = [], 0, 0
    while i < len(a) - 1 or j < len(b) - 1:
        if i < len(a) - 1 and j < len(b) - 1 and ta[i + 1] == tb[j + 1]:
            tris += [(a[i], a[i + 1], b[j + 1]), (a[i], b[j + 1], b[j])]
            i += 1
            j += 1
        elif j == len(b) - 1 or (i < len(a) - 1 and ta[i + 1] < tb[j + 1]):
            tris.append((a[i], a[i + 1], b[j]))
            i += 1
        else:
            tris.append((a[i], b[j + 1], b[j]))
            j += 1
    return np.asarray(tris, np.int64)


def _band_ends(vid, n_theta, rows, turns):
    """Boundary polylines of a band: (bottom row, start wall column, top row,
    end wall column), each ordered by increasing angle or height."""
    imax = turns * n_theta
    bottom = vid(np.arange(n_theta + 1), np.zeros(n_theta + 1, int))
    start = vid(np.zeros(rows + 1, int), np.arange(rows + 1))
    top = vid(np.arange(imax - n_theta, imax + 1), np.full(n_theta + 1, rows))
    end = vid(np.full(rows + 1, imax), np.arange(rows + 1))
    return bottom, start, top, end


def screw_body(n_theta: int, rows: int, turns: int, u0: float, pitch: float = SCREW_PITCH,
               major: float = SCREW_RADIUS, thread_origin: float = 0.0):
    """Solid threaded screw along y: the thread band closed at each end by a
    helicoid ramp (one turn of the band's end row joined to the y axis) and
    a wall in the plane of angle 0 (the band's end column joined to the
    axis). Ramp and wall share their axis points (n_theta must be a multiple
    of rows, so the wall's heights are among the ramp's); the ramp's first
    quad is split so that no ramp triangle lies in the wall's plane.
    Returns (V, F) with F oriented outward."""
    if n_theta % rows:
        raise ValueError("n_theta must be a multiple of rows")
    V, vid, quads = _helical_band(n_theta, rows, turns, u0, pitch,
                                  lambda u: thread_radius((u - thread_origin) / pitch, major, pitch))
    F = [_quads_to_tris(quads)]
    Vs = [V]
    base = V.shape[0]
    bottom, start, top, end = _band_ends(vid, n_theta, rows, turns)
    for ramp, wall, flip in ((bottom, start, False), (top, end, True)):
        ys = V[ramp, 1]
        axis = base + np.arange(ys.size)
        Vs.append(np.stack([np.zeros_like(ys), ys, np.zeros_like(ys)], -1))
        base += ys.size
        a0, a1, b0, b1 = ramp[:-1], ramp[1:], axis[:-1], axis[1:]
        r = np.concatenate([np.stack([a0, a1, b1], -1), np.stack([a0, b1, b0], -1)])
        m = ys.size - 1  # the first quad: split along a1 - b0 instead
        r[0] = (a0[0], a1[0], b0[0])
        r[m] = (a1[0], b1[0], b0[0])
        w = _stitch(wall, axis, V[wall, 1], ys)
        part = np.concatenate([r[:, ::-1], w])
        F.append(part[:, ::-1] if flip else part)
    return np.concatenate(Vs), np.concatenate(F)[:, ::-1]


def nut_body(n_theta_in: int, n_theta_out: int, rows: int, turns: int, u0: float, clearance: float,
             outer: float, pitch: float = SCREW_PITCH, major: float = SCREW_RADIUS, thread_origin: float = 0.0):
    """Threaded nut along y: inner band = the thread surface of a screw whose
    thread starts at helical coordinate `thread_origin`, pushed out radially
    by `clearance`, outer band a cylinder of radius `outer`,
    joined by ramps (stitched by angle) and walls at angle 0. Returns (V, F)
    with F oriented outward."""
    Vi, vi, qi = _helical_band(n_theta_in, rows, turns, u0, pitch,
                               lambda u: thread_radius((u - thread_origin) / pitch, major, pitch) + clearance)
    Vo, vo, qo = _helical_band(n_theta_out, rows, turns, u0, pitch, lambda u: np.full_like(u, outer))
    off = Vi.shape[0]
    V = np.concatenate([Vi, Vo])
    F = [_quads_to_tris(qi)[:, ::-1], _quads_to_tris(qo + off)]
    bi, si, ti, ei = _band_ends(vi, n_theta_in, rows, turns)
    bo, so, to, eo = _band_ends(vo, n_theta_out, rows, turns)
    tin = np.arange(n_theta_in + 1) / n_theta_in
    tout = np.arange(n_theta_out + 1) / n_theta_out
    F.append(_stitch(bi, bo + off, tin, tout))
    F.append(_stitch(ti, to + off, tin, tout)[:, ::-1])
    F.append(_quads_to_tris(np.stack([si[:-1], so[:-1] + off, so[1:] + off, si[1:]], -1)))
    F.append(_quads_to_tris(np.stack([ei[:-1], eo[:-1] + off, eo[1:] + off, ei[1:]], -1))[:, ::-1])
    return V, np.concatenate(F)[:, ::-1]


def _weld(V: np.ndarray, F: np.ndarray):
    """Merge vertices with equal coordinates (to 1e-12), keeping first-seen order."""
    _, first, inv = np.unique(np.round(V, 12), axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(order.size)
    return V[first[order]], remap[inv.reshape(-1)][F]


SCREW_DT = 0.01  # ScrewAndNut's h (affine_body_dynamics.py:764)
SCREW_MOTOR = -np.pi  # rad/s about y (motors={0: (0, -pi, 0)}, affine_body_dynamics.py:770)


def screw_and_nut(seed: int = 0, n_theta: int = 512, n_theta_nut: int = 460, n_theta_outer: int = 64,
                  rows: int = 32, screw_turns: int = 4, nut_turns: int = 3, clearance: float = 1.0e-3,
                  outer: float = 1.4, shift: float = 2.0e-4, h: float = SCREW_DT, omega: float = SCREW_MOTOR):
    """Threaded screw (body 0) inside a matching fixed nut (body 1), as affine bodies.

    The screw is a solid of major radius 1 with an ISO-like 60-degree thread
    of pitch 0.2 over `screw_turns` turns; the nut's threaded bore is the
    screw's thread surface pushed out radially by `clearance` (a normal gap
    of clearance/2 on the flanks), `nut_turns` turns long and centred, so
    the screw's thread shows half a turn beyond each end. Mesh rows follow
    the helix (`rows` per pitch, a multiple of 16 so the profile's kinks lie
    on rows); the screw, the nut's bore and the nut's outside use
    `n_theta`, `n_theta_nut` and `n_theta_outer` columns per turn, so the
    two threaded surfaces are not tessellated alike.

    Seeded placement: the screw sits off the nut's axis by U(-shift, shift)
    in x, y and z (every gap stays positive). The step turns the screw about
    its axis by omega * h (the reference's motor, affine_body_dynamics.py:
    760-770), which drives its flanks into the nut's within the step.

    Returns (soup, q float32 [2 * 4, 3] = (p, A[:,0], A[:,1], A[:,2]) per
    body, dq float32 like q (q1 = q - dq), dx float32 [V, 3] (x1 = x - dx),
    fixed_bodies int32 [2] = (0, 1)). Body-local rest coordinates are in
    soup.xcs; soup.x = A xcs + p.
    """
    if rows % 16:
        raise ValueError("rows must be a multiple of 16")
    rng = np.random.default_rng(seed)
    P = SCREW_PITCH
    u0_s = -(screw_turns + 1) * P / 2.0
    Vs, Fs = screw_body(n_theta, rows, screw_turns, u0_s, thread_origin=u0_s)
    u0_n = u0_s + P * (screw_turns - nut_turns) / 2.0
    Vn, Fn = nut_body(n_theta_nut, n_theta_outer, rows, nut_turns, u0_n, clearance, outer, thread_origin=u0_s)
    soup = assemble([(Vs, Fs, None), (Vn, Fn, None)])
    # body frames: origin on the axis at the middle of each body, A = I
    p = np.array([rng.uniform(-shift, shift, 3), [0.0, 0.0, 0.0]])
    local = soup.xcs.astype(np.float64)
    soup.xcs = local.astype(np.float32)
    q = np.zeros((2, 4, 3))
    q[:, 0] = p
    q[:, 1:] = np.eye(3)[None]
    q32 = q.astype(np.float32)
    qd = q32.astype(np.float64)
    b = soup.body
    rest = soup.xcs.astype(np.float64)
    x = qd[b, 0] + np.einsum("vij,vj->vi", qd[b, 1:].transpose(0, 2, 1), rest)
    soup.x = x.astype(np.float32)
    ang = omega * h
    c, s = np.cos(ang), np.sin(ang)
    Ry = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    q1 = qd.copy()
    q1[0, 1:] = (Ry @ qd[0, 1:].T).T  # A1 = Ry A0 (rows of q hold A's columns)
    q1 = q1.astype(np.float32)
    q1d = q1.astype(np.float64)
    x1 = q1d[b, 0] + np.einsum("vij,vj->vi", q1d[b, 1:].transpose(0, 2, 1), rest)
    dq = (q32 - q1).reshape(-1, 3)
    dx = (soup.x - x1.astype(np.float32)).astype(np.float32)
    return soup, q32.reshape(-1, 3), dq.astype(np.float32), dx, np.array([0, 1], np.int32)
