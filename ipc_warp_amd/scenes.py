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
                center_spread: float = 1.0):
    """Synthetic active contacts with 4 private vertices each.

    Distances d/dhat are log-uniform on [1e-3, 1) with 15% forced below
    1e-2 (the smallest distance, 1e-6 at dhat = 1e-3, stays well above the
    fp32 spacing of coordinates in the unit cube). PT feature mix: 50% face, 30% edge, 20% vertex. EE mix: 70%
    interior, 20% endpoint, 10% near-parallel. Primitives have size
    ~`scale`, positions spread over a cube of side `center_spread`.
    Returns (x float32 [4n,3], contacts int32 [n,4]) with contacts[i] =
    4i .. 4i+3 in the reference's slot order (PT: p, t1, t0, t2).
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
    return x.reshape(-1, 3).astype(np.float32), contacts


def ccd_stream(kind: str, n: int, seed: int = 0, hit_fraction: float = 0.3, coord: float = 10.0,
               scale: float = 0.05):
    """Synthetic CCD candidates with 4 private vertices each.

    Primitives of size ~`scale` centred uniformly in [-coord, coord]^3. A
    fraction `hit_fraction` of the candidates move into contact at a time
    t ~ U(0.05, 0.95) of the step (PT: the point crosses the triangle's
    plane inside the moving triangle; EE: edge b sweeps through an interior
    point of edge a); the rest move apart (the point leaves along the face
    normal, edge b leaves along the common normal) and report no impact.
    Returns (x0, x1 float32 [4n,3], quads int32 [n,4], t float64 [n] with
    t = 1 for the misses); quads[i] = 4i .. 4i+3 in the TOI argument order
    (PT: p, t0, t1, t2; EE: a0, a1, b0, b1).
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
    x0 = X0.reshape(-1, 3).astype(np.float32)
    x1 = X1.reshape(-1, 3).astype(np.float32)
    return x0, x1, np.arange(4 * n, dtype=np.int32).reshape(n, 4), t
