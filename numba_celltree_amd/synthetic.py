"""Seeded synthetic meshes and query sets for benchmarks and parity tests (numpy only).

Every array draws from its own generator ``np.random.default_rng([seed, k])`` so that the
draw order of one array cannot change another.  Workloads:

* ``jittered_grid``: an ``nv x nv`` vertex grid on [0, L]^2; interior vertices jittered by
  U(+-jitter*h) (boundary vertices stay on the boundary, so the mesh covers the square);
  each quad is kept, or split into two CCW triangles with probability ``split``.
  Faces are int32 of width 4 (padded with -1) or 3 when every quad is split.
* ``hex_voronoi``: Voronoi cells of a jittered hexagonal lattice (``nx * ny`` points with
  spacing dx, dy, odd rows shifted by dx/2), each cell the hexagon of the circumcentres of
  the six lattice triangles around its point, clipped to [0, L]^2; width padded with -1.
* query sets: points (a fraction uniform inside, the rest in a frame around the domain),
  axis-aligned boxes with log-uniform sizes, segments with log-uniform lengths.
"""

import numpy as np


def _rng(seed: int, k: int) -> np.random.Generator:
    return np.random.default_rng([seed, k])


def jittered_grid(nv: int, length: float, jitter: float, split: float, seed: int):
    h = length / (nv - 1)
    g = np.linspace(0.0, length, nv)
    x, y = np.meshgrid(g, g, indexing="xy")
    xy = np.column_stack([x.ravel(), y.ravel()])
    interior = np.zeros((nv, nv), dtype=bool)
    interior[1:-1, 1:-1] = True
    interior = interior.ravel()
    jit = _rng(seed, 0).uniform(-jitter * h, jitter * h, size=xy.shape)
    xy[interior] += jit[interior]

    nq = nv - 1
    jj, ii = np.meshgrid(np.arange(nq), np.arange(nq), indexing="ij")
    v0 = (jj * nv + ii).ravel()
    v1, v2, v3 = v0 + 1, v0 + 1 + nv, v0 + nv  # CCW quad: (x,y) (x+h,y) (x+h,y+h) (x,y+h)
    quads = np.column_stack([v0, v1, v2, v3]).astype(np.int32)
    if split >= 1.0:
        t = np.empty((2 * len(quads), 3), dtype=np.int32)
        t[0::2] = quads[:, [0, 1, 2]]
        t[1::2] = quads[:, [0, 2, 3]]
        return xy, t
    mask = np.zeros(len(quads), dtype=bool)
    n_split = int(round(split * len(quads)))
    mask[_rng(seed, 1).choice(len(quads), size=n_split, replace=False)] = True
    keep = quads[~mask]
    sq = quads[mask]
    tri = np.full((2 * len(sq), 4), -1, dtype=np.int32)
    tri[0::2, :3] = sq[:, [0, 1, 2]]
    tri[1::2, :3] = sq[:, [0, 2, 3]]
    faces = np.concatenate([keep, tri])
    order = _rng(seed, 2).permutation(len(faces))
    return xy, faces[order]


def _circumcentres(a, b, c):
    ax, ay = a[..., 0], a[..., 1]
    bx, by = b[..., 0] - ax, b[..., 1] - ay
    cx, cy = c[..., 0] - ax, c[..., 1] - ay
    d = 2.0 * (bx * cy - by * cx)
    b2 = bx * bx + by * by
    c2 = cx * cx + cy * cy
    ux = (cy * b2 - by * c2) / d
    uy = (bx * c2 - cx * b2) / d
    return np.stack([ax + ux, ay + uy], axis=-1)


def _clip_to_box(poly: np.ndarray, length: float) -> np.ndarray:
    """Clip a convex CCW polygon to [0, length]^2 (half-plane at a time)."""
    for axis, bound, keep_below in ((0, 0.0, False), (0, length, True), (1, 0.0, False), (1, length, True)):
        out = []
        n = len(poly)
        for i in range(n):
            p, q = poly[i - 1], poly[i]
            pin = p[axis] <= bound if keep_below else p[axis] >= bound
            qin = q[axis] <= bound if keep_below else q[axis] >= bound
            if qin:
                if not pin:
                    t = (bound - p[axis]) / (q[axis] - p[axis])
                    out.append(p + t * (q - p))
                out.append(q)
            elif pin:
                t = (bound - p[axis]) / (q[axis] - p[axis])
                out.append(p + t * (q - p))
        poly = np.array(out) if out else np.zeros((0, 2))
        if len(poly) < 3:
            return poly
    return poly


def hex_voronoi(nx: int, ny: int, length: float, jitter: float, seed: int):
    dx, dy = length / nx, length / ny
    # lattice with one ring of ghost points so every real point has six neighbours
    jg, ig = np.meshgrid(np.arange(-1, ny + 1), np.arange(-1, nx + 1), indexing="ij")
    px = dx * (ig + 0.25 + 0.5 * (jg % 2))
    py = dy * (jg + 0.5)
    pts = np.stack([px, py], axis=-1).astype(np.float64)
    pts += _rng(seed, 0).uniform(-jitter * min(dx, dy), jitter * min(dx, dy), size=pts.shape)
    J, I = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    odd = (J - 1) % 2 == 1  # real row index parity
    # neighbours in CCW order: E, NE, NW, W, SW, SE
    sh = np.where(odd, 1, 0)
    nbr = [(J, I + 1), (J + 1, I + sh), (J + 1, I + sh - 1), (J, I - 1), (J - 1, I + sh - 1), (J - 1, I + sh)]
    P = pts[J, I]
    N = [pts[a, b] for a, b in nbr]
    hexes = np.stack([_circumcentres(P, N[k], N[(k + 1) % 6]) for k in range(6)], axis=-2)
    hexes = hexes.reshape(-1, 6, 2)
    inside = (hexes[..., 0].min(1) >= 0) & (hexes[..., 0].max(1) <= length) & (
        hexes[..., 1].min(1) >= 0) & (hexes[..., 1].max(1) <= length)
    polys = [hexes[k] if inside[k] else _clip_to_box(hexes[k], length) for k in range(len(hexes))]
    polys = [p for p in polys if len(p) >= 3]
    width = max(len(p) for p in polys)
    counts = np.array([len(p) for p in polys])
    vertices = np.concatenate(polys).astype(np.float64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    faces = np.full((len(polys), width), -1, dtype=np.int32)
    for k in range(width):
        has = counts > k
        faces[has, k] = starts[has] + k
    return vertices, faces


def points(n: int, length: float, inside_fraction: float, frame: float, seed: int) -> np.ndarray:
    n_in = int(round(inside_fraction * n))
    out = np.empty((n, 2))
    out[:n_in] = _rng(seed, 10).uniform(0.0, length, size=(n_in, 2))
    need = n - n_in
    r = _rng(seed, 11)
    got = []
    while need > 0:
        c = r.uniform(-frame, length + frame, size=(2 * need + 16, 2))
        c = c[(c < 0).any(1) | (c > length).any(1)][:need]
        got.append(c)
        need -= len(c)
    if got:
        out[n_in:] = np.concatenate(got)
    return out[_rng(seed, 12).permutation(n)]


def boxes(n: int, length: float, h: float, lo: float, hi: float, seed: int) -> np.ndarray:
    c = _rng(seed, 20).uniform(0.0, length, size=(n, 2))
    wh = np.exp(_rng(seed, 21).uniform(np.log(lo * h), np.log(hi * h), size=(n, 2)))
    return np.column_stack([c[:, 0] - wh[:, 0] / 2, c[:, 0] + wh[:, 0] / 2, c[:, 1] - wh[:, 1] / 2, c[:, 1] + wh[:, 1] / 2])


def segments(n: int, length: float, h: float, lo: float, hi: float, seed: int) -> np.ndarray:
    a = _rng(seed, 30).uniform(0.0, length, size=(n, 2))
    ln = np.exp(_rng(seed, 31).uniform(np.log(lo * h), np.log(hi * h), size=n))
    th = _rng(seed, 32).uniform(0.0, 2 * np.pi, size=n)
    b = a + np.column_stack([ln * np.cos(th), ln * np.sin(th)])
    return np.stack([a, b], axis=1)


def sinusoid_fields(xy: np.ndarray, n_fields: int, seed: int) -> np.ndarray:
    """(n, n_fields): each field a sum of 8 random sinusoids a sin(kx x + ky y + phi)."""
    r = _rng(seed, 40)
    out = np.zeros((len(xy), n_fields))
    for f in range(n_fields):
        a = r.uniform(0.5, 1.5, 8)
        k = r.uniform(-0.05, 0.05, (8, 2))
        ph = r.uniform(0, 2 * np.pi, 8)
        out[:, f] = (a[None, :] * np.sin(xy @ k.T + ph[None, :])).sum(1)
    return out


# ------------------------------------------------------------------- named workloads
def triangle_grid_20k(seed: int = 0):
    """20,000 triangles on [0,10]^2 (101x101 vertices, jitter 0.3h); 1e5 points over [-0.5,10.5]^2."""
    v, f = jittered_grid(101, 10.0, 0.3, 1.0, seed)
    p = _rng(seed, 50).uniform(-0.5, 10.5, size=(100_000, 2))
    return v, f, p


def mixed_mesh_1m(nv: int = 867, seed: int = 0):
    """Mixed mesh: nv x nv vertices on [0,1000]^2, jitter 0.25h, a third of the quads split
    (867 -> 999,941 cells: 499,970 triangles + 499,971 quads), faces int32 width 4."""
    return jittered_grid(nv, 1000.0, 0.25, 1.0 / 3.0, seed)


def mixed_mesh_queries(n_points: int, n_boxes: int, n_segments: int, nv: int = 867, seed: int = 0):
    """Queries on the mixed mesh's domain: points (95% inside, the rest in a 50-wide frame),
    boxes (centres uniform, widths log-uniform in [0.1h, 20h]) and segments (uniform
    start, length log-uniform in [0.1h, 50h], uniform angle)."""
    h = 1000.0 / (nv - 1)
    return (
        points(n_points, 1000.0, 0.95, 50.0, seed),
        boxes(n_boxes, 1000.0, h, 0.1, 20.0, seed),
        segments(n_segments, 1000.0, h, 0.1, 50.0, seed),
    )


def overlap_meshes(nv: int = 708, n_target_side: int = 500, seed: int = 0):
    """Source: nv x nv vertices on [0,1000]^2, jitter 0.25h, all quads split (708 -> 999,698
    triangles).  Target: n_target_side^2 hex-lattice Voronoi cells (500 -> 250,000, 3-6
    vertices after clipping).  Field: one sum of 8 sinusoids at source centroids (seed 1)."""
    sv, sf = jittered_grid(nv, 1000.0, 0.25, 1.0, seed)
    tv, tf = hex_voronoi(n_target_side, n_target_side, 1000.0, 0.1, seed + 7)
    centroids = sv[sf].mean(axis=1)
    field = sinusoid_fields(centroids, 1, 1)[:, 0]
    return sv, sf, tv, tf, field


def batch_meshes(n_meshes: int = 512, seed: int = 0, cells=(1e3, 2e5), n_points=(1e3, 1e6),
            n_segments=(1e2, 1e5)):
    """Batch of independent meshes.  Mesh m: cell count log-uniform in `cells` (jittered grid,
    jitter 0.25h, a third of the quads split: half triangles, half quads, faces width 4) on its
    own square [x0, x0 + L] x [y0, y0 + L] (L log-uniform in [10, 1e4], origin uniform in
    [-1e5, 1e5]^2); point count log-uniform in `n_points` (95% inside, rest in a 5% frame);
    segment count log-uniform in `n_segments` (lengths log-uniform in [0.1h, 50h]).
    Returns a list of (vertices, faces, points, segments)."""
    r = _rng(seed, 60)
    lc = np.exp(r.uniform(np.log(cells[0]), np.log(cells[1]), n_meshes))
    npt = np.exp(r.uniform(np.log(n_points[0]), np.log(n_points[1]), n_meshes)).astype(np.int64)
    nsg = np.exp(r.uniform(np.log(n_segments[0]), np.log(n_segments[1]), n_meshes)).astype(np.int64)
    length = np.exp(r.uniform(np.log(10.0), np.log(1e4), n_meshes))
    origin = r.uniform(-1e5, 1e5, (n_meshes, 2))
    out = []
    for m in range(n_meshes):
        nv = max(3, int(round(np.sqrt(lc[m] * 0.75))) + 1)
        L = float(length[m])
        v, f = jittered_grid(nv, L, 0.25, 1.0 / 3.0, seed * 100_003 + m)
        h = L / (nv - 1)
        p = points(int(npt[m]), L, 0.95, 0.05 * L, seed * 100_003 + m)
        s = segments(int(nsg[m]), L, h, 0.1, 50.0, seed * 100_003 + m)
        o = origin[m]
        out.append((v + o, f, p + o, s + o[None, None, :]))
    return out


def graded_corner(levels: int, ratio: float = 0.5):
    """A unit square refined geometrically toward the origin: at each of `levels` levels the
    current square [0, s]^2 is cut into four quadrants of side h = ratio * s ... (ratio 0.5:
    the usual quadtree corner refinement); the three quadrants away from the origin become two
    triangles each (sharing their bounding box), and the refinement continues in the quadrant
    at the origin, whose last square becomes two triangles too.  6 * levels + 2 triangles; the
    cell tree of this mesh is about `levels` deep, with overlapping child boxes at every level."""
    verts, faces = [], []

    def v(x, y):
        verts.append((x, y))
        return len(verts) - 1

    s = 1.0
    for _ in range(levels):
        h = s * ratio
        for x0, y0, w, hh in ((h, 0.0, s - h, h), (0.0, h, h, s - h), (h, h, s - h, s - h)):
            a, b, c, d = v(x0, y0), v(x0 + w, y0), v(x0 + w, y0 + hh), v(x0, y0 + hh)
            faces += [(a, b, c), (a, c, d)]
        s = h
    a, b, c, d = v(0.0, 0.0), v(s, 0.0), v(s, s), v(0.0, s)
    faces += [(a, b, c), (a, c, d)]
    return np.array(verts, dtype=np.float64), np.array(faces, dtype=np.int32)
