"""Procedural triangle meshes for the wind-tunnel workloads, in the packed
12-float-per-triangle layout LBM_SetSolidMesh takes (v0 xyz pad, v1 xyz pad,
v2 xyz pad; simulation/lib/model_bounds.h:14-18) and the world box of the
reference (x in [-4,4], y, z in [-2,2]).

Every body is star-shaped about a centre point, so one closed, consistently
wound surface is produced by placing the vertices of a latitude/longitude
sphere at the body's radial distance r(direction); the +X ray-parity solid
test of LBM_SetSolidMesh then sees a single closed surface (a union of two
separately meshed closed bodies would have even parity where they overlap).
"""
from __future__ import annotations

import numpy as np


def _sphere_dirs(n_lat: int, n_lon: int):
    th = np.linspace(0.0, np.pi, n_lat + 1)
    ph = np.linspace(0.0, 2 * np.pi, n_lon, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.cos(T), np.sin(T) * np.cos(P), np.sin(T) * np.sin(P)], -1)  # x along the pole axis
    return d


def _triangulate(verts: np.ndarray) -> np.ndarray:
    """verts[n_lat+1, n_lon, 3] on a lat/lon grid -> packed triangles."""
    nl, nm, _ = verts.shape
    tris = []
    for i in range(nl - 1):
        for j in range(nm):
            a, b = verts[i, j], verts[i, (j + 1) % nm]
            c, d = verts[i + 1, j], verts[i + 1, (j + 1) % nm]
            if i > 0:
                tris.append((a, c, b))
            if i < nl - 2:
                tris.append((b, c, d))
    t = np.zeros((len(tris), 12), np.float32)
    for k, (p, q, r) in enumerate(tris):
        t[k, 0:3], t[k, 4:7], t[k, 8:11] = p, q, r
    return t


def radial_mesh(radius_fn, centre, n_lat: int, n_lon: int) -> np.ndarray:
    d = _sphere_dirs(n_lat, n_lon)
    r = radius_fn(d.reshape(-1, 3)).reshape(d.shape[:2])
    return _triangulate(np.asarray(centre, np.float64) + r[..., None] * d)


def _box_radius(d, half):
    with np.errstate(divide="ignore"):
        t = np.min(np.where(np.abs(d) > 1e-12, np.asarray(half) / np.abs(d), np.inf), axis=-1)
    return t


def _ellipsoid_radius(d, axes, offset):
    """Distance from the origin along d to the surface of an ellipsoid centred
    at `offset` (the origin must lie inside)."""
    a = np.asarray(axes, np.float64)
    o = -np.asarray(offset, np.float64)
    A = np.sum((d / a) ** 2, -1)
    B = 2 * np.sum(d * o / a ** 2, -1)
    C = np.sum((o / a) ** 2) - 1.0
    return (-B + np.sqrt(B * B - 4 * A * C)) / (2 * A)


def car_proxy(seed: int = 0, length: float = 2.4, centre=(-1.0, 0.0, 0.0), n_lat: int = 100,
              n_lon: int = 100):
    """Seeded random ellipsoid-plus-box car proxy: the union of a box (the body)
    and an ellipsoid (the cabin) raised above it, both containing `centre`.
    Returns (triangles [n, 12], bbox (minX, minY, minZ, maxX, maxY, maxZ))."""
    rng = np.random.default_rng(seed)
    half = np.array([length / 2, rng.uniform(0.40, 0.55) * length / 2, rng.uniform(0.22, 0.30) * length / 2])
    axes = np.array([rng.uniform(0.55, 0.75) * length / 2, half[1] * rng.uniform(0.75, 0.95),
                     half[2] * rng.uniform(1.5, 2.1)])
    off = np.array([rng.uniform(-0.15, 0.10) * length / 2, 0.0, half[2] * rng.uniform(0.3, 0.6)])

    def r(d):
        return np.maximum(_box_radius(d, half), _ellipsoid_radius(d, axes, off))

    tris = radial_mesh(r, centre, n_lat, n_lon)
    v = tris.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 3)
    return tris, (*v.min(0).tolist(), *v.max(0).tolist())


def ahmed_body(nx: int, ny: int, nz: int, cells=(64, 24, 18), front_radius_cells: float = 6.0,
               slant_deg: float = 25.0, clearance_cells: float = 2.0, n_lat: int = 100, n_lon: int = 100,
               x_front: float = -2.5):
    """Ahmed-like bluff body sized in lattice cells of an nx*ny*nz grid: length,
    width, height = `cells`, rounded front edges of radius `front_radius_cells`,
    a `slant_deg` rear slant over the top, lifted `clearance_cells` above the
    floor of the domain.  Returns (triangles, bbox)."""
    sx, sy, sz = nx / 8.0, ny / 4.0, nz / 4.0
    L, W, H = cells[0] / sx, cells[1] / sy, cells[2] / sz
    R = front_radius_cells / sx
    z0 = -2.0 + clearance_cells / sz
    lo = np.array([x_front, -W / 2, z0])
    hi = np.array([x_front + L, W / 2, z0 + H])
    c = (lo + hi) / 2
    slant_len = min(0.35 * L, H / np.tan(np.radians(slant_deg)) * 0.6)
    n_sl = np.array([np.sin(np.radians(slant_deg)), 0.0, np.cos(np.radians(slant_deg))])
    p_sl = np.array([hi[0] - slant_len, 0.0, hi[2]])

    def sdf(p):
        q = p - c
        h = (hi - lo) / 2
        # front half: box with edges rounded by R; rear half: sharp box
        rr = np.where(q[:, :1] < 0, R, 0.0)
        d = np.abs(q) - (h - rr)
        outside = np.linalg.norm(np.maximum(d, 0), axis=-1)
        inside = np.minimum(np.max(d, -1), 0)
        box = outside + inside - rr[:, 0]
        return np.maximum(box, (p - p_sl) @ n_sl)

    def radius(d):
        lo_r = np.zeros(len(d))
        hi_r = np.full(len(d), 4.0)
        for _ in range(60):
            mid = (lo_r + hi_r) / 2
            inside = sdf(c + mid[:, None] * d) < 0
            lo_r = np.where(inside, mid, lo_r)
            hi_r = np.where(inside, hi_r, mid)
        return (lo_r + hi_r) / 2

    tris = radial_mesh(radius, c, n_lat, n_lon)
    v = tris.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 3)
    return tris, (*v.min(0).tolist(), *v.max(0).tolist())


def random_body(seed: int, n_lat: int = 40, n_lon: int = 40):
    """Seeded random bluff body for parameter sweeps and warm-start data: a
    box, an ellipsoid or a box with a slanted rear (20-35 deg), 10-20 % of
    the tunnel height tall (the tunnel spans 8 x 4 x 4 world units), length
    1.5-2.5 and width 0.8-1.2 times the height, centred near x = -2.
    Returns (triangles, bbox, kind)."""
    rng = np.random.default_rng(seed)
    kind = int(rng.integers(0, 3))
    h = rng.uniform(0.10, 0.20) * 4.0
    L = h * rng.uniform(1.5, 2.5)
    W = h * rng.uniform(0.8, 1.2)
    c = np.array([-2.0 + rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)])
    if kind == 0:
        def r(d):
            return _box_radius(d, np.array([L, W, h]) / 2)
    elif kind == 1:
        def r(d):
            return _ellipsoid_radius(d, np.array([L, W, h]) / 2, np.zeros(3))
    else:
        ang = np.radians(rng.uniform(20, 35))
        n_sl = np.array([np.sin(ang), 0.0, np.cos(ang)])
        p0 = np.array([L / 2 - 0.4 * L, 0.0, h / 2])

        def r(d):
            box = _box_radius(d, np.array([L, W, h]) / 2)
            dn = d @ n_sl
            with np.errstate(divide="ignore", invalid="ignore"):
                t = np.where(dn > 1e-9, (p0 @ n_sl) / dn, np.inf)
            return np.minimum(box, t)
    tris = radial_mesh(r, c, n_lat, n_lon)
    v = tris.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 3)
    return tris, (*v.min(0).tolist(), *v.max(0).tolist()), ["box", "ellipsoid", "slanted-rear"][kind]
