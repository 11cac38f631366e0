"""Deterministic synthetic workloads for the acceleration-structure path.

Every generator is seeded and uses numpy's PCG64 stream, so the same arrays
are produced on any host. Geometry is returned as C-contiguous float32
vertices (V, 3) and uint32 index arrays, the layouts `Mesh`, `Tetrahedron`
and `Particles` take (render/render_prim_triangle.h:22-56,
render/render_prim_tetra.h:30-37, render/render_prim_particle.h:19-27).

Workloads (sizes as named):
  heightfield(512)          512x512 grid, z = sum of 8 seeded sines + N(0, 0.01),
                            2 triangles per quad -> 522,242 triangles
  tet_grid(64)              64^3 hexes jittered U(-0.2, 0.2), 6 tets per hex
                            -> 1,572,864 tets; per-vertex scalar = 16 Gaussians
  plummer_spheres(n)        Plummer-sphere centres, lognormal(-6, 0.5) radii
  gaussian_volume(256)      256^3 float32, sum of 64 anisotropic Gaussians in [0,1]
  transfer_function(256)    256-entry RGBA table, smooth ramps, alpha <= 0.05
  pinhole_rays(...)         one primary ray per pixel centre, 8x8-tile order
  heightfield_cameras(4)    4 frames of the heightfield camera, 90 degrees apart
  batch_scene(64)           64 meshes, log-uniform 10k..2M triangles, alternating
                            perturbed geodesic spheres and heightfields, each
                            with a random rigid transform; 16 orbit cameras
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20261202

RAY_DTYPE = np.dtype([("org", "<f4", (3,)), ("flags", "<u4"),
                      ("dir", "<f4", (3,)), ("prev_prim_id", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"),
                      ("prim_id", "<u4")])
TET_EXT_DTYPE = np.dtype([("d", "<f4", (4,)), ("scalar", "<f4"),
                          ("subface", "<u4"), ("t_leave", "<f4"),
                          ("pad", "<f4")])
# srf_solid_hit_ext: interpolate() weights, crossing position and normal,
# the varying sum, and which crossing (0 entry, 1 exit) was reported
SOL_EXT_DTYPE = np.dtype([("d", "<f4", (8,)), ("pos", "<f4", (3,)), ("scalar", "<f4"),
                          ("nrm", "<f4", (3,)), ("side", "<u4")])
# srf_line_hit_ext: position (org + t dir), cap flag, normal
LIN_EXT_DTYPE = np.dtype([("pos", "<f4", (3,)), ("hit_cap", "<u4"), ("nrm", "<f4", (3,)),
                          ("pad", "<f4")])
# srf_isect: the BuildIntersection fields of a closest hit (80 bytes)
ISECT_DTYPE = np.dtype([("position", "<f4", (3,)), ("f0", "<u4"), ("geometric", "<f4", (3,)),
                        ("f1", "<u4"), ("normal", "<f4", (3,)), ("f2", "<u4"),
                        ("tangent", "<f4", (3,)), ("f3", "<u4"), ("binormal", "<f4", (3,)),
                        ("pad", "<f4")])
NO_PRIM = np.uint32(0xFFFFFFFF)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def heightfield(grid=512, seed=BASE_SEED + 1, jitter=0.01):
    """Grid heightfield over [0,1]^2: V = grid^2, N = 2 (grid-1)^2 triangles."""
    rng = _rng(seed)
    amp = rng.uniform(0.02, 0.06, 8)
    freq = rng.uniform(0.5, 6.0, (8, 2))
    phase = rng.uniform(0.0, 2.0 * np.pi, 8)
    xs = np.linspace(0.0, 1.0, grid)
    x, y = np.meshgrid(xs, xs, indexing="xy")
    z = np.zeros_like(x)
    for k in range(8):
        z += amp[k] * np.sin(2.0 * np.pi * (freq[k, 0] * x + freq[k, 1] * y) + phase[k])
    z += rng.normal(0.0, jitter, z.shape)
    V = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(grid - 1)
    jj, ii = np.meshgrid(i, i, indexing="ij")
    v00 = (jj * grid + ii).ravel()
    v10 = v00 + 1
    v01 = v00 + grid
    v11 = v01 + 1
    t0 = np.stack([v00, v10, v11], axis=-1)
    t1 = np.stack([v00, v11, v01], axis=-1)
    F = np.stack([t0, t1], axis=1).reshape(-1, 3).astype(np.uint32)
    return np.ascontiguousarray(V), np.ascontiguousarray(F)


# Kuhn split of the unit cube into 6 tets along the 0-7 diagonal; corner c
# has offset (c & 1, (c >> 1) & 1, (c >> 2) & 1).
_KUHN = np.array([[0, 1, 3, 7], [0, 3, 2, 7], [0, 2, 6, 7],
                  [0, 6, 4, 7], [0, 4, 5, 7], [0, 5, 1, 7]], dtype=np.int64)


def tet_grid(cells=64, seed=BASE_SEED + 2, jitter=0.2, n_gauss=16):
    """Jittered (cells+1)^3 lattice on [0, cells]^3, 6 tets per hex.

    Returns V (float32), F (uint32, 4 per tet) and a per-vertex float32
    scalar (sum of `n_gauss` anisotropic Gaussians). Tets are oriented with
    det(v1 - v0, v2 - v0, v3 - v0) < 0, the orientation for which
    RayTetraPluecker (render/render_accel_tetra.cc:243-567) labels the face
    a ray crosses first as the entry face.
    """
    rng = _rng(seed)
    n1 = cells + 1
    g = np.arange(n1, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    P = P + rng.uniform(-jitter, jitter, P.shape)
    c = np.arange(cells)
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    base = (cz * n1 * n1 + cy * n1 + cx).ravel()
    corner = np.array([(k & 1) + ((k >> 1) & 1) * n1 + ((k >> 2) & 1) * n1 * n1
                       for k in range(8)], dtype=np.int64)
    F = (base[:, None, None] + corner[_KUHN][None, :, :]).reshape(-1, 4)
    a, b, cc, d = (P[F[:, k]] for k in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, cc - a), d - a)
    pos = vol > 0
    F[pos, 1], F[pos, 2] = F[pos, 2].copy(), F[pos, 1].copy()
    centres = rng.uniform(0.0, cells, (n_gauss, 3))
    sig = rng.uniform(0.05, 0.25, (n_gauss, 3)) * cells
    w = rng.uniform(0.2, 1.0, n_gauss)
    s = np.zeros(len(P))
    for k in range(n_gauss):
        q = (P - centres[k]) / sig[k]
        s += w[k] * np.exp(-0.5 * np.einsum("ij,ij->i", q, q))
    return (np.ascontiguousarray(P.astype(np.float32)),
            np.ascontiguousarray(F.astype(np.uint32)),
            np.ascontiguousarray(s.astype(np.float32)))


# hexahedron faces (outward, counter-clockwise seen from outside), the
# order of kHexaFaces in render/render_accel_solid.cc:99
_HEX_FACES = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5],
                       [2, 3, 7, 6], [0, 4, 7, 3]])


def solid_grid(cells=32, nv=8, seed=BASE_SEED + 8, jitter=0.2, n_gauss=8):
    """Jittered (cells+1)^3 lattice on [0, cells]^3 split into solids.

    nv = 8: one hexahedron per cell (vertices 0-3 the lower face
    counter-clockwise seen from above, 4-7 above them); nv = 6: two prisms
    per cell (lower triangle 0-2, upper 3-5); nv = 5: six pyramids per cell,
    one per face with the cell's centroid as apex (a vertex of its own after
    the lattice). Returns V (float32), F (uint32, nv per solid) and one
    float32 scalar per index position (the reference's solid varyings are
    looked up at nv * prim + k).
    """
    rng = _rng(seed + nv)
    n1 = cells + 1
    g = np.arange(n1, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    P = P + rng.uniform(-jitter, jitter, P.shape)
    c = np.arange(cells)
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    base = (cz * n1 * n1 + cy * n1 + cx).ravel()
    # corners in hexahedron order: (0,0,0) (1,0,0) (1,1,0) (0,1,0), then z+1
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    corner = np.array([dx + dy * n1 + dz * n1 * n1 for dx, dy, dz in off], dtype=np.int64)
    H = base[:, None] + corner[None, :]
    if nv == 8:
        F = H
    elif nv == 6:
        F = np.concatenate([H[:, [0, 1, 2, 4, 5, 6]], H[:, [0, 2, 3, 4, 6, 7]]], axis=1)
        F = F.reshape(-1, 6)
    elif nv == 5:
        centre = P[H].mean(axis=1)
        cid = len(P) + np.arange(len(H))
        P = np.concatenate([P, centre])
        # base (a, d, c, b) for the outward face (a, b, c, d): counter-clockwise
        # seen from the apex, the base order of kPyramidFaces[0] = {0,3,2,1}
        fa = H[:, _HEX_FACES]
        F = np.concatenate([fa[:, :, [0, 3, 2, 1]], np.repeat(cid[:, None, None], 6, 1)], axis=2)
        F = F.reshape(-1, 5)
    else:
        raise ValueError("nv must be 5, 6 or 8")
    centres = rng.uniform(0.0, cells, (n_gauss, 3))
    sig = rng.uniform(0.1, 0.3, (n_gauss, 3)) * cells
    w = rng.uniform(0.2, 1.0, n_gauss)
    q = P[F.ravel()]
    s = np.zeros(len(q))
    for k in range(n_gauss):
        d = (q - centres[k]) / sig[k]
        s += w[k] * np.exp(-0.5 * np.einsum("ij,ij->i", d, d))
    return (np.ascontiguousarray(P.astype(np.float32)),
            np.ascontiguousarray(F.astype(np.uint32)),
            np.ascontiguousarray(s.astype(np.float32)))


def solid_rays(cells=32, width=256, height=256, tiled=True):
    """Pinhole rays at a solid grid from outside one corner."""
    c = float(cells)
    return pinhole_rays((1.6 * c, -0.7 * c, 1.3 * c), (0.5 * c, 0.5 * c, 0.45 * c), (0, 0, 1),
                        45.0, width, height, tiled=tiled)


def polylines(n_curves=2000, seg_per_curve=64, seed=BASE_SEED + 11, step=0.02, radius=None,
              extent=1.0):
    """Random-walk polylines (streamline / hair-like) in [-extent, extent]^3:
    positions (float32), segments (uint32, 2 vertex ids per line, joining
    consecutive vertices of a curve) and, with radius="vary", one radius per
    vertex (float32), else None."""
    rng = _rng(seed)
    start = rng.uniform(-extent, extent, (n_curves, 1, 3))
    d = rng.normal(size=(n_curves, 3))
    steps = np.empty((n_curves, seg_per_curve, 3))
    for k in range(seg_per_curve):
        d = d + rng.normal(scale=0.6, size=d.shape)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        steps[:, k] = d * step
    P = np.concatenate([start, start + np.cumsum(steps, axis=1)], axis=1).reshape(-1, 3)
    base = (np.arange(n_curves) * (seg_per_curve + 1))[:, None] + np.arange(seg_per_curve)[None]
    segs = np.stack([base, base + 1], axis=-1).reshape(-1, 2)
    R = None
    if radius == "vary":
        R = rng.uniform(0.002, 0.006, len(P)).astype(np.float32)
    return (np.ascontiguousarray(P.astype(np.float32)),
            np.ascontiguousarray(segs.astype(np.uint32)), R)


def line_rays(width=256, height=256, tiled=True):
    return pinhole_rays((0.3, -3.2, 0.8), (0.0, 0.0, 0.0), (0, 0, 1), 40.0, width, height,
                        tiled=tiled)


def plummer_spheres(n, seed=BASE_SEED + 5, a=1.0, mu=-6.0, sigma=0.5):
    """Plummer-sphere centres (scale a) and lognormal(mu, sigma) radii."""
    rng = _rng(seed)
    m = rng.uniform(1e-6, 0.999, n)
    r = a / np.sqrt(m ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, 50.0 * a)
    cth = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sth = np.sqrt(1.0 - cth * cth)
    P = np.stack([r * sth * np.cos(phi), r * sth * np.sin(phi), r * cth], axis=-1)
    R = rng.lognormal(mu, sigma, n)
    return (np.ascontiguousarray(P.astype(np.float32)),
            np.ascontiguousarray(R.astype(np.float32)))


def gaussian_volume(dim=256, seed=BASE_SEED + 4, n_gauss=64):
    """dim^3 float32 volume (x fastest), sum of anisotropic Gaussians in [0,1]."""
    rng = _rng(seed)
    cen = rng.uniform(0.15, 0.85, (n_gauss, 3))
    sig = rng.uniform(0.03, 0.15, (n_gauss, 3))
    w = rng.uniform(0.3, 1.0, n_gauss)
    ax = np.linspace(0.0, 1.0, dim, dtype=np.float32)
    vol = np.zeros((dim, dim, dim), dtype=np.float32)
    for k in range(n_gauss):
        gx = np.exp(-0.5 * ((ax - cen[k, 0]) / sig[k, 0]) ** 2).astype(np.float32)
        gy = np.exp(-0.5 * ((ax - cen[k, 1]) / sig[k, 1]) ** 2).astype(np.float32)
        gz = np.exp(-0.5 * ((ax - cen[k, 2]) / sig[k, 2]) ** 2).astype(np.float32)
        vol += np.float32(w[k]) * (gz[:, None, None] * gy[None, :, None] * gx[None, None, :])
    vol /= vol.max()
    return np.ascontiguousarray(vol.reshape(-1))


def transfer_function(size=256, seed=BASE_SEED + 6, max_alpha=0.05):
    """size x RGBA float32 table: smooth colour ramps, alpha in [0, max_alpha]."""
    rng = _rng(seed)
    x = np.linspace(0.0, 1.0, size)
    tf = np.zeros((size, 4))
    for ch in range(3):
        ph = rng.uniform(0, 2 * np.pi)
        tf[:, ch] = 0.5 + 0.5 * np.sin(2 * np.pi * rng.uniform(0.5, 2.0) * x + ph)
    tf[:, 3] = max_alpha * np.clip(x - 0.1, 0, None) ** 1.5 / 0.9 ** 1.5
    return np.ascontiguousarray(tf.astype(np.float32))


def tile_order(width, height, tile=8):
    """Pixel index (y * width + x) of each ray when rays are laid out in
    tile x tile blocks (a 64-ray wavefront covers one 8x8 block)."""
    ty, tx = np.meshgrid(np.arange(0, height, tile), np.arange(0, width, tile), indexing="ij")
    oy, ox = np.meshgrid(np.arange(tile), np.arange(tile), indexing="ij")
    py = (ty.ravel()[:, None] + oy.ravel()[None, :]).ravel()
    px = (tx.ravel()[:, None] + ox.ravel()[None, :]).ravel()
    keep = (py < height) & (px < width)
    return (py[keep] * width + px[keep]).astype(np.int64)


def pinhole_rays(eye, target, up, fov_deg, width, height, tiled=True,
                 double_sided=True):
    """One ray per pixel centre of a pinhole camera.

    Directions are formed in float64, normalised and rounded to float32. With
    tiled=True the rays come in 8x8-tile order (see tile_order).
    """
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    h = np.tan(np.radians(fov_deg) * 0.5)
    px = (np.arange(width) + 0.5) / width * 2.0 - 1.0
    py = 1.0 - (np.arange(height) + 0.5) / height * 2.0
    X, Y = np.meshgrid(px * h * width / height, py * h, indexing="xy")
    d = fwd[None, None, :] + X[..., None] * right + Y[..., None] * upv
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    d = d.reshape(-1, 3)
    if tiled:
        d = d[tile_order(width, height)]
    rays = np.zeros(len(d), RAY_DTYPE)
    rays["org"] = eye.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["flags"] = 1 if double_sided else 0
    rays["prev_prim_id"] = NO_PRIM
    return rays


def heightfield_rays(width=1024, height=1024, tiled=True):
    """Camera framing the [0,1]^2 heightfield from above at an angle."""
    return pinhole_rays((0.5, -0.1, 1.0), (0.5, 0.45, 0.0), (0, 0, 1), 48.0,
                        width, height, tiled)


def heightfield_cameras(count=4, width=1024, height=1024, tiled=True):
    """`count` frames of the heightfield camera turned about the grid's
    vertical centre line by 360/count degrees each (the first is
    heightfield_rays), concatenated: rays of several cameras for one launch."""
    frames = [heightfield_rays(width, height, tiled)]
    for k in range(1, count):
        a = 2.0 * np.pi * k / count
        c, s = np.cos(a), np.sin(a)
        ex, ey = 0.0, -0.6  # eye (0.5, -0.1) relative to the centre (0.5, 0.5)
        tx, ty = 0.0, -0.05  # target (0.5, 0.45)
        eye = (0.5 + c * ex - s * ey, 0.5 + s * ex + c * ey, 1.0)
        target = (0.5 + c * tx - s * ty, 0.5 + s * tx + c * ty, 0.0)
        frames.append(pinhole_rays(eye, target, (0, 0, 1), 48.0, width, height, tiled))
    return np.concatenate(frames)


def tet_rays(cells=64, width=1024, height=1024, tiled=True):
    """Camera looking at the [0, cells]^3 tet grid across a corner."""
    c = float(cells)
    return pinhole_rays((1.6 * c, -0.6 * c, 1.3 * c), (0.5 * c, 0.5 * c, 0.5 * c),
                        (0, 0, 1), 36.0, width, height, tiled)


def sphere_rays(width=2048, height=2048, tiled=True):
    """Camera framing the core of a unit Plummer sphere."""
    return pinhole_rays((0.0, -6.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1), 30.0,
                        width, height, tiled)


def volume_rays(width=1024, height=1024, tiled=True):
    """Camera framing the unit-cube volume (texture space)."""
    return pinhole_rays((1.8, -1.2, 1.5), (0.5, 0.5, 0.5), (0, 0, 1), 35.0,
                        width, height, tiled)


def deep_chain_points(dup=40, radius=1e-3):
    """Particles whose Morton codes run 1, 2, 4, ..., 2^29 (each position
    repeated `dup` times): the radix tree is a 30-deep caterpillar whose
    subtrees exceed the depth cap 22, so Build32 emits depth-capped leaves
    (render/render_accel_triangle.cc:909)."""
    pts = [(0.0, 0.0, 0.0), (1023.0, 1023.0, 1023.0)]
    for p in range(30):
        b, axis = p // 3, 2 - (p % 3)
        q = [0.0, 0.0, 0.0]
        q[axis] = float(1 << b)
        pts.append(tuple(q))
    P = np.repeat(np.array(pts, np.float32), dup, axis=0)
    R = np.full(len(P), radius, np.float32)
    return np.ascontiguousarray(P), R


_ICO_V = None


def clustered_spheres(n_bg=20_000, clusters=6, per_cluster=3_000, spread=2e-3,
                      seed=BASE_SEED + 8):
    """Background Plummer spheres plus dense clusters whose centres share one
    30-bit Morton cell, so Build32 makes leaves of thousands of spheres; each
    cluster also holds exact duplicates (same centre and radius) and shared
    centres with different radii, which tie in t."""
    rng = _rng(seed)
    P, R = plummer_spheres(n_bg, seed=seed + 1, mu=-7.0)
    cs = rng.uniform(-0.6, 0.6, (clusters, 3))
    parts_p, parts_r = [P], [R]
    for k in range(clusters):
        p = cs[k] + rng.uniform(-spread, spread, (per_cluster, 3))
        r = rng.lognormal(np.log(spread * 0.05), 0.6, per_cluster)
        dup = rng.choice(per_cluster, per_cluster // 10, replace=False)
        p[dup[1::2]] = p[dup[0::2]][: len(dup[1::2])]
        r[dup[1::4]] = r[dup[0::4]][: len(dup[1::4])]
        parts_p.append(p)
        parts_r.append(r)
    P = np.concatenate(parts_p).astype(np.float32)
    R = np.concatenate(parts_r).astype(np.float32)
    perm = rng.permutation(len(P))
    return np.ascontiguousarray(P[perm]), np.ascontiguousarray(R[perm]), cs.astype(np.float32)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return V, F


def geodesic_sphere(freq, seed=BASE_SEED + 7, radius=0.5, bumps=6, amp=0.06):
    """Icosahedron with every face split into freq^2 triangles, projected to a
    sphere and perturbed radially by seeded spherical bumps: 20 freq^2
    triangles (face vertices are not shared across icosahedron edges)."""
    rng = _rng(seed)
    IV, IF = _icosahedron()
    i, j = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    keep = (i + j) <= freq
    bi, bj = i[keep], j[keep]                      # barycentric grid of one face
    local = {(a, b): k for k, (a, b) in enumerate(zip(bi.tolist(), bj.tolist()))}
    tris = []
    for a in range(freq):
        for b in range(freq - a):
            tris.append((local[(a, b)], local[(a + 1, b)], local[(a, b + 1)]))
            if a + b + 1 < freq:
                tris.append((local[(a + 1, b)], local[(a + 1, b + 1)], local[(a, b + 1)]))
    tris = np.array(tris, np.int64)
    nloc = len(bi)
    w1, w2 = bi / freq, bj / freq
    w0 = 1.0 - w1 - w2
    P = (w0[None, :, None] * IV[IF[:, 0]][:, None, :] + w1[None, :, None] * IV[IF[:, 1]][:, None, :]
         + w2[None, :, None] * IV[IF[:, 2]][:, None, :]).reshape(-1, 3)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    centres = rng.normal(size=(bumps, 3))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    r = np.ones(len(P))
    for c in centres:
        r += amp * rng.uniform(-1, 1) * np.exp(-8.0 * np.sum((P - c) ** 2, axis=1))
    V = (radius * r[:, None] * P).astype(np.float32)
    F = (tris[None, :, :] + nloc * np.arange(20)[:, None, None]).reshape(-1, 3).astype(np.uint32)
    return np.ascontiguousarray(V), np.ascontiguousarray(F)


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def batch_scene(count=64, lo=10_000, hi=2_000_000, seed=BASE_SEED + 3, spacing=1.5):
    """Many meshes for the batched build and the two-level trace.

    Sizes N_i = floor(exp(U(ln lo, ln hi))); even i is a perturbed geodesic
    sphere with 20 f^2 ~ N_i triangles (radius 0.5), odd i a heightfield with
    2 (g-1)^2 ~ N_i triangles centred on the origin. Each mesh gets a random
    rotation and a translation on a jittered cube lattice of the given
    spacing. Returns (meshes [(V, F)], T) with T[i] the 4x4 float64
    object-to-world matrix in the row-vector convention p' = p T
    (render/render_matrix.h:165-173: rows 0-2 the linear part, row 3 the
    translation)."""
    rng = _rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    side = int(np.ceil(count ** (1.0 / 3.0)))
    meshes, T = [], []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            f = max(1, int(round((n / 20.0) ** 0.5)))
            V, F = geodesic_sphere(f, seed=seed + 1000 + i)
        else:
            g = max(2, int(round((n / 2.0) ** 0.5)) + 1)
            V, F = heightfield(g, seed=seed + 1000 + i)
            V = V - np.array([0.5, 0.5, 0.0], np.float32)
        meshes.append((np.ascontiguousarray(V), np.ascontiguousarray(F)))
        cell = np.array([i % side, (i // side) % side, i // (side * side)], np.float64)
        M = np.eye(4)
        M[:3, :3] = _rotation(rng).T  # row-vector form of a rotation
        M[3, :3] = spacing * (cell + rng.uniform(-0.15, 0.15, 3))
        T.append(M)
    return meshes, np.array(T)


def mixed_scene(kinds=("tri", "tet", "sph", "sol", "lin"), copies=2, seed=BASE_SEED + 14,
                spacing=1.5):
    """A two-level scene mixing primitive kinds, `copies` elements of each:
    "tri" a perturbed geodesic sphere (2,880 triangles, radius 0.5), "tet"
    an 8^3-cell tetrahedral grid (3,072 tets) and "sol" an 11^3 hexahedral
    grid (1,331 hexahedra), both scaled into the unit cube about the origin,
    "sph" 4,000 Plummer spheres (a = 0.25, lognormal radii mu = -5.5), "lin"
    100 random-walk polylines of 32 segments (radius 0.01). Each element
    gets a random rotation and a place on a jittered lattice as in
    batch_scene. Returns (elements, T): elements are dicts with "kind" and
    the arrays the matching build takes (V, F, R, rconst, nv), T [n, 4, 4]
    in the row-vector convention."""
    rng = _rng(seed)
    elems = []
    for c in range(copies):
        for kind in kinds:
            s = seed + 100 * c + len(elems)
            if kind == "tri":
                V, F = geodesic_sphere(12, seed=s)
                elems.append(dict(kind="tri", V=V, F=F))
            elif kind == "tet":
                V, F, _ = tet_grid(8, seed=s)
                V = (V / np.float32(8.0) - np.float32(0.5)).astype(np.float32)
                elems.append(dict(kind="tet", V=np.ascontiguousarray(V), F=F))
            elif kind == "sol":
                V, F, _ = solid_grid(11, nv=8, seed=s)
                V = (V / np.float32(11.0) - np.float32(0.5)).astype(np.float32)
                elems.append(dict(kind="sol", V=np.ascontiguousarray(V), F=F, nv=8))
            elif kind == "sph":
                P, R = plummer_spheres(4000, seed=s, a=0.25, mu=-5.5)
                elems.append(dict(kind="sph", V=P, R=R, rconst=0.0))
            else:
                P, S, _ = polylines(100, 32, seed=s, extent=0.4)
                elems.append(dict(kind="lin", V=P, F=S, R=None, rconst=0.01, cap=1))
    n = len(elems)
    side = int(np.ceil(n ** (1.0 / 3.0)))
    T = []
    for i in range(n):
        cell = np.array([i % side, (i // side) % side, i // (side * side)], np.float64)
        M = np.eye(4)
        M[:3, :3] = _rotation(rng).T
        M[3, :3] = spacing * (cell + rng.uniform(-0.15, 0.15, 3))
        T.append(M)
    return elems, np.array(T)


def scene_cameras(T, count=16, width=1024, height=1024, fov=40.0, tiled=True):
    """`count` pinhole cameras on a horizontal orbit around the scene's
    centre, all looking at it; returns a list of ray arrays."""
    centre = T[:, 3, :3].mean(axis=0)
    extent = np.ptp(T[:, 3, :3], axis=0).max() + 1.0
    out = []
    for k in range(count):
        a = 2.0 * np.pi * k / count
        eye = centre + extent * np.array([1.4 * np.cos(a), 1.4 * np.sin(a), 0.6])
        out.append(pinhole_rays(tuple(eye), tuple(centre), (0, 0, 1), fov, width, height,
                                tiled=tiled))
    return out


def ao_rays_from_hits(rays, hits, P, R, per_hit=4, eps=1e-3, dist=0.05, seed=BASE_SEED + 12,
                      every=1):
    """Ambient-occlusion rays in numpy (a fixed input for tests): for every
    `every`-th ray i with a sphere hit, `per_hit` cosine-weighted directions
    about the outward normal at the hit point, origin p + eps r n, told to
    skip the hit sphere (prev_prim_id). Returns (rays, t_max)."""
    idx = np.nonzero(hits["prim_id"] != NO_PRIM)[0][::every]
    c = hits["prim_id"][idx].astype(np.int64)
    p = (rays["org"][idx].astype(np.float64)
         + hits["t"][idx, None].astype(np.float64) * rays["dir"][idx].astype(np.float64))
    n = p - P[c].astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    org = p + eps * R[c, None].astype(np.float64) * n
    rng = _rng(seed)
    k = len(idx)
    u1 = rng.random((k, per_hit))
    u2 = rng.random((k, per_hit))
    sg = np.where(n[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sg + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    b1 = np.stack([1.0 + sg * n[:, 0] * n[:, 0] * a, sg * b, -sg * n[:, 0]], 1)
    b2 = np.stack([b, sg + n[:, 1] * n[:, 1] * a, -n[:, 1]], 1)
    r, phi = np.sqrt(u1), 2.0 * np.pi * u2
    d = ((r * np.cos(phi))[..., None] * b1[:, None] + (r * np.sin(phi))[..., None] * b2[:, None]
         + np.sqrt(np.maximum(0.0, 1.0 - u1))[..., None] * n[:, None])
    out = np.zeros(k * per_hit, RAY_DTYPE)
    out["org"] = np.repeat(org, per_hit, axis=0).astype(np.float32)
    out["dir"] = d.reshape(-1, 3).astype(np.float32)
    out["flags"] = 1
    out["prev_prim_id"] = np.repeat(hits["prim_id"][idx], per_hit)
    return out, np.full(k * per_hit, np.float32(dist))


def sparse_blocks(seed=BASE_SEED + 13, grid=(4, 3, 2), block=16, components=1, dtype=np.float32,
                  overlap=False):
    """A sparse volume: a grid of blocks of `block`^3 cells in the global
    grid, each with its own voxel resolution (some coarser, as LoD levels
    are), a few blocks left out; with overlap some blocks are shifted into
    their neighbours. Returns (blocks int32 [nb, 9]: offset, extent, size;
    list of voxel arrays; global_dim)."""
    rng = _rng(seed)
    out, data = [], []
    for z in range(grid[2]):
        for y in range(grid[1]):
            for x in range(grid[0]):
                if rng.random() < 0.15:
                    continue
                off = np.array([x, y, z]) * block
                if overlap and rng.random() < 0.3:
                    off = off + rng.integers(-block // 3, block // 3, 3)
                size = rng.choice([block, block // 2, block // 4], 3)
                out.append(np.concatenate([off, [block] * 3, size]).astype(np.int32))
                n = int(np.prod(size)) * components
                if dtype == np.uint8:
                    data.append(rng.integers(0, 256, n).astype(np.uint8))
                else:
                    data.append(rng.uniform(-1, 1, n).astype(dtype))
    gdim = np.array(grid, np.int32) * block
    return np.stack(out), data, gdim
