"""Procedural low-order meshes of the solve's workloads (normals out of the body into the fluid).

Arrays are (P, 4, 3) float64, one quadrilateral per panel, a triangle written as a quadrilateral
whose last vertex repeats the third. The panel layout (vertex order, orientation, panel order and
the coordinates to the last bit) is that of the reference's test meshes (tests/python/meshes.py:
box_panels, hemisphere_panels, cylinder_panels), which the solve's workloads name; the fixtures of
tests/golden/meshes_layout.npz hold it and tests/test_meshes.py requires it bit for bit.

The construction here is array-wise: each surface is a grid of vertices (or, for the flat faces of
a box, a grid of panel origins with the edge vectors added), the quadrilaterals are read off the
grid as (i, j), (i, j+1), (i+1, j+1), (i+1, j), and one vectorised test against the surface's
outward direction reverses the ones whose diagonal cross product points into the body.
"""

from __future__ import annotations

import numpy as np

__all__ = ["box_panels", "hemisphere_panels", "cylinder_panels", "expand_symmetry", "place", "interior_graded"]

# the order of a quadrilateral with its orientation reversed (the first vertex kept)
_REVERSED = [0, 3, 2, 1]


def _outward(quads, direction):
    """Each quadrilateral whose diagonal cross product (v2 - v0) x (v3 - v1) has a positive
    component along its outward `direction` ([P, 3] or [3]) as it is, the others reversed."""
    d1 = quads[:, 2] - quads[:, 0]
    d2 = quads[:, 3] - quads[:, 1]
    n = np.stack([d1[:, 1] * d2[:, 2] - d1[:, 2] * d2[:, 1],
                  d1[:, 2] * d2[:, 0] - d1[:, 0] * d2[:, 2],
                  d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0]], axis=1)
    keep = np.einsum("pk,pk->p", n, np.broadcast_to(direction, n.shape)) > 0
    return np.where(keep[:, None, None], quads, quads[:, _REVERSED])


def _grid_quads(V):
    """Quadrilaterals of a vertex grid V [I + 1][J + 1][3], row-major over (i, j)."""
    q = np.stack([V[:-1, :-1], V[:-1, 1:], V[1:, 1:], V[1:, :-1]], axis=2)
    return q.reshape(-1, 4, 3)


def _close_last_ring(quads, per_ring):
    """The last ring's quadrilaterals (a collapsed edge at the pole or the axis) as triangles:
    the fourth vertex replaced by the third."""
    quads[-per_ring:, 3] = quads[-per_ring:, 2]
    return quads


def _flat_face(origin, du, dv, nu, nv, direction):
    """nu x nv panels of the parallelogram origin + s du + t dv, panel (i, j) at
    origin + (du/nu) i + (dv/nv) j with the corners p, p + a, p + a + b, p + b (i-major)."""
    origin, du, dv = (np.asarray(x, dtype=float) for x in (origin, du, dv))
    a, b = du / nu, dv / nv
    i = np.arange(nu)[:, None, None]
    j = np.arange(nv)[None, :, None]
    p = (origin + a * i + b * j).reshape(-1, 3)
    pa = p + a
    q = np.stack([p, pa, pa + b, p + b], axis=1)
    return _outward(q, np.asarray(direction, dtype=float))


def box_panels(a=1.0, b=0.5, d=0.5, n=(4, 2, 2), quarter=False, lid=False):
    """Surface-piercing box |x| <= a, |y| <= b, -d <= z <= 0: the bottom (nx ny panels), the faces
    x = a (ny nz), y = b (nx nz) and, for the whole box, x = -a and y = -b; quarter: the part
    x >= 0, y >= 0 (for the symmetry planes x = 0 and y = 0); lid: the interior water plane z = 0
    (nx ny panels, normals up) after the hull."""
    nx, ny, nz = n
    x0, y0 = (0.0, 0.0) if quarter else (-a, -b)
    lx, ly = a - x0, b - y0
    faces = [((x0, y0, -d), (lx, 0, 0), (0, ly, 0), nx, ny, (0, 0, -1)),
             ((a, y0, -d), (0, ly, 0), (0, 0, d), ny, nz, (1, 0, 0)),
             ((x0, b, -d), (lx, 0, 0), (0, 0, d), nx, nz, (0, 1, 0))]
    if not quarter:
        faces += [((-a, y0, -d), (0, ly, 0), (0, 0, d), ny, nz, (-1, 0, 0)),
                  ((x0, -b, -d), (lx, 0, 0), (0, 0, d), nx, nz, (0, -1, 0))]
    if lid:
        faces.append(((x0, y0, 0.0), (lx, 0, 0), (0, ly, 0), nx, ny, (0, 0, 1)))
    return np.concatenate([_flat_face(*f) for f in faces])


def hemisphere_panels(radius=1.0, n_theta=8, n_phi=8, quarter=True, graded=False):
    """Floating hemisphere z <= 0 with its vertices on the sphere: n_theta rings in polar angle
    from the waterline to the bottom pole (graded: refined towards the waterline, s -> 1 - cos(pi s
    / 2)), n_phi divisions in azimuth over the whole circle or, quarter, over x >= 0, y >= 0; the
    polar ring holds triangles."""
    phi = np.linspace(0.0, np.pi / 2 if quarter else 2 * np.pi, n_phi + 1)
    s = np.linspace(0.0, 1.0, n_theta + 1)
    if graded:
        s = 1.0 - np.cos(0.5 * np.pi * s)
    theta = np.pi / 2 + 0.5 * np.pi * s  # polar angle from +z
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    cp, sp = np.cos(phi)[None, :], np.sin(phi)[None, :]
    V = radius * np.stack([st * cp, st * sp, np.broadcast_to(ct, st.shape[:1] + cp.shape[1:])], axis=2)
    q = _close_last_ring(_grid_quads(V), n_phi)
    return _outward(q, q.mean(axis=1))  # outward: away from the centre


def _cosine_spacing(n):
    """n intervals on [0, 1] refined at both ends: (1 - cos(pi k / n)) / 2, k = 0..n."""
    if n == 0:
        return np.zeros(1)
    return 0.5 * (1.0 - np.cos(np.pi * np.arange(n + 1) / n))


def cylinder_panels(radius=1.0, draft=0.5, n_phi=8, n_z=2, n_r=2, quarter=True, cosine=False):
    """Floating vertical circular cylinder (a polygon of n_phi sides over the whole circle or,
    quarter, over x >= 0, y >= 0): the side (n_z rings from the waterline down), then the bottom
    (n_r rings from the rim to the axis, the innermost of triangles); cosine: the side and bottom
    rings cosine-spaced, refined towards both ends."""
    phi = np.linspace(0.0, np.pi / 2 if quarter else 2 * np.pi, n_phi + 1)
    spacing = _cosine_spacing if cosine else (lambda m: np.linspace(0.0, 1.0, m + 1))
    z = -draft * spacing(n_z)
    r = radius * (1.0 - spacing(n_r))
    cp, sp = np.cos(phi), np.sin(phi)
    # side: ring i at depth z[i], vertex j at angle phi[j]
    side = np.stack([np.broadcast_to(radius * cp, (n_z + 1, n_phi + 1)),
                     np.broadcast_to(radius * sp, (n_z + 1, n_phi + 1)),
                     np.broadcast_to(z[:, None], (n_z + 1, n_phi + 1))], axis=2)
    qs = _grid_quads(side)
    mid = qs.mean(axis=1)
    qs = _outward(qs, np.stack([mid[:, 0], mid[:, 1], np.zeros(len(mid))], axis=1))  # outward: radial
    # bottom: ring i at radius r[i]
    bottom = np.stack([r[:, None] * cp[None, :], r[:, None] * sp[None, :],
                       np.full((n_r + 1, n_phi + 1), -draft)], axis=2)
    qb = _outward(_close_last_ring(_grid_quads(bottom), n_phi), np.array([0.0, 0.0, -1.0]))
    return np.concatenate([qs, qb])


def expand_symmetry(quadrant):
    """A quadrant mesh with the symmetry planes x = 0 and y = 0 as the whole mesh: the stored
    panels, then their x-, y- and xy-reflections, the orientation reversed where a reflection
    reverses it (geometry.rs:209-234 reflections / reflect_vertices)."""
    q = np.asarray(quadrant, dtype=float)
    out = [q]
    for sx, sy in ((-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0)):
        r = q * np.array([sx, sy, 1.0])
        out.append(r[:, _REVERSED] if sx * sy < 0 else r)
    return np.concatenate(out)


def place(vertices, position=(0.0, 0.0, 0.0)):
    """The panels translated to a body position."""
    return np.asarray(vertices, dtype=float) + np.asarray(position, dtype=float)


def _waterline(vertices, z0, tol):
    """The waterline segments of the wetted panels: the edges in z = z0 of the non-horizontal
    panels, each oriented with the body on its left seen from above (the normals point out of the
    body), as [(p, q)] of 2D points."""
    segs = []
    for quad in np.asarray(vertices, dtype=float):
        n = np.cross(quad[2] - quad[0], quad[3] - quad[1])
        if np.hypot(n[0], n[1]) <= 1e-6 * np.linalg.norm(n):
            continue  # horizontal: no waterline edge
        for i in range(4):
            a, b = quad[i], quad[(i + 1) % 4]
            if abs(a[2] - z0) > tol or abs(b[2] - z0) > tol or np.hypot(*(b - a)[:2]) <= tol:
                continue
            p, q = a[:2].copy(), b[:2].copy()
            d = q - p
            if -d[1] * -n[0] + d[0] * -n[1] < 0.0:  # the body (-n) on the left of p -> q
                p, q = q, p
            segs.append((p, q))
    return segs


def interior_graded(vertices, sym_x=False, sym_y=False, z0=0.0):
    """The graded interior free surface of a body (interior.rs:437-515, theory §5.1): the
    waterplane region inside the waterline of the wetted panels `vertices` [P][4][3] (the stored
    sector: x >= 0 where sym_x, y >= 0 where sym_y), closed along those symmetry planes, is fanned
    from its area centroid (moved onto each symmetry plane the region touches) to every waterline
    segment; the rays are divided at t_k = sin(pi k / (2 M)), k = 0..M, denser towards the
    waterline, with M = ceil(mean ray length / mean segment length); each segment gives one
    triangle at the centroid and M - 1 quadrilaterals, which together cover the polygonal
    waterplane exactly. Returns the panels [m][4][3] in z = z0, counter-clockwise seen from above
    (normals up; a triangle repeats its third vertex), segment by segment along the waterline. One
    region without openings, star-shaped from its centroid; otherwise ValueError."""
    v = np.asarray(vertices, dtype=float).reshape(-1, 4, 3)
    extent = max(1.0, float(np.abs(v).max()))
    tol = 1e-9 * extent
    segs = _waterline(v, z0, tol)
    if not segs:
        raise ValueError("interior surface 'graded': no waterline")
    # chain the segments (end of one = start of the next); an open chain of a sector runs from one
    # symmetry plane to another
    key = lambda p: (round(p[0] / tol), round(p[1] / tol))
    start = {key(p): i for i, (p, q) in enumerate(segs)}
    ends = {key(q) for p, q in segs}
    heads = [i for i, (p, q) in enumerate(segs) if key(p) not in ends]
    if len(heads) > 1:
        raise ValueError("interior surface 'graded': the waterline is not one chain")
    i = heads[0] if heads else 0
    order, seen = [], set()
    while i is not None and i not in seen:
        seen.add(i)
        order.append(i)
        i = start.get(key(segs[i][1]))
    if len(order) != len(segs):
        raise ValueError("interior surface 'graded': the waterline is not one chain")
    chain = [segs[i] for i in order]
    poly = [p for p, q in chain] + [chain[-1][1]]
    closed = np.linalg.norm(poly[0] - poly[-1]) <= tol
    if closed:
        poly = poly[:-1]
    else:
        a, b = poly[0], poly[-1]
        on = lambda p, ax: (ax == 0 and sym_x and abs(p[0]) <= tol) or (ax == 1 and sym_y and abs(p[1]) <= tol)
        if not any(on(a, ax) for ax in (0, 1)) or not any(on(b, ax) for ax in (0, 1)):
            raise ValueError("interior surface 'graded': the waterline does not close")
        if not any(on(a, ax) and on(b, ax) for ax in (0, 1)):
            poly.append(np.zeros(2))  # along both symmetry planes through the origin
    P = np.array(poly)
    Q = np.roll(P, -1, axis=0)
    cr = P[:, 0] * Q[:, 1] - P[:, 1] * Q[:, 0]
    area = 0.5 * cr.sum()
    c = np.array([((P[:, 0] + Q[:, 0]) * cr).sum(), ((P[:, 1] + Q[:, 1]) * cr).sum()]) / (6.0 * area)
    for p, q in zip(P, Q):  # the closing edges on the symmetry planes
        if sym_x and abs(p[0]) <= tol and abs(q[0]) <= tol:
            c[0] = 0.0
        if sym_y and abs(p[1]) <= tol and abs(q[1]) <= tol:
            c[1] = 0.0
    fan, ray, seg = [], 0.0, 0.0
    for p, q in chain:
        s = (p[0] - c[0]) * (q[1] - c[1]) - (p[1] - c[1]) * (q[0] - c[0])
        if abs(s) <= tol * np.linalg.norm(q - p):
            continue
        if s < 0.0:
            raise ValueError("interior surface 'graded': a ray from the centroid meets the waterline twice")
        fan.append((p, q))
        ray += 0.5 * (np.linalg.norm(c - p) + np.linalg.norm(c - q))
        seg += np.linalg.norm(q - p)
    M = max(1, int(np.ceil(ray / seg - 1e-9)))
    t = np.array([np.sin(0.5 * np.pi * k / M) for k in range(M)] + [1.0])
    at = lambda p, s: np.array([c[0] + s * (p[0] - c[0]), c[1] + s * (p[1] - c[1]), z0])
    c3 = np.array([c[0], c[1], z0])
    out = []
    for p, q in fan:
        out.append([c3, at(p, t[1]), at(q, t[1]), at(q, t[1])])
        for k in range(1, M):
            out.append([at(p, t[k]), at(p, t[k + 1]), at(q, t[k + 1]), at(q, t[k])])
    return np.array(out)
