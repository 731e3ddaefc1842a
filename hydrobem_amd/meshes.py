"""Procedural low-order meshes of the solve's workloads (normals out of the body into the fluid).

The generators give the panel layout of the reference's test meshes vertex for vertex
(tests/python/meshes.py: orient :10-13, face_panels :16-24, box_panels :27-46,
hemisphere_panels :49-71, _cosine :74-78, cylinder_panels :81-112): the same vertex order,
the same orientation flips, triangles as a repeated last vertex. Arrays are (P, 4, 3) float64.
"""

from __future__ import annotations

import numpy as np


def orient(quad, normal):
    """The quad as given if its diagonal cross product points along `normal`, else reversed."""
    q = np.asarray(quad, dtype=float)
    n = np.cross(q[2] - q[0], q[3] - q[1])
    return q if np.dot(n, normal) > 0 else q[[0, 3, 2, 1]]


def face_panels(origin, du, dv, nu, nv, normal):
    origin, du, dv = (np.asarray(v, dtype=float) for v in (origin, du, dv))
    out = []
    a, b = du / nu, dv / nv
    for i in range(nu):
        for j in range(nv):
            p = origin + a * i + b * j
            out.append(orient([p, p + a, p + a + b, p + b], normal))
    return out


def box_panels(a=1.0, b=0.5, d=0.5, n=(4, 2, 2), quarter=False):
    """Surface-piercing box |x| <= a, |y| <= b, -d <= z <= 0: bottom nx*ny, x faces ny*nz,
    y faces nx*nz panels (quarter: the part x >= 0, y >= 0)."""
    nx, ny, nz = n
    x0, y0 = (0.0, 0.0) if quarter else (-a, -b)
    lx, ly = a - x0, b - y0
    panels = []
    panels += face_panels((x0, y0, -d), (lx, 0, 0), (0, ly, 0), nx, ny, (0, 0, -1))
    panels += face_panels((a, y0, -d), (0, ly, 0), (0, 0, d), ny, nz, (1, 0, 0))
    panels += face_panels((x0, b, -d), (lx, 0, 0), (0, 0, d), nx, nz, (0, 1, 0))
    if not quarter:
        panels += face_panels((-a, y0, -d), (0, ly, 0), (0, 0, d), ny, nz, (-1, 0, 0))
        panels += face_panels((x0, -b, -d), (lx, 0, 0), (0, 0, d), nx, nz, (0, -1, 0))
    return np.array(panels)


def hemisphere_panels(radius=1.0, n_theta=8, n_phi=8, quarter=True):
    """Floating hemisphere z <= 0: n_theta polar divisions from the waterline to the pole,
    n_phi azimuthal divisions over the part; the bottom ring is triangles."""
    phis = np.linspace(0.0, np.pi / 2 if quarter else 2 * np.pi, n_phi + 1)
    s = np.linspace(0.0, 1.0, n_theta + 1)
    ths = np.pi / 2 + 0.5 * np.pi * s

    def pt(t, p):
        return radius * np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])

    out = []
    for i in range(n_theta):
        for j in range(n_phi):
            a, b = pt(ths[i], phis[j]), pt(ths[i], phis[j + 1])
            c, d = pt(ths[i + 1], phis[j + 1]), pt(ths[i + 1], phis[j])
            quad = [a, b, c, d] if i + 1 < n_theta else [a, b, c, c]
            mid = np.mean(quad, axis=0)
            out.append(orient(quad, mid))
    return np.array(out)


def _cosine(n):
    """n intervals on [0, 1], refined at both ends."""
    if n == 0:
        return np.zeros(1)
    return 0.5 * (1.0 - np.cos(np.pi * np.arange(n + 1) / n))


def cylinder_panels(radius=1.0, draft=0.5, n_phi=8, n_z=2, n_r=2, quarter=True, cosine=False):
    """Floating vertical circular cylinder (polygonal), side then bottom; cosine spacing of the
    vertical and radial rings refines them towards the edges."""
    phis = np.linspace(0.0, np.pi / 2 if quarter else 2 * np.pi, n_phi + 1)
    zs = -draft * (_cosine(n_z) if cosine else np.linspace(0.0, 1.0, n_z + 1))
    rs = radius * (1.0 - (_cosine(n_r) if cosine else np.linspace(0.0, 1.0, n_r + 1)))
    out = []
    for i in range(n_z):
        for j in range(n_phi):
            p0, p1 = phis[j], phis[j + 1]
            q = [
                [radius * np.cos(p0), radius * np.sin(p0), zs[i]],
                [radius * np.cos(p1), radius * np.sin(p1), zs[i]],
                [radius * np.cos(p1), radius * np.sin(p1), zs[i + 1]],
                [radius * np.cos(p0), radius * np.sin(p0), zs[i + 1]],
            ]
            mid = np.mean(q, axis=0)
            out.append(orient(q, [mid[0], mid[1], 0.0]))
    for i in range(n_r):
        for j in range(n_phi):
            p0, p1 = phis[j], phis[j + 1]
            r0, r1 = rs[i], rs[i + 1]
            q = [
                [r0 * np.cos(p0), r0 * np.sin(p0), -draft],
                [r0 * np.cos(p1), r0 * np.sin(p1), -draft],
                [r1 * np.cos(p1), r1 * np.sin(p1), -draft],
                [r1 * np.cos(p0), r1 * np.sin(p0), -draft],
            ]
            if i + 1 == n_r:
                q = q[:3] + [q[2]]
            out.append(orient(q, [0.0, 0.0, -1.0]))
    return np.array(out)


def expand_symmetry(quadrant):
    """A quadrant mesh with the symmetry planes x = 0 and y = 0 as the whole mesh: the stored
    panels, then their x-, y- and xy-reflections, the orientation reversed where a reflection
    reverses it (geometry.rs:209-234 reflections / reflect_vertices)."""
    q = np.asarray(quadrant, dtype=float)
    out = [q]
    for sx, sy in ((-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0)):
        r = q * np.array([sx, sy, 1.0])
        out.append(r[:, [0, 3, 2, 1]] if sx * sy < 0 else r)
    return np.concatenate(out)


def place(vertices, position=(0.0, 0.0, 0.0)):
    """The panels translated to a body position."""
    return np.asarray(vertices, dtype=float) + np.asarray(position, dtype=float)
