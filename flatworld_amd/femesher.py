"""Small mesh generators (flatworld/femesher.py:115-149) plus the structured grid.

``createCircle`` builds the 16-segment fan used by the reference's contact
tests; ``structured_grid`` builds the square grid with the node numbering of
test2D/test_2Dfem_elastic10000.py:23-32 (node (i, j) -> i*(s+1)+j at
(x0 + h*i, y0 + h*j), triangles [a, a+1, b] and [a+1, b+1, b]).
"""

import numpy as np

from .mesh import Mesh


class FEMesher:
    def __init__(self, d):
        self.d = d

    def createTriangle(self, pt1, pt2, pt3):
        coords = np.array([pt1, pt2, pt3], dtype=np.float32)
        conn = np.array([[0, 1, 2]], dtype=np.int32)
        return Mesh(2, conn, coords)

    def createRectangle(self, lb, up):
        lx, ly = lb
        ux, uy = up
        verts = [[lx, ly], [ux, ly], [ux, uy], [lx, uy]]
        conn = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
        return Mesh(2, conn, np.array(verts, dtype=np.float32))

    def createCircle(self, origin, radius):
        x, y = origin
        segments = 16
        verts = []
        for i in range(segments):
            theta = 2.0 * np.pi * i / segments
            verts.append([x + radius * np.cos(theta), y + radius * np.sin(theta)])
        center_idx = len(verts)
        verts.append([x, y])
        conn = [[i, (i + 1) % segments, center_idx] for i in range(segments)]
        return Mesh(2, np.array(conn, dtype=np.int32), np.array(verts, dtype=np.float32))


def structured_grid(side, lower=(0.0, 0.0), extent=1.0, jitter=0.0, seed=0):
    """Square grid of ``side`` x ``side`` cells split into 2*side^2 triangles.

    Returns ``(conn int32[2s^2, 3], coords float32[(s+1)^2, 2])``.  With
    ``jitter > 0`` every node moves by U(-jitter*h, +jitter*h) in x and y from a
    PCG64 generator seeded with ``seed``.
    """
    s = int(side)
    h = float(extent) / s
    ii, jj = np.meshgrid(np.arange(s + 1), np.arange(s + 1), indexing="ij")
    coords = np.empty(((s + 1) * (s + 1), 2), dtype=np.float64)
    coords[:, 0] = (h * ii).reshape(-1)
    coords[:, 1] = (h * jj).reshape(-1)
    if jitter > 0.0:
        rng = np.random.Generator(np.random.PCG64(seed))
        coords += rng.uniform(-jitter * h, jitter * h, size=coords.shape)
    coords[:, 0] += lower[0]
    coords[:, 1] += lower[1]
    ci, cj = np.meshgrid(np.arange(s), np.arange(s), indexing="ij")
    a = (ci * (s + 1) + cj).reshape(-1)
    b = ((ci + 1) * (s + 1) + cj).reshape(-1)
    conn = np.empty((2 * s * s, 3), dtype=np.int32)
    conn[0::2, 0] = a
    conn[0::2, 1] = a + 1
    conn[0::2, 2] = b
    conn[1::2, 0] = a + 1
    conn[1::2, 1] = b + 1
    conn[1::2, 2] = b
    return conn, coords.astype(np.float32)
