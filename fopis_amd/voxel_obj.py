"""Write voxel spacecraft geometry as a MagicaVoxel-style .obj file.

The reference reads objects from .obj files of unit quads on voxel faces
(src/Spacecraft/ObjectUtils.cpp:9-169): vertex coordinates are integers
relative to the grid centre (the loader adds nx/2), and every face is an
axis-aligned quad. This writes one unit quad per exposed voxel face of a
union of axis-aligned boxes, with vertices shared across faces.
"""
from __future__ import annotations

import numpy as np

# face normals in MagicaVoxel order: -x, +x, -y, +y, -z, +z
_NORMALS = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def voxels_from_boxes(boxes, lo, hi) -> np.ndarray:
    """Boolean voxel grid over cells [lo, hi) per axis; boxes are ((x0,y0,z0),(x1,y1,z1)) half-open."""
    shape = tuple(h - l for l, h in zip(lo, hi))
    v = np.zeros(shape, bool)
    for (a, b) in boxes:
        sl = tuple(slice(a[q] - lo[q], b[q] - lo[q]) for q in range(3))
        v[sl] = True
    return v


def write_obj(path: str, voxels: np.ndarray, origin) -> int:
    """Write the exposed faces of `voxels` (cell (i,j,k) spans origin+ijk .. +1). Returns the face count."""
    verts: dict = {}
    vlist = []
    faces = []

    def vid(p):
        if p not in verts:
            verts[p] = len(vlist) + 1
            vlist.append(p)
        return verts[p]

    nx, ny, nz = voxels.shape
    pad = np.zeros((nx + 2, ny + 2, nz + 2), bool)
    pad[1:-1, 1:-1, 1:-1] = voxels
    ox, oy, oz = origin
    for i, j, k in zip(*np.nonzero(voxels)):
        x, y, z = ox + int(i), oy + int(j), oz + int(k)
        for n, (dx, dy, dz) in enumerate(_NORMALS):
            if pad[i + 1 + dx, j + 1 + dy, k + 1 + dz]:
                continue
            if dx != 0:
                xf = x + (1 if dx > 0 else 0)
                quad = [(xf, y, z), (xf, y + 1, z), (xf, y + 1, z + 1), (xf, y, z + 1)]
            elif dy != 0:
                yf = y + (1 if dy > 0 else 0)
                quad = [(x, yf, z), (x + 1, yf, z), (x + 1, yf, z + 1), (x, yf, z + 1)]
            else:
                zf = z + (1 if dz > 0 else 0)
                quad = [(x, y, zf), (x + 1, y, zf), (x + 1, y + 1, zf), (x, y + 1, zf)]
            faces.append(([vid(p) for p in quad], n + 1))
    with open(path, "w") as f:
        f.write("# voxel object: one unit quad per exposed voxel face\n\n")
        for nvec in _NORMALS:
            f.write("vn %d %d %d\n" % nvec)
        f.write("\nvt 0.5 0.5\n\n")
        for p in vlist:
            f.write("v %d %d %d\n" % p)
        f.write("\n")
        for ids, n in faces:
            f.write("f " + " ".join("%d/1/%d" % (v, n) for v in ids) + "\n")
    return len(faces)


def write_boxes(path: str, boxes) -> int:
    """Boxes in grid-centred integer cell coordinates, ((x0,y0,z0),(x1,y1,z1)) half-open."""
    lo = [min(b[0][q] for b in boxes) for q in range(3)]
    hi = [max(b[1][q] for b in boxes) for q in range(3)]
    return write_obj(path, voxels_from_boxes(boxes, lo, hi), lo)
