"""The host side of the LDS-tiled stage: which blocks of a mesh are interior tiles.

fvh_tile_table runs the search fvg_set_mesh runs (no GPU): 8 x 8 x 8 blocks of 512
consecutive leaves of one level, from any position in the leaf list, laid out in Morton
order (local index = bits x0 y0 z0 x1 y1 z1 x2 y2 z2), whose six face-neighbour blocks are
such tiles too.  Each row is checked here against the flattened mesh's own neighbour table,
independently of the search: the +x / +y / +z neighbour of every leaf inside the block is
the Morton one, and across each face the boundary layer continues into the named
neighbour tile at the Morton positions of its opposite layer.  The counts are the
geometric ones: a periodic 32^3 box is 4^3 tiles, all interior; with outflow walls only
the 2^3 inner ones are interior; refining the last 2^3 leaves removes one tile and the
interior status of its six neighbours, and refining the first 2^3 leaves does the same
while shifting every later block by 56 leaves off the multiples of 512.
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from helpers import amr, uniform

T = np.arange(512)
X = (T & 1) | ((T >> 2) & 2) | ((T >> 4) & 4)
Y = ((T >> 1) & 1) | ((T >> 3) & 2) | ((T >> 5) & 4)
Z = ((T >> 2) & 1) | ((T >> 4) & 2) | ((T >> 6) & 4)
XYZ = np.stack([X, Y, Z], axis=1)
FACE_AXIS = [1, 0, 1, 0, 2, 2]
FACE_PLUS = [False, True, True, False, False, True]


def morton(xyz):
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    return ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) |
            ((z & 2) << 4) | ((x & 4) << 4) | ((y & 4) << 5) | ((z & 4) << 6))


def check_rows(flat, rows):
    A = flat.arrays()
    N = flat.ncell
    nbr = A["nbr"].reshape(6, N)
    cf = A["cflag"].astype(np.int64)
    level = (cf >> 16) & 0xFF
    assert np.array_equal(morton(XYZ), T)
    for row in rows:
        b, nb, lev = int(row[0]), row[1:7], int(row[7])
        assert 0 <= b and b + 512 <= N
        cells = b + T
        assert np.all(level[cells] == lev)
        for f in range(6):
            a, plus = FACE_AXIS[f], FACE_PLUS[f]
            # inside the block: the neighbour across f is the Morton one
            q = XYZ.copy()
            q[:, a] += 1 if plus else -1
            inside = (q[:, a] >= 0) & (q[:, a] < 8)
            assert np.all(((cf[cells] >> (2 * f)) & 3) == 0)  # same level across every face
            assert np.array_equal(nbr[f][cells[inside]], b + morton(q[inside]))
            # across the face: the neighbour tile's opposite layer
            q[:, a] %= 8
            assert np.array_equal(nbr[f][cells[~inside]], int(nb[f]) + morton(q[~inside]))
            assert np.all(level[int(nb[f]) + T] == lev)


@pytest.mark.parametrize("kind,expected", [
    ("periodic32", 64), ("walls32", 8), ("periodic8", 1), ("walls16", 0), ("amr", 0)])
def test_tile_table(kind, expected):
    if kind == "periodic32":
        F = uniform(5, 5, periodic=(1, 1, 1), bc=(0,) * 6)
    elif kind == "walls32":
        F = uniform(5, 5)
    elif kind == "periodic8":  # one tile, its own neighbour across every face
        F = uniform(3, 5, periodic=(1, 1, 1), bc=(0,) * 6)
    elif kind == "walls16":  # 8 tiles, none with six tile neighbours
        F = uniform(4, 5)
    else:
        F = amr(3, 5)
    flat = fv.FlatMesh(F.run)
    rows = flat.tile_table()
    assert rows.shape == (expected, 8)
    check_rows(flat, rows)
    if kind == "periodic8":
        assert np.all(rows[0, 1:7] == rows[0, 0])
    flat.close()


@pytest.mark.parametrize("corner", ["last", "first"])
def test_tile_table_after_refinement(corner):
    F = uniform(5, 5, periodic=(1, 1, 1), bc=(0,) * 6)
    xc, _ = F.leaf_info()
    sel = (xc > 1.0 - 1.0 / 16) if corner == "last" else (xc < 1.0 / 16)
    F.adapt(sel.all(axis=1).astype(np.int8), 7)
    flat = fv.FlatMesh(F.run)
    rows = flat.tile_table()
    assert F.nleaves == 32768 + 56 and len(rows) == 57
    check_rows(flat, rows)
    starts = set(rows[:, 0].tolist())
    if corner == "last":
        assert 32768 - 512 not in starts  # the refined block
        assert all(b % 512 == 0 for b in starts)
    else:
        assert 0 not in starts and all(b % 512 == 56 for b in starts)
    flat.close()
