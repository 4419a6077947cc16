"""Restart and export formats on either side of the device path (CPU, no kernels).

- binary restart of this library: exact round trip of the forest and S;
- the reference's text restart (restart_save.c:17-45, restart_load.c:30-58): round trip
  to the 11 significant digits of %15.10e, the tag of spawntree.c's worked example
  (spawntree.c:17-25), rejection of incomplete or malformed files;
- Tecplot FEPOINT (exportfield_linked_V3.c): the single-cell node order hand-derived
  from the reference's face/node traversal, node count and node means on a uniform
  grid in closed form, and on AMR / periodic forests the node set against the
  coincident corner points found by brute force.
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from fv_140823_amd import configs
from fv_140823_amd._capi import FIELD_SG
from helpers import amr, uniform


def _flat_arrays(F):
    fl = fv.FlatMesh(F.run)
    A = {k: np.array(v, copy=True) for k, v in fl.arrays().items()}
    fl.close()
    return A


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 1)])
def test_binary_restart_round_trip(tmp_path, periodic):
    F = amr(3, 5, periodic=periodic)
    S = np.random.default_rng(1).random((5, F.nleaves))
    F.set_state(S)
    F.save(tmp_path / "r.bin", 123, 0.375)
    G, ts, tm = fv.Forest.load(tmp_path / "r.bin")
    assert (ts, tm) == (123, 0.375)
    assert G.geometry() == F.geometry()
    assert np.array_equal(G.get_state(), S)
    for a, b in zip(F.leaf_info(), G.leaf_info()):
        assert np.array_equal(a, b)
    A, B = _flat_arrays(F), _flat_arrays(G)
    assert A.keys() == B.keys()
    for k in A:
        assert np.array_equal(A[k], B[k]), k


def _keys(xc, lev):
    """(level, i, j, k) of leaves of the unit box with one root"""
    h = 1.0 / (1 << (lev.astype(np.int64) - 1))
    ijk = np.round(xc / h[:, None] - 0.5).astype(np.int64)
    return [(int(l), int(a), int(b), int(c)) for l, (a, b, c) in zip(lev, ijk)]


def test_incremental_adaptation_equals_a_full_rebuild(tmp_path):
    """At a size where the host splits the kid allocation, the leaf list and the leaf
    refresh over threads (tens of thousands of leaves): the leaves carry the split and
    merge states, and the forest after adaptations, whose neighbours and geometry were
    refreshed only around the changed nodes, flattens to the same mesh as the forest
    rebuilt whole from its restart file."""
    F, _, meta = configs.blast(5, 2)
    rng = np.random.default_rng(3)
    for k in range(2):
        xc, lev = F.leaf_info()
        S = rng.random((5, F.nleaves))
        F.set_state(S)
        old = {key: q for q, key in enumerate(_keys(xc, lev))}
        r = np.sqrt(((xc - 0.5) ** 2).sum(axis=1))
        f = np.zeros(F.nleaves, np.int8)
        f[(np.abs(r - 0.2 - 0.03 * k) < 0.04) & (lev < meta["max_level"])] = 1
        f[r < 0.12] = -1
        F.adapt(f, meta["max_level"])
        # the kids made on the host's threads carry their parent's state (split.c:492-502),
        # merged parents the mean of their 8 kids in kid order (merge.c:101)
        xc2, lev2 = F.leaf_info()
        S2 = F.get_state()
        nsplit = nmerge = 0
        for q, (L, i, j, kk) in enumerate(_keys(xc2, lev2)):
            if (L, i, j, kk) in old:
                want = S[:, old[(L, i, j, kk)]]
            elif (L - 1, i >> 1, j >> 1, kk >> 1) in old:
                want = S[:, old[(L - 1, i >> 1, j >> 1, kk >> 1)]]
                nsplit += 1
            else:
                want = np.zeros(5)
                for c in range(8):
                    want = want + S[:, old[(L + 1, 2 * i + (c & 1), 2 * j + ((c >> 1) & 1),
                                            2 * kk + (c >> 2))]] / 8.0
                nmerge += 1
            assert np.array_equal(S2[:, q], want), (L, i, j, kk)
        assert nsplit > 8 * 512 and nmerge > 0
    assert F.nleaves > 40000
    F.save(tmp_path / "r.bin", 0, 0.0)
    G, _, _ = fv.Forest.load(tmp_path / "r.bin")
    A, B = _flat_arrays(F), _flat_arrays(G)
    for k in A:
        assert np.array_equal(A[k], B[k]), k


def test_binary_restart_rejects_other_files(tmp_path):
    (tmp_path / "x.bin").write_bytes(b"not a restart file at all" * 8)
    with pytest.raises(fv.FvError):
        fv.Forest.load(tmp_path / "x.bin")
    with pytest.raises(fv.FvError):
        fv.Forest.load(tmp_path / "missing.bin")


def _bare(F):
    g = F.geometry()
    return fv.Forest(g["nroot"], g["x0"], g["x1"], 0, g["bc"], g["periodic"], g["neq"])


@pytest.mark.parametrize("periodic", [(0, 0, 0), (0, 1, 0)])
def test_text_restart_round_trip(tmp_path, periodic):
    F = amr(2, 5, periodic=periodic)
    S = 1.0 + np.random.default_rng(2).random((5, F.nleaves))
    F.set_state(S)
    F.save_text(tmp_path / "restart00000042.dat", 42, 1.5)
    head = (tmp_path / "restart00000042.dat").read_text().splitlines()[0].split()
    assert head == ["42", "1.5000000000e+00", str(F.nleaves), "1", "5"]
    G = _bare(F)
    assert G.load_text(tmp_path / "restart00000042.dat") == (42, 1.5)
    for a, b in zip(F.leaf_info(), G.leaf_info()):
        assert np.array_equal(a, b)
    assert np.allclose(G.get_state(), S, rtol=6e-11, atol=0)
    # the reloaded forest flattens to the same mesh
    A, B = _flat_arrays(F), _flat_arrays(G)
    for k in A:
        assert np.array_equal(A[k], B[k]), k


def test_text_restart_tag_of_worked_example(tmp_path):
    """spawntree.c:17-25: tag 24221 is the leaf reached by kids 1,2,2,4,2 (1-based, from
    level 2 down to level 6); kid c sits at corner c of its parent (split.c:124-128)."""
    F = uniform(5, 5)   # leaves at level 6
    ijk = np.zeros(3, int)
    for c in (0, 1, 1, 3, 1):
        ijk = 2 * ijk + np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1])
    xc, lev = F.leaf_info()
    h = 1.0 / 32
    target = int(np.nonzero(np.all(np.abs(xc - (ijk + 0.5) * h) < 1e-12, axis=1))[0][0])
    F.save_text(tmp_path / "r.txt")
    line = (tmp_path / "r.txt").read_text().splitlines()[1 + target].split()
    assert line[:2] == ["0", "24221"]


def test_text_restart_rejects_bad_files(tmp_path):
    F = uniform(1, 5)
    F.save_text(tmp_path / "ok.txt")
    lines = (tmp_path / "ok.txt").read_text().splitlines()
    # a leaf missing (header still says 8)
    (tmp_path / "short.txt").write_text("\n".join(lines[:-1]) + "\n")
    with pytest.raises(fv.FvError):
        _bare(F).load_text(tmp_path / "short.txt")
    # a digit outside 1..8
    bad = lines[:1] + [lines[1].replace("0 1 ", "0 9 ", 1)] + lines[2:]
    (tmp_path / "digit.txt").write_text("\n".join(bad) + "\n")
    with pytest.raises(fv.FvError):
        _bare(F).load_text(tmp_path / "digit.txt")
    # the forest must be bare roots
    with pytest.raises(fv.FvError):
        uniform(1, 5).load_text(tmp_path / "ok.txt")


def _read_tecplot(path, nvar):
    lines = path.read_text().splitlines()
    assert lines[0].startswith(' VARIABLES =  "X" "Y" "Z" "iel" "part" "rho"')
    zone = lines[1].strip()
    assert zone.startswith('ZONE T="1",N=') and zone.endswith(",F=FEPOINT,ET=BRICK,DT=(double,double,double)")
    n = int(zone.split("N=")[1].split(",")[0])
    e = int(zone.split("E=")[1].split(",")[0])
    nodes = np.array([[float(t) for t in l.split()] for l in lines[2:2 + n]])
    assert nodes.shape == (n, 5 + nvar)
    elems = np.array([[int(t) for t in l.split()] for l in lines[2 + n:2 + n + e]])
    assert elems.shape == (e, 8) and len(lines) == 2 + n + e
    return nodes, elems


# Tecplot BRICK order of the hexahedron corners (corner bits x, y, z)
_BRICK = [0, 1, 3, 2, 4, 5, 7, 6]


def _corners(F):
    xc, lev = F.leaf_info()
    g = F.geometry()
    h = (np.array(g["x1"]) - np.array(g["x0"])) / (np.array(g["nroot"]) * 2.0 ** (lev[:, None] - 1))
    bits = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], float)
    return xc[:, None, :] + (bits[None, :, :] - 0.5) * h[:, None, :]


def test_tecplot_single_cell_node_order(tmp_path):
    """One hexahedron, all faces on the boundary: links open in the order face 0 nodes
    (0,1,5,4), face 1 new nodes (3,7), face 2 new nodes (2,6), so the BRICK line
    (corners 0,1,3,2,4,5,7,6) reads 1 2 5 7 4 3 6 8."""
    F = uniform(0, 5)
    F.set_state(np.array([[1.0], [0.1], [0.2], [0.3], [2.5]]))
    F.export_tecplot(tmp_path / "t.dat", 1, np.array([0.9]))
    nodes, elems = _read_tecplot(tmp_path / "t.dat", 5 + 1 + 4)
    assert elems.tolist() == [[1, 2, 5, 7, 4, 3, 6, 8]]
    order = [0, 1, 5, 4, 3, 7, 2, 6]
    want = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in order], float)
    assert np.array_equal(nodes[:, :3], want)
    assert np.allclose(nodes[:, 5:11], [1.0, 0.1, 0.2, 0.3, 2.5, 0.9], rtol=1e-6)


def test_tecplot_uniform_means(tmp_path):
    """4^3 cells: 5^3 nodes, and a node's values are the means over the cells around it."""
    F = uniform(2, 5)
    N = F.nleaves
    rng = np.random.default_rng(3)
    S = 1.0 + rng.random((5, N))
    SG = rng.random((6, 3, N))
    p = 1.0 + rng.random(N)
    F.set_state(S)
    F.set_field(FIELD_SG, SG)
    F.export_tecplot(tmp_path / "t.dat", 1, p)
    nodes, elems = _read_tecplot(tmp_path / "t.dat", 10)
    assert nodes.shape[0] == 125 and elems.shape[0] == N
    cor = _corners(F)
    for e in range(N):
        assert np.allclose(nodes[elems[e] - 1, :3], cor[e, _BRICK], atol=1e-6)
    key = np.round(nodes[:, :3] * 4).astype(int)
    vals = np.concatenate([S, p[None], np.sqrt((SG[[0, 1, 2, 5]] ** 2).sum(axis=1))])
    acc = {}
    ck = np.round(cor * 4).astype(int)
    for e in range(N):
        for c in range(8):
            acc.setdefault(tuple(ck[e, c]), []).append(vals[:, e])
    want = np.array([np.mean(acc[tuple(k)], axis=0) for k in key])
    assert np.allclose(nodes[:, 5:], want, rtol=2e-6)


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 1, 1)])
def test_tecplot_amr_nodes_are_the_coincident_corners(tmp_path, periodic):
    """On a 2:1 mesh the links join exactly the element corners that coincide (hanging
    nodes stay with the finer cells only because the coarse cell has no corner there);
    periodic sides are not joined."""
    F = amr(2, 5, periodic=periodic)
    N = F.nleaves
    F.set_state(np.ones((5, N)))
    F.export_tecplot(tmp_path / "t.dat", 1, np.ones(N))
    nodes, elems = _read_tecplot(tmp_path / "t.dat", 10)
    cor = _corners(F)
    q = 2.0 ** 6
    pts = {tuple(np.round(c * q).astype(int)) for c in cor.reshape(-1, 3)}
    assert nodes.shape[0] == len(pts)
    assert len({tuple(np.round(x * q).astype(int)) for x in nodes[:, :3]}) == len(pts)
    for e in range(N):
        assert np.allclose(nodes[elems[e] - 1, :3], cor[e, _BRICK], atol=1e-6)
