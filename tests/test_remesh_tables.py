"""Per-level tables of the device re-flatten, host half (CPU, no kernels).

The device re-flatten names, for every table entry, the leaf that sets it first in list
order (the rule below, restated here from the host flattener's scan); the host then
evaluates the entries at those leaves only (fvh_tables_from_owners).  Checked: the
owners restated in numpy from the host-flattened mesh give tables bitwise equal to
fvh_flatten_run's own, on uniform, 2:1, periodic, multi-root and adapted-blast forests;
and an owner of the wrong level is rejected.
"""
import ctypes as C

import numpy as np
import pytest

import fv_140823_amd as fv
from fv_140823_amd import configs
from fv_140823_amd._capi import lib
from helpers import amr, uniform

AXF = ((1, 3), (0, 2), (4, 5))  # faces of each axis
FINER = 1


def owners_of(A):
    """h row L: first leaf of level L or L+1; volume L: first leaf of level L; area (lf, a):
    first leaf with a face of axis a whose finer side has level lf."""
    cf = A["cflag"].astype(np.int64)
    lev = (cf >> 16) & 0xFF
    own = np.full(160, -1, np.int32)

    def first(mask):
        idx = np.nonzero(mask)[0]
        return int(idx[0]) if idx.size else -1

    for L in range(1, 32):
        own[L] = first((lev == L) | (lev == L + 1))
        own[32 + L] = first(lev == L)
    for a in range(3):
        fin = [((cf >> (2 * f)) & 3) == FINER for f in AXF[a]]
        some_fine = fin[0] | fin[1]
        some_coarse = ~fin[0] | ~fin[1]
        for L in range(1, 32):
            cand = ((lev == L) & some_coarse) | ((lev + 1 == L) & some_fine)
            own[64 + 3 * L + a] = first(cand)
    return own


def tables_from_owners(F, own):
    h = np.zeros(96)
    dist = np.zeros(32 * 18)
    area = np.zeros(96)
    vol = np.zeros(32)
    nl = C.c_int32()
    d = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib().fvh_tables_from_owners(lib().fvh_forest_leaves(F._h), F.nleaves,
                                      own.ctypes.data_as(C.POINTER(C.c_int32)), d(h), d(dist),
                                      d(area), d(vol), C.byref(nl))
    return rc, nl.value, h.reshape(32, 3), dist.reshape(32, 3, 3, 2), area.reshape(32, 3), vol


def skewed_amr():
    """A box whose cell sizes are not dyadic: the node differences of the cells of one
    level then differ in their last bits, so the table entries depend on which leaf sets
    them."""
    F = fv.Forest((1, 1, 1), (0.1, 0.2, 0.3), (0.8, 1.3, 1.0), 3, (2, 2, 2, 2, 2, 2), (0, 0, 0), 5)
    for k in range(2):
        xc, _ = F.leaf_info()
        r = np.sqrt(((xc - np.array([0.4, 0.7, 0.6])) ** 2).sum(axis=1))
        F.adapt((r < 0.3 * (1.0 - 0.35 * k)).astype(np.int8), 6)
    return F


def _forests():
    yield "skewed amr", skewed_amr()
    yield "uniform", uniform(3, 5)
    yield "amr", amr(3, 5)
    yield "amr periodic", amr(3, 5, periodic=(1, 0, 1))
    yield "two roots", fv.Forest((2, 1, 1), (0, 0, 0), (2, 1, 1), 2, (2, 2, 1, 2, 2, 2), (0, 0, 0), 5)
    F, _, meta = configs.blast(3, 2)
    yield "blast", F


@pytest.mark.parametrize("name,F", list(_forests()), ids=lambda v: v if isinstance(v, str) else "")
def test_tables_from_owners_equal_the_flattener(name, F):
    fl = fv.FlatMesh(F.run)
    A = fl.arrays()
    nl_ref = fl.mesh.nlevel
    fl.close()
    rc, nl, h, dist, area, vol = tables_from_owners(F, owners_of(A))
    assert rc == 0
    assert nl == nl_ref
    assert np.array_equal(h[:nl], A["h"])
    assert np.array_equal(dist[:nl], A["dist"])
    assert np.array_equal(area[:nl], A["area"])
    assert np.array_equal(vol[:nl], A["vol"])


def test_tables_reject_an_owner_of_the_wrong_level():
    F = amr(3, 5)
    fl = fv.FlatMesh(F.run)
    own = owners_of(fl.arrays())
    fl.close()
    lev = F.leaf_info()[1]
    L = int(lev.max())
    own[32 + L] = int(np.nonzero(lev != L)[0][0])
    assert tables_from_owners(F, own)[0] < 0


def test_owner_rule_matters_on_a_skewed_box():
    """The check above is sharp: taking the last leaf of each class instead of the first
    changes some table entry on the skewed box."""
    F = skewed_amr()
    fl = fv.FlatMesh(F.run)
    A = fl.arrays()
    fl.close()
    rev = {k: (v[::-1].copy() if k == "cflag" else v) for k, v in A.items()}
    own = owners_of(rev)
    n = F.nleaves
    own = np.where(own >= 0, n - 1 - own, -1).astype(np.int32)
    rc, nl, h, dist, area, vol = tables_from_owners(F, own)
    assert rc == 0
    assert not (np.array_equal(h[:nl], A["h"]) and np.array_equal(area[:nl], A["area"])
                and np.array_equal(vol[:nl], A["vol"]))
