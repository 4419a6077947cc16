"""The host structs of include/fv_strdata.h have the reference's layout (strdata.h:60-540):
every field at the reference's byte offset with the reference's size, and every struct of
the reference's size, so that a RUN built by the reference's host code can be passed to
RHS(run) / RK(irk, run) / faceinterp(run) / gradcalc(run) of this library.

The expected layout is the committed fixture tests/golden/strdata_layout.json, written by
tests/golden/strdata_layout.py from the reference's own declarations (compiled with gcc
and the MPICH mpi.h of this image); where the reference is present the fixture is checked
against a fresh run of that script too."""
import json
import os
import pathlib
import sys

import pytest

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent
sys.path.insert(0, str(HERE / "golden"))
import strdata_layout as SL  # noqa: E402

FIXTURE = json.load(open(HERE / "golden" / "strdata_layout.json"))


def test_every_field_has_the_reference_offset_and_size():
    fields = FIXTURE["fields"]
    prelude = '#include "fv_strdata.h"'
    ours = SL.run_probe(SL.probe_program(prelude, fields), [str(REPO / "include")])
    ref = FIXTURE["layout"]
    assert set(ours) == set(ref)
    bad = {k: (ours[k], ref[k]) for k in ref if ours[k] != ref[k]}
    assert not bad, bad
    assert ref["CON"][0] > 2_000_000  # the statistics arrays make CON ~2 MB, as in the reference


def test_fields_the_library_reads_are_in_the_fixture():
    """The fields the flattener, the binding, the oracle and the export read."""
    used = {
        "RUN": ["con", "par", "topo", "info"],
        "CON": ["neq", "model", "ns", "ischeme", "iprimtv", "limiter", "ymin", "eos", "turb", "dt"],
        "PAR": ["matervisc", "materpinf", "matergama"],
        "FOREST": ["locl", "glob", "nleaves", "ntrees"],
        "BRANCH": ["level", "type", "nlfc", "nkids", "prnt", "lnxt", "neigtr", "neigfc",
                   "lfc_neigs", "nsfc", "kids", "el", "cl", "tag", "root"],
        "CELL": ["nx", "ny", "nz", "nxt1", "nyt1", "nzt1", "nxt2", "nyt2", "nzt2", "Vol", "Area",
                 "xc", "yc", "zc", "nd", "fc"],
        "LEAF": ["RHS", "S", "SN", "SF", "SG", "flux_flag"],
        "FC": ["bc"],
        "ND": ["x", "y", "z"],
    }
    for s, names in used.items():
        for n in names:
            assert n in FIXTURE["fields"][s], (s, n)


@pytest.mark.skipif(not os.path.exists("/root/reference/strdata.h"), reason="reference not present")
def test_fixture_matches_the_reference_declarations():
    fields, layout = SL.reference_layout("/root/reference/strdata.h")
    assert fields == FIXTURE["fields"]
    assert layout == FIXTURE["layout"]


def _amr_forest():
    from helpers import amr, state
    F = amr(3, 5)
    xc, _ = F.leaf_info()
    F.set_state(state(1, xc, seed=4))
    return F


def test_flattener_export_and_oracle_read_a_reference_shaped_run(tmp_path):
    """A RUN whose structs are each a malloc of exactly the reference's size (as
    createbranch.c:29 makes them) with guard bytes behind, and garbage in the RUN's
    other fields (a RUN on the reference's stack): the flattener builds the same mesh,
    the Tecplot export writes the same file, the oracle computes the same residual as on
    the library's own forest, and no byte past any struct is written."""
    import ctypes as C

    import numpy as np

    import fv_140823_amd as fv
    import oracle
    F = _amr_forest()
    R = oracle.ReferenceRun(F)
    a = fv.FlatMesh(F.run).arrays()
    b = fv.FlatMesh(R.run).arrays()
    assert set(a) == set(b)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert R.guards() == 0
    p_cell = np.linspace(1.0, 2.0, F.nleaves)
    f1, f2 = tmp_path / "own.dat", tmp_path / "ref.dat"
    F.export_tecplot(str(f1), 1, p_cell)
    assert fv.lib().fvh_export_tecplot(R.run, str(f2).encode(), 1,
                                       p_cell.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert f1.read_bytes() == f2.read_bytes()
    prm = fv.params(1)
    ref = oracle.full_rhs(F, prm)
    per = np.zeros(3)
    L = oracle.lib()
    pp = C.cast(C.byref(prm), C.c_void_p)
    assert L.orc_faceinterp(R.run, pp, per.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert L.orc_rhs(R.run, pp) == 0
    assert np.array_equal(R.get_field(fv.FIELD_RHS), ref)
    assert R.guards() == 0
    R.close()
