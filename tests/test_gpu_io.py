"""GPU checks of restart and export from the device state.

- the cell pressure of the device EOS (the export's P column) against the oracle's EOS;
- a run interrupted by a binary restart (device state -> file -> new forest and
  solver) continues bit for bit like the uninterrupted run, plain and mixed;
- the Tecplot file written from the device state carries the device's S, pressure and
  gradient magnitudes as node means.
"""
import numpy as np
import pytest

import fv_140823_amd as fv
import oracle
from fv_140823_amd import driver
from fv_140823_amd._capi import FIELD_SG
from helpers import amr, state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", [1, 2, 3])
def test_device_pressure_matches_oracle_eos(model):
    print(f"case device pressure model={model}", flush=True)
    neq = fv.neq_of(model)
    F = amr(3, neq)
    xc, _ = F.leaf_info()
    S = state(model, xc, seed=4)
    P = fv.params(model)
    g = fv.Solver.from_forest(F, P)
    g.set_state(S)
    p = g.pressure()
    g.close()
    want = np.array([oracle.eos(P, S[:, i])[0] for i in range(S.shape[1])])
    err = np.abs(p - want).max() / np.abs(want).max()
    print("max rel diff", err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_restart_continues_bitwise(tmp_path, precision):
    print(f"case restart continuity {precision}", flush=True)
    F = amr(3, 5)
    xc, _ = F.leaf_info()
    F.set_state(state(1, xc, seed=8))
    P = fv.params(1, precision=precision)
    g = fv.Solver.from_forest(F, P)
    dt = g.cfl_dt(0.4)
    g.steps(10, dt, 3)
    S_ref = g.get_state()
    g.close()

    g = fv.Solver.from_forest(F, P)
    g.steps(4, dt, 3)
    driver.save_restart(F, g, tmp_path / "r.bin", 4, 4 * dt)
    g.close()
    F2, g2, ts, tm = driver.load_restart(tmp_path / "r.bin", P)
    assert (ts, tm) == (4, 4 * dt)
    g2.steps(6, dt, 3)
    assert np.array_equal(g2.get_state(), S_ref)
    g2.close()


def test_export_from_device_state(tmp_path):
    print("case tecplot export from device", flush=True)
    F = amr(2, 5)
    xc, _ = F.leaf_info()
    F.set_state(state(1, xc, seed=9))
    g = fv.Solver.from_forest(F, fv.params(1))
    g.steps(3, g.cfl_dt(0.4), 3)
    driver.export_tecplot(F, g, tmp_path / "t.dat")
    S, p = g.get_state(), g.pressure()
    g.gradcalc()
    SG = g.get_field(FIELD_SG)
    g.close()
    lines = (tmp_path / "t.dat").read_text().splitlines()
    n = int(lines[1].split("N=")[1].split(",")[0])
    nodes = np.array([[float(t) for t in l.split()] for l in lines[2:2 + n]])
    elems = np.array([[int(t) for t in l.split()] for l in lines[2 + n:]])
    vals = np.concatenate([S, p[None], np.sqrt((SG[[0, 1, 2, 5]] ** 2).sum(axis=1))])
    members = [[] for _ in range(n)]
    for e in range(elems.shape[0]):
        for k in elems[e]:
            members[k - 1].append(e)
    want = np.array([vals[:, m].mean(axis=1) for m in members])
    scale = np.abs(vals).max(axis=1)
    assert np.all(np.abs(nodes[:, 5:] - want) <= 2e-6 * scale)
