"""fvg_steps as captured HIP graphs against eager launches, bit for bit.

fvg_steps runs whole groups of six time steps as one captured graph (six steps bring
the rotating state buffers, and the FP32 copies of the mixed form, back to where they
started) and the last step of a call eagerly.  Each case here replays the graph several
times, with eager work of the program between the replays (host reads and writes of the
state, a change of dt, which recaptures), and compares every call's state and residual
with a solver that launches everything eagerly (the default, fvg_set_graphs(0)).
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from helpers import amr, state, uniform

pytestmark = pytest.mark.gpu

# id: (mesh, model, ns, precision, method, prandtl, tiling)
CASES = {
    "tgv_like_tiles_plain": ("periodic32", 1, 1, "plain", 3, 0.71, True),
    "tgv_like_tiles_mixed": ("periodic32", 1, 1, "mixed", 3, 0.71, True),
    "euler_amr_plain_rk2": ("amr", 1, 0, "plain", 2, 0.0, False),
    "euler_amr_mixed_rk3": ("amr", 1, 0, "mixed", 3, 0.0, False),
    "euler_walls_plain_euler": ("walls32", 1, 0, "plain", 1, 0.0, True),
    "model3_amr_rk3": ("amr", 3, 0, "plain", 3, 0.0, False),
}


def _mesh(kind, neq):
    if kind == "periodic32":
        return uniform(5, neq, periodic=(1, 1, 1), bc=(0,) * 6)
    if kind == "walls32":
        return uniform(5, neq)
    return amr(3, neq)


@pytest.mark.parametrize("name", list(CASES))
def test_graph_steps_bitwise_equal_eager(name):
    mesh, model, ns, precision, method, prandtl, tiling = CASES[name]
    neq = fv.neq_of(model)
    kw = dict(ns=ns, precision=precision)
    if ns:
        kw.update(mu=0.05, visc_work=1)
    if prandtl:
        kw["prandtl"] = prandtl
    p = fv.params(model, **kw)
    F = _mesh(mesh, neq)
    xc, _ = F.leaf_info()
    S0 = state(model, xc, seed=31 + len(name))
    print(f"case graphs {name}: {F.nleaves} leaves, method {method}", flush=True)
    solvers = []
    for graphs in (True, False):
        g = fv.Solver.from_forest(F, p)
        g.set_graphs(graphs)
        if tiling:
            g.set_tiling(True)
        g.set_state(S0)
        solvers.append(g)
    dt = solvers[1].cfl_dt(0.4)
    # 13 steps: two graph replays + one eager step; 7: one replay; a host write between;
    # then a new dt (recapture) and 12 steps
    plan = [(13, dt, None), (7, dt, None), (7, dt, "write"), (12, 0.8 * dt, None), (6, 0.8 * dt, None)]
    for n, h, between in plan:
        outs = []
        for g in solvers:
            if between == "write":
                S = g.get_state()
                g.set_state(S * 1.0)  # eager host round trip of the state
            g.steps(n, h, method)
            outs.append((g.get_state(), g.get_field(fv.FIELD_RHS)))
        (Sa, Ra), (Sb, Rb) = outs
        d = np.abs(Sa - Sb).max()
        print(f"  {n} steps at dt {h:.3e}: max |graph - eager| {d:.3e}", flush=True)
        assert np.all(np.isfinite(Sa))
        assert np.array_equal(Sa, Sb)
        assert np.array_equal(Ra, Rb)
    for g in solvers:
        g.close()
