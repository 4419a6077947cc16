"""The LDS-tiled stage (k_tile_stage) against the face-list path, bit for bit.

On a mesh with interior tiles (8 x 8 x 8 same-level blocks of consecutive cells whose
six neighbour blocks are tiles too) the out-of-place stages of fvg_steps update the
tiles' cells with one kernel that evaluates every face of the tile from LDS and keeps
the face fluxes there; the face-list kernels skip the blocks that hold only those
faces and cells.  A face is evaluated with the face-list kernel's arithmetic in the
face-list orientation and each cell's six fluxes are summed in the gather's order, so
the states and the last stage's residual are bitwise those of the face-list path
(fvg_set_tiling(0)).  The face-list path itself is compared with the oracle in
test_gpu_stage_parity, which also runs the tile meshes here against the oracle.

Meshes: 32^3 periodic (every tile interior: the face kernel and the gather are not
launched at all), 32^3 with outflow walls (the 8 inner tiles; blocks of both kinds), a
32^3 mesh refined in a sphere (tiles beside level changes and hanging faces), and the
periodic box with its first 2^3 leaves refined (tiles off the 512 alignment: blocks of
cells covered in part, which the gather and the gradient kernel mask cell by cell).
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from fv_140823_amd.configs import prim_to_cons
from helpers import amr, state, uniform

pytestmark = pytest.mark.gpu


def _mesh(kind, neq):
    if kind == "periodic32":
        return uniform(5, neq, periodic=(1, 1, 1), bc=(0,) * 6)
    if kind == "walls32":
        return uniform(5, neq)
    if kind == "amr32":
        return amr(5, neq, passes=1, centre=(0.3, 0.3, 0.3), radius=0.12)
    if kind == "shifted32":
        # the first 2^3 leaves refined: every later tile starts 56 leaves off a multiple
        # of 512, so cell blocks are covered in part (per-cell masks)
        F = uniform(5, neq, periodic=(1, 1, 1), bc=(0,) * 6)
        xc, _ = F.leaf_info()
        F.adapt(((xc < 1.0 / 16).all(axis=1)).astype(np.int8), 7)
        return F
    raise ValueError(kind)


# id: (model, ns, limiter, precision, method, options)
CONFIGS = {
    "m1_euler_minmod_plain_rk3": (1, 0, "minmod", "plain", 3, {}),
    "m1_euler_vanalbada_plain_rk3": (1, 0, "vanalbada", "plain", 3, {}),
    "m1_euler_vanalbada_mixed_rk3": (1, 0, "vanalbada", "mixed", 3, {}),
    "m1_euler_minmod_mixed_rk2": (1, 0, "minmod", "mixed", 2, {}),
    "m1_euler_minmod_plain_euler": (1, 0, "minmod", "plain", 1, {}),
    "m1_hypersonic_mixed_rk3": (1, 0, "minmod", "mixed", 3, {"hypersonic": True}),
    "m1_ns_plain_rk3": (1, 1, "minmod", "plain", 3, {}),
    "m1_ns_mixed_rk3": (1, 1, "minmod", "mixed", 3, {}),
    "m1_ns_heat_plain_rk3": (1, 1, "minmod", "plain", 3, {"prandtl": 0.71}),
    "m1_ns_heat_mixed_rk3": (1, 1, "vanalbada", "mixed", 3, {"prandtl": 0.71}),
    "m2_plain_rk3": (2, 0, "minmod", "plain", 3, {}),
    "m2_mixed_rk3": (2, 0, "vanalbada", "mixed", 3, {}),
    "m3_ns_plain_rk3": (3, 1, "minmod", "plain", 3, {}),
    "m3_euler_mixed_rk3": (3, 0, "minmod", "mixed", 3, {}),
}

MESHES = ["periodic32", "walls32", "amr32", "shifted32"]


def _setup(mesh, name):
    model, ns, limiter, precision, method, opt = CONFIGS[name]
    neq = fv.neq_of(model)
    kw = dict(ns=ns, limiter=limiter, precision=precision)
    if ns:
        kw.update(mu=0.05 if model == 1 else 2.0, visc_work=1)
    if opt.get("prandtl"):
        kw["prandtl"] = opt["prandtl"]
    p = fv.params(model, **kw)
    F = _mesh(mesh, neq)
    xc, _ = F.leaf_info()
    S = state(model, xc, seed=7 + sum(map(ord, name)))
    if opt.get("hypersonic"):
        # Mach ~40 along x: the FP32 inversion is rejected and the FP64 rebuild runs
        rho = S[0]
        u, v, w = S[1] / rho, S[2] / rho, S[3] / rho
        pr = 0.4 * (S[4] - 0.5 * rho * (u * u + v * v + w * w))
        S = prim_to_cons(rho, u + 50.0, v, w, pr)
    return F, p, S, method, neq


def _run(F, p, S, method, dt, tiling, nsteps=3):
    g = fv.Solver.from_forest(F, p)
    g.set_tiling(tiling)
    g.set_state(S)
    ntile = g.tile_count
    g.steps(nsteps, dt, method)
    out = g.get_state(), g.get_field(fv.FIELD_RHS), g.check_state()
    g.close()
    return ntile, out


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("name", list(CONFIGS))
def test_tiled_stage_bitwise_equals_face_list(mesh, name):
    F, p, S, method, neq = _setup(mesh, name)
    print(f"case tiles {mesh} {name}: {F.nleaves} leaves", flush=True)
    g = fv.Solver.from_forest(F, p)
    g.set_state(S)
    dt = g.cfl_dt(0.4)
    g.close()
    nt_on, (S_on, R_on, bad_on) = _run(F, p, S, method, dt, True)
    nt_off, (S_off, R_off, bad_off) = _run(F, p, S, method, dt, False)
    print(f"  {nt_on} interior tiles", flush=True)
    assert nt_on == nt_off
    if mesh == "periodic32":
        assert nt_on == 64
    elif mesh == "walls32":
        assert nt_on == 8
    elif mesh == "shifted32":
        assert nt_on == 57
    assert bad_on == -1 and bad_off == -1
    assert np.all(np.isfinite(S_on))
    assert np.abs(S_on - S).max() > 0
    d = np.abs(S_on - S_off).max(axis=1)
    print(f"  max |S_tiled - S_face_list| per variable {d}", flush=True)
    assert np.array_equal(S_on, S_off)
    assert np.array_equal(R_on, R_off)


def test_tiles_follow_the_mesh():
    """A refined mesh rebuilds the tile table: the periodic 32^3 box has 64 interior
    tiles (4 x 4 x 4); refining the 8 cells at the far corner (the last in Morton
    order) takes away the last tile and leaves its 6 face neighbours (periodic) without
    a tile across one face: 57 remain, and the stage still matches the face-list path
    bitwise."""
    F = uniform(5, 5, periodic=(1, 1, 1), bc=(0,) * 6)
    p = fv.params(1)
    g = fv.Solver.from_forest(F, p)
    assert g.tile_count == 64
    g.close()
    xc, _ = F.leaf_info()
    flags = ((xc > 1.0 - 1.0 / 16).all(axis=1)).astype(np.int8)
    F.adapt(flags, 7)  # leaves of the 32^3 box are level 6 (the root is level 1)
    xc, _ = F.leaf_info()
    S = state(1, xc, seed=3)
    g = fv.Solver.from_forest(F, p)
    n = g.tile_count
    g.set_state(S)
    dt = g.cfl_dt(0.4)
    g.close()
    print(f"case tiles after refinement: {F.nleaves} leaves, {n} interior tiles", flush=True)
    assert F.nleaves == 32768 + 56 and n == 57
    _, (S_on, R_on, _) = _run(F, p, S, 3, dt, True)
    _, (S_off, R_off, _) = _run(F, p, S, 3, dt, False)
    assert np.array_equal(S_on, S_off)
    assert np.array_equal(R_on, R_off)


def test_tiling_modes():
    """fvg_set_tiling: the automatic choice keeps the face-list kernels for a mesh with few
    tiles (64 here, below 256) and the forced mode runs the tiled kernel on it; kernel id
    4 of fvg_time_steps is the tiled stage's duration, 0 when it is not launched.  A mode
    outside the three is refused."""
    F = uniform(5, 5, periodic=(1, 1, 1), bc=(0,) * 6)
    xc, _ = F.leaf_info()
    g = fv.Solver.from_forest(F, fv.params(1))
    g.set_state(state(1, xc, seed=4))
    dt = g.cfl_dt(0.4)
    _, k_auto = g.time_steps(2, dt, 3)
    g.set_tiling(True)
    _, k_on = g.time_steps(2, dt, 3)
    g.set_tiling(False)
    _, k_off = g.time_steps(2, dt, 3)
    print(f"case tiling modes: tile kernel ms auto {k_auto[4]}, on {k_on[4]}, off {k_off[4]}", flush=True)
    assert g.tile_count == 64
    assert k_auto[4] == 0 and k_off[4] == 0 and k_on[4] > 0
    assert k_on[2] == 0 and k_auto[2] > 0  # every face in a tile: no face-list launch when forced
    with pytest.raises(fv.FvError):
        g.set_tiling(5)
    g.close()
