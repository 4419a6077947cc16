"""The benchmark workloads at their full sizes, checked through properties that hold at
any size (the oracle is too slow there; the per-step parity runs on the small configs).

- TGV 128^3, periodic Navier-Stokes: mass, momentum and energy totals conserved to
  round-off (every face flux enters its two cells with opposite signs), plain and mixed;
- Sod 128^3: the tube stays one-dimensional to round-off: v and w, and the spread of
  every x-plane, stay at the rounding of the arithmetic (1e-14 of the scale plain, 1e-6
  mixed, whose face flux is FP32; the face-normal frames are not bitwise alike across
  axes, so they are not exactly zero);
- blast 32^3 + 2 levels: 200 steps with the density-gradient adaptation every 10 steps
  (device criterion, device re-flatten with the state carried over): mass and energy
  conserved to round-off while the blast stays inside the box (split copies, merge
  averages over 8 equal volumes);
- cavitation 128^3 (barotropic model): the centred bubble keeps the symmetries of the
  cube (x <-> y exchange, x -> -x reflection) to round-off.
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from fv_140823_amd import configs, driver

pytestmark = pytest.mark.gpu


def _drift(t0, t1, scale):
    return np.abs(t1 - t0) / scale


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_tgv_128_conserves_totals(precision):
    print(f"case tgv 128^3 totals {precision}", flush=True)
    F, p, meta = configs.tgv(7)
    p.precision = fv.FVG_MIXED if precision == "mixed" else fv.FVG_PLAIN
    g = fv.Solver.from_forest(F, p)
    t0 = g.totals()
    dt = g.cfl_dt(0.5)
    g.steps(20, dt, 3)
    t1 = g.totals()
    S = g.get_state()
    vol = t0[5]
    # momentum totals are ~0: judged against the total of |rho u| (times the volume)
    mscale = np.abs(S[1:4]).sum(axis=1).max() * vol / S.shape[1]
    d = _drift(t0, t1, np.array([abs(t0[0]), mscale, mscale, mscale, abs(t0[4]), vol]))
    print("drift", d, flush=True)
    g.close()
    assert np.all(d < 1e-13), d


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_sod_128_stays_one_dimensional(precision):
    print(f"case sod 128^3 one-dimensional {precision}", flush=True)
    F, p, meta = configs.sod(7)
    p.precision = fv.FVG_MIXED if precision == "mixed" else fv.FVG_PLAIN
    xc, _ = F.leaf_info()
    g = fv.Solver.from_forest(F, p)
    dt = g.cfl_dt(0.5)
    g.steps(20, dt, 3)
    S = g.get_state()
    g.close()
    tol = 1e-14 if precision == "plain" else 1e-6
    scale = np.abs(S).max(axis=1)
    mom = scale[1]
    print("max |rho v|, |rho w| / max |rho u|", np.abs(S[2:4]).max(axis=1) / mom, flush=True)
    assert np.all(np.abs(S[2:4]).max(axis=1) <= tol * mom)
    ix = np.round(xc[:, 0] * 128 - 0.5).astype(int)
    order = np.argsort(ix, kind="stable")
    planes = S[:, order].reshape(S.shape[0], 128, 128 * 128)
    spread = np.abs(planes - planes[:, :, :1]).max(axis=(1, 2)) / np.maximum(scale, mom)
    print("x-plane spread", spread, flush=True)
    assert np.all(spread <= tol)
    # the shock has moved: the planes differ from the initial two states
    assert len(np.unique(planes[0, :, 0])) > 10


def test_blast_conserves_mass_and_energy_through_adaptation():
    print("case blast 200 steps with adaptation, totals", flush=True)
    F, p, meta = configs.blast(5, 2)
    g = fv.Solver.from_forest(F, p, device_flatten=True)
    t0 = g.totals()
    flags = dict(refine=0.01, coarsen=0.002, base_level=6, kind=0)
    st = driver.evolve(F, g, 200, meta["cfl"], 3, meta["adapt_every"], flags, meta["max_level"])
    t1 = g.totals()
    nleaves = F.nleaves
    S = g.get_state()
    g.close()
    xc, _ = F.leaf_info()
    print("leaves", nleaves, "remeshes", len(st["remesh"]), flush=True)
    assert len(st["remesh"]) == 19 and nleaves > 100000
    d = _drift(t0[[0, 4, 5]], t1[[0, 4, 5]], np.abs(t0[[0, 4, 5]]))
    print("drift of mass, energy, volume", d, flush=True)
    assert np.all(d < 1e-13), d
    # the blast has stayed inside: the cells near the walls still hold the ambient
    # state to round-off, so no mass or energy has left through the walls
    edge = np.any((xc < 0.05) | (xc > 0.95), axis=1)
    dev = np.abs(S[0, edge] - 1.0).max(), np.abs(S[1:4, edge]).max()
    print("wall cells", int(edge.sum()), "max |rho - 1|, |rho u|", dev, flush=True)
    assert edge.sum() > 1000 and max(dev) < 1e-10
    assert S[0].max() > 1.5


def test_cavitation_128_cube_symmetry():
    print("case cavitation 128^3 symmetry", flush=True)
    F, p, meta = configs.cavitation(7)
    xc, _ = F.leaf_info()
    g = fv.Solver.from_forest(F, p)
    dt = g.cfl_dt(0.5)
    g.steps(50, dt, 3)
    S = g.get_state()
    g.close()
    n = 128
    ijk = np.round((xc + 1.0) * n / 2 - 0.5).astype(int)
    grid = np.empty((S.shape[0], n, n, n))
    grid[:, ijk[:, 0], ijk[:, 1], ijk[:, 2]] = S
    scale = np.abs(grid).reshape(S.shape[0], -1).max(axis=1)
    mom = np.abs(grid[1:4]).max()
    # x <-> y: rho and w unchanged, rho u <-> rho v
    sw = grid.transpose(0, 2, 1, 3)
    e_rho = np.abs(grid[0] - sw[0]).max() / scale[0]
    e_u = np.abs(grid[1] - sw[2]).max() / mom
    e_w = np.abs(grid[3] - sw[3]).max() / mom
    # x -> -x: rho even, rho u odd
    fl = grid[:, ::-1]
    e_rx = np.abs(grid[0] - fl[0]).max() / scale[0]
    e_ux = np.abs(grid[1] + fl[1]).max() / mom
    print("symmetry errors", e_rho, e_u, e_w, e_rx, e_ux, "max |rho u|", mom, flush=True)
    assert mom > 0
    assert max(e_rho, e_rx) < 1e-12 and max(e_u, e_w, e_ux) < 1e-10
