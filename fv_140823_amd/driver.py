"""Host time loop around the device path: the stage sequence of the reference's
(undefined) advancetimeexplicit -- RK(0), then per stage faceinterp, gradcalc,
RHS, RK(k) (loop.c:35-74, RK.c:16-51) -- with a CFL time step and a host mesh
adaptation every `adapt_every` steps (meshadaptation.c:23-57): the device
state is copied into the leaves, the forest is adapted, re-flattened and
uploaded again.
"""
from __future__ import annotations

import time
from typing import Callable, Optional

import numpy as np

from ._capi import FIELD_SG
from .core import FlatMesh, Forest, Solver


def remesh(forest: Forest, solver: Solver, flags_fn, max_level: int, device_flatten: bool = True) -> dict:
    """Adapt the forest on the device state and rebuild the device mesh.  Returns timings.

    flags_fn is either a host function (forest, SG, S, max_level) -> int8 flags, for
    which the gradients are downloaded, or a dict dict(refine=, coarsen=,
    base_level=, kind=0) naming the density-gradient criterion, which then runs on
    the device (fvg_adapt_flags) so that only the flags cross to the host.

    device_flatten: the mesh is rebuilt on the device from the adapted forest's leaf
    keys (fvg_remesh_*).  Once the device mesh has keys (after its first such rebuild)
    the state is carried over on the device too, and S never crosses to the host;
    the forest's own leaf states are then stale until the next download (restart and
    export download them).  Otherwise the host flattens and the state makes the round
    trip through the forest's leaves."""
    t0 = time.perf_counter()
    carry = device_flatten and solver.has_keys
    if isinstance(flags_fn, dict):
        if not carry:
            forest.set_state(solver.get_state())
        c = flags_fn
        flags = solver.adapt_flags(c["refine"], c["coarsen"], c["base_level"], max_level, c.get("kind", 0))
    else:
        S = solver.get_state()
        forest.set_state(S)
        solver.gradcalc()
        SG = solver.get_field(FIELD_SG)
        flags = flags_fn(forest, SG, S, max_level)
    t1 = time.perf_counter()
    forest.adapt(flags, max_level)
    t2 = time.perf_counter()
    if device_flatten:
        solver.remesh(forest, carry_state=carry)
        if not carry:
            solver.set_state(forest.get_state())
    else:
        flat = FlatMesh(forest.run)
        solver.set_mesh(flat)
        solver.set_state(forest.get_state())
    t3 = time.perf_counter()
    return dict(download_s=t1 - t0, adapt_s=t2 - t1, flatten_upload_s=t3 - t2, nleaves=forest.nleaves,
                device_flatten=bool(device_flatten), carried_on_device=bool(carry))


def save_restart(forest: Forest, solver: Solver, path: str, tstep: int = 0, time: float = 0.0,
                 text: bool = False) -> None:
    """Restart from the device state: S is downloaded into the leaves and written in the
    binary format (exact) or the reference's text layout (restart_save.c)."""
    forest.set_state(solver.get_state())
    (forest.save_text if text else forest.save)(path, tstep, time)


def load_restart(path: str, p, device: int = 0):
    """Forest and solver from a binary restart, the state on the device.
    Returns (forest, solver, tstep, time)."""
    forest, tstep, time = Forest.load(path)
    solver = Solver.from_forest(forest, p, device)
    return forest, solver, tstep, time


def export_tecplot(forest: Forest, solver: Solver, path: str) -> None:
    """Tecplot FEPOINT file of the device state (exportfield_linked_V3.c): S, gradcalc's
    SG and the FP64 cell pressure are downloaded into the leaves and written."""
    solver.gradcalc()
    forest.set_state(solver.get_state())
    forest.set_field(FIELD_SG, solver.get_field(FIELD_SG))
    forest.export_tecplot(path, solver.params.model, solver.pressure())


def evolve(forest: Forest, solver: Solver, steps: int, cfl: float, method: int = 3,
           adapt_every: int = 0, flags_fn: Optional[Callable] = None, max_level: int = 0,
           dt_fixed: Optional[float] = None, log: Optional[Callable] = None,
           device_flatten: bool = True) -> dict:
    t = 0.0
    stats = dict(steps=0, cell_updates=0, remesh=[])
    for n in range(steps):
        if adapt_every and flags_fn is not None and n > 0 and n % adapt_every == 0:
            stats["remesh"].append(remesh(forest, solver, flags_fn, max_level, device_flatten))
            if log:
                log(f"step {n}: remeshed to {forest.nleaves} leaves")
        dt = dt_fixed if dt_fixed is not None else solver.cfl_dt(cfl)
        solver.step(dt, method)
        t += dt
        stats["steps"] += 1
        stats["cell_updates"] += solver.ncell * method
    stats["time"] = t
    return stats
