"""Benchmark of the RHS + RK hot path on one MI355X.

Headline workload (`value`): the compressible Navier-Stokes Taylor-Green vortex on
a uniform 128^3 periodic octree (2,097,152 leaves; Re 1600, Mach 0.1, Prandtl 0.71,
MUSCL-minmod + HLLC, Green-Gauss gradients, viscous fluxes and heat conduction), the
largest single-GPU configuration.  K SSP-RK3 steps are timed with every input resident in HBM; a step
is RK(0) + 3 x (gradcalc, faceinterp, RHS, RK(k)) and the metric is cell-updates/s =
leaves x RK stages / second.  The plain FP64 form gives `value`; the mixed
FP32-flux / FP64-accumulate form is timed on the same state and reported beside it
with its relative L1 error against the plain form after --acc-steps steps.

Beside the headline, on the same run:
  roofline      the RK stage's kernels against the HBM roofline: the reference data
                flow's algorithmic bytes per cell-update (`achieved`, `frac`) and the
                HBM bytes the kernels really move, from the committed rocprofv3
                counters (`traffic`, `traffic_frac`), over the kernels' mean launch
                durations (HIP events on the solver's stream);
  hbm_copy      the rate of a plain streaming copy kernel (the achievable ceiling);
  cpu_baseline  the oracle (the FP64 C restatement of the reference path) on the host
                cores, on a bounded sample of the same mesh;
  adjoint       the isentropic-vortex rollout (496,896 leaves, 50 steps): adjoint /
                forward time with checkpoints every 10 steps and with the rollout
                stored, and the worst relative difference from central differences;
                the mixed form's forward and adjoint timed beside the plain ones;
  extra         the 3D blast with AMR (base 32^3 + 2 levels, evolved 200 steps with
                adaptation every 10), timed the same way on its final mesh.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
KERNELS = ["recon", "grad", "face", "gather_rk", "tile_stage", "tile_grad"]  # fvg_time_steps kernel ids
METRIC = "cell-updates/s, FP64 plain vs FP32-flux mixed @≤1e-6 rel L1; adjoint/fwd ratio"


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tgv", choices=["tgv", "cavitation", "sod", "blast", "vortex"])
    ap.add_argument("--level", type=int, default=None,
                    help="uniform level (tgv/cavitation/sod) or base level (blast/vortex)")
    ap.add_argument("--extra", default="blast", help="comma-separated extra workloads ('' for none)")
    ap.add_argument("--evolve", type=int, default=200, help="untimed AMR steps before timing (blast)")
    ap.add_argument("--refine", type=float, default=0.01, help="blast: split where h|grad rho|/rho exceeds this")
    ap.add_argument("--acc-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--tiling", default="auto", choices=["auto", "on", "off"],
                    help="LDS-tiled stage for interior 8^3 tiles (fvg_set_tiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 6 streaming copies of MIB MiB (calibrates rocprof byte counters)")
    return ap.parse_args()


def log(msg, quiet=False):
    if not quiet:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ref_bytes_per_cell(part, neq, f, w, ns_visc):
    """Algorithmic bytes per cell-update of the reference's data flow, one read and
    one write of each of its data products per stage (RHS()/RK() split into K1
    recon: S -> SF, K2 face flux: SF -> face flux, K3 gather: face flux -> RHS,
    K4 RK: S, SN, RHS, 1/V -> S; K0 gradient S -> SG for Navier-Stokes), with
    f faces per cell, 8-byte states and w-byte face intermediates (8 plain, 4
    mixed).  K1..K4 sum to neq (48 + 6w) + 4 w neq f + 12 f + 56; Navier-Stokes
    adds K0 and the SG reads of K2, 8 neq + 24 + 6 w (neq + 1) in all (56 neq + 72
    plain, 32 neq + 48 mixed).  The fused kernels of this build are charged the
    reference kernels they replace: the face kernel K1 + K2, the gather K3 + K4."""
    k1 = 8 * neq + 24 + 6 * w * neq
    k2 = 12 * f + 3 * w * neq * f
    k3 = 24 + w * neq * f + 8 * neq
    k4 = 32 * neq + 8
    k0 = 0.0
    if ns_visc:
        k0 = 8 * neq + 24 + 3 * w * (neq + 1)       # K0: S + neighbours -> SG
        k2 += 3 * w * (neq + 1)                     # K2 also reads SG (both sides)
    return {"grad": k0, "face": k1 + k2, "gather_rk": k3 + k4,
            "stage": k0 + k1 + k2 + k3 + k4}[part]


def compulsory_bytes(kernel, m, neq, rs, ns_visc, first_stage, heat=False, sfp32=False):
    """Bytes a launch of this build's fused kernels must move at least: each input
    read once, each output written once (MUSCL inside the face kernel).  rs is the
    byte width of the face flux and the gradients (8 plain, 4 mixed).  For the
    LDS-tiled stage (m["tiles"] interior 8^3 tiles) each tile reads its cells and two
    cell layers beyond each face (1,280 cells) once, the viscous rows of its cells and,
    for the first of those layers (384 cells), the rows of that layer's axis, SN, and
    writes the new state (and, mixed, its FP32 copy, the region then being read from
    that copy).  Viscous rows per cell: the six symmetric combinations of the momentum
    gradients (three per axis), plus the three components of grad(p / rho) with heat
    conduction (one per axis).  sfp32: the gradient kernels read the FP32 copy of the
    state (the mixed ideal-gas form)."""
    N, F, Q, Fr = m["N"], m["F"], m["Q"], m["Freg"]
    meta_c = 4 * N + 24 * N + 16 * Q          # cflag + nbr + quad neighbours
    grows = 9 if heat else 6                   # viscous rows per cell
    arows = 4 if heat else 3                   # rows an axis's faces read
    sb = 4 if sfp32 else 8                     # bytes of a state value the gradients read
    if kernel == "grad":                       # S rows -> momentum (and p / rho) gradients
        return sb * (neq if heat else 3) * N + meta_c + grows * rs * N
    if kernel == "tile_grad":                  # tile + one layer beyond each face -> gradients
        return m.get("tiles", 0) * (896 * sb * (neq if heat else 3) + 512 * grows * rs + 32)
    if kernel == "tile_stage":
        nt = m.get("tiles", 0)
        region = 1280 * neq * (4 + 8 * 512 / 1280 if rs == 4 else 8)
        per = region + (512 * grows + 384 * arows) * rs * (1 if ns_visc else 0) + 8 * neq * 512 * 2 + 32
        if rs == 4:
            per += 4 * neq * 512
        return nt * per
    if kernel == "face":
        # S, mesh, face records (L, R, meta), far stencil cells of regular faces, flux
        b = 8 * neq * N + meta_c + 12 * F + 8 * Fr + neq * rs * F
        if ns_visc:
            b += grows * rs * N
        return b
    if kernel == "gather_rk":
        b = 4 * N + 24 * N + neq * rs * F + 8 * neq * N * (1 if first_stage else 2) + 8 * neq * N
        return b
    raise KeyError(kernel)


def build_workload(name, evolve, quiet, level=None, refine=0.01):
    import fv_140823_amd as fv
    from fv_140823_amd import configs, driver
    if name == "blast":
        F, p, meta = configs.blast(level or 5, 2)
        g = fv.Solver.from_forest(F, p, device_flatten=True)
        t0 = time.perf_counter()
        base = (level or 5) + 1
        # the density-gradient criterion on the device (the same flags as
        # configs.density_gradient_flags on the downloaded gradients)
        flags = dict(refine=refine, coarsen=refine / 5, base_level=base, kind=0)
        st = driver.evolve(F, g, evolve, meta["cfl"], 3, meta["adapt_every"], flags, meta["max_level"],
                           log=lambda s: log(s, quiet), device_flatten=True)
        # final adaptation so the timed mesh matches the state
        rm = st["remesh"] + [driver.remesh(F, g, flags, meta["max_level"], True)]
        per = {k: round(1e3 * sum(r[k] for r in rm) / len(rm), 3)
               for k in ("download_s", "adapt_s", "flatten_upload_s")}
        meta = dict(meta, evolved_steps=evolve, evolved_time=st["time"],
                    evolve_wall_s=round(time.perf_counter() - t0, 3), remesh_count=len(rm),
                    remesh_ms_mean={"flags": per["download_s"], "host_adapt": per["adapt_s"],
                                    "flatten": per["flatten_upload_s"]})
        F.set_state(g.get_state())
        g.close()
        return F, p, meta
    builder = getattr(configs, name)
    return builder(level) if level else builder()


def measured_traffic(workload, kernel, grid=None):
    """HBM bytes per launch of `kernel` on `workload` (its name, e.g. "tgv_128^3") from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary profiles/traffic.json,
    already scaled by the copy-kernel calibration; None if that workload was not
    profiled or was profiled on a mesh of another size (launch grid `grid`, threads)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get(workload, {}).get("kernels", {}).get(kernel)
    if k is None or (grid is not None and k.get("grid") != grid):
        return None
    return int(k["read_bytes"] + k["write_bytes"])


def mesh_facts(flat, F):
    A = flat.arrays()
    _, lev = F.leaf_info()
    rnd = lambda n: (n + 255) // 256 * 256
    # the gather's grid is padded to a multiple of 8 blocks of 256 (one range per XCD)
    rnd8 = lambda n: ((n + 255) // 256 + 7) // 8 * 8 * 256
    return dict(N=flat.ncell, F=flat.nface, Q=int(flat.mesh.nquad),
                Freg=int(sum(flat.mesh.nface_reg)), Fb=int((((A["face_meta"] >> 6) & 3) != 0).sum()),
                levels={int(l): int(c) for l, c in zip(*np.unique(lev, return_counts=True))},
                grid_face=sum(rnd(int(n)) for n in flat.mesh.nface_axis), grid_cell=rnd8(flat.ncell),
                grid_grad=rnd(flat.ncell))


TILING = "auto"  # --tiling


def _solver(flat, S0, p):
    import fv_140823_amd as fv
    g = fv.Solver(p)
    g.set_mesh(flat)
    g.set_tiling({"auto": "auto", "on": True, "off": False}[TILING])
    g.set_state(S0)
    return g


def time_form(flat, S0, p, steps, warmup, dt):
    """Wall time of `steps` steps (after `warmup`), and the mean duration of each stage
    kernel from HIP events on the solver's stream (fvg_time_steps)."""
    import torch
    g = _solver(flat, S0, p)
    if warmup:
        g.steps(warmup, dt, 3)
    g.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.steps(steps, dt, 3)
    g.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _, kms = g.time_steps(max(3, steps // 3), dt, 3)
    g.close()
    return wall, kms


def l1_errors(flat, F, S0, p_plain, p_mixed, steps, dt):
    """Relative L1 (volume-weighted) of density and pressure, mixed vs plain, after
    `steps` steps at the fixed dt."""
    from fv_140823_amd.configs import pressure
    out = {}
    for tag, p in (("plain", p_plain), ("mixed", p_mixed)):
        g = _solver(flat, S0, p)
        g.steps(steps, dt, 3)
        out[tag] = g.get_state()
        g.close()
    _, lev = F.leaf_info()
    V = 8.0 ** (-(lev - 1).astype(float))
    a, b = out["mixed"], out["plain"]
    l1 = lambda x, y: float((np.abs(x - y) * V).sum() / (np.abs(y) * V).sum())
    res = dict(density=l1(a[0], b[0]), steps=steps)
    pa, pb = pressure(a, p_plain), pressure(b, p_plain)
    if pa is not None:
        res["pressure"] = l1(pa, pb)
    return res


def _oracle_rate(F, p, dt, budget_s, threads):
    """cell-updates/s of the oracle's RK stages ([faceinterp], [gradcalc], RHS, RK) on
    `threads` threads: one stage to size the sample, then as many stages as fit in
    about budget_s seconds (at least the one already timed)."""
    import oracle
    oracle.set_threads(threads)
    N = F.nleaves
    irk = [1, 2, 3]

    def stage(k):
        if p.ischeme == 2:
            oracle.faceinterp(F, p)
        if p.ns:
            oracle.gradcalc(F, p)
        oracle.rhs(F, p)
        oracle.rk(F, irk[k % 3], dt)

    oracle.rk(F, 0, dt)
    t0 = time.perf_counter()
    stage(0)
    one = time.perf_counter() - t0
    extra = min(300, int(budget_s / max(one, 1e-4))) - 1
    stages, wall = 1, one
    if extra > 0:
        t0 = time.perf_counter()
        for k in range(1, 1 + extra):
            stage(k)
        stages, wall = extra, time.perf_counter() - t0
    oracle.set_threads(1)
    return N * stages / wall, stages, wall


def cpu_baseline(F, p, dt, budget_s, quiet):
    """The oracle (the FP64 C restatement of the reference path) timed on the host
    cores of this machine: all the cores this process may use (OpenMP over leaves,
    like the reference's one MPI rank per core) and, beside it, one core."""
    N = F.nleaves
    S0 = F.get_state()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, 64))
    vt, stt, wt = _oracle_rate(F, p, dt, budget_s * 0.6, threads)
    F.set_state(S0)
    v1, st1, w1 = _oracle_rate(F, p, dt, budget_s * 0.4, 1)
    F.set_state(S0)
    log(f"cpu oracle: {threads} threads {stt} stages in {wt:.2f}s; 1 thread {st1} stages in {w1:.2f}s", quiet)
    return dict(value=round(vt, 1), unit="cell-updates/s", cores=threads, kind="port",
                value_1core=round(v1, 1),
                sample=f"{stt} RK stages ([faceinterp], [gradcalc], RHS, RK) of the oracle on the same "
                       f"{N}-leaf mesh on {threads} threads, {wt:.1f} s ({st1} stages on 1 thread, "
                       f"{w1:.1f} s, for value_1core)")


def hbm_copy(quiet, mib=2048, reps=10):
    """Rate of the library's streaming copy kernel over `mib` MiB (read + written)."""
    import ctypes as C

    import torch

    import fv_140823_amd as fv
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    ms = C.c_double()
    rc = fv.lib().fvg_stream_copy(torch.cuda.current_device(), C.c_void_p(y.data_ptr()),
                                  C.c_void_p(x.data_ptr()), n * 8, reps, C.byref(ms))
    if rc:
        raise RuntimeError(f"fvg_stream_copy returned {rc}")
    ok = bool(torch.equal(x, y))
    del x, y
    gbs = 2 * n * 8 / (ms.value * 1e-3) / 1e9
    log(f"stream copy {mib} MiB: {ms.value:.3f} ms, {gbs:.0f} GB/s", quiet)
    return dict(bytes_per_copy=2 * n * 8, ms=round(ms.value, 4), gbs=round(gbs, 1),
                frac=round(gbs / HBM_PEAK_GBS, 4), copy_exact=ok)


def time_workload(F, p, meta, a, quiet, with_kernels=True):
    """Plain and mixed timings, per-kernel durations, roofline and mixed-vs-plain error
    of one workload."""
    import fv_140823_amd as fv
    flat = fv.FlatMesh(F.run)
    mf = mesh_facts(flat, F)
    log(f"{meta['name']}: {mf['N']} leaves, {mf['F']} faces, {mf['Q']} hanging quads", quiet)
    neq = fv.neq_of(p.model)
    S0 = F.get_state()
    g = _solver(flat, S0, p)
    dt = g.cfl_dt(meta["cfl"])
    mf["tiles"] = g.tile_count if TILING != "off" else 0
    mf["grid_tile"] = 512 * g.tile_count
    g.close()
    log(f"{mf['tiles']} interior 8^3 tiles ({512 * mf['tiles']} leaves) on the LDS-tiled stage", quiet)
    pm = type(p).from_buffer_copy(p)
    pm.precision = fv.FVG_MIXED
    log(f"timing plain FP64: {a.steps} steps after {a.warmup} warm-up", quiet)
    wall_p, k_p = time_form(flat, S0, p, a.steps, a.warmup, dt)
    log("timing mixed", quiet)
    wall_m, k_m = time_form(flat, S0, pm, a.steps, a.warmup, dt)
    cu = mf["N"] * 3 * a.steps
    val_p, val_m = cu / wall_p, cu / wall_m
    log(f"plain {val_p:.4g} cell-upd/s, mixed {val_m:.4g} cell-upd/s", quiet)
    acc = l1_errors(flat, F, S0, p, pm, a.acc_steps, dt)
    log(f"mixed vs plain after {a.acc_steps} steps: {acc}", quiet)
    fpc = mf["F"] / mf["N"]
    fname = "k_face_fused" if p.ischeme == 2 else "k_face"
    # the stage's launches: those that ran (with interior tiles the face kernel and the
    # gather may cover nothing and are not launched)
    stage_kernels = [k for k in ("grad", "tile_grad", "face", "gather_rk", "tile_stage")
                     if k_p[KERNELS.index(k)] > 0]
    names = {"grad": ("k_grad", mf["grid_grad"]), "face": (fname, mf["grid_face"]),
             "gather_rk": ("k_gather_rk", mf["grid_cell"]), "tile_stage": ("k_tile_stage", mf["grid_tile"]),
             "tile_grad": ("k_tile_grad", mf["grid_tile"])}
    heat = bool(p.ns and p.prandtl > 0)
    # the roofline of one RK stage of the plain form: its launches' mean durations
    t_stage = sum(k_p[KERNELS.index(k)] for k in stage_kernels)
    B = ref_bytes_per_cell("stage", neq, fpc, 8, p.ns)
    achieved = B * mf["N"] / (t_stage * 1e-3) / 1e9
    wkey = meta["name"] + ("" if TILING == "auto" else f" tiling={TILING}")
    tr = [measured_traffic(wkey, *names[k]) for k in stage_kernels]
    traffic = sum(tr) if all(t is not None for t in tr) else None
    comp = sum(compulsory_bytes(k, mf, neq, 8, p.ns, False, heat) for k in stage_kernels)
    roof = dict(bound="hbm", kernel="RK stage: " + " + ".join(names[k][0] for k in stage_kernels),
                achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_gbs=round(traffic / (t_stage * 1e-3) / 1e9, 1) if traffic else None,
                traffic_frac=round(traffic / (t_stage * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                alg_bytes_per_launch=int(B * mf["N"]), alg_bytes_per_cell_update=round(B, 1),
                avg_launch_ms=round(t_stage, 5), compulsory_bytes_per_launch=int(comp),
                compulsory_frac=round(comp / (t_stage * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                faces_per_cell=round(fpc, 4))
    stage_roof = {}
    for tag, val, w in (("plain", val_p, 8), ("mixed", val_m, 4)):
        Bw = ref_bytes_per_cell("stage", neq, fpc, w, p.ns)
        stage_roof[tag] = dict(bytes_per_cell_update=round(Bw, 1), achieved_gbs=round(Bw * val / 1e9, 1),
                               frac=round(Bw * val / 1e9 / HBM_PEAK_GBS, 4))
    kern = {}
    for i, k in enumerate(KERNELS):
        if (k_p[i] <= 0 and k_m[i] <= 0) or k not in names:
            continue
        e = dict(name=names[k][0], plain_ms=round(k_p[i], 5), mixed_ms=round(k_m[i], 5))
        tb = measured_traffic(wkey, *names[k])
        if tb is not None:
            e["plain_traffic_bytes"] = tb
            e["plain_traffic_gbs"] = round(tb / (k_p[i] * 1e-3) / 1e9, 1)
        for tag, ms, rs in (("plain", k_p[i], 8), ("mixed", k_m[i], 4)):
            if ms:
                cb = compulsory_bytes(k, mf, neq, rs, p.ns, False, heat,
                                      sfp32=rs == 4 and p.model == 1 and p.ischeme == 2)
                e[f"{tag}_compulsory_bytes"] = int(cb)
                e[f"{tag}_compulsory_gbs"] = round(cb / (ms * 1e-3) / 1e9, 1)
        kern[k] = e
    cfg = {"workload": meta["name"], "leaves": mf["N"], "faces": mf["F"], "hanging_quads": mf["Q"],
           "regular_faces": mf["Freg"], "levels": mf["levels"], "interior_tiles": mf["tiles"],
           "tiling": TILING, "tiled_stage_ran": bool(k_p[KERNELS.index("tile_stage")] > 0),
           "model": {1: f"ideal gas gamma {p.gamma:g}", 2: "barotropic (model 2)",
                     3: "two-phase barotropic (model 3)"}[p.model] + (" Navier-Stokes" if p.ns else ""),
           "scheme": ("MUSCL-" + ("minmod" if p.limiter == 0 else "van Albada") if p.ischeme == 2
                      else "first order") + " + HLLC(Davis), SSP-RK3", "dt": dt,
           **{k: v for k, v in meta.items() if k != "name"}}
    if p.ns:
        cfg.update(mu=p.mu, viscous_work=bool(p.visc_work), prandtl=p.prandtl or None)
    out = dict(value=round(val_p, 1), ms_per_step=round(wall_p / a.steps * 1e3, 4), config=cfg,
               mixed={"value": round(val_m, 1), "ms_per_step": round(wall_m / a.steps * 1e3, 4),
                      "speedup_vs_plain": round(val_m / val_p, 3), "l1_rel_err_vs_plain": acc},
               roofline=roof, stage_roofline=stage_roof)
    if with_kernels:
        out["kernels"] = kern
    flat.close()
    return out, dt


def adjoint_block(quiet):
    """The isentropic-vortex rollout: 64^3 base + two-level patch (496,896 leaves),
    van Albada MUSCL, 50 SSP-RK3 steps, J = time-integrated probe pressure.  Forward =
    the plain rollout (fvg_steps); adjoint = the reverse sweep (fvg_adj_backward into
    HBM, recomputing checkpointed segments), after a forward that keeps checkpoints
    every 10 steps, or every stage state (spacing chosen from free memory).  The mixed
    form's rollout and reverse sweep (checkpoints every 10 steps) are timed beside them,
    with the relative L2 difference of its gradient from the plain one."""
    import torch

    import fv_140823_amd as fv
    from fv_140823_amd import configs
    F, p, meta = configs.vortex()
    xc, _ = F.leaf_info()
    probe = ((np.abs(xc[:, 0] - 6.5) < 1.0) & (np.abs(xc[:, 1] - 6.5) < 1.0)).astype(np.uint8)
    g = fv.Solver.from_forest(F, p)
    g.adj_set_probe(probe)
    S0 = F.get_state()
    n = meta["steps"]
    dt = g.cfl_dt(meta["cfl"])
    g.adj_forward(2, 1, dt)  # warm-up
    g.set_state(S0)
    g.steps(2, dt, 3)
    fwd = []
    for _ in range(3):
        g.set_state(S0)
        g.sync()
        t0 = time.perf_counter()
        g.steps(n, dt, 3)
        g.sync()
        fwd.append(time.perf_counter() - t0)
    t_fwd = min(fwd)
    lam_dev = torch.empty((g.neq, g.ncell), dtype=torch.float64, device="cuda")
    res, grads = {}, {}
    for ck, key in ((meta["ckpt"], "ckpt10"), (0, "store")):
        best = None
        for _ in range(2):
            g.set_state(S0)
            g.sync()
            t0 = time.perf_counter()
            g.adj_forward(n, ck, dt)
            t_f = time.perf_counter() - t0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.adj_backward_device(lam_dev.data_ptr())
            t_b = time.perf_counter() - t0
            best = (t_f, t_b) if best is None or t_b < best[1] else best
        grads[key] = lam_dev.cpu().numpy()
        res[key] = dict(spacing=g.adj_checkpoint_spacing, forward_with_checkpoints_ms=round(best[0] * 1e3, 3),
                        adjoint_ms=round(best[1] * 1e3, 3), ratio=round(best[1] / t_fwd, 3))
    lam = grads["ckpt10"]
    eps = 1e-6
    worst = 0.0
    for k in range(3):
        r = np.random.default_rng(500 + k)
        c = np.array([r.uniform(5.5, 7.5), r.uniform(5.5, 7.5), r.uniform(0.0, 10.0)])
        w = r.uniform(0.5, 1.5)
        b = np.exp(-((xc - c) ** 2).sum(axis=1) / w ** 2)
        v = np.zeros(S0.shape)
        v[:4] = b * (r.choice([-1.0, 1.0], 4) * r.uniform(0.5, 1.5, 4))[:, None]
        Js = []
        for sgn in (1.0, -1.0):
            g.set_state(S0 + sgn * eps * v)
            g.adj_forward(n, meta["ckpt"], dt)
            Js.append(g.adj_objective())
        (ph, pl), (mh, ml) = Js
        fd = ((ph - mh) + (pl - ml)) / (2 * eps)
        worst = max(worst, abs(float((lam * v).sum()) - fd) / abs(fd))
    g.close()
    # the mixed form (FP32 face states, fluxes and their reverse; FP64 accumulation) on the
    # same rollout, checkpoints every 10 steps, timed against the plain form
    pm = type(p).from_buffer_copy(p)
    pm.precision = fv.FVG_MIXED
    gm = fv.Solver.from_forest(F, pm)
    gm.adj_set_probe(probe)
    gm.set_state(S0)
    gm.adj_forward(2, 1, dt)  # warm-up
    fwd_m, adj_m = [], []
    for _ in range(2):
        gm.set_state(S0)
        gm.sync()
        t0 = time.perf_counter()
        gm.steps(n, dt, 3)
        gm.sync()
        fwd_m.append(time.perf_counter() - t0)
        gm.set_state(S0)
        gm.adj_forward(n, meta["ckpt"], dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gm.adj_backward_device(lam_dev.data_ptr())
        adj_m.append(time.perf_counter() - t0)
    lam_m = lam_dev.cpu().numpy()
    gm.close()
    mixed = dict(forward_ms=round(min(fwd_m) * 1e3, 3), adjoint_ms_ckpt10=round(min(adj_m) * 1e3, 3),
                 forward_speedup_vs_plain=round(t_fwd / min(fwd_m), 3),
                 adjoint_speedup_vs_plain=round(res["ckpt10"]["adjoint_ms"] / (min(adj_m) * 1e3), 3),
                 grad_rel_l2_vs_plain=float(f"{np.linalg.norm(lam_m[:4] - lam[:4]) / np.linalg.norm(lam[:4]):.3e}"))
    out = dict(workload=meta["name"], leaves=F.nleaves, steps=n, dt=dt, forward_ms=round(t_fwd * 1e3, 3),
               ratio_ckpt10=res["ckpt10"]["ratio"], ratio_store=res["store"]["ratio"],
               ckpt10=res["ckpt10"], store=res["store"],
               store_grad_equal_ckpt10=bool(np.array_equal(grads["store"], grads["ckpt10"])),
               fd_worst=float(f"{worst:.3e}"), fd_directions=3, fd_eps=eps, mixed=mixed)
    log(f"adjoint: {out}", quiet)
    return out


def main():
    a = _args()
    global TILING
    TILING = a.tiling
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the path has no "
              "multi-GPU halo exchange", file=sys.stderr)
        sys.exit(2)
    import torch

    import fv_140823_amd as fv
    if a.copy_calibration:
        import ctypes as C
        n = a.copy_calibration * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        ms = C.c_double()
        rc = fv.lib().fvg_stream_copy(torch.cuda.current_device(), C.c_void_p(y.data_ptr()),
                                      C.c_void_p(x.data_ptr()), n * 8, 5, C.byref(ms))
        print(json.dumps({"copy_calibration_mib": a.copy_calibration, "copies": 6, "rc": rc,
                          "bytes_read": n * 8, "bytes_written": n * 8, "ms": ms.value}))
        return
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    t_start = time.perf_counter()
    log(f"building workload {a.workload}", a.quiet)
    F, p, meta = build_workload(a.workload, a.evolve, a.quiet, a.level, a.refine)
    head, dt = time_workload(F, p, meta, a, a.quiet)
    copy = hbm_copy(a.quiet)
    extra = {}
    for name in [x for x in a.extra.split(",") if x and x != a.workload]:
        log(f"extra workload {name}", a.quiet)
        Fx, px, mx = build_workload(name, a.evolve, a.quiet, None, a.refine)
        extra[name], _ = time_workload(Fx, px, mx, a, a.quiet, with_kernels=False)
        Fx.close()
    adj = None if a.no_adjoint else adjoint_block(a.quiet)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(F, p, dt, a.cpu_budget, a.quiet)
    res = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": head["config"],
        "mixed": head["mixed"],
        "kernels": head["kernels"],
        "roofline": head["roofline"],
        "stage_roofline": head["stage_roofline"],
        "hbm_copy": copy,
        "cpu_baseline": cpu,
        "adjoint": adj,
        "extra": extra,
        "wall_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
