"""Benchmark of the RHS + RK hot path on one MI355X.

Workload (default): the 3D spherical blast with AMR (ideal gas, gamma 1.4, unit
cube, base 32^3 refined two levels by a density-gradient criterion with 2:1
balance, re-adapted every 10 steps).  The blast is evolved with host AMR for
--evolve steps (untimed) so the mesh has the size and hanging-face mix of the
running case; then K SSP-RK3 steps are timed on that mesh with every input
resident in HBM.  A step = RK(0) + 3 x (faceinterp, RHS, RK(k)); the metric is
cell-updates/s = leaves x RK stages / second.

The plain FP64 form gives `value`; the mixed FP32-flux / FP64-accumulate form is
timed on the same state and reported beside it with its relative L1 error
against the plain form after --acc-steps steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
KERNELS = ["recon", "grad", "face", "gather_rk"]


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="blast", choices=["blast", "sod", "cavitation", "tgv", "vortex"])
    ap.add_argument("--evolve", type=int, default=200, help="untimed AMR steps before timing (blast)")
    ap.add_argument("--host-flatten", action="store_true",
                    help="blast evolution: flatten on the host after each adaptation (default: on the device)")
    ap.add_argument("--level", type=int, default=None, help="uniform level (sod/cavitation/tgv) or base level (blast/vortex)")
    ap.add_argument("--refine", type=float, default=0.01, help="blast: split where h|grad rho|/rho exceeds this")
    ap.add_argument("--acc-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (calibrates rocprof byte counters)")
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


def compulsory_bytes(kernel, m, neq, rs, ns_visc, first_stage):
    """Bytes a launch of this build's fused kernels must move at least: each
    input read once, each output written once (MUSCL inside the face kernel)."""
    N, F, Q = m["N"], m["F"], m["Q"]
    meta_c = 4 * N + 24 * N + 16 * Q          # cflag + nbr + quad neighbours
    if kernel == "grad":
        return 8 * 3 * N + meta_c + 9 * rs * N
    if kernel == "face":
        b = 8 * neq * N + meta_c + 12 * F + neq * rs * F
        if ns_visc:
            b += 9 * rs * N
        return b
    if kernel == "gather_rk":
        b = 4 * N + 24 * N + neq * rs * F + 8 * neq * N * (1 if first_stage else 2) + 8 * neq * N
        if first_stage:
            b += 8 * neq * N
        return b
    raise KeyError(kernel)


def build_workload(name, evolve, quiet, level=None, refine=0.01, device_flatten=True):
    import fv_140823_amd as fv
    from fv_140823_amd import configs, driver
    if name == "blast":
        F, p, meta = configs.blast(level or 5, 2)
        g = fv.Solver.from_forest(F, p, device_flatten=device_flatten)
        t0 = time.perf_counter()
        base = (level or 5) + 1
        # the density-gradient criterion on the device (the same flags as
        # configs.density_gradient_flags on the downloaded gradients)
        flags = dict(refine=refine, coarsen=refine / 5, base_level=base, kind=0)
        st = driver.evolve(F, g, evolve, meta["cfl"], 3, meta["adapt_every"], flags, meta["max_level"],
                           log=lambda s: log(s, quiet), device_flatten=device_flatten)
        # final adaptation so the timed mesh matches the state
        rm = st["remesh"] + [driver.remesh(F, g, flags, meta["max_level"], device_flatten)]
        per = {k: round(1e3 * sum(r[k] for r in rm) / len(rm), 3)
               for k in ("download_s", "adapt_s", "flatten_upload_s")}
        meta = dict(meta, evolved_steps=evolve, evolved_time=st["time"],
                    evolve_wall_s=time.perf_counter() - t0, remesh_count=len(rm),
                    remesh_ms_mean={"flags": per["download_s"], "host_adapt": per["adapt_s"],
                                    "flatten": per["flatten_upload_s"]},
                    flatten="device" if device_flatten else "host")
        F.set_state(g.get_state())
        g.close()
        return F, p, meta
    builder = getattr(configs, name)
    F, p, meta = builder(level) if level else builder()
    return F, p, meta


def measured_traffic(workload, kernel, grid=None):
    """HBM bytes per launch of `kernel` on `workload` (the workload name, e.g.
    "blast_32^3+2") from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    profiles/traffic.json, already scaled by the copy-kernel calibration; None if
    that workload was not profiled or was profiled on a mesh of another size
    (launch grid `grid`, threads)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get(workload, {}).get("kernels", {}).get(kernel)
    if k is None or (grid is not None and k.get("grid") != grid):
        return None
    return int(k["read_bytes"] + k["write_bytes"])


def mesh_facts(F):
    import fv_140823_amd as fv
    fl = fv.FlatMesh(F.run)
    A = fl.arrays()
    Fb = int(((A["face_meta"] >> 6) & 3 != 0).sum())
    _, lev = F.leaf_info()
    rnd = lambda n: (n + 255) // 256 * 256
    # the gather's grid is padded to a multiple of 8 blocks of 256 (one range per XCD)
    rnd8 = lambda n: ((n + 255) // 256 + 7) // 8 * 8 * 256
    return dict(N=fl.ncell, F=fl.nface, Q=int(fl.mesh.nquad), Fb=Fb,
                levels={int(l): int(c) for l, c in zip(*np.unique(lev, return_counts=True))},
                grid_face=sum(rnd(int(n)) for n in fl.mesh.nface_axis), grid_cell=rnd8(fl.ncell))


def time_form(F, p, steps, warmup, dt):
    import torch

    import fv_140823_amd as fv
    g = fv.Solver.from_forest(F, p)
    if warmup:
        g.steps(warmup, dt, 3)
    g.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.steps(steps, dt, 3)
    g.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel durations from HIP events on the context's stream, around batches
    # of back-to-back launches of each stage kernel (fvg_time_steps)
    ev_total, kms = g.time_steps(max(3, steps // 3), dt, 3)
    g.close()
    return wall, kms


def l1_rel_density(F, p_plain, p_mixed, steps, dt):
    """Relative L1 (volume-weighted) of density and pressure, mixed vs plain, after
    `steps` steps at the fixed dt."""
    import fv_140823_amd as fv
    from fv_140823_amd.configs import pressure
    out = {}
    for tag, p in (("plain", p_plain), ("mixed", p_mixed)):
        g = fv.Solver.from_forest(F, p)
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
    """cell-updates/s of the oracle's RK stages (faceinterp, [gradcalc], RHS, RK)
    on `threads` threads, over about budget_s seconds of stages."""
    import oracle
    oracle.set_threads(threads)
    N = F.nleaves
    t0 = time.perf_counter()
    oracle.faceinterp(F, p)
    oracle.rhs(F, p)
    one = time.perf_counter() - t0
    stages = max(3, min(300, int(budget_s / max(one, 1e-4))))
    oracle.rk(F, 0, dt)
    t0 = time.perf_counter()
    irk = [1, 2, 3]
    for k in range(stages):
        if p.ischeme == 2:
            oracle.faceinterp(F, p)
        if p.ns:
            oracle.gradcalc(F, p)
        oracle.rhs(F, p)
        oracle.rk(F, irk[k % 3], dt)
    wall = time.perf_counter() - t0
    oracle.set_threads(1)
    return N * stages / wall, stages, wall


def cpu_baseline(F, p, dt, budget_s, quiet):
    """The oracle (the FP64 C restatement of the reference path) timed on the host
    cores of this machine: all the cores this process may use (OpenMP over leaves,
    like the reference's one MPI rank per core) and, beside it, one core."""
    N = F.nleaves
    S0 = F.get_state()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    v1, st1, w1 = _oracle_rate(F, p, dt, budget_s / 2, 1)
    F.set_state(S0)
    vt, stt, wt = _oracle_rate(F, p, dt, budget_s / 2, threads)
    F.set_state(S0)
    log(f"cpu oracle: 1 thread {st1} stages in {w1:.2f}s; {threads} threads {stt} stages in {wt:.2f}s", quiet)
    return dict(value=vt, unit="cell-updates/s", cores=threads, kind="port",
                value_1core=v1,
                sample=f"{stt} RK stages (faceinterp+RHS+RK) of the oracle on the same {N}-leaf mesh on "
                       f"{threads} threads ({st1} stages on 1 thread for value_1core)")


def main():
    a = _args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the path has no "
              "multi-GPU halo exchange; see the sharding notes in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch

    import fv_140823_amd as fv
    if a.copy_calibration:
        n = a.copy_calibration * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": a.copy_calibration, "bytes_read": n * 8,
                          "bytes_written": n * 8}))
        return
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    log(f"building workload {a.workload}", a.quiet)
    F, p, meta = build_workload(a.workload, a.evolve, a.quiet, a.level, a.refine, not a.host_flatten)
    mf = mesh_facts(F)
    log(f"mesh: {mf}", a.quiet)
    neq = fv.neq_of(p.model)
    g = fv.Solver.from_forest(F, p)
    dt = g.cfl_dt(meta["cfl"])
    g.close()
    pm = type(p).from_buffer_copy(p)
    pm.precision = fv.FVG_MIXED
    log(f"timing plain FP64: {a.steps} steps after {a.warmup} warm-up", a.quiet)
    wall_p, k_p = time_form(F, p, a.steps, a.warmup, dt)
    log(f"timing mixed: {a.steps} steps", a.quiet)
    wall_m, k_m = time_form(F, pm, a.steps, a.warmup, dt)
    stages = 3
    cu = mf["N"] * stages * a.steps
    val_p, val_m = cu / wall_p, cu / wall_m
    log(f"plain {val_p:.4g} cell-upd/s, mixed {val_m:.4g} cell-upd/s", a.quiet)
    # Roofline of one RK stage of the plain form (its two launches, face + gather)
    # against the reference data flow's bytes per cell-update (B, ref_bytes_per_cell):
    # achieved = B x leaves / (mean face + mean gather duration, HIP events);
    # traffic = the two launches' HBM bytes from the committed rocprof counters.
    fpc = mf["F"] / mf["N"]
    stage_kernels = ["grad", "face", "gather_rk"] if p.ns else ["face", "gather_rk"]
    t_stage = sum(k_p[KERNELS.index(k)] for k in stage_kernels)
    B = ref_bytes_per_cell("stage", neq, fpc, 8, p.ns)
    achieved = B * mf["N"] / (t_stage * 1e-3) / 1e9
    fname = "k_face_fused" if p.ischeme == 2 else "k_face"
    tf = measured_traffic(meta["name"], fname, mf["grid_face"])
    tg = measured_traffic(meta["name"], "k_gather_rk", mf["grid_cell"])
    traffic = tf + tg if (tf is not None and tg is not None and not p.ns) else None
    comp = sum(compulsory_bytes(k, mf, neq, 8, p.ns, False) for k in stage_kernels)
    roof = dict(bound="hbm", kernel=f"RK stage: {fname} + k_gather_rk" + (" + k_grad" if p.ns else ""),
                achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_gbs=round(traffic / (t_stage * 1e-3) / 1e9, 1) if traffic else None,
                alg_bytes_per_launch=int(B * mf["N"]), alg_bytes_per_cell_update=round(B, 1),
                avg_launch_ms=round(t_stage, 5), compulsory_bytes_per_launch=int(comp),
                faces_per_cell=round(fpc, 4))
    # the same against the wall clock of the timed steps, plain and mixed
    stage_roof = {}
    for tag, val, w in (("plain", val_p, 8), ("mixed", val_m, 4)):
        Bw = ref_bytes_per_cell("stage", neq, fpc, w, p.ns)
        stage_roof[tag] = dict(bytes_per_cell_update=round(Bw, 1), achieved_gbs=round(Bw * val / 1e9, 1),
                               frac=round(Bw * val / 1e9 / HBM_PEAK_GBS, 4))
    # per kernel: mean durations, the actual HBM rate (rocprof bytes / duration) and
    # the rate on the kernel's compulsory bytes
    kern = {}
    for i in range(4):
        if k_p[i] <= 0:
            continue
        k = KERNELS[i]
        e = dict(plain_ms=round(k_p[i], 5), mixed_ms=round(k_m[i], 5))
        kn = {"face": (fname, mf["grid_face"]), "gather_rk": ("k_gather_rk", mf["grid_cell"])}.get(k)
        tb = measured_traffic(meta["name"], *kn) if kn else None
        if tb is not None:
            e["plain_traffic_bytes"] = tb
            e["plain_traffic_gbs"] = round(tb / (k_p[i] * 1e-3) / 1e9, 1)
        if k != "recon":
            for tag, ms, rs in (("plain", k_p[i], 8), ("mixed", k_m[i], 4)):
                if ms:
                    e[f"{tag}_compulsory_gbs"] = round(compulsory_bytes(k, mf, neq, rs, p.ns, False) / (ms * 1e-3) / 1e9, 1)
        kern[k] = e
    acc = l1_rel_density(F, p, pm, a.acc_steps, dt)
    log(f"mixed vs plain after {a.acc_steps} steps: {acc}", a.quiet)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(F, p, dt, a.cpu_budget, a.quiet)
    res = {
        "metric": "cell-updates/s (RHS+RK, per RK stage), FP64 plain; FP32-flux mixed beside it",
        "value": round(val_p, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall_p / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": meta["name"], "leaves": mf["N"], "faces": mf["F"],
                   "hanging_quads": mf["Q"], "levels": mf["levels"],
                   "model": {1: f"ideal gas gamma {p.gamma:g}", 2: "barotropic (model 2)",
                             3: "two-phase barotropic (model 3)"}[p.model] + (" Navier-Stokes" if p.ns else ""),
                   "scheme": ("MUSCL-" + ("minmod" if p.limiter == 0 else "van Albada") if p.ischeme == 2
                              else "first order") + " + HLLC(Davis), SSP-RK3", "dt": dt,
                   **{k: v for k, v in meta.items() if k != "name"}},
        "mixed": {"value": round(val_m, 1), "ms_per_step": round(wall_m / a.steps * 1e3, 4),
                  "speedup_vs_plain": round(val_m / val_p, 3), "l1_rel_err_vs_plain": acc},
        "kernels": kern,
        "roofline": roof,
        "stage_roofline": stage_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
