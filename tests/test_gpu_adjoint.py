"""GPU checks of the discrete adjoint of RHS + RK rollouts (fvg_adj_*), Euler and
Navier-Stokes:
  - <dJ/dS0, v> against the forward-mode tangent (fvg_tlm) of the same rollout,
    two independent derivative codes (hand-written reverse vs dual numbers);
  - against central finite differences of J (FP64, eps = 1e-6 relative);
  - the torch.autograd.Function wrapper returns the same gradient;
  - checkpoint spacing does not change the gradient (given, chosen from the free device
    memory, or wider than the rollout).
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from fv_140823_amd.configs import vortex_state
from helpers import amr, state, uniform

pytestmark = pytest.mark.gpu


def _vortex(level=4, patch=True):
    box = 10.0
    F = fv.Forest((1, 1, 1), (0, 0, 0), (box, box, box), level, (0,) * 6, (1, 1, 1), 5)
    if patch:
        xc, _ = F.leaf_info()
        r = np.hypot(xc[:, 0] - 5, xc[:, 1] - 5)
        F.adapt((r < 2.0).astype(np.int8), level + 2)
    xc, _ = F.leaf_info()
    F.set_state(vortex_state(xc, box))
    probe = ((np.abs(xc[:, 0] - 6.0) < 1.5) & (np.abs(xc[:, 1] - 5.0) < 1.5)).astype(np.uint8)
    return F, probe


def _solver(F, probe, **kw):
    p = fv.params(1, **kw)
    g = fv.Solver.from_forest(F, p)
    g.adj_set_probe(probe)
    return g


@pytest.mark.parametrize("limiter,ischeme,patch", [(1, 2, True), (0, 2, False), (1, 1, True)])
def test_adjoint_matches_tangent(limiter, ischeme, patch):
    print(f"case adjoint vs tangent limiter={limiter} ischeme={ischeme} patch={patch}", flush=True)
    F, probe = _vortex(4, patch)
    g = _solver(F, probe, limiter=limiter, ischeme=ischeme)
    S0 = F.get_state()
    dt = 0.5 * g.cfl_dt(0.4)
    J = g.adj_forward(6, 2, dt)
    lam = g.adj_backward()
    rng = np.random.default_rng(0)
    for k in range(3):
        v = rng.standard_normal(S0.shape) * np.abs(S0).max(axis=1, keepdims=True) * 1e-2
        g.set_state(S0)
        J2, dJ = g.tlm(v, 6, dt)
        assert abs(J2 - J) <= 1e-13 * abs(J)
        adj = float((lam * v).sum())
        print(f"  dir {k}: adjoint {adj:.15e} tangent {dJ:.15e}", flush=True)
        assert abs(adj - dJ) <= 1e-10 * max(abs(dJ), 1e-300), (adj, dJ)


def _smooth_direction(k, xc, shape):
    """Direction k (seed = k): a smooth perturbation of density and momentum, a
    Gaussian bump of random width centred at a random point of the probe box, with
    amplitudes of random sign and magnitude 0.5..1.5 (energy held fixed)."""
    rng = np.random.default_rng(k)
    c = np.array([rng.uniform(4.5, 7.5), rng.uniform(3.5, 6.5), rng.uniform(0.0, 10.0)])
    w = rng.uniform(0.5, 2.0)
    b = np.exp(-((xc - c) ** 2).sum(axis=1) / w ** 2)
    amp = rng.choice([-1.0, 1.0], 4) * rng.uniform(0.5, 1.5, 4)
    v = np.zeros(shape)
    v[:4] = b * amp[:, None]
    return v


def test_adjoint_matches_central_fd():
    """20 seeded random directions, central differences with eps = 1e-6 (FP64).
    The directions are smooth bumps: white-noise directions perturb every face
    and cross the kinks of the Davis wave speeds (max/min) and of sonic faces,
    where the discrete map is not differentiable; the adjoint is checked there
    against the tangent instead (test_adjoint_matches_tangent).
    J is read as the double-double pair (hi, lo) that the library accumulates:
    rounded to one double, J(S0 +- eps v) would carry about 1e-16 |J| each, and
    the difference quotient an error of about 1e-16 |J| / eps (the `floor`
    printed), which is 1e-6 of a derivative only 1e6 times larger."""
    print("case adjoint vs central FD", flush=True)
    F, probe = _vortex(4, True)
    g = _solver(F, probe, limiter=1)
    S0 = F.get_state()
    xc, _ = F.leaf_info()
    dt = 0.5 * g.cfl_dt(0.4)
    g.set_state(S0)
    J0 = g.adj_forward(8, 4, dt)
    lam = g.adj_backward()
    eps = 1e-6
    floor = 1e-16 * abs(J0) / eps
    worst = 0.0

    def J(S):
        g.set_state(S)
        g.adj_forward(8, 4, dt)
        return g.adj_objective()

    for k in range(20):
        v = _smooth_direction(k, xc, S0.shape)
        (ph, pl), (mh, ml) = J(S0 + eps * v), J(S0 - eps * v)
        fd = ((ph - mh) + (pl - ml)) / (2 * eps)
        adj = float((lam * v).sum())
        rel = abs(adj - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"  dir {k}: adjoint {adj:.12e} fd {fd:.12e} rel {rel:.2e} (fd/floor {abs(fd) / floor:.1e})",
              flush=True)
        assert rel <= 1e-6, (adj, fd)
    print("worst relative difference", worst, flush=True)


def test_checkpoint_spacing_invariant():
    print("case checkpoint spacing", flush=True)
    F, probe = _vortex(4, True)
    g = _solver(F, probe, limiter=1)
    S0 = F.get_state()
    dt = 0.5 * g.cfl_dt(0.4)
    out = []
    # 0: the spacing chosen from the free device memory, the whole rollout at this size
    for ck in (1, 3, 9, 0, 20):
        g.set_state(S0)
        out.append((g.adj_forward(9, ck, dt), g.adj_backward()))
        assert g.adj_checkpoint_spacing == (9 if ck in (0, 20) else ck)
    for J, lam in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(lam, out[0][1])


def test_adjoint_transmissive_amr_boundary():
    """2:1 mesh with outflow sides: ghost handling and quad gathers in the reverse."""
    print("case adjoint amr outflow", flush=True)
    F = amr(3, 5)
    xc, _ = F.leaf_info()
    F.set_state(state(1, xc, seed=21, jump=False))
    probe = (xc[:, 0] > 0.6).astype(np.uint8)
    g = _solver(F, probe, limiter=1)
    S0 = F.get_state()
    dt = g.cfl_dt(0.3)
    J = g.adj_forward(4, 2, dt)
    lam = g.adj_backward()
    v = np.random.default_rng(5).standard_normal(S0.shape) * 1e-3
    g.set_state(S0)
    _, dJ = g.tlm(v, 4, dt)
    adj = float((lam * v).sum())
    assert abs(adj - dJ) <= 1e-10 * abs(dJ), (adj, dJ)


def test_torch_autograd_wrapper():
    print("case torch autograd", flush=True)
    import torch

    from fv_140823_amd.adjoint import rollout_objective
    F, probe = _vortex(4, True)
    g = _solver(F, probe, limiter=1)
    S0 = F.get_state()
    dt = 0.5 * g.cfl_dt(0.4)
    g.set_state(S0)
    g.adj_forward(5, 5, dt)
    lam = g.adj_backward()
    dev = torch.device("cuda:0")
    rho = torch.tensor(S0[0], dtype=torch.float64, device=dev, requires_grad=True)
    mom = torch.tensor(S0[1:4], dtype=torch.float64, device=dev, requires_grad=True)
    E = torch.tensor(S0[4], dtype=torch.float64, device=dev)
    J = rollout_objective(rho, mom, E, g, 5, 5, dt)
    J.backward()
    assert np.allclose(rho.grad.cpu().numpy(), lam[0], rtol=0, atol=1e-14 * np.abs(lam).max())
    assert np.allclose(mom.grad.cpu().numpy(), lam[1:4], rtol=0, atol=1e-14 * np.abs(lam).max())


def test_mixed_adjoint_close_to_plain():
    print("case mixed adjoint", flush=True)
    F, probe = _vortex(4, True)
    gp = _solver(F, probe, limiter=1)
    gm = _solver(F, probe, limiter=1, precision="mixed")
    dt = 0.5 * gp.cfl_dt(0.4)
    Jp = gp.adj_forward(6, 3, dt)
    Jm = gm.adj_forward(6, 3, dt)
    lp, lm = gp.adj_backward(), gm.adj_backward()
    assert abs(Jm - Jp) <= 1e-6 * abs(Jp)
    e = np.abs(lm - lp).max() / np.abs(lp).max()
    print("mixed adjoint rel Linf", e, flush=True)
    assert e < 1e-4


NS_CASES = [
    # (mesh, limiter, ischeme, visc_work)
    ("vortex_patch", 1, 2, 1),   # periodic, 2:1 hanging faces, van Albada
    ("amr_outflow", 0, 2, 1),    # outflow sides: ghost terms of the gradient transpose
    ("vortex_patch", 1, 1, 0),   # first order, no viscous work
]


def _ns_problem(mesh):
    if mesh == "vortex_patch":
        F, probe = _vortex(4, True)
    else:
        F = amr(3, 5)
        xc, _ = F.leaf_info()
        F.set_state(state(1, xc, seed=23, jump=False))
        probe = (xc[:, 0] > 0.6).astype(np.uint8)
    return F, probe


@pytest.mark.parametrize("mesh,limiter,ischeme,work", NS_CASES)
def test_ns_adjoint_matches_tangent(mesh, limiter, ischeme, work):
    """Navier-Stokes rollouts: the adjoint (HLLC + viscous VJP, gradient transpose)
    against the dual-number tangent of the same rollout, three random directions.
    mu is raised so the viscous terms carry a sizeable share of the flux."""
    print(f"case NS adjoint vs tangent mesh={mesh} limiter={limiter} ischeme={ischeme} work={work}",
          flush=True)
    F, probe = _ns_problem(mesh)
    mu = 0.05 if mesh == "vortex_patch" else 2e-3
    g = _solver(F, probe, limiter=limiter, ischeme=ischeme, ns=1, mu=mu, visc_work=work)
    S0 = F.get_state()
    dt = 0.3 * g.cfl_dt(0.4)
    J = g.adj_forward(5, 2, dt)
    lam = g.adj_backward()
    rng = np.random.default_rng(2)
    for k in range(3):
        v = rng.standard_normal(S0.shape) * np.abs(S0).max(axis=1, keepdims=True) * 1e-2
        g.set_state(S0)
        J2, dJ = g.tlm(v, 5, dt)
        assert abs(J2 - J) <= 1e-13 * abs(J)
        adj = float((lam * v).sum())
        print(f"  dir {k}: adjoint {adj:.15e} tangent {dJ:.15e}", flush=True)
        assert abs(adj - dJ) <= 1e-10 * max(abs(dJ), 1e-300), (adj, dJ)


def test_ns_adjoint_matches_central_fd():
    """Navier-Stokes rollout on the vortex patch mesh, 8 seeded smooth directions,
    central differences with eps = 1e-6 on the double-double objective."""
    print("case NS adjoint vs central FD", flush=True)
    F, probe = _vortex(4, True)
    g = _solver(F, probe, limiter=1, ns=1, mu=0.05, visc_work=1)
    S0 = F.get_state()
    xc, _ = F.leaf_info()
    dt = 0.3 * g.cfl_dt(0.4)
    g.set_state(S0)
    g.adj_forward(6, 3, dt)
    lam = g.adj_backward()
    eps = 1e-6

    def J(S):
        g.set_state(S)
        g.adj_forward(6, 3, dt)
        return g.adj_objective()

    for k in range(8):
        v = _smooth_direction(100 + k, xc, S0.shape)
        (ph, pl), (mh, ml) = J(S0 + eps * v), J(S0 - eps * v)
        fd = ((ph - mh) + (pl - ml)) / (2 * eps)
        adj = float((lam * v).sum())
        rel = abs(adj - fd) / abs(fd)
        print(f"  dir {k}: adjoint {adj:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6, (adj, fd)


def test_vortex_workload_full_size():
    """The adjoint rollout workload at its size: periodic 10^3 box, 64^3 base with a
    two-level patch over the vortex core (496,896 leaves), van Albada MUSCL, 50 SSP-RK3
    steps at CFL 0.5, checkpoints every 10 steps, objective = time-integrated pressure
    over a probe box beside the core.  The gradient is checked against the tangent along
    3 seeded random directions (<= 1e-10), against central differences along 3 seeded
    smooth directions (<= 1e-6, eps 1e-6, double-double J), and the gradient with
    checkpoints every 10 steps equals the one with the whole rollout stored, bit for bit."""
    from fv_140823_amd import configs
    F, p, meta = configs.vortex()
    assert F.nleaves == 496896
    print(f"case vortex full size: {F.nleaves} leaves, {meta['steps']} steps, ckpt {meta['ckpt']}",
          flush=True)
    xc, _ = F.leaf_info()
    probe = ((np.abs(xc[:, 0] - 6.5) < 1.0) & (np.abs(xc[:, 1] - 6.5) < 1.0)).astype(np.uint8)
    g = fv.Solver.from_forest(F, p)
    g.adj_set_probe(probe)
    S0 = F.get_state()
    dt = g.cfl_dt(meta["cfl"])
    n, ck = meta["steps"], meta["ckpt"]
    J = g.adj_forward(n, ck, dt)
    assert g.adj_checkpoint_spacing == ck
    lam = g.adj_backward()
    g.set_state(S0)
    J0 = g.adj_forward(n, 0, dt)
    lam0 = g.adj_backward()
    print(f"  J {J:.15e}; spacing chosen from free memory {g.adj_checkpoint_spacing}", flush=True)
    assert J0 == J
    assert np.array_equal(lam0, lam)
    rng = np.random.default_rng(50)
    for k in range(3):
        v = rng.standard_normal(S0.shape) * np.abs(S0).max(axis=1, keepdims=True) * 1e-2
        g.set_state(S0)
        J2, dJ = g.tlm(v, n, dt)
        assert abs(J2 - J) <= 1e-13 * abs(J)
        adj = float((lam * v).sum())
        print(f"  tangent dir {k}: adjoint {adj:.15e} tangent {dJ:.15e}", flush=True)
        assert abs(adj - dJ) <= 1e-10 * abs(dJ), (adj, dJ)
    eps = 1e-6

    def Jdd(S):
        g.set_state(S)
        g.adj_forward(n, ck, dt)
        return g.adj_objective()

    for k in range(3):
        # smooth bumps centred near the probe box, so that they reach it within the rollout
        r = np.random.default_rng(500 + k)
        c = np.array([r.uniform(5.5, 7.5), r.uniform(5.5, 7.5), r.uniform(0.0, 10.0)])
        w = r.uniform(0.5, 1.5)
        b = np.exp(-((xc - c) ** 2).sum(axis=1) / w ** 2)
        v = np.zeros(S0.shape)
        v[:4] = b * (r.choice([-1.0, 1.0], 4) * r.uniform(0.5, 1.5, 4))[:, None]
        (ph, pl), (mh, ml) = Jdd(S0 + eps * v), Jdd(S0 - eps * v)
        fd = ((ph - mh) + (pl - ml)) / (2 * eps)
        adj = float((lam * v).sum())
        rel = abs(adj - fd) / abs(fd)
        print(f"  fd dir {k}: adjoint {adj:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6, (adj, fd)
    g.close()


@pytest.mark.parametrize("ns", [0, 1])
def test_adjoint_on_tiled_stages(ns):
    """On a mesh of interior tiles the out-of-place stages of the rollout (its last
    segment, and every checkpointed segment the reverse sweep recomputes) run on the
    LDS-tiled kernels: the objective and the gradient are bitwise those of the
    face-list kernels."""
    print(f"case adjoint tiled vs face-list ns={ns}", flush=True)
    F = uniform(5, 5, periodic=(1, 1, 1), bc=(0,) * 6)
    xc, _ = F.leaf_info()
    F.set_state(state(1, xc, seed=29, jump=False))
    probe = (xc[:, 0] > 0.6).astype(np.uint8)
    kw = dict(limiter=1, ns=ns)
    if ns:
        kw.update(mu=2e-3, visc_work=1)
    out = []
    for tiling in (True, False):
        g = _solver(F, probe, **kw)
        g.set_tiling(tiling)
        assert g.tile_count == 64
        dt = 0.5 * g.cfl_dt(0.4)
        J = g.adj_forward(6, 2, dt)
        out.append((J, g.adj_backward()))
        g.close()
    (Ja, la), (Jb, lb) = out
    print(f"  J tiled {Ja!r} face-list {Jb!r}", flush=True)
    assert Ja == Jb
    assert np.array_equal(la, lb)
