"""Synthetic meshes and initial data for the five benchmark workloads.

Each builder returns (forest, params, meta): a host forest whose leaves already
hold the initial conserved state, the fvg_params of the workload, and a dict of
workload facts (name, steps, cfl, ...).  Initial data are point values at leaf
centres.  Primitive -> conserved follows Init_S.c:3-17 (E = rho (e + |u|^2/2),
e = (p + gamma pinf) / (rho (gamma - 1))).

  sod        uniform 2^L cube, Sod tube along x, transmissive sides
  blast      base 2^L cube + 2 AMR levels around an overpressured sphere
  cavitation [-1,1]^3, model 2 barotropic liquid with a vapour bubble
  tgv        periodic [0,2 pi]^3 Taylor-Green vortex, Navier-Stokes, Prandtl 0.71
  vortex     periodic 10^3 isentropic vortex, base level + finer core patch
"""
from __future__ import annotations

import math

import numpy as np

from .core import Forest, params


def prim_to_cons(rho, u, v, w, p, gamma=1.4, pinf=0.0):
    e_int = (p + gamma * pinf) / (rho * (gamma - 1.0))
    E = rho * (e_int + 0.5 * (u * u + v * v + w * w))
    return np.stack([rho, rho * u, rho * v, rho * w, E])


def cons_to_prim(S, gamma=1.4, pinf=0.0):
    rho = S[0]
    u, v, w = S[1] / rho, S[2] / rho, S[3] / rho
    ei = S[4] / rho - 0.5 * (u * u + v * v + w * w)
    p = rho * (gamma - 1.0) * ei - gamma * pinf
    return rho, u, v, w, p


def pressure(S, p):
    """Cell pressure of a conserved state [neq][N] for the EOS of params p: the
    ideal/stiffened gas (model 1) or the barotropic law (model 2, barotropic_eos.c:
    13-22); None for the two-phase model 3 (its pressure needs the mixture root)."""
    if p.model == 1:
        rho = S[0]
        ke = 0.5 * (S[1] ** 2 + S[2] ** 2 + S[3] ** 2) / rho
        return (p.gamma - 1.0) * (S[4] - ke) - p.gamma * p.pinf
    if p.model == 2:
        rho = S[0]
        out = np.empty_like(rho)
        liq = rho >= p.baro_rsat
        out[liq] = p.baro_psat + p.baro_beta * ((rho[liq] / p.baro_rsat) ** p.baro_nexp - 1.0)
        out[~liq] = p.baro_psat + p.baro_cbar * (1.0 / p.baro_rsat - 1.0 / rho[~liq])
        return out
    return None


def sod(level: int = 5, nroot=(1, 1, 1)):
    F = Forest(nroot, (0, 0, 0), (1, 1, 1), level, (2, 2, 2, 2, 2, 2), (0, 0, 0), 5)
    xc, _ = F.leaf_info()
    left = xc[:, 0] < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    F.set_state(prim_to_cons(rho, z, z, z, p))
    return F, params(1, ischeme=2), dict(name=f"sod_{2**level}^3", steps=100, cfl=0.5, method=3)


def _blast_state(xc, r0=0.1, centre=(0.5, 0.5, 0.5)):
    r = np.sqrt(((xc - np.asarray(centre)) ** 2).sum(axis=1))
    rho = np.ones(len(xc))
    p = np.where(r < r0, 10.0, 0.1)
    z = np.zeros_like(rho)
    return prim_to_cons(rho, z, z, z, p)


def blast_refine_flags(forest: Forest, max_level: int, r0: float = 0.1, band: float = 2.0):
    """Geometric pre-refinement around the sphere r = r0 (main.c:139-166 refines by
    geometry before the first step): leaves within `band` own-cell sizes of it."""
    xc, lev = forest.leaf_info()
    h = 1.0 / (2.0 ** (lev - 1))
    r = np.sqrt(((xc - 0.5) ** 2).sum(axis=1))
    near = np.abs(r - r0) < band * h * math.sqrt(3.0)
    return np.where(near & (lev < max_level), 1, 0).astype(np.int8)


def blast(base_level: int = 5, extra: int = 2):
    F = Forest((1, 1, 1), (0, 0, 0), (1, 1, 1), base_level, (2, 2, 2, 2, 2, 2), (0, 0, 0), 5)
    max_level = base_level + 1 + extra  # reference levels: roots are level 1
    for _ in range(extra):
        F.adapt(blast_refine_flags(F, max_level), max_level)
    xc, _ = F.leaf_info()
    F.set_state(_blast_state(xc))
    return F, params(1, ischeme=2), dict(name=f"blast_{2**base_level}^3+{extra}", steps=200, cfl=0.5,
                                         method=3, max_level=max_level, adapt_every=10)


def density_gradient_flags(forest: Forest, SG: np.ndarray, S: np.ndarray, max_level: int,
                           refine: float = 0.05, coarsen: float = 0.01, base_level: int = 6):
    """Split where h |grad rho| / rho exceeds `refine`, merge below `coarsen`
    (the density-gradient criterion of criterionsplt.c:60-69, made scale-free)."""
    xc, lev = forest.leaf_info()
    h = 1.0 / (2.0 ** (lev - 1))
    g = np.sqrt((SG[0] ** 2).sum(axis=0)) * h / S[0]
    flags = np.zeros(len(lev), np.int8)
    flags[(g > refine) & (lev < max_level)] = 1
    flags[(g < coarsen) & (lev > base_level)] = -1
    return flags


def barotropic_cavitation_params(rho_l=998.0, c_l=1480.0, rho_v=0.02, c_v=3.0, psat=2339.0,
                                 nexp=7.15, **kw):
    """Model-2 EOS constants for a liquid of (rho_l, c_l) and a vapour of (rho_v, c_v):
    beta = c_l^2 rho_l / n gives c(rho_l) = c_l on the liquid branch, and
    cbar = (c_v rho_v)^2 gives c(rho_v) = c_v on the cavitating branch
    (barotropic_eos.c:20-27)."""
    return params(2, baro_nexp=nexp, baro_beta=c_l * c_l * rho_l / nexp, baro_cbar=(c_v * rho_v) ** 2,
                  baro_psat=psat, baro_rsat=rho_l, **kw)


def cavitation(level: int = 7, radius: float = 0.2, ratio: float = 100.0):
    p = barotropic_cavitation_params(ischeme=2)
    F = Forest((1, 1, 1), (-1, -1, -1), (1, 1, 1), level, (2, 2, 2, 2, 2, 2), (0, 0, 0), 4)
    xc, _ = F.leaf_info()
    h = 2.0 / 2 ** level
    r = np.sqrt((xc ** 2).sum(axis=1))
    alpha = 0.5 * (1.0 - np.tanh((r - radius) / (2.0 * h)))
    # liquid density at ratio * psat from the liquid branch p = psat + beta ((rho/rsat)^n - 1)
    pl = ratio * p.baro_psat
    rho_l = p.baro_rsat * (1.0 + (pl - p.baro_psat) / p.baro_beta) ** (1.0 / p.baro_nexp)
    rho = alpha * 0.02 + (1.0 - alpha) * rho_l
    z = np.zeros_like(rho)
    F.set_state(np.stack([rho, z, z, z]))
    return F, p, dict(name=f"cavitation_{2**level}^3", steps=500, cfl=0.5, method=3)


def tgv(level: int = 7, Re: float = 1600.0, Mach: float = 0.1, gamma: float = 1.4,
        prandtl: float = 0.71):
    """Taylor-Green vortex, Navier-Stokes with viscous work and heat conduction at
    Prandtl `prandtl` (0: none, as flux_viscous.c)."""
    L = 2.0 * math.pi
    F = Forest((1, 1, 1), (0, 0, 0), (L, L, L), level, (0, 0, 0, 0, 0, 0), (1, 1, 1), 5)
    xc, _ = F.leaf_info()
    x, y, z = xc[:, 0], xc[:, 1], xc[:, 2]
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    p0 = 1.0 / (gamma * Mach * Mach)
    p = p0 + (1.0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    rho = np.ones_like(u)
    F.set_state(prim_to_cons(rho, u, v, w, p, gamma))
    return F, params(1, ischeme=2, ns=1, mu=1.0 / Re, visc_work=1, gamma=gamma, prandtl=prandtl), dict(
        name=f"tgv_{2**level}^3", steps=1000, cfl=0.5, method=3)


def vortex_state(xc, box=10.0, beta=5.0, gamma=1.4, mach=0.5, centre=None):
    """Isentropic vortex (strength beta, unit radius) in a free stream of Mach `mach`
    along the x-y diagonal; axis along z; rho_inf = p_inf = 1."""
    c = np.asarray(centre if centre is not None else (box / 2, box / 2, box / 2))
    dx = xc[:, 0] - c[0]
    dy = xc[:, 1] - c[1]
    r2 = dx * dx + dy * dy
    ainf = math.sqrt(gamma)
    U = mach * ainf / math.sqrt(2.0)
    f = np.exp(0.5 * (1.0 - r2))
    u = U - beta / (2 * math.pi) * dy * f
    v = U + beta / (2 * math.pi) * dx * f
    w = np.zeros_like(u)
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * math.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return prim_to_cons(rho, u, v, w, p, gamma)


def vortex(base_level: int = 6, patch_radius: float = 2.0, box: float = 10.0):
    F = Forest((1, 1, 1), (0, 0, 0), (box, box, box), base_level, (0, 0, 0, 0, 0, 0), (1, 1, 1), 5)
    xc, lev = F.leaf_info()
    r = np.sqrt((xc[:, 0] - box / 2) ** 2 + (xc[:, 1] - box / 2) ** 2)
    F.adapt(np.where(r < patch_radius, 1, 0).astype(np.int8), base_level + 2)
    xc, _ = F.leaf_info()
    F.set_state(vortex_state(xc, box))
    return F, params(1, ischeme=2, limiter=1), dict(name=f"vortex_{2**base_level}^3+patch", steps=50,
                                                    cfl=0.5, method=3, ckpt=10)


BUILDERS = {"sod": sod, "blast": blast, "cavitation": cavitation, "tgv": tgv, "vortex": vortex}
