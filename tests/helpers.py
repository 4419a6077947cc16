"""Shared test helpers: small meshes, seeded physical states, comparisons."""
from __future__ import annotations

import math

import numpy as np

import fv_140823_amd as fv
from fv_140823_amd.configs import prim_to_cons


def uniform(level=3, neq=5, periodic=(0, 0, 0), bc=(2, 2, 2, 2, 2, 2), box=((0, 0, 0), (1, 1, 1))):
    return fv.Forest((1, 1, 1), box[0], box[1], level, bc, periodic, neq)


def amr(level=3, neq=5, periodic=(0, 0, 0), passes=2, centre=(0.4, 0.55, 0.5), radius=0.3):
    """Base 2^level cube refined in `passes` nested spheres -> 2:1 hanging faces."""
    F = fv.Forest((1, 1, 1), (0, 0, 0), (1, 1, 1), level, (2, 2, 2, 2, 2, 2), periodic, neq)
    for k in range(passes):
        xc, lev = F.leaf_info()
        r = np.sqrt(((xc - np.asarray(centre)) ** 2).sum(axis=1))
        flags = (r < radius * (1.0 - 0.35 * k)).astype(np.int8)
        F.adapt(flags, level + 1 + passes)
    return F


def smooth(xc, seed, k=1.0):
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * math.pi, 9)
    x, y, z = xc[:, 0], xc[:, 1], xc[:, 2]
    t = 2 * math.pi * k
    return (np.sin(t * x + ph[0]) * np.cos(t * y + ph[1]) * np.sin(t * z + ph[2]),
            np.cos(t * x + ph[3]) * np.sin(t * y + ph[4]) * np.cos(t * z + ph[5]),
            np.sin(t * x + ph[6]) * np.sin(t * y + ph[7]) * np.cos(t * z + ph[8]))


def state(model, xc, seed=0, jump=True):
    """A physical conserved state with smooth variation plus a jump at x = 0.5."""
    a, b, c = smooth(xc, seed)
    j = (xc[:, 0] > 0.5).astype(float) if jump else 0.0
    if model == 1:
        rho = 1.0 + 0.3 * a - 0.4 * j
        p = 1.0 + 0.3 * b - 0.5 * j
        return prim_to_cons(rho, 0.3 * b, -0.2 * c, 0.25 * a, p)
    if model == 2:
        rho = 998.2 + 2.0 * a - 1.5 * j
        return np.stack([rho, rho * 0.3 * b, rho * -0.2 * c, rho * 0.1 * a])
    if model == 3:
        rho = 900.0 + 60.0 * a - 100.0 * j
        Y = 1e-3 * (1.0 + 0.5 * b) + 2e-3 * j
        return np.stack([rho, rho * 0.4 * b, rho * -0.3 * c, rho * 0.2 * a, rho * Y])
    raise ValueError(model)


def rel_linf(x, ref, nvar=None):
    """Per-variable max |x - ref| / max |ref| (rows = variables).  With nvar = neq the
    rows are read as blocks of nvar variables and the momentum components 1..3 of a
    block share one scale (the largest of the three): a momentum component that is
    zero in exact arithmetic (y, z in a 1D problem) is then judged against the
    momentum magnitude, not against its own round-off."""
    ref = np.asarray(ref)
    r = ref.reshape(ref.shape[0], -1)
    x = np.asarray(x).reshape(r.shape)
    den = np.abs(r).max(axis=1)
    if nvar is not None and nvar >= 4:
        blk = den.reshape(-1, nvar)
        mom = blk[:, 1:4].max(axis=1, keepdims=True)
        blk[:, 1:4] = mom
        den = blk.reshape(-1)
    den = np.where(den == 0, 1.0, den)
    return np.abs(x - r).max(axis=1) / den
