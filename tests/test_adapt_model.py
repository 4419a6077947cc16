"""The host forest's adaptation against a set-based model of the same rules.

The model keeps the leaves as a dict (level, i, j, k) -> (S, SG) and the internal
nodes as a set, and applies the reference's adaptation rules in their plainest
form: split by injection of the parent state (split.c:492-502); 2:1 face balance
by repeated full passes that split every leaf with a face neighbour two levels
finer; merge of a parent whose 8 kids are leaves flagged -1 on the old tree, kept
only if no kid has a finer face neighbour on the post-split tree, with the mean of
the kids (merge.c:92-107).  The library's leaf list (depth-first kid order 0..7,
rebuildlist.c:34-91), levels and states must match it exactly after random
sequences of adaptations, on outflow, periodic and multi-root forests.
"""
import numpy as np
import pytest

import fv_140823_amd as fv
from fv_140823_amd._capi import FIELD_SG

AXIS = (1, 0, 1, 0, 2, 2)   # face f -> axis (operatetp.c face numbering)
SIGN = (-1, 1, 1, -1, -1, 1)


def _kid(node, c):
    l, i, j, k = node
    return (l + 1, 2 * i + (c & 1), 2 * j + ((c >> 1) & 1), 2 * k + ((c >> 2) & 1))


class Model:
    def __init__(self, nroot, level, periodic, neq):
        self.nroot, self.periodic, self.neq = nroot, periodic, neq
        self.roots = [(1, i, j, k) for k in range(nroot[2]) for j in range(nroot[1]) for i in range(nroot[0])]
        self.leaves = {r: None for r in self.roots}
        self.internal = set()
        for _ in range(level):
            for b in list(self.order()):
                self.split(b)

    def ncell(self, l, a):
        return self.nroot[a] << (l - 1)

    def across(self, node, f):
        l, *n = node
        a = AXIS[f]
        n[a] += SIGN[f]
        N = self.ncell(l, a)
        if n[a] < 0 or n[a] >= N:
            if not self.periodic[a]:
                return None
            n[a] %= N
        return (l, *n)

    def order(self):
        out = []

        def walk(b):
            if b in self.internal:
                for c in range(8):
                    walk(_kid(b, c))
            else:
                out.append(b)
        for r in self.roots:
            walk(r)
        return out

    def split(self, b):
        val = self.leaves.pop(b)
        self.internal.add(b)
        for c in range(8):
            self.leaves[_kid(b, c)] = None if val is None else (val[0].copy(), val[1].copy())

    def facing_kids(self, nb, f):
        """kids of node nb on its side facing the cell across face f"""
        a = AXIS[f]
        side = 0 if SIGN[f] > 0 else 1
        return [_kid(nb, c) for c in range(8) if ((c >> a) & 1) == side]

    def needs_balance(self, b):
        for f in range(6):
            nb = self.across(b, f)
            if nb is not None and nb in self.internal:
                if any(k in self.internal for k in self.facing_kids(nb, f)):
                    return True
        return False

    def merge_allowed(self, p):
        for c in range(8):
            for f in range(6):
                nb = self.across(_kid(p, c), f)
                if nb is not None and nb in self.internal:
                    return False
        return True

    def adapt(self, flags, max_level):
        order = self.order()
        split = [b for b, fl in zip(order, flags) if fl > 0 and b[0] < max_level]
        mflag = {b for b, fl in zip(order, flags) if fl < 0}
        cand = []
        for b in order:
            if b[0] <= 1:
                continue
            p = (b[0] - 1, b[1] >> 1, b[2] >> 1, b[3] >> 1)
            if _kid(p, 0) == b and all(_kid(p, c) in mflag for c in range(8)):
                cand.append(p)
        for b in split:
            self.split(b)
        changed = True
        while changed:
            changed = False
            for b in self.order():
                if self.needs_balance(b):
                    self.split(b)
                    changed = True
        merge = [p for p in cand if all(_kid(p, c) in self.leaves for c in range(8)) and self.merge_allowed(p)]
        for p in merge:
            kids = [self.leaves.pop(_kid(p, c)) for c in range(8)]
            S = np.zeros(self.neq)
            for k in kids:
                S = S + k[0] / 8.0
            SG = sum(k[1] for k in kids) / 8.0  # volume-weighted; equal volumes up to round-off
            self.internal.discard(p)
            self.leaves[p] = (S, SG)


def _leaf_keys(F, nroot, x0):
    xc, lev = F.leaf_info()
    keys = []
    for c, l in zip(xc, lev):
        h = 1.0 / (1 << (l - 1))   # box = nroot unit roots
        keys.append((int(l), *(int(round((c[a] - x0[a]) / h - 0.5)) for a in range(3))))
    return keys


@pytest.mark.parametrize("nroot,level,periodic,bc,seed", [
    ((1, 1, 1), 2, (0, 0, 0), (2, 2, 2, 2, 2, 2), 11),
    ((1, 1, 1), 2, (1, 1, 1), (2, 2, 2, 2, 2, 2), 12),
    ((2, 1, 3), 1, (0, 1, 0), (2, 1, 2, 2, 2, 1), 13),
])
def test_adapt_matches_set_model(nroot, level, periodic, bc, seed):
    neq, max_level = 5, 5
    x0 = (0.0, 0.0, 0.0)
    F = fv.Forest(nroot, x0, tuple(float(n) for n in nroot), level, bc, periodic, neq)
    M = Model(nroot, level, periodic, neq)
    rng = np.random.default_rng(seed)
    assert _leaf_keys(F, nroot, x0) == M.order()
    for it in range(5):
        order = M.order()
        N = len(order)
        S = 1.0 + rng.random((neq, N))
        SG = rng.random((neq + 1, 3, N))
        F.set_state(S)
        F.set_field(FIELD_SG, SG)
        for q, b in enumerate(order):
            M.leaves[b] = (S[:, q].copy(), SG[:, :, q].copy())
        u = rng.random(N)
        # splits sparse, merges dense so that whole sibling groups get flagged
        flags = np.where(u < 0.1, 1, np.where(u > 0.2, -1, 0)).astype(np.int8)
        n = F.adapt(flags, max_level)
        M.adapt(flags, max_level)
        order = M.order()
        assert n == len(order) == F.nleaves
        assert _leaf_keys(F, nroot, x0) == order, it
        Sm = np.stack([M.leaves[b][0] for b in order], axis=1)
        SGm = np.stack([M.leaves[b][1] for b in order], axis=2)
        assert np.array_equal(F.get_state(), Sm), it
        # gradients of the conserved variables are injected / volume-averaged; the
        # pressure-gradient row SG[neq] is not carried (merge.c:92 loops iv < neq)
        SGf = F.get_field(FIELD_SG)[:neq]
        assert np.all(np.isfinite(SGf))
        assert np.allclose(SGf, SGm[:neq], rtol=1e-13, atol=0), it
    fl = fv.FlatMesh(F.run)   # the adapted forest still flattens (every face resolved)
    assert fl.ncell == F.nleaves
    fl.close()
    F.close()


def test_merge_into_parent_that_never_was_a_leaf():
    """A forest created at level 2 has internal nodes that were never leaves; merging
    their kids must still weight the gradients by the parent's volume (merge.c:85-107)."""
    F = fv.Forest((1, 1, 1), (0, 0, 0), (1, 1, 1), 2, (2,) * 6, (0, 0, 0), 5)
    SG = np.arange(6 * 3 * 64, dtype=float).reshape(6, 3, 64)
    F.set_field(FIELD_SG, SG)
    assert F.adapt(-np.ones(64, np.int8), 4) == 8
    got = F.get_field(FIELD_SG)[:5]
    assert np.allclose(got, SG[:5].reshape(5, 3, 8, 8).mean(axis=3), rtol=1e-14, atol=0)
    F.close()


def test_adaptation_in_a_forked_child():
    """The host's worker threads do not survive a fork: a child that adapts a forest large
    enough for several threads (after the parent has started its workers) must finish
    with its own, and agree with the parent."""
    import os
    import time
    from fv_140823_amd import configs

    def run():
        F, _, meta = configs.blast(5, 2)
        xc, lev = F.leaf_info()
        r = np.sqrt(((xc - 0.5) ** 2).sum(axis=1))
        f = np.where((np.abs(r - 0.2) < 0.05) & (lev < meta["max_level"]), 1,
                     np.where(r < 0.1, -1, 0)).astype(np.int8)
        return F.adapt(f, meta["max_level"])

    n_parent = run()
    pid = os.fork()
    if pid == 0:
        code = 1
        try:
            code = 0 if run() == n_parent else 2
        finally:
            os._exit(code)
    t0 = time.time()
    while True:
        done, status = os.waitpid(pid, os.WNOHANG)
        if done:
            break
        if time.time() - t0 > 120:
            os.kill(pid, 9)
            os.waitpid(pid, 0)
            raise AssertionError("the forked child did not finish its adaptation")
        time.sleep(0.05)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0, status
