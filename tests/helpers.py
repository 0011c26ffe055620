"""Shared fixtures of the test suite: seeded states and set/tolerance comparisons."""
import numpy as np

from flowmol_amd import config
from oracle.oracle import Oracle


def canonical_pairs(i, j, s):
    """Canonical neighbour set: unshifted pairs unordered, shifted entries directed.

    The reference's half list holds each unshifted pair once (either direction) and each pair
    that crosses the periodic boundary twice, once per end against the other's image
    (src/linklist.f90:660-804); a full list holds the unshifted pairs twice.
    """
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    s = np.asarray(s, np.int64)
    lo = np.where(s == 13, np.minimum(i, j), i)
    hi = np.where(s == 13, np.maximum(i, j), j)
    key = (lo << 32) | (hi << 5) | s
    return np.unique(key)


def full_from_half(i, j, s):
    """Directed full list from the reference half list: unshifted entries in both directions."""
    m = s == 13
    fi = np.concatenate([i, j[m]])
    fj = np.concatenate([j, i[m]])
    fs = np.concatenate([s, s[m]])
    return fi, fj, fs


def equilibrated(cfg, steps, **kw):
    """State after `steps` oracle NVE steps from the seeded lattice (forces at t=0 are zero on an
    FCC WCA lattice, so parity inputs are taken from a disordered state)."""
    p, r, v = cfg(**kw)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    if steps:
        o.run_vv(1, steps)
    o.rebuild()
    r2, v2, a2 = o.state()
    return p, r2, v2, a2


def rel_force_err(a_test, a_ref, scale):
    """max_i |a_test_i - a_ref_i|_inf / max(sum_j |f_ij|, tiny): the scale-aware relative error of
    a per-molecule net force."""
    d = np.abs(a_test - a_ref).max(axis=1)
    return float((d / np.maximum(scale, 1e-300)).max())
