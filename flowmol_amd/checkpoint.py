"""Checkpoint and resume of a device context (no counterpart in the reference, whose restart is
the final_state file: src/parallel_io.MPI2.f90:1495-1724 write, 729-1352 read).

A snapshot is taken at a step boundary after the rebuild sequence of src/md.f90 (sendmols wraps
and renumbers, the cells and lists are rebuilt, the displacement sums start again), so a context
resumed from it continues exactly as the context that took it:
  * positions, velocities and accelerations (the accelerations of the last force call: a resumed
    context does not recompute them, so its first step uses the same values);
  * tether sites, tags, molecule types and monomer fields in the current numbering;
  * the next iteration number and simtime;
  * with MOP planes, the tally accumulated so far (a resumed context starts its own tally from
    zero; add the two for the whole run's).
The resumed context bins the same positions in the same numbering, which gives the same slot
order, lists and summation order: the two runs are bit-identical from there on.  A snapshot
stores as .npz without pickled objects (numpy.load with allow_pickle=False reads it).
"""
import numpy as np

from ._abi import FM_ENSEMBLE_TAG_MOVE, FmParams
from .md import MDContext


class Snapshot:
    def __init__(self, params, next_iter, simtime, r, v, a, tag=None, rtether=None, moltype=None, monomers=None,
                 mop=None):
        self.params = params
        self.mop = mop
        self.next_iter = int(next_iter)
        self.simtime = float(simtime)
        self.r, self.v, self.a = r, v, a
        self.tag, self.rtether, self.moltype, self.monomers = tag, rtether, moltype, monomers

    @property
    def np(self):
        return self.r.shape[0]

    def save(self, path):
        arrays = dict(params=np.frombuffer(bytes(self.params), np.uint8), next_iter=np.int64(self.next_iter),
                      simtime=np.float64(self.simtime), r=self.r, v=self.v, a=self.a)
        for k in ("tag", "rtether", "moltype", "mop"):
            if getattr(self, k) is not None:
                arrays[k] = getattr(self, k)
        if self.monomers is not None:
            for k, x in zip(("chainID", "subchainID", "funcy", "bin_bflag", "glob_no"), self.monomers):
                arrays["mono_" + k] = x
        np.savez(path, **arrays)

    @staticmethod
    def load(path):
        d = np.load(path, allow_pickle=False)
        p = FmParams.from_buffer_copy(d["params"].tobytes())
        mono = None
        if "mono_chainID" in d:
            mono = tuple(d["mono_" + k] for k in ("chainID", "subchainID", "funcy", "bin_bflag", "glob_no"))
        g = lambda k: d[k] if k in d else None
        return Snapshot(p, int(d["next_iter"]), float(d["simtime"]), d["r"], d["v"], d["a"], g("tag"), g("rtether"),
                        g("moltype"), mono, g("mop"))


def take(ctx, next_iter):
    """Rebuild (src/md.f90:150-157) and read the context back; ctx keeps running afterwards.
    next_iter: the iteration number of the context's next step."""
    ctx.rebuild()
    r, v, a = ctx.state()
    tag, mt, mono = ctx.labels()
    rt = ctx.rtether() if ctx.params.ensemble == FM_ENSEMBLE_TAG_MOVE else None
    mop = ctx.mop() if 1 <= ctx.params.vflux_outflag <= 3 else None
    return Snapshot(type(ctx.params).from_buffer_copy(ctx.params), next_iter, ctx.simtime, r, v, a, tag, rt, mt, mono,
                    mop)


def resume(snap):
    """A new context holding the snapshot's state, with its cells and lists built and without a
    force call; run it from snap.next_iter."""
    ctx = MDContext(snap.params, snap.r, snap.v, snap.a)
    if snap.moltype is not None:
        ctx.set_moltype(snap.moltype)
    if snap.tag is not None:
        ctx.set_tags(snap.tag, snap.rtether)
    if snap.monomers is not None:
        ctx.set_monomers(*snap.monomers)
    ctx.simtime = snap.simtime
    ctx.assign_to_cell()
    ctx.messenger_updateborders(1)
    ctx.assign_to_neighbourlist()
    return ctx
