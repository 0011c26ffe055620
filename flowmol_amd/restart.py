"""The reference's `final_state` restart format: reader and single-rank writer.

Layout (src/parallel_io.MPI2.f90:1495-1724; readers 729-1352):
  per molecule, float64, concatenated rank by rank:
    tag, r_global(3), v(3),
    [rtrue(3)        if rtrue_flag == 1]
    [rtether(3)      if tag is a tether tag: 3, 5, 6, 7, 10]
    [moltype         if mie_potential == 1]
    [glob_no         if global_numbering == 1]
    [chainID, subchainID, funcy, glob_no, bin_bflag(4)   if potential_flag == 1]
  header (native endian):
    int32  globalnp, initialnunits(3), iter, tplot, seed(2), periodic(3), potential_flag,
           rtrue_flag, solvent_flag, nmonomers, npx, npy, npz, procnp(nproc), proctethernp(nproc)
    float64 globaldomain(3), density, rcutoff, delta_t, elapsedtime, simtime, k_c, R_0, eps_pp,
           eps_ps, eps_ss, delta_rneighbr
    int32  mie_potential, global_numbering
  int64 offset of the header (last 8 bytes).
"""
import struct

import numpy as np

TETHER_TAGS = (3, 5, 6, 7, 10)

_I = "<i"
_D = "<d"


def _header_fields(nproc):
    ints1 = ["globalnp", "nunits0", "nunits1", "nunits2", "iter", "tplot", "seed0", "seed1", "periodic0",
             "periodic1", "periodic2", "potential_flag", "rtrue_flag", "solvent_flag", "nmonomers", "npx", "npy",
             "npz"]
    return ints1


def read_final_state(path):
    """Decode a final_state file into numpy arrays (molecules in file order)."""
    b = open(path, "rb").read()
    if len(b) < 8:
        raise ValueError("final_state too short")
    (hpos,) = struct.unpack("<q", b[-8:])
    if not (0 < hpos < len(b) - 8):
        raise ValueError(f"bad header offset {hpos}")
    h = b[hpos:len(b) - 8]
    o = 0

    def ints(n):
        nonlocal o
        v = struct.unpack(f"<{n}i", h[o:o + 4 * n])
        o += 4 * n
        return v

    def dbls(n):
        nonlocal o
        v = struct.unpack(f"<{n}d", h[o:o + 8 * n])
        o += 8 * n
        return v

    hd = {}
    (hd["globalnp"],) = ints(1)
    hd["initialnunits"] = ints(3)
    (hd["iter"], hd["tplot"]) = ints(2)
    hd["seed"] = ints(2)
    hd["periodic"] = ints(3)
    (hd["potential_flag"], hd["rtrue_flag"], hd["solvent_flag"], hd["nmonomers"], hd["npx"], hd["npy"],
     hd["npz"]) = ints(7)
    nproc = hd["npx"] * hd["npy"] * hd["npz"]
    hd["procnp"] = ints(nproc)
    hd["proctethernp"] = ints(nproc)
    hd["globaldomain"] = dbls(3)
    (hd["density"], hd["rcutoff"], hd["delta_t"], hd["elapsedtime"], hd["simtime"], hd["k_c"], hd["R_0"],
     hd["eps_pp"], hd["eps_ps"], hd["eps_ss"], hd["delta_rneighbr"]) = dbls(11)
    (hd["mie_potential"], hd["global_numbering"]) = ints(2)
    if o != len(h):
        raise ValueError(f"header length {len(h)} does not match the fields read ({o})")
    body = np.frombuffer(b[:hpos], dtype="<f8")
    n = hd["globalnp"]
    tag = np.zeros(n, np.int32)
    r = np.zeros((n, 3))
    v = np.zeros((n, 3))
    rtrue = np.zeros((n, 3)) if hd["rtrue_flag"] == 1 else None
    rtether = np.zeros((n, 3))
    moltype = np.zeros(n, np.int32) if hd["mie_potential"] == 1 else None
    glob_no = np.zeros(n, np.int32) if hd["global_numbering"] == 1 else None
    mono = np.zeros((n, 8), np.int32) if hd["potential_flag"] == 1 else None
    p = 0
    for m in range(n):
        tag[m] = int(round(body[p]))
        r[m] = body[p + 1:p + 4]
        v[m] = body[p + 4:p + 7]
        p += 7
        if rtrue is not None:
            rtrue[m] = body[p:p + 3]
            p += 3
        if tag[m] in TETHER_TAGS:
            rtether[m] = body[p:p + 3]
            p += 3
        if moltype is not None:
            moltype[m] = int(round(body[p]))
            p += 1
        if glob_no is not None:
            glob_no[m] = int(round(body[p]))
            p += 1
        if mono is not None:
            mono[m] = np.rint(body[p:p + 8]).astype(np.int32)
            p += 8
    if p != body.size:
        raise ValueError(f"molecule records use {p} doubles, file holds {body.size}")
    return dict(header=hd, tag=tag, r=r, v=v, rtrue=rtrue, rtether=rtether, moltype=moltype, glob_no=glob_no,
                monomers=mono)


def write_final_state(path, r, v, *, nunits, density, rcutoff, delta_t, delta_rneighbr, domain, tag=None,
                      rtether=None, moltype=None, monomers=None, iter_=0, tplot=1, simtime=0.0, k_c=0.0,
                      R_0=0.0, seed=(0, 0), periodic=(1, 1, 1), nmonomers=0):
    """Write a single-rank final_state (npx=npy=npz=1) in the reference's layout.  monomers rows
    are in the file's order: chainID, subchainID, funcy, glob_no, bin_bflag(4)."""
    n = r.shape[0]
    tag = np.zeros(n, np.int32) if tag is None else np.asarray(tag, np.int32)
    recs = []
    ntether = 0
    for m in range(n):
        rec = [float(tag[m]), *map(float, r[m]), *map(float, v[m])]
        if int(tag[m]) in TETHER_TAGS:
            rec += list(map(float, rtether[m]))
            ntether += 1
        if moltype is not None:
            rec.append(float(moltype[m]))
        if monomers is not None:
            rec += list(map(float, monomers[m]))
        recs.append(rec)
    body = np.array([x for rec in recs for x in rec], dtype="<f8").tobytes()
    hdr = struct.pack("<i", n) + struct.pack("<3i", *nunits) + struct.pack("<2i", iter_, tplot)
    hdr += struct.pack("<2i", *seed) + struct.pack("<3i", *periodic)
    hdr += struct.pack("<7i", 1 if monomers is not None else 0, 0, 0, int(nmonomers), 1, 1, 1)
    hdr += struct.pack("<i", n) + struct.pack("<i", ntether)
    hdr += struct.pack("<3d", *map(float, domain))
    hdr += struct.pack("<11d", density, rcutoff, delta_t, 0.0, simtime, k_c, R_0, 0.0, 0.0, 0.0, delta_rneighbr)
    hdr += struct.pack("<2i", 1 if moltype is not None else 0, 0)
    with open(path, "wb") as f:
        f.write(body)
        f.write(hdr)
        f.write(struct.pack("<q", len(body)))


def write_final_state_from_context(ctx, path, *, nunits, iter_, seed=(0, 0), nmonomers=0):
    """parallel_io_final_state for one device context (src/parallel_io.MPI2.f90:1495-1704): the
    rebuild sequence first (linklist_deallocateall, sendmols, sort_mols, assign_to_cell,
    messenger_updateborders, assign_to_neighbourlist: molecules wrapped into the box and
    renumbered), then the state, tags, tether sites, molecule types and monomer fields in the
    current numbering.  The device keeps running from the rebuilt lists, as the reference does
    after a rescue snapshot."""
    from ._abi import FM_ENSEMBLE_TAG_MOVE

    p = ctx.params
    ctx.rebuild()
    r, v, _ = ctx.state()
    tag, mt, mono = ctx.labels()
    rt = ctx.rtether() if p.ensemble == FM_ENSEMBLE_TAG_MOVE else None
    monomers = None
    if mono is not None:
        chain, sub, funcy, bflag, glob = mono
        monomers = np.column_stack([chain, sub, funcy, glob, bflag]).astype(np.int32)
    write_final_state(path, r, v, nunits=nunits, density=p.density, rcutoff=p.rcutoff, delta_t=p.delta_t,
                      delta_rneighbr=p.delta_rneighbr[0], domain=[p.domain[0], p.domain[1], p.domain[2]], tag=tag,
                      rtether=rt, moltype=mt, monomers=monomers, iter_=iter_, tplot=p.tplot, simtime=ctx.simtime,
                      k_c=p.k_c if p.potential_flag == 1 else 0.0, R_0=p.R_0 if p.potential_flag == 1 else 0.0,
                      seed=seed, nmonomers=nmonomers)
