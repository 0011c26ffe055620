"""Record-step outputs in the reference's file formats, written from the device sums.

  results/macroscopic_properties   macroscopic_properties_header / _record
                                   (src/parallel_io.MPI2.f90:4588-4641): one header line, then one
                                   ';'-separated line per record step in the reference's Fortran
                                   edit descriptors (quirks kept: with potential_flag 1 the
                                   temperature takes f15.4 and the kinetic energy f10.4)
  results/pplane                   MOP_stress_io (src/parallel_io.MPI2.f90:3987-4070): Pxy_plane
                                   divided by Nstress_ave and by 4*Lx*Lz, written as raw native
                                   doubles at record (iter/(tplot*Nmass_ave) - 1) of 3*nplanes
                                   doubles; like the reference, only the first nplanes doubles of
                                   the column-major (3, nplanes) array are written into each record
  kinetic MOP term                 cumulative_momentum_flux, vflux_outflag 1..3
                                   (src/simulation_record.f90:5556-5630): every molecule whose
                                   coordinate normal to the planes crossed a plane between
                                   r - dt*v and r adds m*v to that plane

The configurational MOP tally (pressure_tensor_forces_MOP, src/simulation_record.f90:7068-7112)
is accumulated on the device by the force launches of record steps; momentum_flux_averaging
(src/simulation_record.f90:5698-5745) adds the kinetic term at the same record steps and, every
Nvflux_ave samples, writes the record and zeroes Pxy_plane.
"""
import math
import os

import numpy as np

# ---- Fortran fixed-point edit descriptors -----------------------------------------------------------


def fortran_f(x, w, d):
    """The Fortran Fw.d edit descriptor: right-justified in w characters with d decimals, the
    leading zero of |x| < 1 dropped only when the field is one character short, and w asterisks
    when the number does not fit."""
    x = float(x)
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
        if len(s) > w:
            s = "Inf" if x > 0 else "-Inf"
    else:
        s = f"{x:.{d}f}"
    if len(s) > w and (s.startswith("0.") or s.startswith("-0.")):
        s = s.replace("0.", ".", 1)
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def fortran_i(n, w):
    s = str(int(n))
    return "*" * w if len(s) > w else s.rjust(w)


# ---- results/macroscopic_properties -----------------------------------------------------------------
HEADER_LJ = " iter; simtime; VSum; V^2Sum; Temp; KE; PE; TE; Pressure"
HEADER_POLY = " iter; simtime; VSum; V^2Sum; Temp; KE; PE (LJ); PE (POLY); PE (Tot); TE; Pressure; Etevtcf; R_g "


def macroscopic_properties_line(iter_, simtime, m, potential_flag=0, etevtcf=0.0, R_g=0.0):
    """One record line; m holds the sums of fm_macros (vsum, mv2sum, temperature, kinenergy,
    potenergy, totenergy, pressure and, with potential_flag 1, potenergysum_LJ/_POLY)."""
    if potential_flag == 0:
        fields = [(m["vsum"], 15, 4), (m["mv2sum"], 15, 4), (m["temperature"], 10, 4), (m["kinenergy"], 19, 15),
                  (m["potenergy"], 19, 15), (m["totenergy"], 19, 15), (m["pressure"], 10, 4)]
    else:
        n = m["np"]
        pe_lj = (0.5 * m["potenergysum_LJ"]) / n
        pe_poly = (0.5 * m["potenergysum_POLY"]) / n
        fields = [(m["vsum"], 15, 4), (m["mv2sum"], 15, 4), (m["temperature"], 15, 4), (m["kinenergy"], 10, 4),
                  (pe_lj, 19, 15), (pe_poly, 19, 15), (m["potenergy"], 19, 15), (m["totenergy"], 19, 15),
                  (m["pressure"], 19, 15), (etevtcf, 10, 4), (R_g, 10, 4)]
    out = " " + fortran_i(iter_, 8) + ";" + fortran_f(simtime, 15, 4)
    for x, w, d in fields:
        out += ";" + fortran_f(x, w, d)
    return out


class MacroscopicProperties:
    """results/macroscopic_properties under prefix_dir, written as the reference writes it."""

    def __init__(self, prefix_dir, potential_flag=0):
        self.path = os.path.join(prefix_dir, "results", "macroscopic_properties")
        self.potential_flag = int(potential_flag)
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        with open(self.path, "w") as f:
            f.write((HEADER_LJ if self.potential_flag == 0 else HEADER_POLY) + "\n")

    def record(self, iter_, simtime, macros, np_=None, etevtcf=0.0, R_g=0.0):
        m = dict(macros)
        if np_ is not None:
            m["np"] = np_
        with open(self.path, "a") as f:
            f.write(macroscopic_properties_line(iter_, simtime, m, self.potential_flag, etevtcf, R_g) + "\n")


def read_macroscopic_properties(path):
    """Columns of a macroscopic_properties file as a dict of float arrays (header names)."""
    with open(path) as f:
        names = [h.strip() for h in f.readline().split(";") if h.strip()]
        rows = [[float(x) for x in line.split(";")] for line in f if line.strip()]
    a = np.array(rows, dtype=np.float64).reshape(-1, len(names))
    return {n: a[:, k] for k, n in enumerate(names)}


# ---- MOP: the kinetic term and results/pplane -----------------------------------------------------
def mop_planes(domain, nplanes, planespacing, pnxyz=2):
    """planes(n) = planespacing*(n-1) + shift - halfdomain and shift of set_parameters_allocate
    (src/setup_set_parameters.f90:2298-2340, one process)."""
    L = float(domain[pnxyz - 1])
    shift = 0.5 * (L - planespacing * (nplanes - 1))
    n = np.arange(1, nplanes + 1, dtype=np.float64)
    return planespacing * (n - 1) + shift - 0.5 * L, shift


def mop_kinetic(r, v, mass, domain, nplanes, planespacing, delta_t, pnxyz=2):
    """cumulative_momentum_flux for vflux_outflag 1..3 (src/simulation_record.f90:5593-5628):
    returns the (3, nplanes) increment m*v at the plane crossed between r - dt*v and r."""
    r = np.asarray(r, np.float64).reshape(-1, 3)
    v = np.asarray(v, np.float64).reshape(-1, 3)
    m = np.broadcast_to(np.asarray(mass, np.float64), (r.shape[0],))
    ax = pnxyz - 1
    L = float(domain[ax])
    half = 0.5 * L
    shift = 0.5 * (L - planespacing * (nplanes - 1))
    x = r[:, ax]
    c1 = np.ceil((x + half - shift) / planespacing)
    c0 = np.ceil((x - delta_t * v[:, ax] + half - shift) / planespacing)
    cross = c1 - c0
    out = np.zeros((3, nplanes))
    sel = np.nonzero(cross != 0.0)[0]
    for n in sel:  # few molecules cross a plane in one step; keep the reference's order of sums
        heav = math.ceil(math.copysign(0.5, cross[n]))
        pno = int(c1[n]) - heav + 1
        pno = min(max(pno, 1), nplanes)
        out[:, pno - 1] += m[n] * v[n]
    return out


def pplane_normalise(Pxy_plane, Nstress_ave, domain):
    """MOP_stress_io: Pxy_plane/Nstress_ave, then /(4*domain(1)*domain(3))
    (src/parallel_io.MPI2.f90:4003-4006)."""
    P = np.asarray(Pxy_plane, np.float64) / float(Nstress_ave)
    return P / (4 * float(domain[0]) * float(domain[2]))


def write_pplane(prefix_dir, iter_, tplot, Nmass_ave, Pxy_plane_normalised):
    """The MPI_FILE_WRITE_ALL of MOP_stress_io on one process: record m = iter/(tplot*Nmass_ave)
    starts at byte (m-1)*3*nplanes*8 and receives the first nplanes doubles of Pxy_plane in
    Fortran (column-major) order."""
    P = np.asarray(Pxy_plane_normalised, np.float64).reshape(3, -1)
    nplanes = P.shape[1]
    m = iter_ // (tplot * Nmass_ave)
    path = os.path.join(prefix_dir, "results", "pplane")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    flat = np.asfortranarray(P).ravel(order="F")[:nplanes]
    mode = "r+b" if os.path.exists(path) else "w+b"
    with open(path, mode) as f:
        f.seek((m - 1) * 3 * nplanes * 8)
        f.write(flat.astype("<f8").tobytes())
    return path


class MOPRecorder:
    """momentum_flux_averaging for vflux_outflag 1..3 on the host: at each record step add the
    device's configurational tally since its last reset and the kinetic crossings; every
    Nvflux_ave samples write the normalised record and zero the planes."""

    def __init__(self, prefix_dir, params, Nvflux_ave=1, Nstress_ave=None, Nmass_ave=1):
        self.prefix = prefix_dir
        self.p = params
        self.Nvflux_ave = int(Nvflux_ave)
        self.Nstress_ave = int(Nstress_ave if Nstress_ave is not None else Nvflux_ave)
        self.Nmass_ave = int(Nmass_ave)
        self.nplanes = int(params.nplanes)
        self.P = np.zeros((3, self.nplanes))
        self.count = 0

    def sample(self, ctx, iter_, r, v, mass=1.0):
        """Call at a record step after the step's positions and velocities are final."""
        p = self.p
        dom = [p.domain[0], p.domain[1], p.domain[2]]
        self.P += ctx.mop().T
        ctx.reset_mop()
        self.P += mop_kinetic(r, v, mass, dom, self.nplanes, p.planespacing, p.delta_t, p.vflux_outflag)
        self.count += 1
        if self.count == self.Nvflux_ave:
            norm = pplane_normalise(self.P, self.Nstress_ave, dom)
            write_pplane(self.prefix, iter_, p.tplot, self.Nmass_ave, norm)
            self.P[:] = 0.0
            self.count = 0
            return norm
        return None


# ---- a time loop with the record-step outputs -----------------------------------------------------
def run_with_records(ctx, iter0, nsteps, prefix_dir, Nvflux_ave=1, Nmass_ave=1, mass=1.0):
    """Advance ctx from iter0 for nsteps as src/md.f90 does and, at every record step
    (mod(iter, tplot) == 0), write what simulation_record writes for the macroscopic properties
    and, with vflux_outflag 1..3, the MOP planes:
      velocity Verlet: the record follows pass 2 (src/md.f90:159-161), so the fused loop runs up
        to and including the record step and the record reads its final state;
      leap-frog: the record follows the force call and precedes the move (src/md.f90:114-128), so
        the record step runs routine by routine.
    Returns (macros per record step, number of rebuilds)."""
    from ._abi import FM_INTEGRATION_LFV

    p = ctx.params
    tplot = int(p.tplot)
    mp = MacroscopicProperties(prefix_dir, p.potential_flag)
    mop = MOPRecorder(prefix_dir, p, Nvflux_ave=Nvflux_ave, Nmass_ave=Nmass_ave) if 1 <= p.vflux_outflag <= 3 else None
    lfv = p.integration_algorithm == FM_INTEGRATION_LFV
    out, nrb = [], 0
    it, end = int(iter0), int(iter0) + int(nsteps)

    def record(k, m):
        m = dict(m)
        m["np"] = ctx.np
        mp.record(k, ctx.simtime, m)
        if mop is not None:
            # the kinetic term reads r and v as they stand (under leap-frog v is half a step back)
            r, v, _ = ctx.state()
            mop.sample(ctx, k, r, v, mass)
        out.append(m)

    while it < end:
        nxt = -(-it // tplot) * tplot  # first record step >= it
        if nxt >= end:
            if lfv:
                _, n = ctx.run_lfv(it, end - it)
            else:
                _, n = ctx.run_vv(it, end - it)
            nrb += n
            break
        if lfv:
            if nxt > it:
                _, n = ctx.run_lfv(it, nxt - it)
                nrb += n
            ctx.simulation_compute_forces(nxt)
            record(nxt, ctx.evaluate_macroscopic_properties())
            ctx.simulation_move_particles_lfv()
            ctx.messenger_updateborders(0)
            if ctx.simulation_checkrebuild(nxt):
                ctx.rebuild()
                nrb += 1
        else:
            ms, n = ctx.run_vv(it, nxt - it + 1, n_macros=1)
            nrb += n
            record(nxt, ms[0])
        it = nxt + 1
    return out, nrb
