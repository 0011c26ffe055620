"""Parameters and synthetic initial states for the force / neighbour / integration path.

The derived quantities follow the reference's setup (paths relative to the reference's root):
  globaldomain, initialunitsize   src/setup_set_parameters.f90:1543-1569
  FCC lattice positions           src/setup_initialise_microstate.f90:244-304
  Mie constant C and potshift     src/setup_set_parameters.f90:934-946
  MOP plane spacing / count       src/setup_set_parameters.f90:2298-2340
Velocities are this package's own seeded generator (Gaussian, zero net momentum, rescaled to the
target temperature); the reference's gfortran random_number stream is not reproduced.
"""
import math

import numpy as np

from ._abi import (FM_MAX_TYPES, FM_PREC_FP64, FM_REBUILD_DISPLACEMENT, FM_ENSEMBLE_NVE,
                   FM_INTEGRATION_VV, FmParams)

WCA_CUT = 2.0 ** (1.0 / 6.0)


def fcc_globaldomain(nunits, density):
    """globaldomain(ixyz) = initialnunits(ixyz)/((density/4)**(1/nd))."""
    f = (density / 4.0) ** (1.0 / 3)
    return np.array([n / f for n in nunits], dtype=np.float64)


def fcc_lattice(nunits, density):
    """Molecule positions of setup_initialise_lattice on one rank, (np, 3) float64.

    Loop order nz, nx, ny; 4-site basis; c = (n - 0.75)*a; r = rc - halfdomain.
    """
    gd = fcc_globaldomain(nunits, density)
    a = gd / np.array(nunits, dtype=np.float64)
    half = 0.5 * gd
    out = np.empty((4 * nunits[0] * nunits[1] * nunits[2], 3), dtype=np.float64)
    n = 0
    # vectorised over ny and the basis; loop order is preserved by construction
    ny = np.arange(1, nunits[1] + 1, dtype=np.float64)
    c2 = (ny - 0.75) * a[1]
    for nz in range(1, nunits[2] + 1):
        c3 = (nz - 0.75) * a[2]
        for nx in range(1, nunits[0] + 1):
            c1 = (nx - 0.75) * a[0]
            block = np.empty((nunits[1], 4, 3))
            block[:, :, 0] = c1
            block[:, :, 1] = c2[:, None]
            block[:, :, 2] = c3
            block[:, 1, 0] = c1 + 0.5 * a[0]
            block[:, 1, 2] = c3 + 0.5 * a[2]
            block[:, 2, 1] = c2 + 0.5 * a[1]
            block[:, 2, 2] = c3 + 0.5 * a[2]
            block[:, 3, 0] = c1 + 0.5 * a[0]
            block[:, 3, 1] = c2 + 0.5 * a[1]
            block -= half
            m = 4 * nunits[1]
            out[n:n + m] = block.reshape(m, 3)
            n += m
    return out


def seeded_velocities(np_, temperature, seed, mass=None):
    """Gaussian velocities from a seeded PCG64 stream, zero net momentum, scaled so that
    sum(m v^2)/(3N) == temperature."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((np_, 3))
    m = np.ones(np_) if mass is None else np.asarray(mass, dtype=np.float64)
    p = (m[:, None] * v).sum(axis=0) / m.sum()
    v -= p
    mv2 = float((m[:, None] * v * v).sum())
    v *= math.sqrt(temperature * 3 * np_ / mv2)
    return np.ascontiguousarray(v)


def mie_C(lr, la):
    """C_lookup = (lr/(lr-la)) * (lr/la)**(la/(lr-la))."""
    return (lr / (lr - la)) * (lr / la) ** (la / (lr - la))


def mie_energy(invrij2, C, eps, lr, la, alpha=1.0, potshift=0.0):
    return C * eps * (invrij2 ** (0.5 * lr) - alpha * invrij2 ** (0.5 * la)) - potshift


def make_params(domain, density, rcutoff, delta=0.3, dt=0.005, tplot=10, teval=None,
                rebuild_criteria=FM_REBUILD_DISPLACEMENT, fixed_rebuild=20, precision=FM_PREC_FP64,
                potential_flag=0, k_c=30.0, R_0=1.5, ensemble=FM_ENSEMBLE_NVE,
                vflux_outflag=0, nplanes=None, planespacing=None, pressure_outflag=0,
                max_neighbours=0, binspercell=1.0, integration=None, sort_flag=0, sort_freq=1,
                sortblocksize=2):
    p = FmParams()
    for d in range(3):
        p.domain[d] = float(domain[d])
        p.delta_rneighbr[d] = float(delta)
    p.density = float(density)
    p.rcutoff = float(rcutoff)
    p.delta_t = float(dt)
    p.tplot = int(tplot)
    p.teval = int(tplot if teval is None else teval)
    p.rebuild_criteria = int(rebuild_criteria)
    p.fixed_rebuild = int(fixed_rebuild)
    p.potential_flag = int(potential_flag)
    p.mie_potential = 0
    p.ntypes = 1
    p.precision = int(precision)
    for t in range(FM_MAX_TYPES):
        p.mass_lookup[t] = 1.0
    p.k_c = float(k_c)
    p.R_0 = float(R_0)
    p.ensemble = int(ensemble)
    p.integration_algorithm = FM_INTEGRATION_VV if integration is None else int(integration)
    p.vflux_outflag = int(vflux_outflag)
    p.pressure_outflag = int(pressure_outflag)
    if 1 <= vflux_outflag <= 3:
        pn = vflux_outflag - 1
        if planespacing is None:
            # reference: planespacing = cellsidelength(2), nplanes = floor(domain/planespacing)
            ncy = math.floor(domain[1] / (rcutoff + delta))
            planespacing = domain[1] / ncy
        if nplanes is None:
            nplanes = math.floor(domain[pn] / planespacing)
        p.planespacing = float(planespacing)
        p.nplanes = int(nplanes)
    for d in range(3):  # BIN2CELLRATIO, used by the control-volume tally (vflux_outflag 4)
        p.binspercell[d] = float(binspercell[d] if hasattr(binspercell, "__len__") else binspercell)
    p.max_neighbours = int(max_neighbours)
    p.sort_flag, p.sort_freq, p.sortblocksize = int(sort_flag), int(sort_freq), int(sortblocksize)
    return p


def set_mie_types(p, types, rcutoff=None):
    """Fill the Mie lookup tables from {(ti, tj): dict(eps, lr, la, alpha, rc)} and masses.

    potshift_lookup(i,j) = get_energy(1/rc**2, i, j) with the pair's own cutoff.
    """
    p.mie_potential = 1
    ntypes = 1 + max(max(k) for k in types)
    p.ntypes = ntypes
    for (ti, tj), d in types.items():
        for (a, b) in ((ti, tj), (tj, ti)):
            k = a * FM_MAX_TYPES + b
            lr, la = float(d["lr"]), float(d["la"])
            eps = float(d.get("eps", 1.0))
            alpha = float(d.get("alpha", 1.0))
            C = mie_C(lr, la)
            rc = float(d.get("rc", rcutoff if rcutoff is not None else p.rcutoff))
            p.epsilon_lookup[k] = eps
            p.lambdar_lookup[k] = lr
            p.lambdaa_lookup[k] = la
            p.C_lookup[k] = C
            p.alpha_lookup[k] = alpha
            p.potshift_lookup[k] = mie_energy(1.0 / rc ** 2, C, eps, lr, la, alpha)
            p.rcutoff2_lookup[k] = rc * rc
    return p


def config_fcc(nunits, density, rcutoff, seed, temperature=1.0, **kw):
    """FCC lattice + seeded velocities + parameters for one periodic domain."""
    domain = fcc_globaldomain(nunits, density)
    r = fcc_lattice(nunits, density)
    v = seeded_velocities(r.shape[0], temperature, seed)
    p = make_params(domain, density, rcutoff, **kw)
    return p, r, v


def config_c1(**kw):
    """2048-molecule WCA NVE case (FCC 8^3, density 0.8, rc 2^(1/6), delta 0.3, dt 0.005)."""
    kw.setdefault("tplot", 100)
    return config_fcc((8, 8, 8), 0.8, WCA_CUT, seed=1, **kw)


def config_c2(nunits=64, **kw):
    """1,048,576-molecule LJ NVE case (FCC 64^3, density 0.8, rc 2.5, delta 0.3, dt 0.005)."""
    kw.setdefault("tplot", 10)
    return config_fcc((nunits, nunits, nunits), 0.8, 2.5, seed=2, **kw)


# ---- polymer melt: dense FENE chains on the FCC lattice ---------------------------------------------
def chain_topology(nchains, nmono):
    """polymer_info_MD monomer fields for linear chains: chainID, subchainID, funcy, bin_bflag(4).

    Consecutive molecules form chains of nmono beads (setup_lattice_dense_FENE_info,
    src/setup_initialise_microstate.f90:347-425); bead s is bonded to s-1 and s+1, recorded as
    bits of bin_bflag (get_bondflag, src/modules.f90:597-610: group ceil(s/31), bit mod(s,31)-1).
    """
    n = nchains * nmono
    chain = np.repeat(np.arange(1, nchains + 1), nmono).astype(np.int32)
    sub = np.tile(np.arange(1, nmono + 1), nchains).astype(np.int32)
    funcy = np.where((sub == 1) | (sub == nmono), 1, 2).astype(np.int32)
    if nmono == 1:
        funcy[:] = 0
    bflag = np.zeros((n, 4), np.int64)
    for nb in (-1, 1):
        t = sub + nb
        ok = (t >= 1) & (t <= nmono)
        g = (t - 1) // 31
        b = (t - 1) % 31
        for grp in range(4):
            m = ok & (g == grp)
            bflag[m, grp] |= (1 << b[m])
    return chain, sub, funcy, bflag.astype(np.int32)


def random_walk_chains(domain, nchains, nmono, seed, bond=0.97, min_sep=0.8, max_tries=200):
    if not min_sep < bond:
        raise ValueError("min_sep must be below the bond length")
    """Seeded random-walk chains in a periodic box.  Chain starts are distinct sites of the FCC
    lattice that fills the box at the same density (drawn with a seeded permutation); bead s+1
    lies at distance `bond` from bead s in a uniformly random direction, redrawn until it is at
    least `min_sep` from every bead placed before it (all chains grow one bead per round, and
    candidates of one round are also checked against each other).  A bead that finds no such
    direction in max_tries draws (2-3 % of the beads at density 0.85) is placed with the
    separation lowered by 0.05 at a time.  Returns (nchains*nmono, 3) positions in [-L/2, L/2) with chain c's beads
    at rows c*nmono .. c*nmono+nmono-1."""
    from scipy.spatial import cKDTree

    L = np.asarray(domain, np.float64)
    rng = np.random.default_rng(seed)
    ntot = nchains * nmono
    nunits = [int(round(x)) for x in L * ((ntot / (L[0] * L[1] * L[2])) / 4.0) ** (1.0 / 3.0)]
    sites = fcc_lattice(nunits, 4.0 * nunits[0] * nunits[1] * nunits[2] / (L[0] * L[1] * L[2]))
    if sites.shape[0] < nchains:
        raise ValueError("box too small for the chain starts")
    pos = np.zeros((nchains, nmono, 3))
    pos[:, 0] = sites[rng.permutation(sites.shape[0])[:nchains]]

    def wrap(x):
        return x - L * np.floor((x + 0.5 * L) / L)

    for s in range(1, nmono):
        tree = cKDTree(wrap(pos[:, :s].reshape(-1, 3)) + 0.5 * L, boxsize=L)
        extra = np.zeros((0, 3))  # beads of this round placed since the tree was built
        todo = np.arange(nchains)
        sep = min_sep
        while todo.size and sep > 0.5 * min_sep:
            for _ in range(max_tries):
                d = rng.standard_normal((todo.size, 3))
                d /= np.linalg.norm(d, axis=1)[:, None]
                cand = pos[todo, s - 1] + bond * d
                cw = wrap(cand) + 0.5 * L
                cw = np.where(cw >= L, cw - L, cw)
                # every placed bead (the chain's own previous one lies at `bond` > sep)
                dn, _ = tree.query(cw)
                ok = dn >= sep - 1e-12
                if extra.shape[0]:
                    de, _ = cKDTree(extra, boxsize=L).query(cw)
                    ok &= de >= sep - 1e-12
                # candidates of one draw against each other: the later one of a close pair redraws
                for a, b in cKDTree(cw, boxsize=L).query_pairs(sep - 1e-12):
                    ok[max(a, b)] = False
                pos[todo[ok], s] = cand[ok]
                extra = np.concatenate([extra, cw[ok]])
                if extra.shape[0] > 2048:
                    tree = cKDTree(np.concatenate([tree.data, extra]), boxsize=L)
                    extra = np.zeros((0, 3))
                todo = todo[~ok]
                if todo.size == 0:
                    break
            sep -= 0.05
        if todo.size:
            raise RuntimeError(f"{todo.size} chains could not place bead {s + 1}")
    return np.ascontiguousarray(wrap(pos.reshape(ntot, 3)))


def relax_overlaps(r, domain, nmono, bond=0.97, target=0.85, floor=0.8, max_iter=200):
    """Move beads of chains (nmono consecutive rows each) until no two non-bonded beads are closer
    than `floor`: every non-bonded pair closer than `target` is pushed apart symmetrically to
    `target`, then the bonds are pulled back towards `bond` (a few symmetric sweeps), repeatedly.
    Ends with every bond within 0.05 of `bond`; raises if max_iter sweeps do not get there."""
    from scipy.spatial import cKDTree

    L = np.asarray(domain, np.float64)
    r = np.array(r, dtype=np.float64)
    n = r.shape[0]
    bi = np.arange(n - 1)
    bi = bi[(bi % nmono) != nmono - 1]  # bead i bonded to i+1

    def wrap(x):
        return x - L * np.floor((x + 0.5 * L) / L)

    def min_image(d):
        return d - L * np.round(d / L)

    for _ in range(max_iter):
        pairs = cKDTree(wrap(r) + 0.5 * L, boxsize=L).query_pairs(target, output_type="ndarray")
        if pairs.size:
            i, j = pairs[:, 0], pairs[:, 1]
            keep = ~((j == i + 1) & ((i % nmono) != nmono - 1))
            i, j = i[keep], j[keep]
        else:
            i = j = np.zeros(0, np.int64)
        d = min_image(r[i] - r[j])
        dist = np.linalg.norm(d, axis=1)
        bd = np.linalg.norm(min_image(r[bi + 1] - r[bi]), axis=1)
        if (dist.size == 0 or dist.min() >= floor) and np.abs(bd - bond).max() <= 0.05:
            return wrap(r)
        push = (0.5 * (target - dist) / np.maximum(dist, 1e-12))[:, None] * d
        dr = np.zeros_like(r)
        np.add.at(dr, i, push)
        np.add.at(dr, j, -push)
        r = r + dr
        for _ in range(4):
            d = min_image(r[bi + 1] - r[bi])
            ln = np.linalg.norm(d, axis=1)
            c = (0.5 * (ln - bond) / ln)[:, None] * d
            db = np.zeros_like(r)
            np.add.at(db, bi, c)
            np.add.at(db, bi + 1, -c)
            r = r + db
    raise RuntimeError("overlaps remain after relaxation")


def config_polymer(nunits=(6, 8, 6), density=0.85, nmono=8, seed=4, k_c=30.0, R_0=1.5, dt=0.002,
                   mie_fraction=0.5, tplot=10, init="lattice", **kw):
    """FENE chains of nmono beads with WCA between all beads, and a second species (whole chains,
    chosen with a seeded shuffle) whose mutual interaction is Mie 14-6 (eps 1, rc 2.5); species 0
    pairs and mixed pairs are WCA.  The box holds the FCC lattice of nunits at the density.
    init "lattice": consecutive lattice sites form the chains (4*nunits[1] divisible by nmono);
    init "random_walk": seeded random-walk chains (random_walk_chains), then the close non-bonded
    pairs its densest spots leave pushed apart to at least 0.8 (relax_overlaps): a start a FENE
    chain survives (pairs nearer than that release enough energy to stretch bonds past R0)."""
    if init == "lattice" and (4 * nunits[1]) % nmono:
        raise ValueError("4*nunits_y must be divisible by the chain length")
    domain = fcc_globaldomain(nunits, density)
    r = fcc_lattice(nunits, density)
    n = r.shape[0]
    if init == "random_walk":
        r = relax_overlaps(random_walk_chains(domain, n // nmono, nmono, seed), domain, nmono)
    elif init != "lattice":
        raise ValueError(f"unknown init {init!r}")
    nchains = n // nmono
    v = seeded_velocities(n, 1.0, seed)
    p = make_params(domain, density, 2.5, dt=dt, tplot=tplot, potential_flag=1, k_c=k_c, R_0=R_0, **kw)
    set_mie_types(p, {(0, 0): dict(lr=12.0, la=6.0, eps=1.0, rc=WCA_CUT),
                      (0, 1): dict(lr=12.0, la=6.0, eps=1.0, rc=WCA_CUT),
                      (1, 1): dict(lr=14.0, la=6.0, eps=1.0, rc=2.5)})
    rng = np.random.default_rng(seed)
    chain_species = np.zeros(nchains, np.int32)
    chain_species[rng.permutation(nchains)[: int(round(mie_fraction * nchains))]] = 1
    moltype = np.repeat(chain_species, nmono).astype(np.int32)
    mono = chain_topology(nchains, nmono)
    return p, r, v, moltype, mono


# ---- Couette flow between tethered walls -------------------------------------------------------------
def config_couette(nunits=(8, 8, 8), density=0.8, rcutoff=2.5, wall_units=3, wallslidev=1.0, nplanes=64,
                   seed=3, tplot=1, precision=0, **kw):
    """Tethered FCC walls `wall_units` lattice units thick at the bottom and top of y; the top wall
    slides at wallslidev along x (TETHERCOEFFICIENTS 0 5000 5000000 as src/default.in); MOP stress
    on nplanes planes normal to y every record step."""
    from ._abi import FM_ENSEMBLE_TAG_MOVE, FM_TAG_FREE, FM_TAG_TETH, FM_TAG_TETH_SLIDE
    domain = fcc_globaldomain(nunits, density)
    r = fcc_lattice(nunits, density)
    n = r.shape[0]
    a = domain[1] / nunits[1]
    h = 0.5 * domain[1]
    tag = np.full(n, FM_TAG_FREE, np.int32)
    tag[r[:, 1] < -h + wall_units * a] = FM_TAG_TETH
    tag[r[:, 1] > h - wall_units * a] = FM_TAG_TETH_SLIDE
    v = seeded_velocities(n, 1.0, seed)
    v[tag != FM_TAG_FREE] = 0.0
    spacing = domain[1] / nplanes
    p = make_params(domain, density, rcutoff, tplot=tplot, ensemble=FM_ENSEMBLE_TAG_MOVE, vflux_outflag=2,
                    nplanes=nplanes, planespacing=spacing, precision=precision, **kw)
    p.teth_k2, p.teth_k4, p.teth_k6 = 0.0, 5000.0, 5.0e6
    p.wallslidev[0] = wallslidev
    return p, r, v, tag, r.copy()
