/*
 * flowmol_hip.h -- C ABI of the MI355X (gfx950) force / neighbour / integration path of Flowmol.
 *
 * Every entry point below replaces the body of one Fortran routine of the reference (file:line cited
 * per function, paths relative to the reference's root).  The Fortran driver keeps calling the
 * routines by their old names; a thin iso_c_binding shim forwards to these functions.
 *
 * Conventions
 *   - All functions return 0 on success and a nonzero FM_E* code on error; fm_last_error() returns
 *     the message.  The reference aborts through error_abort (src/messenger.MPI.f90:3376-3410);
 *     the shim calls error_abort(fm_last_error()) on a nonzero return.
 *   - Host arrays use the reference's layout: r, v, a are (3, np) column-major, i.e. molecule n's
 *     x,y,z at [3n], [3n+1], [3n+2] (src/modules.f90:389-395).  Molecule numbers are 0-based here
 *     (Fortran molno-1).
 *   - One context = one periodic domain on one GPU with its own HIP stream.  Calls on a context are
 *     stream-ordered and must come from one host thread; different contexts are independent.
 *   - No torch or HIP types cross this boundary.
 */
#ifndef FLOWMOL_HIP_H
#define FLOWMOL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FM_ABI_VERSION 4

#define FM_MAX_TYPES 4      /* moltype values 0..FM_MAX_TYPES-1 (Mie lookup tables) */
#define FM_MAX_FUNCY 4      /* max bonds per monomer, src/modules.f90:468 */
#define FM_MAX_PLANES 1024  /* MOP planes */

/* error codes */
#define FM_OK 0
#define FM_EINVAL 1         /* bad argument / parameter */
#define FM_EHIP 2           /* HIP runtime error */
#define FM_ESTATE 3         /* call out of order (e.g. forces before a neighbour list) */
#define FM_EESCAPE 4        /* escaped molecule, src/linklist.f90:2665-2671 */
#define FM_EOVERFLOW 5      /* neighbour / cell / bond capacity exceeded */
#define FM_EBOND 6          /* broken FENE bond, src/setup_set_parameters.f90:355 */
#define FM_ENOMEM 7
#define FM_EABORT 8         /* max(potenergymol) > 1e4 at a record step, src/simulation_record.f90:1271-1282 */

/* integration_algorithm / ensemble values as in src/modules.f90 */
#define FM_ENSEMBLE_NVE 0
#define FM_ENSEMBLE_TAG_MOVE 6

/* rebuild_criteria, src/simulation_checkrebuild.f90:83-122 */
#define FM_REBUILD_VMAX 0
#define FM_REBUILD_FIXED 1
#define FM_REBUILD_DISPLACEMENT 2

/* precision modes */
#define FM_PREC_FP64 0      /* everything in FP64 (parity path) */
#define FM_PREC_MIXED 1     /* pair law in FP32; positions, image shifts, MOP tests, accumulators FP64 */

/* integration_algorithm, src/modules.f90:112-113 */
#define FM_INTEGRATION_LFV 0
#define FM_INTEGRATION_VV 1

/* per-molecule tags (src/modules.f90:78-88); the tags without a thermostat are on this path */
#define FM_TAG_FREE 0
#define FM_TAG_FIXED 1
#define FM_TAG_FIXED_SLIDE 2
#define FM_TAG_TETH 3
#define FM_TAG_TETH_SLIDE 6

typedef struct fm_params {
    /* domain and lists: set_parameters_global_domain / set_parameters_cells
       (src/setup_set_parameters.f90:1534-1600, 1848-1903) */
    double domain[3];            /* globaldomain(1:3) -- one periodic domain, centred on 0 */
    double density;              /* input density (pressure formula, src/simulation_record.f90:1286) */
    double rcutoff;
    double delta_rneighbr[3];
    double delta_t;
    int32_t tplot;               /* record every tplot steps; also forced rebuild at mod(iter+1,tplot)==0 */
    int32_t teval;               /* virial every teval steps (= tplot unless boundary forces are on) */
    int32_t rebuild_criteria;    /* FM_REBUILD_* */
    int32_t fixed_rebuild;

    /* potentials: src/setup_set_parameters.f90:96-204, 348-380, 1198-1233 */
    int32_t potential_flag;      /* 0: pair potential only; 1: + FENE bonds (simulation_compute_forces_poly) */
    int32_t mie_potential;       /* 0: LJ 12-6 with mass 1; 1: Mie from the per-type tables below */
    int32_t ntypes;              /* used when mie_potential==1 */
    int32_t precision;           /* FM_PREC_* */
    double mass_lookup[FM_MAX_TYPES];
    double epsilon_lookup[FM_MAX_TYPES * FM_MAX_TYPES];   /* [ti*FM_MAX_TYPES+tj] */
    double lambdar_lookup[FM_MAX_TYPES * FM_MAX_TYPES];
    double lambdaa_lookup[FM_MAX_TYPES * FM_MAX_TYPES];
    double C_lookup[FM_MAX_TYPES * FM_MAX_TYPES];
    double alpha_lookup[FM_MAX_TYPES * FM_MAX_TYPES];
    double potshift_lookup[FM_MAX_TYPES * FM_MAX_TYPES];  /* E(rc) per pair, src/setup_set_parameters.f90:946 */
    double rcutoff2_lookup[FM_MAX_TYPES * FM_MAX_TYPES];  /* per-pair cutoff^2; 0 means use rcutoff^2 */
    double k_c, R_0;             /* FENE */

    /* ensemble and walls: src/molecular_properties.f90:833-915 */
    int32_t ensemble;            /* FM_ENSEMBLE_NVE or FM_ENSEMBLE_TAG_MOVE */
    int32_t integration_algorithm; /* FM_INTEGRATION_VV or FM_INTEGRATION_LFV.  Leap-frog applies the
                                    tethers inside the move, as the reference does; velocity Verlet
                                    (no tag_move in the reference) applies them with the forces. */
    double teth_k2, teth_k4, teth_k6;
    double wallslidev[3];        /* slide velocity of FM_TAG_TETH_SLIDE (times sign(y)) and FM_TAG_FIXED_SLIDE molecules */

    /* stress tallies */
    int32_t vflux_outflag;       /* 0 none; 1,2,3: MOP planes normal to x,y,z (the tally always uses y, as
                                    src/simulation_record.f90:7090-7109 does); 4: control-volume surface
                                    stresses Pxyface (src/simulation_record.f90:5281-5317) */
    int32_t pressure_outflag;    /* 1: per-molecule virial tensor rfmol (src/simulation_record.f90:6239-6254) */
    int32_t nplanes;             /* number of MOP planes (<= FM_MAX_PLANES) */
    int32_t _pad1;
    double planespacing;         /* plane positions: planespacing*(n-1)+shift-halfdomain, src/setup_set_parameters.f90:2298-2340 */

    /* capacities and launch shape (0 = automatic) */
    int32_t max_neighbours;      /* per-molecule neighbour list capacity */
    int32_t max_cell_np;         /* per-cell capacity check */
    int32_t brick[3];            /* cells per LDS brick along x,y,z (0,0,0: chosen from the occupancy) */
    int32_t _pad2;

    /* averaging bins of the control-volume tally (vflux_outflag 4): BIN2CELLRATIO bins per cell,
       src/setup_set_parameters.f90:1990-2011 */
    double binspercell[3];
    int32_t cv_conserve;         /* CV_conserve: 0 tally on record steps only; 1 (every step) is rejected */
    int32_t _pad3;

    /* SORT_FLAG, SORT_FREQ, SORTBLOCKSIZE of sort_mols (src/linklist.f90:2744-2865): renumber the
       molecules block by block along a curve through blocks of sortblocksize^3 cells */
    int32_t sort_flag;           /* 0 none; 1 blocks in x-fastest order; 2 blocks along a 3D Hilbert curve */
    int32_t sort_freq;           /* sort at every sort_freq-th call (rebuild) */
    int32_t sortblocksize;
    int32_t _pad4;
} fm_params;

typedef struct fm_ctx fm_ctx;   /* opaque */

/* Macroscopic properties, evaluate_macroscopic_properties, src/simulation_record.f90:1205-1314 */
typedef struct fm_macros {
    double potenergysum;         /* sum of potenergymol (LJ + poly) */
    double potenergysum_LJ;
    double potenergysum_POLY;
    double virial;               /* sum of virialmol */
    double mv2sum;               /* sum m v^2 (VV: v at the current step) */
    double vsum;                 /* sum m v (all components) */
    double msum;
    double kinenergy, potenergy, totenergy, temperature, pressure;
    double max_potenergymol;
} fm_macros;

/* ---- lifecycle: replaces set_parameters_allocate + setup_linklist
        (src/setup_set_parameters.f90:1248-1375, 1924-1959) ---------------------------------- */
int  fm_init(const fm_params* p, int32_t np, int32_t capacity, fm_ctx** out);
void fm_finalize(fm_ctx* ctx);
const char* fm_last_error(void);
int  fm_abi_version(void);

/* Derived geometry the reference computes in set_parameters_cells: ncells(3), cellsidelength(3),
   rneighbr2, rcutoff2, potshift (src/setup_set_parameters.f90:1848-1903, 1222). */
int  fm_get_geometry(const fm_ctx* ctx, int32_t ncells[3], double cellsidelength[3],
                     double* rneighbr2, double* rcutoff2, double* potshift);

/* Capacities of the device lists: neighbour-list entries per molecule, and molecules per staged
   brick region (LDS).  No counterpart in the reference, whose lists are linked lists.  Both grow
   when a build meets more; a fused loop's queued build that meets more repeats itself after
   growing them.  The region capacity is the largest region of the last measured binning raised to
   what the LDS share of a workgroup holds; FLOWMOL_TIGHT_REGIONS=1 in the environment at fm_init
   keeps it at the measured value. */
int  fm_get_capacity(const fm_ctx* ctx, int32_t* list_entries, int32_t* region_molecules);

/* ---- host <-> device state (arrays_MD, src/modules.f90:363-397) ---------------------------
   r, v, a: (3,np) AoS; a may be NULL (zeros).  moltype/tag/rtether may be NULL. */
int  fm_upload_state(fm_ctx* ctx, int32_t np, const double* r, const double* v, const double* a);
int  fm_upload_moltype(fm_ctx* ctx, const int32_t* moltype);
int  fm_upload_tags(fm_ctx* ctx, const int32_t* tag, const double* rtether);
/* polymer_info_MD monomer fields (src/modules.f90:472-481): chainID, subchainID, funcy,
   bin_bflag(4), glob_no.  glob_no may be NULL: molecule n then carries n (1-based), as one
   process numbers them at setup. */
int  fm_upload_monomers(fm_ctx* ctx, const int32_t* chainID, const int32_t* subchainID,
                        const int32_t* funcy, const int32_t* bin_bflag /* (4,np) */, const int32_t* glob_no);
int  fm_download_state(fm_ctx* ctx, double* r, double* v, double* a);
int  fm_download_tags(fm_ctx* ctx, double* rtether);
/* Per-molecule labels in the current molecule numbering, which sendmols and sort_mols change
   (src/messenger.MPI.f90:1460-1904, src/linklist.f90:2820-2860): tag codes, moltype (0-based
   device type), and the monomer fields chainID, subchainID, funcy, bin_bflag (4,np), glob_no.
   Any pointer may be NULL.  A host mirror reads them back after a rebuild. */
int  fm_download_labels(fm_ctx* ctx, int32_t* tag, int32_t* moltype, int32_t* chainID, int32_t* subchainID,
                        int32_t* funcy, int32_t* bin_bflag, int32_t* glob_no);
/* potenergymol_LJ and potenergymol_POLY (src/modules.f90:373-374) as of the last record step */
int  fm_download_potenergy_parts(fm_ctx* ctx, double* potenergymol_LJ, double* potenergymol_POLY);
/* per-molecule potenergymol, virialmol (src/modules.f90:371-378) as of the last record step */
int  fm_download_molecular(fm_ctx* ctx, double* potenergymol, double* virialmol);
/* rfmol(np,3,3) when pressure_outflag==1, molecule-major: [9n + 3*i + j] */
int  fm_download_rfmol(fm_ctx* ctx, double* rfmol);
int  fm_get_np(const fm_ctx* ctx, int32_t* np);
double fm_simtime(const fm_ctx* ctx);
/* simtime of a context resumed from a checkpoint (the final_state header carries it) */
int  fm_set_simtime(fm_ctx* ctx, double simtime);

/* ---- the hot path: one call per reference routine ------------------------------------------ */
/* simulation_move_particles_vv(pass_num), src/simulation_move_particles_vv.f90:44-193 */
int  fm_move_particles_vv(fm_ctx* ctx, int32_t pass_num);
/* simulation_move_particles_lfv, src/simulation_move_particles_lfv.f90:38-177 (NVE) and 308-490
   (tag_move: free, fixed, fixed_slide, teth, teth_slide; tether_force is applied here) */
int  fm_move_particles_lfv(fm_ctx* ctx);
/* simulation_checkrebuild(rebuild), src/simulation_checkrebuild.f90:17-174 (criteria 0,1,2) */
int  fm_checkrebuild(fm_ctx* ctx, int32_t iter, int32_t* rebuild);
/* linklist_deallocateall, src/linklist.f90:2872-2963 */
int  fm_linklist_deallocateall(fm_ctx* ctx);
/* sendmols on one periodic rank, src/messenger.MPI.f90:1383-1421,1460-1904: wrap and renumber */
int  fm_sendmols(fm_ctx* ctx);
/* sort_mols (src/linklist.f90:2744-2865): SORT_FLAG 1 or 2 renumbers the molecules by the block
   curve every SORT_FREQ rebuilds (fm_params.sort_flag, sort_freq, sortblocksize); 0 does nothing */
int  fm_sort_mols(fm_ctx* ctx);
/* assign_to_cell, src/linklist.f90:2645-2676 */
int  fm_assign_to_cell(fm_ctx* ctx);
/* messenger_updateborders(rebuild), src/messenger.MPI.f90:594-628: halo images are implicit
   (image-shift codes in the list), so this is a no-op kept for the call sequence. */
int  fm_updateborders(fm_ctx* ctx, int32_t rebuild);
/* assign_to_neighbourlist (FORCE_LIST 3), src/linklist.f90:2710-2737, 660-804, 1118-1180;
   with potential_flag==1 also rebuilds bond lists (src/linklist.f90:2542-2584) */
int  fm_assign_to_neighbourlist(fm_ctx* ctx);
/* simulation_compute_forces, src/simulation_compute_forces.f90:182-275, 628-789
   (+ tether_force for tagged molecules, src/molecular_properties.f90:833-915) */
int  fm_compute_forces(fm_ctx* ctx, int32_t iter);
/* evaluate_macroscopic_properties, src/simulation_record.f90:1205-1314 */
int  fm_evaluate_macroscopic_properties(fm_ctx* ctx, fm_macros* out);
/* Pxy_plane(3,nplanes) accumulated by pressure_tensor_forces_MOP since the last reset
   (src/simulation_record.f90:7068-7112) */
int  fm_get_mop(fm_ctx* ctx, double* Pxy_plane, int32_t nplanes);
int  fm_reset_mop(fm_ctx* ctx);
/* Control-volume bins: nbins, nhb (halo bins per side) and binsize as set_parameters_allocate
   computes them for one process (src/setup_set_parameters.f90:1990-2011); nbinso = nbins+2*nhb. */
int  fm_get_cv_bins(const fm_ctx* ctx, int32_t nbins[3], int32_t nhb[3], double binsize[3]);
/* Pxyface(nbinso(1),nbinso(2),nbinso(3),3,6), column-major, as control_volume_stresses_opt leaves
   it (src/simulation_record.f90:5281-5317, 4878-5064) since the last reset, before the halo fold
   and scaling of surface_stress_io (src/parallel_io.MPI2.f90:4075-4134).  n = element count. */
int  fm_get_cv_stress(fm_ctx* ctx, double* pxyface, int64_t n);
int  fm_reset_cv_stress(fm_ctx* ctx);

/* ---- record-step sums without allocation ----------------------------------------------------
   No counterpart in the reference: a variant of fm_evaluate_macroscopic_properties for callers
   that cannot allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  It takes caller-held device arrays in molecule order
   (v, and a for the leap-frog half-step shift or NULL; (3,np) like arrays_MD), the per-molecule
   potenergymol_LJ, potenergymol_POLY (or NULL), virialmol and masses (or NULL for 1), and writes
   the eight sums (PE LJ, PE POLY, PE, virial, sum m v^2, sum m v, sum m, max PE) to sums[8] in
   device memory, stream-ordered on `stream` (a hipStream_t, NULL for the default stream), with no
   host synchronisation.  Its partial sums use a static device buffer for at most 16,777,216
   molecules (larger np is refused); every call on one device shares it, so calls that overlap on
   one device race.  fm_derive_macros turns the sums into fm_macros on the host. */
int  fm_macros_sums_noalloc(int32_t np, const double* v, const double* a, double delta_t,
                            const double* potenergymol_LJ, const double* potenergymol_POLY,
                            const double* virialmol, const double* mass, double* sums, void* stream);
int  fm_derive_macros(const double sums[8], int32_t np, double density, int32_t potential_flag, fm_macros* out);

/* ---- fused time loop: md_advance_vv (src/md.f90:143-163) nsteps times from iter0, without a
   host round trip per step.  Same kernels and arithmetic as the call-per-routine sequence.
   If macros != NULL, record-step macros are written to macros[k] for k-th record step
   (capacity n_macros). */
int  fm_run_vv(fm_ctx* ctx, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
               int32_t* n_rebuilds);
/* md_advance_lfv (src/md.f90:109-139) nsteps times from iter0: compute_forces, record (macros on
   record steps, leap-frog kinetic terms), move_particles_lfv, checkrebuild and rebuild. */
int  fm_run_lfv(fm_ctx* ctx, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
                int32_t* n_rebuilds);

/* ---- inspection (parity tests) --------------------------------------------------------------
   Cell membership: cell_of_mol[n] = flattened 1-based-with-halo cell index of molecule n,
   ((kcell-1)*(ncy+2) + (jcell-1))*(ncx+2) + (icell-1).  */
int  fm_download_cells(fm_ctx* ctx, int32_t* cell_of_mol);
/* Full neighbour list as directed entries (molno i, molno j, shift code s) with
   s = (sx+1)*9+(sy+1)*3+(sz+1), the image of j used being r_j + (sx,sy,sz)*domain.
   Returns the entry count in *n_entries; arrays may be NULL to query the count. */
int  fm_download_neighbourlist(fm_ctx* ctx, int32_t* mol_i, int32_t* mol_j, int32_t* shift,
                               int64_t capacity, int64_t* n_entries);
/* Bond lists: bond[b + FM_MAX_FUNCY*n] = partner molno or -1, with shift code as above. */
int  fm_download_bonds(fm_ctx* ctx, int32_t* bond, int32_t* bond_shift);

/* ---- replica batches ---------------------------------------------------------------------------
   Many independent contexts (e.g. ensemble replicas of uneven size) advanced together: each step
   is one packed launch of VV pass 1 + displacement check over all members and one packed launch
   of the forces (+ VV pass 2) over all members' bricks, captured as HIP graphs, with one host
   synchronisation per step.  Results are identical to advancing each context alone (fm_run_vv).
   No counterpart in the reference (one process runs one domain); it replaces running many
   reference processes side by side.  Members share precision, pair law and bond/tether kind and
   use rebuild criterion 1 or 2; while a member is in a batch, drive it only through the batch. */
typedef struct fm_batch fm_batch;   /* opaque */
int  fm_batch_create(fm_batch** out);
void fm_batch_destroy(fm_batch* b);
int  fm_batch_use_graphs(fm_batch* b, int32_t enable);
/* ctx must have its initial neighbour list and forces; iter0 = iteration number of its next step */
int  fm_batch_add(fm_batch* b, fm_ctx* ctx, int32_t iter0);
int  fm_batch_remove(fm_batch* b, fm_ctx* ctx);
/* nsteps velocity Verlet steps of every member, queued in windows of up to 16 steps with one
   host synchronisation per window (rebuild decisions on the device; see fm_batch.hip) */
int  fm_batch_run(fm_batch* b, int32_t nsteps, int32_t* n_rebuilds);
/* totals since fm_batch_create: member list rebuilds, and host synchronisations (windows) */
int  fm_batch_counters(const fm_batch* b, int64_t* rebuilds, int64_t* host_syncs);
int  fm_batch_size(const fm_batch* b, int32_t* n);
int  fm_batch_get_iter(const fm_batch* b, const fm_ctx* ctx, int32_t* iter);
/* device bytes fm_init would allocate (admission control by HBM footprint) */
int  fm_memory_footprint(const fm_params* p, int32_t np, int64_t* bytes);

/* ---- timing / synchronisation -------------------------------------------------------------- */
int  fm_synchronize(fm_ctx* ctx);
/* Average duration in ms of each kernel class over the last fm_run_vv (HIP events on the
   context's stream).  names is a '\n'-separated list of kernel class names. */
int  fm_kernel_timing(fm_ctx* ctx, int32_t enable);
int  fm_get_kernel_timing(fm_ctx* ctx, char* names, int32_t names_cap, double* avg_ms,
                          int64_t* counts, int32_t max_classes, int32_t* n_classes);

#ifdef __cplusplus
}
#endif
#endif /* FLOWMOL_HIP_H */
