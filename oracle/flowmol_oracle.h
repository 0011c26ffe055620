/*
 * flowmol_oracle.h -- CPU restatement of the reference's force / neighbour / integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP implementation: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product path never
 * calls it.
 *
 * It restates, step for step and in the same floating-point operation order, the single-rank
 * behaviour of the reference Fortran (paths relative to the reference's root):
 *   linked cell lists            src/linklist.f90:2645-2702 (head insertion, 1596-1630)
 *   halo images                  src/messenger.MPI.f90:594-628, 880-1377 (faces, edges, corners)
 *   half neighbour list          src/linklist.f90:660-804, 1118-1180, 1636-1668
 *   bond discovery               src/linklist.f90:2542-2584, src/modules.f90:597-610
 *   forces                       src/simulation_compute_forces.f90:182-275, 280-341, 628-789
 *   pair laws                    src/setup_set_parameters.f90:96-204, 348-380, 1198-1233
 *   MOP / virial tensor          src/simulation_record.f90:7068-7112, 6239-6254
 *   velocity Verlet              src/simulation_move_particles_vv.f90:64-69, 132-135
 *   rebuild check                src/simulation_checkrebuild.f90:17-174
 *   sendmols (1 rank)            src/messenger.MPI.f90:1383-1904
 *   macroscopic properties       src/simulation_record.f90:1205-1314
 *   tether force                 src/molecular_properties.f90:833-915
 * The pins that check each function are listed in tests/test_oracle.py.
 */
#ifndef FLOWMOL_ORACLE_H
#define FLOWMOL_ORACLE_H

#include <stdint.h>
#include "../include/flowmol_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state* orc_create(const fm_params* p, int32_t np, int32_t capacity);
void orc_destroy(orc_state* s);
const char* orc_last_error(void);

void orc_geometry(const orc_state* s, int32_t ncells[3], double cellside[3], double* rneighbr2,
                  double* rcutoff2, double* potshift);

int orc_set_state(orc_state* s, int32_t np, const double* r, const double* v, const double* a);
int orc_set_moltype(orc_state* s, const int32_t* moltype);
int orc_set_tags(orc_state* s, const int32_t* tag, const double* rtether);
int orc_set_monomers(orc_state* s, const int32_t* chainID, const int32_t* subchainID,
                     const int32_t* funcy, const int32_t* bin_bflag);
int orc_get_state(const orc_state* s, double* r, double* v, double* a);
int orc_get_tags(const orc_state* s, double* rtether);
int orc_get_molecular(const orc_state* s, double* potenergymol, double* virialmol);
int orc_get_rfmol(const orc_state* s, double* rfmol);
int32_t orc_np(const orc_state* s);

/* reference routines */
int orc_linklist_deallocateall(orc_state* s);
int orc_sendmols(orc_state* s);
int orc_sort_mols(orc_state* s);
int orc_get_hcurve(const orc_state* s, int32_t* hcurve);
int orc_assign_to_cell(orc_state* s);
int orc_updateborders(orc_state* s, int32_t rebuild);
int orc_assign_to_neighbourlist(orc_state* s);
int orc_compute_forces(orc_state* s, int32_t iter);
int orc_compute_forces_allpairs(orc_state* s, int32_t iter);
int orc_move_particles_vv(orc_state* s, int32_t pass_num);
int orc_move_particles_lfv(orc_state* s);
int orc_run_lfv(orc_state* s, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
                int32_t* n_rebuilds);
int orc_checkrebuild(orc_state* s, int32_t iter, int32_t* rebuild);
int orc_macros(orc_state* s, fm_macros* out);
int orc_get_mop(const orc_state* s, double* Pxy_plane);
int orc_reset_mop(orc_state* s);
int orc_get_cv_bins(const orc_state* s, int32_t nbins[3], int32_t nhb[3], double binsize[3]);
int orc_get_cv_stress(const orc_state* s, double* pxyface);
int orc_reset_cv_stress(orc_state* s);
int orc_get_planes(const orc_state* s, double* planes);

/* md_advance_vv (src/md.f90:143-163) nsteps times; macros of record steps into macros[] */
int orc_run_vv(orc_state* s, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
               int32_t* n_rebuilds);

/* inspection */
int orc_get_cells(const orc_state* s, int32_t* cell_of_mol);
/* neighbour entries in list-walk order: owner molno, partner source molno, shift code
   (sx+1)*9+(sy+1)*3+(sz+1) of the partner image (0 shift for interior partners). */
int64_t orc_neighbour_count(const orc_state* s);
int orc_get_neighbourlist(const orc_state* s, int32_t* mol_i, int32_t* mol_j, int32_t* shift);
int orc_get_bonds(const orc_state* s, int32_t* bond, int32_t* bond_shift);
int32_t orc_halo_np(const orc_state* s);
/* per-molecule sum over in-cutoff pair entries of |accijmag * rij| (force scale for tolerances) */
int orc_get_force_scale(const orc_state* s, double* scale);

/* Brute-force O(N^2 * 27) neighbour set with the same image arithmetic, for pinning the cell
   search: every directed (i, j, s) with |r_i - (r_j + s*L)|^2 < rneighbr2 where j's cell is
   adjacent to i's cell through s.  Returns the count; arrays may be NULL. */
int64_t orc_bruteforce_pairs(const orc_state* s, int32_t* mol_i, int32_t* mol_j, int32_t* shift,
                             int64_t capacity);

#ifdef __cplusplus
}
#endif
#endif
