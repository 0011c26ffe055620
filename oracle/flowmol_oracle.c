/*
 * flowmol_oracle.c -- CPU restatement of the reference's single-rank force / neighbour /
 * integration path.  TEST INFRASTRUCTURE ONLY (see flowmol_oracle.h): the checker for the HIP
 * implementation, never part of the product path.
 *
 * Every routine follows the reference's loop order and floating-point operation order, cited
 * file:line (relative to the reference's root).  It is compiled with -ffp-contract=off so that no
 * multiply-add is fused, as gfortran -O3 on x86-64 without -mfma does not fuse them.
 *
 * Integer powers follow gcc's __builtin_powi expansion (the code gfortran emits for x**n with a
 * constant integer n): x2=x*x, x3=x*x2, x4=x2*x2, x6=x3*x3, x7=x3*x4, x12=x6*x6.
 */
#include "flowmol_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

static int fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return 1;
}

const char* orc_last_error(void) { return g_err; }

/* gcc powi expansion of integer powers */
static inline double p3(double x) { return x * (x * x); }
static inline double p4(double x) { double x2 = x * x; return x2 * x2; }
static inline double p6(double x) { double x3 = x * (x * x); return x3 * x3; }
static inline double p7(double x) { double x2 = x * x; return (x * x2) * (x2 * x2); }
static inline double p12(double x) { double x3 = x * (x * x); double x6 = x3 * x3; return x6 * x6; }

/* Fortran intrinsics */
static inline int f_ceiling(double x) { return (int)ceil(x); }
static inline int f_nint(double x) { return (int)lround(x); }
static inline double f_sign(double a, double b) { return copysign(fabs(a), b); }
static inline int f_isign(int a, int b) { return b >= 0 ? abs(a) : -abs(a); }
static inline double dp_heaviside(double x) { return ceil(f_sign(0.5, x)); } /* src/functions_lib.f90:2137-2146 */

typedef struct nbnode { int32_t mol; int32_t next; } nbnode;

struct orc_state {
    fm_params p;
    int np, cap;
    double *r, *v, *a;            /* (3, cap) */
    double *pe, *pe_lj, *pe_poly, *vir, *fscale;
    double *rfmol;                /* (cap, 3, 3) molecule-major */
    int32_t* moltype;
    int32_t* tag;
    double *rtether, *slidev;
    /* monomer info (cap) */
    int32_t *chainID, *subchainID, *funcy, *bflag;   /* bflag: 4 per molecule */
    int32_t *bond, *bondcount;    /* bond[b + FM_MAX_FUNCY*n] (0-based molno, -1 = none) */
    /* geometry */
    int ncells[3], nx2, ny2, nz2, ncelltot;
    double halfdomain[3], cellside[3], rneighbr2, rcutoff2, potshift;
    /* cell lists: head insertion, walk from head through next */
    int32_t *cell_head, *cell_np, *mol_next;
    int cells_valid;
    /* halo */
    int halo_np;
    int32_t* halo_src;           /* image -> source molecule */
    int8_t* halo_shift;          /* 3 per image */
    /* neighbour lists */
    int32_t *nb_head, *nb_n;
    nbnode* nodes;
    int64_t nnodes, nodecap;
    int nlist_valid;
    /* rebuild check */
    double* rdisp;
    int rdisp_alloc;
    double rmax;
    /* MOP */
    int nplanes;
    double* planes;
    double* Pxy_plane;           /* (3, nplanes) */
    /* control-volume surface stresses */
    int nbins[3], nhb[3], nbinso[3];
    double binsize[3];
    double* Pxyface;             /* (nbinso(1),nbinso(2),nbinso(3),3,6) column-major */
    /* sort_mols */
    int sort_nblocks[3];
    double sort_bsl[3];
    int32_t* hcurve;             /* Hcurve(nblocks(1),nblocks(2),nblocks(3)), 1-based curve positions */
    int sort_nrebuilds;          /* the routine's saved counter, starts at 1 */
    double sort_sf1;             /* saved safety factors, 1.5 and 5 */
    int sort_sf2;
    double simtime;
};

static inline int cidx(const orc_state* s, int i, int j, int k) { /* 1-based with halo */
    return ((k - 1) * s->ny2 + (j - 1)) * s->nx2 + (i - 1);
}

void orc_destroy(orc_state* s) {
    if (!s) return;
    free(s->r); free(s->v); free(s->a);
    free(s->pe); free(s->pe_lj); free(s->pe_poly); free(s->vir); free(s->fscale); free(s->rfmol);
    free(s->moltype); free(s->tag); free(s->rtether); free(s->slidev);
    free(s->chainID); free(s->subchainID); free(s->funcy); free(s->bflag);
    free(s->bond); free(s->bondcount);
    free(s->cell_head); free(s->cell_np); free(s->mol_next);
    free(s->halo_src); free(s->halo_shift);
    free(s->nb_head); free(s->nb_n); free(s->nodes);
    free(s->rdisp); free(s->planes); free(s->Pxy_plane); free(s->Pxyface); free(s->hcurve);
    free(s);
}

/* set_parameters_cells (src/setup_set_parameters.f90:1848-1903), potshift (1222), MOP planes
   (2298-2340) */
/* hilbert3 of build_hilbert (src/functions_lib.f90:2610-2636): the 8^n points of the curve of
   order n, built by concatenating eight transformed copies of the curve of order n-1 */
static int hilbert3(int n, double** x, double** y, double** z, size_t* len) {
    if (n <= 0) {
        *x = calloc(1, sizeof(double));
        *y = calloc(1, sizeof(double));
        *z = calloc(1, sizeof(double));
        *len = 1;
        return !*x || !*y || !*z;
    }
    double *xo, *yo, *zo;
    size_t m;
    if (hilbert3(n - 1, &xo, &yo, &zo, &m)) return 1;
    *len = 8 * m;
    *x = malloc(*len * sizeof(double));
    *y = malloc(*len * sizeof(double));
    *z = malloc(*len * sizeof(double));
    if (!*x || !*y || !*z) return 1;
    for (size_t k = 0; k < m; ++k) {
        const double X = xo[k], Y = yo[k], Z = zo[k];
        const double px[8] = {0.5 + Z, 0.5 + Y, -0.5 + Y, -0.5 - X, -0.5 - X, -0.5 - Y, 0.5 - Y, 0.5 + Z};
        const double py[8] = {0.5 + X, 0.5 + Z, 0.5 + Z, 0.5 + Y, -0.5 + Y, -0.5 - Z, -0.5 - Z, -0.5 - X};
        const double pz[8] = {0.5 + Y, -0.5 + X, -0.5 + X, 0.5 - Z, 0.5 - Z, -0.5 + X, -0.5 + X, 0.5 - Y};
        for (int b = 0; b < 8; ++b) {
            (*x)[b * m + k] = 0.5 * px[b];
            (*y)[b * m + k] = 0.5 * py[b];
            (*z)[b * m + k] = 0.5 * pz[b];
        }
    }
    free(xo); free(yo); free(zo);
    return 0;
}

/* Hcurve for SORT_FLAG 1 (src/setup_set_parameters.f90:1141-1153) and 2 (build_hilbert,
   src/functions_lib.f90:2555-2604) over blocks of sortblocksize cells */
static int orc_setup_sort(orc_state* s) {
    const fm_params* p = &s->p;
    if (p->sort_flag != 1 && p->sort_flag != 2) return fail("Incorrect value of sort_flag");
    if (p->sortblocksize < 1 || p->sort_freq < 1) return fail("sortblocksize and sort_freq must be >= 1");
    for (int d = 0; d < 3; ++d) {
        s->sort_bsl[d] = p->sortblocksize * s->cellside[d];
        s->sort_nblocks[d] = f_ceiling(p->domain[d] / s->sort_bsl[d]);
    }
    const int *nb = s->sort_nblocks;
    const size_t nblk = (size_t)nb[0] * nb[1] * nb[2];
    s->hcurve = malloc(nblk * sizeof(int32_t));
    if (!s->hcurve) return fail("out of memory");
#define HC(i, j, k) s->hcurve[((size_t)(k) - 1) * nb[0] * nb[1] + ((size_t)(j) - 1) * nb[0] + ((i) - 1)]
    if (p->sort_flag == 1) {
        for (int i = 1; i <= nb[0]; ++i)
            for (int j = 1; j <= nb[1]; ++j)
                for (int k = 1; k <= nb[2]; ++k) HC(i, j, k) = i + nb[0] * (j - 1) + nb[0] * nb[1] * (k - 1);
        return 0;
    }
    int maxcells = nb[0] > nb[1] ? nb[0] : nb[1];
    if (nb[2] > maxcells) maxcells = nb[2];
    const int hindex = f_ceiling(log((double)maxcells) / log(2.0));
    double *x, *y, *z;
    size_t n;
    if (hilbert3(hindex, &x, &y, &z, &n)) return fail("out of memory");
    if (hindex - log((double)maxcells) / log(2.0) > 0.0001) maxcells = f_ceiling(pow(8.0, (double)hindex / 3.0));
    double mn = x[0];
    for (size_t k = 0; k < n; ++k) {
        if (x[k] < mn) mn = x[k];
        if (y[k] < mn) mn = y[k];
        if (z[k] < mn) mn = z[k];
    }
    const double shift = 0.00001 - mn;
    for (size_t k = 0; k < nblk; ++k) s->hcurve[k] = -666;
    int m = 0;
    for (size_t k = 0; k < n; ++k) {
        const int ic = f_ceiling((x[k] + shift) * maxcells), jc = f_ceiling((y[k] + shift) * maxcells),
                  kc = f_ceiling((z[k] + shift) * maxcells);
        if (ic <= nb[0] && jc <= nb[1] && kc <= nb[2]) HC(ic, jc, kc) = ++m;
    }
    free(x); free(y); free(z);
    for (size_t k = 0; k < nblk; ++k)
        if (s->hcurve[k] == -666) return fail("Error in generating Hilbert curve - change sort flag");
    return 0;
#undef HC
}

int orc_get_hcurve(const orc_state* s, int32_t* out) {
    if (!s->hcurve) return fail("sort_flag is 0");
    memcpy(out, s->hcurve, (size_t)s->sort_nblocks[0] * s->sort_nblocks[1] * s->sort_nblocks[2] * sizeof(int32_t));
    return 0;
}

orc_state* orc_create(const fm_params* p, int32_t np, int32_t capacity) {
    orc_state* s = (orc_state*)calloc(1, sizeof(orc_state));
    if (!s) { fail("out of memory"); return NULL; }
    s->p = *p;
    s->np = np;
    if (capacity < np) capacity = np;
    s->cap = capacity;
    for (int d = 0; d < 3; ++d) s->halfdomain[d] = 0.5 * p->domain[d];
    double rneighbr = p->rcutoff + (p->delta_rneighbr[0] + p->delta_rneighbr[1] + p->delta_rneighbr[2]) / 3.0;
    s->rneighbr2 = rneighbr * rneighbr;
    if (p->potential_flag == 1 && rneighbr < p->R_0) {
        rneighbr = p->R_0;
        s->rneighbr2 = p->R_0 * p->R_0;
    }
    for (int d = 0; d < 3; ++d) {
        s->ncells[d] = (int)floor(p->domain[d] / (p->rcutoff + p->delta_rneighbr[d]));
    }
    if (s->ncells[0] < 3 || s->ncells[1] < 3 || s->ncells[2] < 3) {
        fail("ERROR - DOMAIN SHOULD HAVE AT LEAST 3 CELLS, IN X, Y AND Z");
        free(s);
        return NULL;
    }
    for (int d = 0; d < 3; ++d) s->cellside[d] = p->domain[d] / s->ncells[d];
    s->nx2 = s->ncells[0] + 2; s->ny2 = s->ncells[1] + 2; s->nz2 = s->ncells[2] + 2;
    s->ncelltot = s->nx2 * s->ny2 * s->nz2;
    s->rcutoff2 = p->rcutoff * p->rcutoff;
    s->potshift = 4.0 * (1.0 / p12(p->rcutoff) - 1.0 / p6(p->rcutoff));

    int cap = s->cap;
    s->r = calloc(3 * (size_t)cap, sizeof(double));
    s->v = calloc(3 * (size_t)cap, sizeof(double));
    s->a = calloc(3 * (size_t)cap, sizeof(double));
    s->pe = calloc(cap, sizeof(double));
    s->pe_lj = calloc(cap, sizeof(double));
    s->pe_poly = calloc(cap, sizeof(double));
    s->vir = calloc(cap, sizeof(double));
    s->fscale = calloc(cap, sizeof(double));
    s->rfmol = calloc(9 * (size_t)cap, sizeof(double));
    s->moltype = calloc(cap, sizeof(int32_t));
    s->tag = calloc(cap, sizeof(int32_t));
    s->rtether = calloc(3 * (size_t)cap, sizeof(double));
    s->slidev = calloc(3 * (size_t)cap, sizeof(double));
    s->chainID = calloc(cap, sizeof(int32_t));
    s->subchainID = calloc(cap, sizeof(int32_t));
    s->funcy = calloc(cap, sizeof(int32_t));
    s->bflag = calloc(4 * (size_t)cap, sizeof(int32_t));
    s->bond = malloc(FM_MAX_FUNCY * (size_t)cap * sizeof(int32_t));
    s->bondcount = calloc(cap, sizeof(int32_t));
    s->cell_head = malloc(s->ncelltot * sizeof(int32_t));
    s->cell_np = calloc(s->ncelltot, sizeof(int32_t));
    s->mol_next = malloc(cap * sizeof(int32_t));
    s->halo_src = malloc(cap * sizeof(int32_t));
    s->halo_shift = calloc(3 * (size_t)cap, sizeof(int8_t));
    s->nb_head = malloc(cap * sizeof(int32_t));
    s->nb_n = calloc(cap, sizeof(int32_t));
    s->nodecap = (int64_t)np * 48 + 1024;
    s->nodes = malloc(s->nodecap * sizeof(nbnode));
    if (!s->r || !s->v || !s->a || !s->pe || !s->pe_lj || !s->pe_poly || !s->vir || !s->fscale ||
        !s->rfmol || !s->moltype || !s->tag || !s->rtether || !s->slidev || !s->chainID ||
        !s->subchainID || !s->funcy || !s->bflag || !s->bond || !s->bondcount || !s->cell_head ||
        !s->cell_np || !s->mol_next || !s->halo_src || !s->halo_shift || !s->nb_head || !s->nb_n ||
        !s->nodes) {
        orc_destroy(s);
        fail("out of memory");
        return NULL;
    }
    for (int c = 0; c < s->ncelltot; ++c) s->cell_head[c] = -1;
    for (size_t k = 0; k < FM_MAX_FUNCY * (size_t)cap; ++k) s->bond[k] = -1;
    for (int n = 0; n < cap; ++n) s->nb_head[n] = -1;

    if (p->vflux_outflag >= 1 && p->vflux_outflag <= 3) {
        int pn = p->vflux_outflag - 1;
        s->nplanes = p->nplanes;
        if (s->nplanes < 1 || s->nplanes > FM_MAX_PLANES) {
            orc_destroy(s);
            fail("nplanes out of range");
            return NULL;
        }
        s->planes = calloc(s->nplanes, sizeof(double));
        s->Pxy_plane = calloc(3 * (size_t)s->nplanes, sizeof(double));
        double shift = 0.5 * (p->domain[pn] - p->planespacing * (s->nplanes - 1));
        for (int n = 1; n <= s->nplanes; ++n)
            s->planes[n - 1] = p->planespacing * (n - 1) + shift - s->halfdomain[pn];
    }
    s->sort_nrebuilds = 1;
    s->sort_sf1 = 1.5;
    s->sort_sf2 = 5;
    if (p->sort_flag != 0) {
        if (orc_setup_sort(s)) {
            orc_destroy(s);
            return NULL;
        }
    }
    if (p->vflux_outflag == 4) {
        /* src/setup_set_parameters.f90:1990-2011 with one process per direction: gnbins is the
           real product assigned to an integer (truncation), nbins = nint(gnbins/1) */
        for (int d = 0; d < 3; ++d) {
            s->nbins[d] = (int)(p->binspercell[d] * s->ncells[d]);
            if (s->nbins[d] < 1) {
                orc_destroy(s);
                fail("Error -- bin size less than one, decrease bin to cell ratio or use bigger domain");
                return NULL;
            }
            s->nhb[d] = f_ceiling((double)s->nbins[d] / (double)s->ncells[d]);
            s->nbinso[d] = s->nbins[d] + 2 * s->nhb[d];
            s->binsize[d] = p->domain[d] / s->nbins[d];
        }
        s->Pxyface = calloc((size_t)s->nbinso[0] * s->nbinso[1] * s->nbinso[2] * 18, sizeof(double));
        if (!s->Pxyface) {
            orc_destroy(s);
            fail("out of memory");
            return NULL;
        }
    }
    return s;
}

void orc_geometry(const orc_state* s, int32_t ncells[3], double cellside[3], double* rneighbr2,
                  double* rcutoff2, double* potshift) {
    for (int d = 0; d < 3; ++d) { ncells[d] = s->ncells[d]; cellside[d] = s->cellside[d]; }
    *rneighbr2 = s->rneighbr2;
    *rcutoff2 = s->rcutoff2;
    *potshift = s->potshift;
}

int32_t orc_np(const orc_state* s) { return s->np; }
int32_t orc_halo_np(const orc_state* s) { return s->halo_np; }

/* read_tag (src/molecular_properties.f90:764-830): slide velocity from the tag */
static void read_tag(orc_state* s, int n) {
    for (int d = 0; d < 3; ++d) s->slidev[3 * n + d] = 0.0;
    if (s->p.ensemble != FM_ENSEMBLE_TAG_MOVE) return;
    if (s->tag[n] == FM_TAG_TETH_SLIDE) {
        double sg = f_sign(1.0, s->r[3 * n + 1]);
        for (int d = 0; d < 3; ++d) s->slidev[3 * n + d] = s->p.wallslidev[d] * sg;
    } else if (s->tag[n] == FM_TAG_FIXED_SLIDE) {
        for (int d = 0; d < 3; ++d) s->slidev[3 * n + d] = s->p.wallslidev[d];
    }
}

static int is_tethered(int tag) { return tag == FM_TAG_TETH || tag == FM_TAG_TETH_SLIDE; }
/* the tags of src/modules.f90:78-88 without a thermostat */
static int tag_supported(int tag) {
    return tag == FM_TAG_FREE || tag == FM_TAG_FIXED || tag == FM_TAG_FIXED_SLIDE || tag == FM_TAG_TETH ||
           tag == FM_TAG_TETH_SLIDE;
}

int orc_set_state(orc_state* s, int32_t np, const double* r, const double* v, const double* a) {
    if (np > s->cap) return fail("np %d exceeds capacity %d", np, s->cap);
    s->np = np;
    memcpy(s->r, r, 3 * (size_t)np * sizeof(double));
    memcpy(s->v, v, 3 * (size_t)np * sizeof(double));
    if (a) memcpy(s->a, a, 3 * (size_t)np * sizeof(double));
    else memset(s->a, 0, 3 * (size_t)np * sizeof(double));
    s->cells_valid = 0;
    s->nlist_valid = 0;
    return 0;
}

int orc_set_moltype(orc_state* s, const int32_t* moltype) {
    for (int n = 0; n < s->np; ++n) {
        if (moltype[n] < 0 || moltype[n] >= FM_MAX_TYPES) return fail("moltype out of range");
        s->moltype[n] = moltype[n];
    }
    return 0;
}

int orc_set_tags(orc_state* s, const int32_t* tag, const double* rtether) {
    for (int n = 0; n < s->np; ++n)
        if (!tag_supported(tag[n])) return fail("Invalid molecular Tag %d for molecule %d", tag[n], n + 1);
    memcpy(s->tag, tag, s->np * sizeof(int32_t));
    if (rtether) memcpy(s->rtether, rtether, 3 * (size_t)s->np * sizeof(double));
    for (int n = 0; n < s->np; ++n) read_tag(s, n);
    return 0;
}

int orc_set_monomers(orc_state* s, const int32_t* chainID, const int32_t* subchainID,
                     const int32_t* funcy, const int32_t* bin_bflag) {
    memcpy(s->chainID, chainID, s->np * sizeof(int32_t));
    memcpy(s->subchainID, subchainID, s->np * sizeof(int32_t));
    memcpy(s->funcy, funcy, s->np * sizeof(int32_t));
    memcpy(s->bflag, bin_bflag, 4 * (size_t)s->np * sizeof(int32_t));
    return 0;
}

int orc_get_state(const orc_state* s, double* r, double* v, double* a) {
    if (r) memcpy(r, s->r, 3 * (size_t)s->np * sizeof(double));
    if (v) memcpy(v, s->v, 3 * (size_t)s->np * sizeof(double));
    if (a) memcpy(a, s->a, 3 * (size_t)s->np * sizeof(double));
    return 0;
}

int orc_get_tags(const orc_state* s, double* rtether) {
    memcpy(rtether, s->rtether, 3 * (size_t)s->np * sizeof(double));
    return 0;
}

int orc_get_molecular(const orc_state* s, double* potenergymol, double* virialmol) {
    if (potenergymol) memcpy(potenergymol, s->pe, s->np * sizeof(double));
    if (virialmol) memcpy(virialmol, s->vir, s->np * sizeof(double));
    return 0;
}

int orc_get_rfmol(const orc_state* s, double* rfmol) {
    memcpy(rfmol, s->rfmol, 9 * (size_t)s->np * sizeof(double));
    return 0;
}

int orc_get_force_scale(const orc_state* s, double* scale) {
    memcpy(scale, s->fscale, s->np * sizeof(double));
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* linklist_deallocateall (src/linklist.f90:2872-2963) */
int orc_linklist_deallocateall(orc_state* s) {
    if (s->p.potential_flag == 1) {
        for (size_t k = 0; k < FM_MAX_FUNCY * (size_t)s->cap; ++k) s->bond[k] = -1;
        memset(s->bondcount, 0, s->cap * sizeof(int32_t));
    }
    for (int c = 0; c < s->ncelltot; ++c) { s->cell_head[c] = -1; s->cell_np[c] = 0; }
    for (int n = 0; n < s->cap; ++n) { s->nb_head[n] = -1; s->nb_n[n] = 0; }
    s->nnodes = 0;
    s->cells_valid = 0;
    s->nlist_valid = 0;
    return 0;
}

/* sendmols on one rank (src/messenger.MPI.f90:1383-1421): for i=1..nd, for ixyz, for dir=-1,+1:
   checksendbuild (1426-1456) pushes leavers n=1..np onto the pass list (head insertion);
   sendrecvface (1460-1824) packs them from the head (descending n) and receives them at np+1..
   wrapped by -dir*domain; reorderdata (1830-1904) walks the pass list from the head and fills
   slot molno with r(:,np+new_np), new_np decreasing.  Net effect: the leavers' slots are refilled
   in reversed order. */
static void move_mol(orc_state* s, int dst, int src) {
    for (int d = 0; d < 3; ++d) {
        s->r[3 * dst + d] = s->r[3 * src + d];
        s->v[3 * dst + d] = s->v[3 * src + d];
    }
    if (s->p.ensemble == FM_ENSEMBLE_TAG_MOVE) {
        s->tag[dst] = s->tag[src];
        if (is_tethered(s->tag[src]))
            for (int d = 0; d < 3; ++d) s->rtether[3 * dst + d] = s->rtether[3 * src + d];
    }
    if (s->p.potential_flag == 1) {
        s->chainID[dst] = s->chainID[src];
        s->subchainID[dst] = s->subchainID[src];
        s->funcy[dst] = s->funcy[src];
        for (int g = 0; g < 4; ++g) s->bflag[4 * dst + g] = s->bflag[4 * src + g];
    }
    if (s->p.mie_potential == 1) s->moltype[dst] = s->moltype[src];
}

int orc_sendmols(orc_state* s) {
    int np = s->np;
    int* leav = malloc((np > 0 ? np : 1) * sizeof(int));
    if (!leav) return fail("out of memory");
    for (int it = 0; it < 3; ++it) {
        for (int ixyz = 0; ixyz < 3; ++ixyz) {
            for (int dir = -1; dir <= 1; dir += 2) {
                int k = 0;
                for (int n = 0; n < np; ++n) {
                    double x = s->r[3 * n + ixyz];
                    if (dir == -1 ? (x < -s->halfdomain[ixyz]) : (x >= s->halfdomain[ixyz])) leav[k++] = n;
                }
                if (k == 0) continue;
                if (np + k > s->cap) { free(leav); return fail("sendmols: capacity exceeded"); }
                /* pack from the pass-list head: descending n; receive at np+1..np+k */
                for (int m = 0; m < k; ++m) {
                    int src = leav[k - 1 - m];
                    int dst = np + m;
                    move_mol(s, dst, src);
                    s->r[3 * dst + ixyz] = s->r[3 * dst + ixyz] - dir * s->p.domain[ixyz];
                    if (s->p.ensemble == FM_ENSEMBLE_TAG_MOVE && is_tethered(s->tag[dst]))
                        s->rtether[3 * dst + ixyz] = s->rtether[3 * dst + ixyz] - dir * s->p.domain[ixyz];
                }
                /* reorderdata: walk pass list from the head (descending molno) */
                int new_np = k;
                for (int m = 0; m < k; ++m) {
                    int molno = leav[k - 1 - m];
                    move_mol(s, molno, np + new_np - 1);
                    if (s->p.ensemble == FM_ENSEMBLE_TAG_MOVE) read_tag(s, molno);
                    new_np -= 1;
                }
            }
        }
    }
    free(leav);
    return 0;
}

/* linklist_checkpush (src/linklist.f90:1596-1630): head insertion */
static inline void cell_push(orc_state* s, int c, int n) {
    s->mol_next[n] = s->cell_head[c];
    s->cell_head[c] = n;
    s->cell_np[c] += 1;
}

/* assign_to_cell (src/linklist.f90:2645-2676) */
int orc_assign_to_cell(orc_state* s) {
    for (int n = 0; n < s->np; ++n) {
        int ic = f_ceiling((s->r[3 * n + 0] + s->halfdomain[0]) / s->cellside[0]) + 1;
        int jc = f_ceiling((s->r[3 * n + 1] + s->halfdomain[1]) / s->cellside[1]) + 1;
        int kc = f_ceiling((s->r[3 * n + 2] + s->halfdomain[2]) / s->cellside[2]) + 1;
        if (ic < 1 || jc < 1 || kc < 1 || ic > s->nx2 || jc > s->ny2 || kc > s->nz2)
            return fail("Aborting due to escaped mol. %d", n);
        cell_push(s, cidx(s, ic, jc, kc), n);
    }
    s->cells_valid = 1;
    return 0;
}

/* pack_cell (src/messenger.MPI.f90:647-706) + unpack + shift: append images of a cell's
   molecules (walked from the head) at np+halo_np. */
static int pack_cell(orc_state* s, int i, int j, int k, int sx, int sy, int sz) {
    int c = cidx(s, i, j, k);
    for (int n = s->cell_head[c]; n >= 0; n = s->mol_next[n]) {
        int m = s->np + s->halo_np;
        if (m >= s->cap) return fail("halo overflow: increase capacity");
        int sh[3] = {sx, sy, sz};
        for (int d = 0; d < 3; ++d) {
            double x = s->r[3 * n + d];
            /* faces add/subtract domain; edges/corners add sign(1,..)*domain: same IEEE result */
            s->r[3 * m + d] = sh[d] == 0 ? x : x + sh[d] * s->p.domain[d];
            s->halo_shift[3 * m + d] = (int8_t)sh[d];
        }
        s->halo_src[m] = n;
        if (s->p.potential_flag == 1) {
            s->chainID[m] = s->chainID[n];
            s->subchainID[m] = s->subchainID[n];
            s->funcy[m] = s->funcy[n];
            for (int g = 0; g < 4; ++g) s->bflag[4 * m + g] = s->bflag[4 * n + g];
        }
        if (s->p.mie_potential == 1) s->moltype[m] = s->moltype[n];
        s->halo_np += 1;
    }
    return 0;
}

/* messenger_updateborders (src/messenger.MPI.f90:594-628) on one periodic rank */
int orc_updateborders(orc_state* s, int32_t rebuild) {
    if (!s->cells_valid) return fail("updateborders before assign_to_cell");
    int nx = s->ncells[0], ny = s->ncells[1], nz = s->ncells[2];
    s->halo_np = 0;
    /* updatefacedown(1..3): cells at index 2, image r + domain (880-1000) */
    for (int j = 2; j <= ny + 1; ++j) for (int k = 2; k <= nz + 1; ++k) if (pack_cell(s, 2, j, k, 1, 0, 0)) return 1;
    for (int i = 2; i <= nx + 1; ++i) for (int k = 2; k <= nz + 1; ++k) if (pack_cell(s, i, 2, k, 0, 1, 0)) return 1;
    for (int i = 2; i <= nx + 1; ++i) for (int j = 2; j <= ny + 1; ++j) if (pack_cell(s, i, j, 2, 0, 0, 1)) return 1;
    /* updatefaceup(1..3): cells at index n+1, image r - domain (1002-1142) */
    for (int j = 2; j <= ny + 1; ++j) for (int k = 2; k <= nz + 1; ++k) if (pack_cell(s, nx + 1, j, k, -1, 0, 0)) return 1;
    for (int i = 2; i <= nx + 1; ++i) for (int k = 2; k <= nz + 1; ++k) if (pack_cell(s, i, ny + 1, k, 0, -1, 0)) return 1;
    for (int i = 2; i <= nx + 1; ++i) for (int j = 2; j <= ny + 1; ++j) if (pack_cell(s, i, j, nz + 1, 0, 0, -1)) return 1;
    /* updateedge(1,2), (2,3), (1,3): ixyz = 3, 1, 2 (1147-1293) */
    const int ixyz_order[3] = {3, 1, 2};
    for (int e = 0; e < 3; ++e) {
        int ixyz = ixyz_order[e];
        for (int q = 0; q < 4; ++q) {
            int a1, a2; /* edge1, edge2 for this ixyz */
            int nA, nB; /* ncells of the two other axes */
            if (ixyz == 1) { nA = ny; nB = nz; }
            else if (ixyz == 2) { nA = nx; nB = nz; }
            else { nA = nx; nB = ny; }
            const int e1[4] = {2, 2, nA + 1, nA + 1};
            const int e2[4] = {2, nB + 1, 2, nB + 1};
            a1 = e1[q]; a2 = e2[q];
            int sA = f_isign(1, nA - a1), sB = f_isign(1, nB - a2);
            if (ixyz == 1) {
                for (int i = 2; i <= nx + 1; ++i) if (pack_cell(s, i, a1, a2, 0, sA, sB)) return 1;
            } else if (ixyz == 2) {
                for (int j = 2; j <= ny + 1; ++j) if (pack_cell(s, a1, j, a2, sA, 0, sB)) return 1;
            } else {
                for (int k = 2; k <= nz + 1; ++k) if (pack_cell(s, a1, a2, k, sA, sB, 0)) return 1;
            }
        }
    }
    /* updatecorners (1296-1377) */
    const int ic[8] = {2, 2, 2, 2, nx + 1, nx + 1, nx + 1, nx + 1};
    const int jc[8] = {2, 2, ny + 1, ny + 1, 2, 2, ny + 1, ny + 1};
    const int kc[8] = {2, nz + 1, 2, nz + 1, 2, nz + 1, 2, nz + 1};
    for (int q = 0; q < 8; ++q)
        if (pack_cell(s, ic[q], jc[q], kc[q], f_isign(1, nx - ic[q]), f_isign(1, ny - jc[q]), f_isign(1, nz - kc[q])))
            return 1;
    /* assign_to_halocell(np+1, np+halo_np) (src/linklist.f90:2681-2702) */
    if (rebuild == 1) {
        for (int m = s->np; m < s->np + s->halo_np; ++m) {
            int i = f_ceiling((s->r[3 * m + 0] + s->halfdomain[0]) / s->cellside[0]) + 1;
            int j = f_ceiling((s->r[3 * m + 1] + s->halfdomain[1]) / s->cellside[1]) + 1;
            int k = f_ceiling((s->r[3 * m + 2] + s->halfdomain[2]) / s->cellside[2]) + 1;
            if (i < 1 || j < 1 || k < 1 || i > s->nx2 || j > s->ny2 || k > s->nz2)
                return fail("halo image outside halo cells");
            cell_push(s, cidx(s, i, j, k), m);
        }
    }
    return 0;
}

/* linklist_checkpushneighbr (src/linklist.f90:1636-1668): head insertion into molnoi's list */
static int nb_push(orc_state* s, int owner, int mol) {
    if (owner >= s->np) return fail("neighbour list owner is a halo image");
    if (s->nnodes == s->nodecap) {
        int64_t nc = s->nodecap * 2;
        nbnode* nn = realloc(s->nodes, nc * sizeof(nbnode));
        if (!nn) return fail("out of memory");
        s->nodes = nn;
        s->nodecap = nc;
    }
    nbnode* nd = &s->nodes[s->nnodes];
    nd->mol = mol;
    nd->next = s->nb_head[owner];
    s->nb_head[owner] = (int32_t)s->nnodes;
    s->nnodes += 1;
    s->nb_n[owner] += 1;
    return 0;
}

/* get_bondflag (src/modules.f90:597-610) */
static int get_bondflag(const orc_state* s, int molno, int scID) {
    const int intbits = 31;
    int group = (int)ceilf((float)scID / (float)intbits);
    int expo = scID % intbits - 1;
    if (expo == -1) expo = intbits - 1;
    if (group < 1 || group > 4) return 0;
    return (s->bflag[4 * molno + group - 1] >> expo) & 1;
}

/* check_update_adjacentbeadinfo (src/linklist.f90:2542-2584) */
static int check_bead(orc_state* s, int i, int j) {
    if (s->chainID[i] - s->chainID[j] != 0) return 0;
    int jsc = s->subchainID[j];
    if (get_bondflag(s, i, jsc) == 1) {
        s->bondcount[i] += 1;
        s->bondcount[j] += 1;
        if (s->bondcount[i] > s->funcy[i]) return fail("Error: too many bonds for molno %d", i + 1);
        if (s->bondcount[j] > s->funcy[j]) return fail("Error: too many bonds for molno %d", j + 1);
        s->bond[(s->bondcount[i] - 1) + FM_MAX_FUNCY * i] = j;
        s->bond[(s->bondcount[j] - 1) + FM_MAX_FUNCY * j] = i;
    }
    return 0;
}

static inline double dist2(const double* ri, const double* rj) {
    double x = ri[0] - rj[0], y = ri[1] - rj[1], z = ri[2] - rj[2];
    return (x * x + y * y) + z * z; /* dot_product summed in index order */
}

static const int ICS[13] = {1, 1, 0, -1, 0, 1, 1, 0, -1, -1, -1, 0, 1};
static const int JCS[13] = {0, 1, 1, 1, 0, 0, 1, 1, 1, 0, -1, -1, -1};
static const int KCS[13] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};

/* calculate_cell_interactions_opt (src/linklist.f90:1118-1180) */
static int cell_block(orc_state* s, int i0, int i1, int ish, int j0, int j1, int jsh, int k0, int k1, int ksh) {
    for (int kc = k0; kc <= k1; ++kc)
        for (int jc = j0; jc <= j1; ++jc)
            for (int ic = i0; ic <= i1; ++ic) {
                int c = cidx(s, ic, jc, kc), ca = cidx(s, ic + ish, jc + jsh, kc + ksh);
                for (int mi = s->cell_head[c]; mi >= 0; mi = s->mol_next[mi]) {
                    const double* ri = &s->r[3 * mi];
                    for (int mj = s->cell_head[ca]; mj >= 0; mj = s->mol_next[mj]) {
                        double rij2 = dist2(ri, &s->r[3 * mj]);
                        if (s->p.potential_flag == 1 && check_bead(s, mi, mj)) return 1;
                        if (rij2 < s->rneighbr2 && nb_push(s, mj, mi)) return 1;
                    }
                }
            }
    return 0;
}

/* assign_to_neighbourlist_halfint_opt (src/linklist.f90:660-804) */
int orc_assign_to_neighbourlist(orc_state* s) {
    if (!s->cells_valid) return fail("neighbour list before assign_to_cell");
    int nx = s->ncells[0], ny = s->ncells[1], nz = s->ncells[2];
    for (int n = 0; n < s->cap; ++n) { s->nb_head[n] = -1; s->nb_n[n] = 0; }
    s->nnodes = 0;
    for (int ic = 2; ic <= nx + 1; ++ic)
        for (int jc = 2; jc <= ny + 1; ++jc)
            for (int kc = 2; kc <= nz + 1; ++kc) {
                int c = cidx(s, ic, jc, kc);
                /* own cell, j after i in list order (702-726) */
                for (int mi = s->cell_head[c]; mi >= 0; mi = s->mol_next[mi]) {
                    const double* ri = &s->r[3 * mi];
                    for (int mj = s->mol_next[mi]; mj >= 0; mj = s->mol_next[mj]) {
                        if (mi == mj) continue;
                        double rij2 = dist2(ri, &s->r[3 * mj]);
                        if (s->p.potential_flag == 1 && check_bead(s, mi, mj)) return 1;
                        if (rij2 < s->rneighbr2 && nb_push(s, mi, mj)) return 1;
                    }
                }
                /* 13 forward cells (731-770) */
                for (int q = 0; q < 13; ++q) {
                    int ca = cidx(s, ic + ICS[q], jc + JCS[q], kc + KCS[q]);
                    for (int mi = s->cell_head[c]; mi >= 0; mi = s->mol_next[mi]) {
                        const double* ri = &s->r[3 * mi];
                        for (int mj = s->cell_head[ca]; mj >= 0; mj = s->mol_next[mj]) {
                            double rij2 = dist2(ri, &s->r[3 * mj]);
                            if (s->p.potential_flag == 1 && check_bead(s, mi, mj)) return 1;
                            if (rij2 < s->rneighbr2 && nb_push(s, mi, mj)) return 1;
                        }
                    }
                }
            }
    /* all_cell_interactions (778-802) */
    int x = nx, x1 = nx + 1, x2 = nx + 2, y = ny, y1 = ny + 1, y2 = ny + 2, z = nz, z1 = nz + 1;
    (void)z;
    const int imin[27] = {2, 1, 3, 2, 2, 1, 3, 3, 1, 1, 1, 1, 1, 1, x2, x2, x2, x2, 2, 2, 3, 2, 2, 3, 2, 3, 2};
    const int imax[27] = {x1, x, x2, x1, x1, x, x2, x2, x, 1, 1, 1, 1, 1, x2, x2, x2, x2, x1, x, x1, x1, x, x1, x1, x1, x};
    const int jmin[27] = {2, 2, 2, 1, 3, 1, 1, 3, 3, 2, 2, 1, 1, 3, 1, 2, 1, 3, 1, 1, 1, 1, 1, 1, y2, y2, y2};
    const int jmax[27] = {y1, y1, y1, y, y2, y, y, y2, y2, y1, y1, y, y, y2, y, y1, y, y2, 1, 1, 1, 1, 1, 1, y2, y2, y2};
    const int kmin[27] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
    const int kmax[27] = {1, 1, 1, 1, 1, 1, 1, 1, 1, z1, nz, z1, nz, nz, z1, nz, nz, nz, z1, z1, z1, nz, nz, nz, nz, nz, nz};
    const int shift[27] = {5, 6, 10, 8, 12, 7, 9, 11, 13, 1, 6, 2, 7, 13, 4, 10, 9, 11, 3, 2, 4, 8, 7, 9, 12, 11, 13};
    for (int q = 0; q < 27; ++q) {
        int sh = shift[q] - 1;
        if (cell_block(s, imin[q], imax[q], ICS[sh], jmin[q], jmax[q], JCS[sh], kmin[q], kmax[q], KCS[sh])) return 1;
    }
    s->nlist_valid = 1;
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* pair laws */
static inline int pair_type(const orc_state* s, int i, int j) {
    return s->moltype[i] * FM_MAX_TYPES + s->moltype[j];
}

static inline double mass_of(const orc_state* s, int i) {
    return s->p.mie_potential == 1 ? s->p.mass_lookup[s->moltype[i]] : 1.0; /* LJ_mass = 1 */
}

/* LJ_accijmag (src/setup_set_parameters.f90:105-113), Mie_accijmag (149-168) */
static inline double accijmag(const orc_state* s, double invrij2, int i, int j) {
    if (s->p.mie_potential == 0) return 48.0 * (p7(invrij2) - 0.5 * p4(invrij2));
    int t = pair_type(s, i, j);
    double lr = s->p.lambdar_lookup[t], la = s->p.lambdaa_lookup[t];
    return s->p.C_lookup[t] * s->p.epsilon_lookup[t] *
           (lr * pow(invrij2, 0.5 * lr + 1) - s->p.alpha_lookup[t] * la * pow(invrij2, 0.5 * la + 1));
}

/* LJ_energy (126-134), Mie_energy (182-204) */
static inline double energy(const orc_state* s, double invrij2, int i, int j) {
    if (s->p.mie_potential == 0) return 4.0 * (p6(invrij2) - p3(invrij2)) - s->potshift;
    int t = pair_type(s, i, j);
    double lr = s->p.lambdar_lookup[t], la = s->p.lambdaa_lookup[t];
    return s->p.C_lookup[t] * s->p.epsilon_lookup[t] *
               (pow(invrij2, 0.5 * lr) - s->p.alpha_lookup[t] * pow(invrij2, 0.5 * la)) -
           s->p.potshift_lookup[t];
}

static inline double cutoff2(const orc_state* s, int i, int j) {
    if (s->p.mie_potential == 1) {
        double c = s->p.rcutoff2_lookup[pair_type(s, i, j)];
        if (c > 0.0) return c;
    }
    return s->rcutoff2;
}

/* pressure_tensor_forces_MOP (src/simulation_record.f90:7068-7112) */
static void mop(orc_state* s, int pnxyz, const double* ri, const double* rj, const double* rij, double acc) {
    double spacing = s->p.planespacing;
    int nplanes = s->nplanes;
    double shift = 0.5 * (s->p.domain[pnxyz] - spacing * (nplanes - 1));
    int pi = f_nint((ri[1] + s->halfdomain[pnxyz] - shift) / spacing) + 1;
    if (pi < 1) pi = 1;
    if (pi > nplanes) pi = nplanes;
    int pj = f_nint((rj[1] + s->halfdomain[pnxyz] - shift) / spacing) + 1;
    if (pj < 1) pj = 1;
    if (pj > nplanes) pj = nplanes;
    int step = f_isign(1, pj - pi);
    for (int n = pi;; n += step) {
        double plane = s->planes[n - 1];
        double pyb0 = ri[0] + (rij[0] / rij[1]) * (plane - ri[1]);
        double pyb2 = ri[2] + (rij[2] / rij[1]) * (plane - ri[1]);
        double fx = (f_sign(1.0, ri[1] - plane) - f_sign(1.0, rj[1] - plane));
        double hx = dp_heaviside(s->halfdomain[0] - pyb0) - dp_heaviside(-s->halfdomain[0] - pyb0);
        double hz = dp_heaviside(s->halfdomain[2] - pyb2) - dp_heaviside(-s->halfdomain[2] - pyb2);
        for (int d = 0; d < 3; ++d)
            s->Pxy_plane[3 * (n - 1) + d] = s->Pxy_plane[3 * (n - 1) + d] + acc * rij[d] * fx * hx * hz;
        if (n == pj) break;
    }
}

/* get_bin => bin_from_integer_division (src/simulation_record.f90:108-122): 1-based with nhb halo
   bins below */
static inline void cv_bin(const orc_state* s, const double* r, int bin[3]) {
    for (int d = 0; d < 3; ++d) bin[d] = f_ceiling((r[d] + s->halfdomain[d]) / s->binsize[d]) + s->nhb[d];
}

static inline size_t cv_index(const orc_state* s, const int b[3], int comp, int face) { /* 1-based b, face */
    return (size_t)(b[0] - 1) +
           (size_t)s->nbinso[0] * ((size_t)(b[1] - 1) +
                                   (size_t)s->nbinso[1] * ((size_t)(b[2] - 1) +
                                                           (size_t)s->nbinso[2] * (comp + 3 * (size_t)(face - 1))));
}

/* control_volume_stresses_opt (src/simulation_record.f90:5281-5317): quantity = 2 fij passed to
   cumulative_flux_opt (4878-5064) without an intrinsic surface, whose crossings come from
   get_crossings (308-353).  The segment ri->rj crosses the bin planes between bin1(n) and
   bin2(n); each crossing adds crossdir*quantity to the face n of the bin just above the plane and
   copies that face's running total to face n+3 of the bin below. */
static int cv_stress(orc_state* s, const double* ri, const double* rj, const double* fij) {
    const double eps = (double)1e-12f; /* `real(kind(0.d0)),parameter :: eps = 1e-12` is a default-real literal */
    double q[3], r12[3];
    int b1[3], b2[3];
    for (int d = 0; d < 3; ++d) q[d] = 2.0 * fij[d];
    cv_bin(s, ri, b1);
    cv_bin(s, rj, b2);
    for (int d = 0; d < 3; ++d) {
        r12[d] = ri[d] - rj[d];
        if (fabs(r12[d]) < 0.000001) r12[d] = 0.000001;
    }
    for (int n = 0; n < 3; ++n) {
        if (b1[n] == b2[n]) continue;
        const int t1 = (n + 1) % 3, t2 = (n + 2) % 3;
        const double crossdir = f_sign(1.0, r12[n]);
        const int minbin = b1[n] < b2[n] ? b1[n] : b2[n], maxbin = b1[n] < b2[n] ? b2[n] : b1[n];
        for (int i = minbin; i <= maxbin - 1; ++i) {
            const double pt = (double)(i - s->nhb[n]) * s->binsize[n] - s->halfdomain[n];
            double rc[3];
            rc[n] = pt;
            rc[t1] = ri[t1] + (r12[t1] / r12[n]) * (pt - ri[n]);
            rc[t2] = ri[t2] + (r12[t2] / r12[n]) * (pt - ri[n]);
            rc[n] = rc[n] + eps;
            int cb[3];
            cv_bin(s, rc, cb);
            for (int d = 0; d < 3; ++d)
                if (cb[d] < 1 || cb[d] > s->nbinso[d]) return fail("CV crossing outside the bins (%d,%d,%d)", cb[0], cb[1], cb[2]);
            for (int c = 0; c < 3; ++c) {
                size_t k = cv_index(s, cb, c, n + 1);
                s->Pxyface[k] = s->Pxyface[k] + crossdir * q[c];
            }
            int cm[3] = {cb[0], cb[1], cb[2]};
            cm[n] -= 1;
            if (cm[0] >= 1 && cm[1] >= 1 && cm[2] >= 1)
                for (int c = 0; c < 3; ++c) s->Pxyface[cv_index(s, cm, c, n + 4)] = s->Pxyface[cv_index(s, cb, c, n + 1)];
        }
    }
    return 0;
}

/* pressure_tensor_forces (src/simulation_record.f90:6239-6254) */
static inline void rf_tally(orc_state* s, int molno, const double* rij, double acc) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            s->rfmol[9 * (size_t)molno + 3 * i + j] = s->rfmol[9 * (size_t)molno + 3 * i + j] + acc * rij[i] * rij[j];
}

static inline int fmod_i(int a, int b) { return a % b; } /* Fortran mod: sign of a */

/* simulation_compute_forces_LJ_neigbr_halfint (src/simulation_compute_forces.f90:628-726) */
static int forces_halfint(orc_state* s, int iter) {
    int rec = fmod_i(iter, s->p.tplot) == 0;
    int ev = s->p.teval > 0 && fmod_i(iter, s->p.teval) == 0;
    for (int mi = 0; mi < s->np; ++mi) {
        const double ri[3] = {s->r[3 * mi], s->r[3 * mi + 1], s->r[3 * mi + 2]};
        for (int nd = s->nb_head[mi]; nd >= 0; nd = s->nodes[nd].next) {
            int mj = s->nodes[nd].mol;
            const double* rj = &s->r[3 * mj];
            double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
            double rij2 = (rij[0] * rij[0] + rij[1] * rij[1]) + rij[2] * rij[2];
            if (rij2 < cutoff2(s, mi, mj)) {
                double invrij2 = 1.0 / rij2;
                double acc = accijmag(s, invrij2, mi, mj);
                double mi_ = mass_of(s, mi), mj_ = mass_of(s, mj);
                for (int d = 0; d < 3; ++d) s->a[3 * mi + d] = s->a[3 * mi + d] + acc * rij[d] / mi_;
                for (int d = 0; d < 3; ++d) s->a[3 * mj + d] = s->a[3 * mj + d] - acc * rij[d] / mj_;
                double fmag = fabs(acc) * sqrt(rij2);
                s->fscale[mi] += fmag;
                s->fscale[mj] += fmag;
                if (s->p.vflux_outflag == 4 && (s->p.cv_conserve == 1 || rec)) {
                    /* interior pairs carry 2 fij, pairs with a halo image fij (each is met twice) */
                    double f2 = (mj >= s->np || mi >= s->np) ? acc : 2.0 * acc;
                    double fij[3] = {f2 * rij[0], f2 * rij[1], f2 * rij[2]};
                    if (cv_stress(s, ri, rj, fij)) return 1;
                }
                if (rec) {
                    double e = energy(s, invrij2, mi, mj);
                    s->pe_lj[mi] = s->pe_lj[mi] + e;
                    s->pe_lj[mj] = s->pe_lj[mj] + energy(s, invrij2, mi, mj);
                    s->vir[mi] = s->vir[mi] + acc * rij2;
                    s->vir[mj] = s->vir[mj] + acc * rij2;
                    if (s->p.pressure_outflag == 1) {
                        rf_tally(s, mi, rij, acc);
                        if (mj < s->np) rf_tally(s, mj, rij, acc);
                    }
                    if (s->p.vflux_outflag >= 1 && s->p.vflux_outflag <= 3)
                        mop(s, s->p.vflux_outflag - 1, ri, rj, rij, acc);
                } else if (ev) {
                    s->vir[mi] = s->vir[mi] + acc * rij2;
                    s->vir[mj] = s->vir[mj] + acc * rij2;
                }
            }
        }
    }
    for (int n = 0; n < s->cap; ++n) s->pe[n] = s->pe[n] + s->pe_lj[n];
    return 0;
}

/* simulation_compute_forces_poly (src/simulation_compute_forces.f90:731-789) with FENE
   (src/setup_set_parameters.f90:348-380) */
static int forces_poly(orc_state* s, int iter) {
    int rec = fmod_i(iter, s->p.tplot) == 0;
    int ev = s->p.teval > 0 && fmod_i(iter, s->p.teval) == 0;
    double R02 = s->p.R_0 * s->p.R_0;
    for (int mi = 0; mi < s->np; ++mi) {
        const double* ri = &s->r[3 * mi];
        for (int b = 0; b < s->funcy[mi]; ++b) {
            int mj = s->bond[b + FM_MAX_FUNCY * mi];
            if (mj < 0) continue;
            const double* rj = &s->r[3 * mj];
            double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
            double rij2 = (rij[0] * rij[0] + rij[1] * rij[1]) + rij[2] * rij[2];
            if (rij2 >= R02) return fail("polymer bond broken between %d and %d", mi + 1, mj + 1);
            double acc = -s->p.k_c / (1.0 - (rij2 / R02));
            double m = mass_of(s, mi);
            for (int d = 0; d < 3; ++d) s->a[3 * mi + d] = s->a[3 * mi + d] + acc * rij[d] / m;
            s->fscale[mi] += fabs(acc) * sqrt(rij2);
            if (s->p.vflux_outflag == 4 && (s->p.cv_conserve == 1 || rec)) {
                double fij[3] = {acc * rij[0], acc * rij[1], acc * rij[2]};
                if (cv_stress(s, ri, rj, fij)) return 1;
            }
            if (rec) {
                if (s->p.pressure_outflag == 1) rf_tally(s, mi, rij, acc);
                s->pe_poly[mi] = s->pe_poly[mi] + 0.5 * s->p.k_c * s->p.R_0 * s->p.R_0 * log(1.0 - (rij2 / R02));
                s->vir[mi] = s->vir[mi] + acc * rij2;
            } else if (ev) {
                s->vir[mi] = s->vir[mi] + acc * rij2;
            }
        }
    }
    for (int n = 0; n < s->cap; ++n) s->pe[n] = s->pe[n] + s->pe_poly[n];
    return 0;
}

/* tether_force (src/molecular_properties.f90:833-915), k2/k4/k6 branch; applied once per force
   evaluation to tethered molecules, after the pair and bond forces. */
static void tether_force(orc_state* s, int n) {
    double rio[3];
    for (int d = 0; d < 3; ++d) rio[d] = s->r[3 * n + d] - s->rtether[3 * n + d];
    double mag = sqrt((rio[0] * rio[0] + rio[1] * rio[1]) + rio[2] * rio[2]);
    /* magnitude**2.d0 is an exact square; magnitude**4.d0 is a call of the real power */
    double acctmag = -2.0 * s->p.teth_k2 * mag - 4.0 * s->p.teth_k4 * (mag * mag) -
                     6.0 * s->p.teth_k6 * pow(mag, 4.0);
    for (int d = 0; d < 3; ++d) s->a[3 * n + d] = s->a[3 * n + d] + acctmag * rio[d];
    for (int d = 0; d < 3; ++d)
        s->rtether[3 * n + d] = s->rtether[3 * n + d] + s->slidev[3 * n + d] * s->p.delta_t;
}

/* velocity Verlet has no tag_move in the reference; here the tethers act with the pair forces */
static void forces_tether(orc_state* s) {
    for (int n = 0; n < s->np; ++n)
        if (is_tethered(s->tag[n])) tether_force(s, n);
}

/* simulation_compute_forces (src/simulation_compute_forces.f90:182-275), force_list 3 */
int orc_compute_forces(orc_state* s, int32_t iter) {
    if (!s->nlist_valid) return fail("compute_forces before assign_to_neighbourlist");
    memset(s->a, 0, 3 * (size_t)s->cap * sizeof(double));
    memset(s->pe, 0, s->cap * sizeof(double));
    memset(s->pe_lj, 0, s->cap * sizeof(double));
    memset(s->pe_poly, 0, s->cap * sizeof(double));
    memset(s->vir, 0, s->cap * sizeof(double));
    memset(s->fscale, 0, s->cap * sizeof(double));
    if (s->p.pressure_outflag == 1 && fmod_i(iter, s->p.tplot) == 0)
        memset(s->rfmol, 0, 9 * (size_t)s->cap * sizeof(double));
    if (forces_halfint(s, iter)) return 1;
    if (s->p.potential_flag == 1 && forces_poly(s, iter)) return 1;
    if (s->p.ensemble == FM_ENSEMBLE_TAG_MOVE && s->p.integration_algorithm == FM_INTEGRATION_VV) forces_tether(s);
    return 0;
}

/* simulation_compute_forces_LJ_AP (src/simulation_compute_forces.f90:280-341): all pairs with the
   minimum image convention; an independent check of the list-based routine. */
int orc_compute_forces_allpairs(orc_state* s, int32_t iter) {
    memset(s->a, 0, 3 * (size_t)s->cap * sizeof(double));
    memset(s->pe, 0, s->cap * sizeof(double));
    memset(s->pe_lj, 0, s->cap * sizeof(double));
    memset(s->vir, 0, s->cap * sizeof(double));
    memset(s->fscale, 0, s->cap * sizeof(double));
    int rec = fmod_i(iter, s->p.tplot) == 0;
    int ev = s->p.teval > 0 && fmod_i(iter, s->p.teval) == 0;
    for (int i = 0; i < s->np; ++i) {
        for (int j = i + 1; j < s->np; ++j) {
            double rij[3], rij2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                rij[d] = s->r[3 * i + d] - s->r[3 * j + d];
                if (fabs(rij[d]) > s->halfdomain[d]) rij[d] = rij[d] - f_sign(s->p.domain[d], rij[d]);
                rij2 = rij2 + rij[d] * rij[d];
            }
            if (rij2 < cutoff2(s, i, j)) {
                double invrij2 = 1.0 / rij2;
                double acc = accijmag(s, invrij2, i, j);
                double mi_ = mass_of(s, i), mj_ = mass_of(s, j);
                for (int d = 0; d < 3; ++d) s->a[3 * i + d] = s->a[3 * i + d] + acc * rij[d] / mi_;
                for (int d = 0; d < 3; ++d) s->a[3 * j + d] = s->a[3 * j + d] - acc * rij[d] / mj_;
                double fmag = fabs(acc) * sqrt(rij2);
                s->fscale[i] += fmag;
                s->fscale[j] += fmag;
                if (rec) {
                    s->pe_lj[i] = s->pe_lj[i] + energy(s, invrij2, i, j);
                    s->pe_lj[j] = s->pe_lj[j] + energy(s, invrij2, i, j);
                    s->vir[i] = s->vir[i] + acc * rij2;
                    s->vir[j] = s->vir[j] + acc * rij2;
                } else if (ev) {
                    s->vir[i] = s->vir[i] + acc * rij2;
                    s->vir[j] = s->vir[j] + acc * rij2;
                }
            }
        }
    }
    for (int n = 0; n < s->np; ++n) s->pe[n] = s->pe[n] + s->pe_lj[n];
    return 0;
}

/* simulation_move_particles_vv NVE (src/simulation_move_particles_vv.f90:64-69, 132-135, 191).
   Tethered molecules (ensemble tag_move) are integrated with the same velocity Verlet update;
   their tether force is part of a. */
int orc_move_particles_vv(orc_state* s, int32_t pass_num) {
    double dt = s->p.delta_t;
    if (pass_num == 1) {
        for (int n = 0; n < s->np; ++n)
            for (int d = 0; d < 3; ++d) {
                double* r = &s->r[3 * n + d];
                double* v = &s->v[3 * n + d];
                double a = s->a[3 * n + d];
                *r = *r + dt * *v + 0.5 * (dt * dt) * a;
                *v = *v + 0.5 * dt * a;
            }
    } else if (pass_num == 2) {
        for (int n = 0; n < s->np; ++n)
            for (int d = 0; d < 3; ++d) s->v[3 * n + d] = s->v[3 * n + d] + 0.5 * dt * s->a[3 * n + d];
        s->simtime = s->simtime + dt;
    } else {
        return fail("pass_num must be 1 or 2");
    }
    return 0;
}

/* simulation_move_particles_lfv (src/simulation_move_particles_lfv.f90:38-177): NVE, or tag_move
   through simulation_move_particles_lfv_tag (308-490) for the tags without a thermostat; the
   tethers act here, inside the move. */
int orc_move_particles_lfv(orc_state* s) {
    const double dt = s->p.delta_t;
    if (s->p.integration_algorithm != FM_INTEGRATION_LFV) return fail("integration_algorithm is not leap-frog");
    for (int n = 0; n < s->np; ++n) {
        double* r = &s->r[3 * n];
        double* v = &s->v[3 * n];
        double* a = &s->a[3 * n];
        const double* sv = &s->slidev[3 * n];
        int tag = s->p.ensemble == FM_ENSEMBLE_TAG_MOVE ? s->tag[n] : FM_TAG_FREE;
        switch (tag) {
        case FM_TAG_FREE:
            for (int d = 0; d < 3; ++d) v[d] = v[d] + dt * a[d];
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * v[d];
            break;
        case FM_TAG_FIXED:
            break;
        case FM_TAG_FIXED_SLIDE:
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * sv[d];
            break;
        case FM_TAG_TETH:
            tether_force(s, n);
            for (int d = 0; d < 3; ++d) v[d] = v[d] + dt * a[d];
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * v[d];
            break;
        case FM_TAG_TETH_SLIDE:
            tether_force(s, n);
            for (int d = 0; d < 3; ++d) v[d] = v[d] + dt * a[d];
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * v[d] + dt * sv[d];
            break;
        default:
            return fail("Invalid molecular Tag %d for molecule %d", tag, n + 1);
        }
    }
    s->simtime = s->simtime + dt;
    return 0;
}

/* sort_mols (src/linklist.f90:2744-2865): every sort_freq-th call, renumber the molecules block
   by block in Hcurve order, keeping their relative order inside a block; a block with more
   molecules than the estimate raises the saved safety factors and skips this round.  r, v, tag
   and rtether move with their molecules (a is recomputed before it is used again). */
int orc_sort_mols(orc_state* s) {
    if (s->sort_nrebuilds != s->p.sort_freq) {
        s->sort_nrebuilds = s->sort_nrebuilds + 1;
        return 0;
    }
    s->sort_nrebuilds = 1;
    if (s->p.sort_flag == 0) return 0;
    if (s->p.potential_flag == 1) return fail("Sort should be turned off - Not developed for polymers");
    if (s->p.mie_potential == 1) return fail("Sort should be turned off - Not developed for Mie potential");
    const int* nb = s->sort_nblocks;
    const int blocks = nb[0] * nb[1] * nb[2];
    int ave = s->np / blocks;
    ave = f_ceiling(ave * s->sort_sf1 + s->sort_sf2);
    int* blk = malloc((size_t)(s->np > 0 ? s->np : 1) * sizeof(int));
    int* cnt = calloc((size_t)blocks + 2, sizeof(int));
    if (!blk || !cnt) { free(blk); free(cnt); return fail("out of memory"); }
    for (int n = 0; n < s->np; ++n) {
        int b[3];
        for (int d = 0; d < 3; ++d) {
            b[d] = f_ceiling((s->r[3 * n + d] + s->halfdomain[d]) / s->sort_bsl[d]);
            if (b[d] < 1 || b[d] > nb[d]) { free(blk); free(cnt); return fail("sort_mols: molecule %d outside the blocks", n + 1); }
        }
        const int i = s->hcurve[((size_t)b[2] - 1) * nb[0] * nb[1] + ((size_t)b[1] - 1) * nb[0] + (b[0] - 1)];
        blk[n] = i;
        cnt[i] = cnt[i] + 1;
        if (cnt[i] > ave) {
            s->sort_sf1 = s->sort_sf1 + 0.1;
            s->sort_sf2 = s->sort_sf2 + 5;
            free(blk); free(cnt);
            return 0;
        }
    }
    /* start of each block in the new numbering, then the molecules in their old order */
    int* pos = malloc((size_t)(blocks + 2) * sizeof(int));
    int* newof = malloc((size_t)(s->np > 0 ? s->np : 1) * sizeof(int));
    if (!pos || !newof) { free(blk); free(cnt); free(pos); free(newof); return fail("out of memory"); }
    pos[1] = 0;
    for (int i = 1; i <= blocks; ++i) pos[i + 1] = pos[i] + cnt[i];
    for (int n = 0; n < s->np; ++n) newof[n] = pos[blk[n]]++;
    double* r = malloc(3 * (size_t)(s->np > 0 ? s->np : 1) * sizeof(double));
    double* v = malloc(3 * (size_t)(s->np > 0 ? s->np : 1) * sizeof(double));
    double* rt = malloc(3 * (size_t)(s->np > 0 ? s->np : 1) * sizeof(double));
    int32_t* tg = malloc((size_t)(s->np > 0 ? s->np : 1) * sizeof(int32_t));
    if (!r || !v || !rt || !tg) { free(blk); free(cnt); free(pos); free(newof); free(r); free(v); free(rt); free(tg); return fail("out of memory"); }
    for (int n = 0; n < s->np; ++n) {
        const int m = newof[n];
        for (int d = 0; d < 3; ++d) {
            r[3 * m + d] = s->r[3 * n + d];
            v[3 * m + d] = s->v[3 * n + d];
            rt[3 * m + d] = s->rtether[3 * n + d];
        }
        tg[m] = s->tag[n];
    }
    memcpy(s->r, r, 3 * (size_t)s->np * sizeof(double));
    memcpy(s->v, v, 3 * (size_t)s->np * sizeof(double));
    if (s->p.ensemble == FM_ENSEMBLE_TAG_MOVE) {
        memcpy(s->rtether, rt, 3 * (size_t)s->np * sizeof(double));
        memcpy(s->tag, tg, (size_t)s->np * sizeof(int32_t));
        for (int n = 0; n < s->np; ++n) read_tag(s, n);
    }
    free(blk); free(cnt); free(pos); free(newof); free(r); free(v); free(rt); free(tg);
    return 0;
}

/* simulation_checkrebuild (src/simulation_checkrebuild.f90:17-174), criteria 0, 1, 2 */
int orc_checkrebuild(orc_state* s, int32_t iter, int32_t* rebuild) {
    int rb = 0;
    if (fmod_i(iter + 1, s->p.tplot) == 0) rb = 1;
    double dmin = s->p.delta_rneighbr[0];
    if (s->p.delta_rneighbr[1] < dmin) dmin = s->p.delta_rneighbr[1];
    if (s->p.delta_rneighbr[2] < dmin) dmin = s->p.delta_rneighbr[2];
    double dt = s->p.delta_t;
    switch (s->p.rebuild_criteria) {
    case FM_REBUILD_VMAX: {
        double vmax = 0.0;
        for (int n = 0; n < s->np; ++n) {
            const double* v = &s->v[3 * n];
            double m = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
            if (n == 0 || m > vmax) vmax = m;
        }
        vmax = sqrt(vmax);
        s->rmax = s->rmax + vmax * dt;
        if (s->rmax > 0.5 * dmin) rb = 1;
        break;
    }
    case FM_REBUILD_FIXED:
        if (fmod_i(iter, s->p.fixed_rebuild) == 0) rb = 1;
        break;
    case FM_REBUILD_DISPLACEMENT:
        if (!s->rdisp_alloc) {
            free(s->rdisp);
            s->rdisp = calloc(3 * (size_t)(s->np > 0 ? s->np : 1), sizeof(double));
            if (!s->rdisp) return fail("out of memory");
            s->rdisp_alloc = 1;
        }
        for (int n = 0; n < s->np; ++n) {
            int hit = 0;
            for (int d = 0; d < 3; ++d) {
                s->rdisp[3 * n + d] = s->rdisp[3 * n + d] + s->v[3 * n + d] * dt;
                if (fabs(s->rdisp[3 * n + d]) > 0.5 * dmin) hit = 1;
            }
            if (hit) { rb = 1; break; }
        }
        break;
    default:
        return fail("unknown rebuild criteria");
    }
    if (rb) {
        s->rmax = 0.0;
        if (s->p.rebuild_criteria == FM_REBUILD_DISPLACEMENT) s->rdisp_alloc = 0;
    }
    *rebuild = rb;
    return 0;
}

/* evaluate_macroscopic_properties (src/simulation_record.f90:1205-1314), velocity Verlet branch */
int orc_macros(orc_state* s, fm_macros* o) {
    memset(o, 0, sizeof *o);
    int np = s->np;
    double pe = 0.0, pl = 0.0, pp = 0.0;
    if (s->p.potential_flag == 0) {
        for (int n = 0; n < np; ++n) pe = pe + s->pe[n];
        pl = pe;
    } else {
        for (int n = 0; n < np; ++n) pl = pl + s->pe_lj[n];
        for (int n = 0; n < np; ++n) pp = pp + s->pe_poly[n];
        for (int n = 0; n < np; ++n) pe = pe + (s->pe_lj[n] + s->pe_poly[n]);
    }
    double msum = 0.0, vsum = 0.0, mv2 = 0.0;
    for (int n = 0; n < np; ++n) {
        double m = mass_of(s, n);
        msum = msum + m;
        for (int d = 0; d < 3; ++d) {
            /* leap-frog: velocity shifted half a step, src/simulation_record.f90:1238-1247 */
            double vel = s->p.integration_algorithm == FM_INTEGRATION_LFV
                             ? s->v[3 * n + d] + 0.5 * s->a[3 * n + d] * s->p.delta_t
                             : s->v[3 * n + d];
            vsum = vsum + m * vel;
            mv2 = mv2 + m * (vel * vel);
        }
    }
    double vir = 0.0, mx = 0.0;
    for (int n = 0; n < np; ++n) vir = vir + s->vir[n];
    for (int n = 0; n < np; ++n) if (n == 0 || s->pe[n] > mx) mx = s->pe[n];
    o->potenergysum = pe;
    o->potenergysum_LJ = pl;
    o->potenergysum_POLY = pp;
    o->virial = vir;
    o->mv2sum = mv2;
    o->vsum = vsum;
    o->msum = msum;
    o->kinenergy = (0.5 * mv2) / (double)np;
    o->potenergy = (0.5 * pe) / (double)np;
    o->totenergy = o->kinenergy + o->potenergy;
    o->temperature = mv2 / (double)(3 * np);
    o->pressure = (s->p.density / (double)(np * 3)) * (mv2 + vir / 2.0);
    o->max_potenergymol = mx;
    return 0;
}

int orc_get_mop(const orc_state* s, double* Pxy) {
    if (!s->Pxy_plane) return fail("MOP not enabled");
    memcpy(Pxy, s->Pxy_plane, 3 * (size_t)s->nplanes * sizeof(double));
    return 0;
}

int orc_reset_mop(orc_state* s) {
    if (s->Pxy_plane) memset(s->Pxy_plane, 0, 3 * (size_t)s->nplanes * sizeof(double));
    return 0;
}

int orc_get_cv_bins(const orc_state* s, int32_t nbins[3], int32_t nhb[3], double binsize[3]) {
    if (!s->Pxyface) return fail("CV stresses not enabled (vflux_outflag 4)");
    for (int d = 0; d < 3; ++d) { nbins[d] = s->nbins[d]; nhb[d] = s->nhb[d]; binsize[d] = s->binsize[d]; }
    return 0;
}

int orc_get_cv_stress(const orc_state* s, double* pxyface) {
    if (!s->Pxyface) return fail("CV stresses not enabled (vflux_outflag 4)");
    memcpy(pxyface, s->Pxyface, (size_t)s->nbinso[0] * s->nbinso[1] * s->nbinso[2] * 18 * sizeof(double));
    return 0;
}

int orc_reset_cv_stress(orc_state* s) {
    if (s->Pxyface) memset(s->Pxyface, 0, (size_t)s->nbinso[0] * s->nbinso[1] * s->nbinso[2] * 18 * sizeof(double));
    return 0;
}

int orc_get_planes(const orc_state* s, double* planes) {
    if (!s->planes) return fail("MOP not enabled");
    memcpy(planes, s->planes, s->nplanes * sizeof(double));
    return 0;
}

/* one rebuild as in md_advance_vv (src/md.f90:150-157) */
static int rebuild_all(orc_state* s) {
    if (orc_linklist_deallocateall(s)) return 1;
    if (orc_sendmols(s)) return 1;
    if (orc_sort_mols(s)) return 1;
    if (orc_assign_to_cell(s)) return 1;
    if (orc_updateborders(s, 1)) return 1;
    return orc_assign_to_neighbourlist(s);
}

/* md_advance_vv (src/md.f90:143-163) */
int orc_run_vv(orc_state* s, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
               int32_t* n_rebuilds) {
    int nrb = 0, nm = 0;
    if (!s->nlist_valid) return fail("run_vv before the initial neighbour list");
    for (int iter = iter0; iter < iter0 + nsteps; ++iter) {
        int32_t rb;
        if (orc_move_particles_vv(s, 1)) return 1;
        if (orc_updateborders(s, 0)) return 1;
        if (orc_checkrebuild(s, iter, &rb)) return 1;
        if (rb) {
            if (rebuild_all(s)) return 1;
            nrb++;
        }
        if (orc_compute_forces(s, iter)) return 1;
        if (orc_move_particles_vv(s, 2)) return 1;
        if (fmod_i(iter, s->p.tplot) == 0 && macros && nm < n_macros) {
            orc_macros(s, &macros[nm]);
            nm++;
        }
    }
    if (n_rebuilds) *n_rebuilds = nrb;
    return 0;
}

/* md_advance_lfv (src/md.f90:109-139) nsteps times: forces, record, move, rebuild check */
int orc_run_lfv(orc_state* s, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
                int32_t* n_rebuilds) {
    int nrb = 0, nm = 0;
    if (!s->nlist_valid) return fail("run_lfv before the initial neighbour list");
    for (int iter = iter0; iter < iter0 + nsteps; ++iter) {
        int32_t rb;
        if (orc_compute_forces(s, iter)) return 1;
        if (fmod_i(iter, s->p.tplot) == 0 && macros && nm < n_macros) {
            orc_macros(s, &macros[nm]);
            nm++;
        }
        if (orc_move_particles_lfv(s)) return 1;
        if (orc_updateborders(s, 0)) return 1;
        if (orc_checkrebuild(s, iter, &rb)) return 1;
        if (rb) {
            if (rebuild_all(s)) return 1;
            nrb++;
        }
    }
    if (n_rebuilds) *n_rebuilds = nrb;
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
int orc_get_cells(const orc_state* s, int32_t* cell_of_mol) {
    if (!s->cells_valid) return fail("no cells");
    for (int n = 0; n < s->np; ++n) cell_of_mol[n] = -1;
    for (int c = 0; c < s->ncelltot; ++c)
        for (int n = s->cell_head[c]; n >= 0; n = s->mol_next[n])
            if (n < s->np) cell_of_mol[n] = c;
    return 0;
}

static inline int shift_code(int sx, int sy, int sz) { return (sx + 1) * 9 + (sy + 1) * 3 + (sz + 1); }

static inline void resolve(const orc_state* s, int m, int* src, int* code) {
    if (m < s->np) { *src = m; *code = 13; return; }
    *src = s->halo_src[m];
    *code = shift_code(s->halo_shift[3 * m], s->halo_shift[3 * m + 1], s->halo_shift[3 * m + 2]);
}

int64_t orc_neighbour_count(const orc_state* s) { return s->nlist_valid ? s->nnodes : -1; }

int orc_get_neighbourlist(const orc_state* s, int32_t* mol_i, int32_t* mol_j, int32_t* shift) {
    if (!s->nlist_valid) return fail("no neighbour list");
    int64_t k = 0;
    for (int i = 0; i < s->np; ++i)
        for (int nd = s->nb_head[i]; nd >= 0; nd = s->nodes[nd].next) {
            int src, code;
            resolve(s, s->nodes[nd].mol, &src, &code);
            mol_i[k] = i; mol_j[k] = src; shift[k] = code;
            ++k;
        }
    return 0;
}

int orc_get_bonds(const orc_state* s, int32_t* bond, int32_t* bond_shift) {
    for (int n = 0; n < s->np; ++n)
        for (int b = 0; b < FM_MAX_FUNCY; ++b) {
            int m = s->bond[b + FM_MAX_FUNCY * n];
            if (m < 0) { bond[b + FM_MAX_FUNCY * n] = -1; bond_shift[b + FM_MAX_FUNCY * n] = 13; continue; }
            int src, code;
            resolve(s, m, &src, &code);
            bond[b + FM_MAX_FUNCY * n] = src;
            bond_shift[b + FM_MAX_FUNCY * n] = code;
        }
    return 0;
}

/* Brute force: for every ordered pair (i, j != i) and every shift s in {-1,0,1}^3 whose image
   would be visible from i's cell (the image's cell is i's cell or one of its 26 neighbours under
   the reference's own cell assignment), test the reference's image arithmetic. */
int64_t orc_bruteforce_pairs(const orc_state* s, int32_t* mol_i, int32_t* mol_j, int32_t* shift,
                             int64_t capacity) {
    int64_t cnt = 0;
    int np = s->np;
    int* cc = malloc(3 * (size_t)(np > 0 ? np : 1) * sizeof(int));
    if (!cc) return -1;
    for (int n = 0; n < np; ++n)
        for (int d = 0; d < 3; ++d)
            cc[3 * n + d] = f_ceiling((s->r[3 * n + d] + s->halfdomain[d]) / s->cellside[d]) + 1;
    for (int i = 0; i < np; ++i) {
        const double* ri = &s->r[3 * i];
        for (int j = 0; j < np; ++j) {
            if (j == i) continue;
            for (int sx = -1; sx <= 1; ++sx)
                for (int sy = -1; sy <= 1; ++sy)
                    for (int sz = -1; sz <= 1; ++sz) {
                        int sh[3] = {sx, sy, sz};
                        int ok = 1;
                        double img[3];
                        for (int d = 0; d < 3 && ok; ++d) {
                            int nd = s->ncells[d];
                            int cj = cc[3 * j + d];
                            /* only boundary cells are imaged: +L from cell 2, -L from cell n+1 */
                            if (sh[d] == 1 && cj != 2) ok = 0;
                            if (sh[d] == -1 && cj != nd + 1) ok = 0;
                            int cimg = cj + sh[d] * nd;
                            if (abs(cimg - cc[3 * i + d]) > 1) ok = 0;
                            img[d] = sh[d] == 0 ? s->r[3 * j + d] : s->r[3 * j + d] + sh[d] * s->p.domain[d];
                        }
                        if (!ok) continue;
                        double rij2 = dist2(ri, img);
                        if (rij2 < s->rneighbr2) {
                            if (mol_i && cnt < capacity) {
                                mol_i[cnt] = i; mol_j[cnt] = j; shift[cnt] = shift_code(sx, sy, sz);
                            }
                            ++cnt;
                        }
                    }
        }
    }
    free(cc);
    return cnt;
}
