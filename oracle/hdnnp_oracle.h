/* hdnnp_oracle.h -- CPU restatement of the reference HDNNP evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker: tests/, the smoke check and
 * bench.py's cpu_baseline leg may call it; no product code links it.
 *
 * Plain C, FP64, following the reference Fortran arithmetic step by step
 * (file:line of each routine in hdnnp_oracle.c). Memory layouts follow the
 * reference's atStruct: pos[nat][3] = ats(3,nat); lat[3][3] with row k =
 * lattice vector k = lat(:,k); force like pos; dEdLat[a][b] = dEdLat(b+1,a+1)
 * in Fortran terms, i.e. the same 9 doubles in the same memory order.
 */
#ifndef HDNNP_ORACLE_H
#define HDNNP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXEL 16
#define ORC_MAXLAYERS 8

typedef struct {
    int nel;                     /* number of elements, sorted by Z        */
    int Z[ORC_MAXEL];
    int nsf[ORC_MAXEL];          /* symmetry functions per element         */
    int sf_off[ORC_MAXEL];       /* offset of element e in the sf_* arrays */
    const int *sf_type;          /* 2 = G2, 3 = G4                         */
    const int *sf_e1, *sf_e2;    /* neighbour element(s) as Z              */
    const double *sf_rcut, *sf_eta, *sf_rs, *sf_lambda, *sf_zeta;
    const double *sf_min, *sf_max, *sf_mean;
    int scalesf, centersf;
    double atom_energy[ORC_MAXEL];
    double maxrc;
    int nlayers;                 /* weight layers (hidden + output)        */
    int nodes[ORC_MAXLAYERS];    /* nOut of each layer                     */
    char act[ORC_MAXLAYERS];     /* 'l' 't' 'p' 'r'                        */
    const double *weights[ORC_MAXEL]; /* per layer W(i,j) at i+nOut*j, then bias */
    /* RuNNer keywords the reference does not read */
    int cutoff_type;             /* 0..8; 2 = the reference's tanh^3        */
    int scale_range;             /* scale_min_short / scale_max_short given */
    double scale_min, scale_max;
} orc_model;

/* linalg.f90:264-289 */
void orc_inv3(const double A[9], double B[9]);
/* pointsets.f90:201-217 */
void orc_wrap(int nat, const double lat[9], double *pos);
/* cutoff of RuNNer's cutoff_type; type 2 is symmetryFunctions.f90:310-327 */
void orc_cutoff(int type, double r, double rc, double *fc, double *dfc);

/* Neighbour entries of buildNeighborListGrid (neighborLists.f90:104-230),
 * preceded by the wrap of hdnnpCalcNeighborLists (hdnnps.f90:373-388),
 * in the reference's list order, self slot excluded: ent[n] = (i, j, Lx, Ly,
 * Lz). Returns the number of entries, or -(needed) if cap is too small. */
long orc_neighbor_entries(int nat, const double *pos, int periodic, const double *lat,
                          double rc, long cap, int *ent);

/* hdnnpEnergyAndForces (hdnnps.f90:333-371). dEdLat may be NULL (it is
 * only formed for periodic input). G and dEdG (scaled SFs and dE/dG, per
 * atom, row stride ldg) may be NULL. nthreads <= 0 means OpenMP default.
 * Returns 0 on success. */
int orc_evaluate(const orc_model *m, int nat, const int *Z, const double *pos,
                 int periodic, const double *lat, double *energy, double *force,
                 double *dEdLat, double *G, double *dEdG, int ldg, int nthreads);

/* Like orc_evaluate but only the SF values, atomic energies and dE/dG of the
 * atoms listed in sel (full neighbour list, per-atom work only for sel). */
int orc_atoms(const orc_model *m, int nat, const int *Z, const double *pos, int periodic,
              const double *lat, int nsel, const int *sel, double *Ei, double *G,
              double *dEdG, int ldg);

#ifdef __cplusplus
}
#endif
#endif
