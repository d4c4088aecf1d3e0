/* hdnnp_oracle.c -- plain-C FP64 restatement of the reference HDNNP path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Each routine cites the reference
 * Fortran it restates and keeps its operation order where that order decides
 * results bit for bit (cell wrap, image shifts, the strict distance test).
 * Build with -ffp-contract=off so no multiply-add is fused.
 *
 * Index helper: a Fortran 3x3 array A(r,c) (column-major) is A[r + 3c] with
 * 0-based r, c; lat(:,k) (lattice vector k) is lat[3k .. 3k+2].
 */
#include "hdnnp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define M3(A, r, c) (A)[(r) + 3 * (c)]

/* det3D, linalg.f90:227-238 */
static double det3(const double *M)
{
    double t = M3(M, 0, 0) * M3(M, 1, 1) * M3(M, 2, 2);
    t = t + M3(M, 0, 1) * M3(M, 1, 2) * M3(M, 2, 0);
    t = t + M3(M, 0, 2) * M3(M, 1, 0) * M3(M, 2, 1);
    t = t - M3(M, 0, 0) * M3(M, 1, 2) * M3(M, 2, 1);
    t = t - M3(M, 0, 1) * M3(M, 1, 0) * M3(M, 2, 2);
    t = t - M3(M, 0, 2) * M3(M, 1, 1) * M3(M, 2, 0);
    return t;
}

/* inv3DM, linalg.f90:264-289: adjugate rows divided by det3D */
void orc_inv3(const double A[9], double B[9])
{
    const double a1[3] = {M3(A, 0, 0), M3(A, 0, 1), M3(A, 0, 2)};
    const double a2[3] = {M3(A, 1, 0), M3(A, 1, 1), M3(A, 1, 2)};
    const double a3[3] = {M3(A, 2, 0), M3(A, 2, 1), M3(A, 2, 2)};
    const double det = det3(A);
    const double b1[3] = {a2[1] * a3[2] - a2[2] * a3[1], a1[2] * a3[1] - a1[1] * a3[2],
                          a1[1] * a2[2] - a1[2] * a2[1]};
    const double b2[3] = {a2[2] * a3[0] - a2[0] * a3[2], a1[0] * a3[2] - a1[2] * a3[0],
                          a1[2] * a2[0] - a1[0] * a2[2]};
    const double b3[3] = {a2[0] * a3[1] - a2[1] * a3[0], a1[1] * a3[0] - a1[0] * a3[1],
                          a1[0] * a2[1] - a1[1] * a2[0]};
    for (int c = 0; c < 3; ++c) {
        M3(B, 0, c) = b1[c] / det;
        M3(B, 1, c) = b2[c] / det;
        M3(B, 2, c) = b3[c] / det;
    }
}

/* matMulVec3, linalg.f90:194-203: res(r) = sum(mat(r,:) * vec(:)) */
static void matvec3(const double *M, const double *v, double *out)
{
    double r[3];
    for (int i = 0; i < 3; ++i)
        r[i] = M3(M, i, 0) * v[0] + M3(M, i, 1) * v[1] + M3(M, i, 2) * v[2];
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

/* moveAtomsIntoCell, pointsets.f90:201-217: fractional (toRelativeCoordinates
 * :174-187), modulo(s, 1), back to Cartesian (toAbsoluteCoordinates :189-199).
 * modulo(s,1) is the exact s - floor(s) rounded once, as s - floor(s) is. */
void orc_wrap(int nat, const double lat[9], double *pos)
{
    double inv[9];
    orc_inv3(lat, inv);
    for (int i = 0; i < nat; ++i) {
        double s[3];
        matvec3(inv, pos + 3 * i, s);
        for (int k = 0; k < 3; ++k) s[k] = s[k] - floor(s[k]);
        matvec3(lat, s, pos + 3 * i);
    }
}

/* Cutoff functions, value fc and dfc = fc'(r)/r, by RuNNer's cutoff_type
 * (x = r/rc): 2 is the reference's, symmetryFunctions.f90:310-327,
 * fc = tanh^3(1 - x), dfc = -3/(rc r) (t^2 - t^4). The others are written
 * here from their definitions, as polynomials in x:
 *   0 hard: 1
 *   1 cosine: (cos(pi x) + 1) / 2
 *   3 tanh^3(1 - x) / tanh^3(1)
 *   4 exp(1 - 1/(1 - x^2))
 *   5 1 - 3x^2 + 2x^3
 *   6 1 - 10x^3 + 15x^4 - 6x^5
 *   7 1 - 35x^4 + 84x^5 - 70x^6 + 20x^7
 *   8 1 - 126x^5 + 420x^6 - 540x^7 + 315x^8 - 70x^9
 * fc = dfc = 0 from rc on. */
void orc_cutoff(int type, double r, double rc, double *fc, double *dfc)
{
    if (!(r < rc)) {
        *fc = 0.0;
        *dfc = 0.0;
        return;
    }
    const double x = r / rc;
    double f, d;   /* f(x), f'(x) */
    switch (type) {
    case 0:
        f = 1.0;
        d = 0.0;
        break;
    case 1:
        f = 0.5 * (cos(M_PI * x) + 1.0);
        d = -0.5 * M_PI * sin(M_PI * x);
        break;
    case 2:
    case 3: {
        const double t = tanh(1.0 - r / rc);
        const double tt = t * t;
        const double nrm = type == 3 ? 1.0 / pow(tanh(1.0), 3) : 1.0;
        *fc = nrm * (tt * t);
        *dfc = nrm * -(3.0 / (rc * r) * (tt - tt * tt));
        return;
    }
    case 4: {
        const double u = 1.0 - x * x;
        f = exp(1.0 - 1.0 / u);
        d = f * (-2.0 * x / (u * u));
        break;
    }
    default: {
        /* the expanded monomials cancel terms of up to ~500 near x = 1:
         * summed in extended precision, the result is exact to double */
        const long double X = (long double)r / (long double)rc;
        long double F, D;
        if (type == 5) {
            F = 1.0L - 3.0L * powl(X, 2) + 2.0L * powl(X, 3);
            D = -6.0L * X + 6.0L * powl(X, 2);
        } else if (type == 6) {
            F = 1.0L - 10.0L * powl(X, 3) + 15.0L * powl(X, 4) - 6.0L * powl(X, 5);
            D = -30.0L * powl(X, 2) + 60.0L * powl(X, 3) - 30.0L * powl(X, 4);
        } else if (type == 7) {
            F = 1.0L - 35.0L * powl(X, 4) + 84.0L * powl(X, 5) - 70.0L * powl(X, 6) + 20.0L * powl(X, 7);
            D = -140.0L * powl(X, 3) + 420.0L * powl(X, 4) - 420.0L * powl(X, 5) + 140.0L * powl(X, 6);
        } else {
            F = 1.0L - 126.0L * powl(X, 5) + 420.0L * powl(X, 6) - 540.0L * powl(X, 7) + 315.0L * powl(X, 8) -
                70.0L * powl(X, 9);
            D = -630.0L * powl(X, 4) + 2520.0L * powl(X, 5) - 3780.0L * powl(X, 6) + 2520.0L * powl(X, 7) -
                630.0L * powl(X, 8);
        }
        f = (double)F;
        d = (double)D;
        break;
    }
    }
    *fc = f;
    *dfc = d / (rc * r);
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* getNcells, pointsets.f90:220-243 */
static void get_ncells(const double *L, double rc, int n[3])
{
    double axb[3], axc[3], bxc[3], nrm;
    cross3(L + 0, L + 3, axb);
    cross3(L + 0, L + 6, axc);
    cross3(L + 3, L + 6, bxc);
    nrm = sqrt(axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2]);
    for (int k = 0; k < 3; ++k) axb[k] = axb[k] / nrm;
    nrm = sqrt(axc[0] * axc[0] + axc[1] * axc[1] + axc[2] * axc[2]);
    for (int k = 0; k < 3; ++k) axc[k] = axc[k] / nrm;
    nrm = sqrt(bxc[0] * bxc[0] + bxc[1] * bxc[1] + bxc[2] * bxc[2]);
    for (int k = 0; k < 3; ++k) bxc[k] = bxc[k] / nrm;
    const double p0 = L[0] * bxc[0] + L[1] * bxc[1] + L[2] * bxc[2];
    const double p1 = L[3] * axc[0] + L[4] * axc[1] + L[5] * axc[2];
    const double p2 = L[6] * axb[0] + L[7] * axb[1] + L[8] * axb[2];
    n[0] = (int)ceil(rc / fabs(p0));
    n[1] = (int)ceil(rc / fabs(p1));
    n[2] = (int)ceil(rc / fabs(p2));
}

typedef struct {
    int n, cap;
    int *j;        /* neighbour atom index            */
    int *L;        /* image shift, 3 per entry        */
    double *p;     /* image position, 3 per entry     */
} nlist;

static void nl_push(nlist *l, int j, const int *L, const double *p)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 16;
        l->j = (int *)realloc(l->j, sizeof(int) * l->cap);
        l->L = (int *)realloc(l->L, sizeof(int) * 3 * l->cap);
        l->p = (double *)realloc(l->p, sizeof(double) * 3 * l->cap);
    }
    l->j[l->n] = j;
    memcpy(l->L + 3 * l->n, L, 3 * sizeof(int));
    memcpy(l->p + 3 * l->n, p, 3 * sizeof(double));
    l->n++;
}

static void nl_free(nlist *nl, int nat)
{
    for (int i = 0; i < nat; ++i) {
        free(nl[i].j);
        free(nl[i].L);
        free(nl[i].p);
    }
    free(nl);
}

/* hdnnpCalcNeighborLists (hdnnps.f90:373-388) + buildNeighborListGrid
 * (neighborLists.f90:104-230). selfpos receives the slot-1 position of each
 * atom (the structure handed to buildNeighborListGrid). Entries exclude the
 * self slot. */
static nlist *build_nl(int nat, const double *pos_in, int periodic, const double *lat_in,
                       double rc, double *selfpos)
{
    double lat[9], tlat[9], origin[3] = {0, 0, 0};
    double *ats = (double *)malloc(sizeof(double) * 3 * (nat ? nat : 1));
    double *tmp = (double *)malloc(sizeof(double) * 3 * (nat ? nat : 1));
    double *rel = (double *)malloc(sizeof(double) * 3 * (nat ? nat : 1));
    nlist *nl = (nlist *)calloc(nat ? nat : 1, sizeof(nlist));
    memcpy(ats, pos_in, sizeof(double) * 3 * nat);
    if (periodic) {
        memcpy(lat, lat_in, sizeof lat);
        orc_wrap(nat, lat, ats);                      /* hdnnps.f90:381 */
    } else {
        memset(lat, 0, sizeof lat);
    }
    memcpy(selfpos, ats, sizeof(double) * 3 * nat);   /* neighborLists.f90:125 */
    memcpy(tmp, ats, sizeof(double) * 3 * nat);
    memcpy(tlat, lat, sizeof tlat);
    if (!periodic) {                                   /* :131-138 */
        memset(tlat, 0, sizeof tlat);
        for (int a = 0; a < 3; ++a) {
            double mn = tmp[a];
            for (int i = 1; i < nat; ++i)
                if (tmp[3 * i + a] < mn) mn = tmp[3 * i + a];
            origin[a] = mn - 1.e-4;
            for (int i = 0; i < nat; ++i) tmp[3 * i + a] = tmp[3 * i + a] - origin[a];
            double mx = tmp[a];
            for (int i = 1; i < nat; ++i)
                if (tmp[3 * i + a] > mx) mx = tmp[3 * i + a];
            M3(tlat, a, a) = mx + 1.e-4;
        }
    }
    int ng[3], nc[3];
    double grid[9];
    const double target = rc * 0.6;                   /* :142 */
    for (int a = 0; a < 3; ++a) {
        const double *v = tlat + 3 * a;
        ng[a] = (int)(sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / target) + 1;
        for (int k = 0; k < 3; ++k) grid[3 * a + k] = v[k] / ng[a];
    }
    orc_wrap(nat, tlat, tmp);                         /* :156 */
    double inv[9];
    orc_inv3(tlat, inv);                              /* :159-160 */
    for (int i = 0; i < nat; ++i) matvec3(inv, tmp + 3 * i, rel + 3 * i);

    /* bin atoms (in atom order) -- :161-166; a fractional coordinate that
     * rounds to exactly 1 would index past the grid there; clamp it */
    const long ncell = (long)ng[0] * ng[1] * ng[2];
    int *cellof = (int *)malloc(sizeof(int) * 3 * (nat ? nat : 1));
    long *cstart = (long *)calloc(ncell + 1, sizeof(long));
    int *catoms = (int *)malloc(sizeof(int) * (nat ? nat : 1));
    for (int i = 0; i < nat; ++i) {
        for (int a = 0; a < 3; ++a) {
            int c = (int)(rel[3 * i + a] * ng[a]);
            if (c > ng[a] - 1) c = ng[a] - 1;
            if (c < 0) c = 0;
            cellof[3 * i + a] = c;
        }
        const long id = ((long)cellof[3 * i + 0] * ng[1] + cellof[3 * i + 1]) * ng[2] + cellof[3 * i + 2];
        cstart[id + 1]++;
    }
    for (long c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    {
        long *fill = (long *)malloc(sizeof(long) * (ncell ? ncell : 1));
        memcpy(fill, cstart, sizeof(long) * ncell);
        for (int i = 0; i < nat; ++i) {
            const long id = ((long)cellof[3 * i + 0] * ng[1] + cellof[3 * i + 1]) * ng[2] + cellof[3 * i + 2];
            catoms[fill[id]++] = i;
        }
        free(fill);
    }
    get_ncells(grid, rc, nc);                         /* :169 */
    /* Without periodicity only offsets that land inside the grid matter, so
     * the stencil is clamped to the grid: the same cells are visited, and a
     * flat input (artificial box 2e-4 thick) no longer loops over ~1e10
     * empty offsets. */
    if (!periodic)
        for (int a = 0; a < 3; ++a)
            if (nc[a] > ng[a] - 1) nc[a] = ng[a] - 1;
    const double rc2 = rc * rc;

#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nat; ++i) {                   /* :172-227 */
        const int ix = cellof[3 * i + 0] + 1, iy = cellof[3 * i + 1] + 1, iz = cellof[3 * i + 2] + 1;
        for (int jx = -nc[0]; jx <= nc[0]; ++jx) {
            for (int jy = -nc[1]; jy <= nc[1]; ++jy) {
                for (int jz = -nc[2]; jz <= nc[2]; ++jz) {
                    double dlat[3] = {0.0, 0.0, 0.0};
                    int cx = ix + jx, cy = iy + jy, cz = iz + jz;
                    int sh[3] = {0, 0, 0};
                    if (cz < 1 || cz > ng[2]) {
                        if (!periodic) continue;              /* cycle loop_jz */
                        const int ddl = (int)floor(1.0 * (cz - 1) / ng[2]);
                        cz = cz - ng[2] * ddl;
                        for (int k = 0; k < 3; ++k) dlat[k] = dlat[k] + lat[6 + k] * ddl;
                        sh[2] = ddl;
                    }
                    if (cy < 1 || cy > ng[1]) {
                        if (!periodic) break;                 /* cycle loop_jy */
                        const int ddl = (int)floor(1.0 * (cy - 1) / ng[1]);
                        cy = cy - ng[1] * ddl;
                        for (int k = 0; k < 3; ++k) dlat[k] = dlat[k] + lat[3 + k] * ddl;
                        sh[1] = ddl;
                    }
                    if (cx < 1 || cx > ng[0]) {
                        if (!periodic) goto next_jx;          /* cycle loop_jx */
                        const int ddl = (int)floor(1.0 * (cx - 1) / ng[0]);
                        cx = cx - ng[0] * ddl;
                        for (int k = 0; k < 3; ++k) dlat[k] = dlat[k] + lat[0 + k] * ddl;
                        sh[0] = ddl;
                    }
                    {
                        const long id = ((long)(cx - 1) * ng[1] + (cy - 1)) * ng[2] + (cz - 1);
                        for (long q = cstart[id]; q < cstart[id + 1]; ++q) {
                            const int jat = catoms[q];
                            if (i != jat || !(jx == 0 && jy == 0 && jz == 0)) {
                                double d[3], img[3];
                                for (int k = 0; k < 3; ++k) {
                                    img[k] = tmp[3 * jat + k] + dlat[k];
                                    d[k] = tmp[3 * i + k] - img[k];
                                }
                                if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rc2) {
                                    double p[3];
                                    for (int k = 0; k < 3; ++k) p[k] = img[k] + origin[k];
                                    nl_push(&nl[i], jat, sh, p);
                                }
                            }
                        }
                    }
                }
            }
        next_jx:;
        }
    }
    free(cellof);
    free(cstart);
    free(catoms);
    free(ats);
    free(tmp);
    free(rel);
    return nl;
}

long orc_neighbor_entries(int nat, const double *pos, int periodic, const double *lat,
                          double rc, long cap, int *ent)
{
    double *selfpos = (double *)malloc(sizeof(double) * 3 * (nat ? nat : 1));
    nlist *nl = build_nl(nat, pos, periodic, lat, rc, selfpos);
    long total = 0;
    for (int i = 0; i < nat; ++i) total += nl[i].n;
    if (total <= cap) {
        long q = 0;
        for (int i = 0; i < nat; ++i)
            for (int n = 0; n < nl[i].n; ++n, ++q) {
                ent[5 * q + 0] = i;
                ent[5 * q + 1] = nl[i].j[n];
                ent[5 * q + 2] = nl[i].L[3 * n + 0];
                ent[5 * q + 3] = nl[i].L[3 * n + 1];
                ent[5 * q + 4] = nl[i].L[3 * n + 2];
            }
    }
    nl_free(nl, nat);
    free(selfpos);
    return total <= cap ? total : -total;
}

/* integer power by binary exponentiation (the lowering of real**integer) */
static double powi(double a, int b)
{
    const int recip = b < 0;
    double r = 1.0;
    if (recip) b = -b;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static int elem_index(const orc_model *m, int z)
{
    for (int e = 0; e < m->nel; ++e)
        if (m->Z[e] == z) return e;
    return -1;
}

/* calcAllSymmetryFunctions, symmetryFunctions.f90:158-308, centre = slot 0.
 * P[nn][3] slot positions, els[nn] Z of each slot; sf[nsf], dsf[nsf][nn][3]. */
static int calc_all(const orc_model *m, int e, int nn, const double *P, const int *els,
                    double *sf, double *dsf, double *scratch)
{
    const int nsf = m->nsf[e], o = m->sf_off[e];
    int sf3start = -1;
    const double rcut = m->sf_rcut[o];
    for (int s = 0; s < nsf; ++s) {                   /* :183-199 (type 9 sorts after 3) */
        if (m->sf_type[o + s] != 2 && sf3start < 0) sf3start = s;
        else if (m->sf_type[o + s] == 2 && sf3start >= 0) return 3;   /* not sorted */
        const double dr = m->sf_rcut[o + s] - rcut;
        if (dr * dr > 1.e-10) return 4;                                /* mixed rc */
        sf[s] = 0.0;
    }
    if (sf3start < 0) sf3start = nsf;   /* no G4: all SFs are G2 */
    memset(dsf, 0, sizeof(double) * (size_t)nsf * nn * 3);
    double *dijs = scratch, *rijs = scratch + 3 * nn, *fcs = rijs + nn, *dfcs = fcs + nn;
    for (int j = 1; j < nn; ++j) {                    /* :201-208 */
        for (int k = 0; k < 3; ++k) dijs[3 * j + k] = P[k] - P[3 * j + k];
        const double *d = dijs + 3 * j;
        rijs[j] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        orc_cutoff(m->cutoff_type, rijs[j], rcut, &fcs[j], &dfcs[j]);
    }
    for (int j = 1; j < nn; ++j) {
        const double *dij = dijs + 3 * j;
        const double rij = rijs[j], fcij = fcs[j], dfcij = dfcs[j];
        for (int s = 0; s < sf3start; ++s) {         /* G2, :220-230 */
            if (m->sf_e1[o + s] != els[j]) continue;
            const double eta = m->sf_eta[o + s], rs = m->sf_rs[o + s];
            const double ex = exp(-(eta * ((rij - rs) * (rij - rs))));
            double *dj = dsf + ((size_t)s * nn + j) * 3, *di = dsf + (size_t)s * nn * 3;
            for (int k = 0; k < 3; ++k) {
                dj[k] = -(ex * dfcij * dij[k]) + 2.0 * eta * (rij - rs) * dij[k] / rij * ex * fcij;
                di[k] = di[k] - dj[k];
            }
            sf[s] = sf[s] + ex * fcij;
        }
        for (int kk = j + 1; kk < nn; ++kk) {         /* G4, :231-305 */
            const double *dik = dijs + 3 * kk;
            const double rik = rijs[kk], fcik = fcs[kk], dfcik = dfcs[kk];
            double djk[3];
            for (int k = 0; k < 3; ++k) djk[k] = P[3 * j + k] - P[3 * kk + k];
            const double rjk = sqrt(djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2]);
            double fcjk, dfcjk;
            orc_cutoff(m->cutoff_type, rjk, rcut, &fcjk, &dfcjk);
            const double t1 = -1.0 / (rij * rik);
            const double cost = 0.5 * (rjk * rjk - rij * rij - rik * rik) * t1;
            const double t2 = dij[0] * dik[0] + dij[1] * dik[1] + dij[2] * dik[2];
            double dcj[3], dck[3], dfj[3], dfk[3];
            const double rij3 = powi(rij, 3), rik3 = powi(rik, 3);
            for (int k = 0; k < 3; ++k) {
                dcj[k] = t1 * dik[k] + dij[k] * t2 / (rij3 * rik);
                dck[k] = t1 * dij[k] + dik[k] * t2 / (rik3 * rij);
            }
            const double fcpart = fcij * fcik * fcjk;
            for (int k = 0; k < 3; ++k) {
                dfj[k] = fcik * (dfcjk * djk[k] * fcij - dfcij * dij[k] * fcjk);
                dfk[k] = -(fcij * (dfcjk * djk[k] * fcik + dfcik * dik[k] * fcjk));
            }
            double lasteta = 0.0, ex = 1.0, dej[3] = {0, 0, 0}, dek[3] = {0, 0, 0};
            for (int s = sf3start; s < nsf; ++s) {
                const int e1 = m->sf_e1[o + s], e2 = m->sf_e2[o + s];
                if (!((els[j] == e1 && els[kk] == e2) || (els[j] == e2 && els[kk] == e1))) continue;
                if (m->sf_type[o + s] == 9) {
                    /* RuNNer's wide angular function G5, which the reference does
                     * not have: 2^(1-zeta) (1 + lambda cos)^zeta
                     * exp(-eta (r_ij^2 + r_ik^2)) fc(r_ij) fc(r_ik), restated from
                     * that definition with the G4 terms above: no fc(r_jk), no
                     * r_jk^2 in the exponential */
                    const double zeta = m->sf_zeta[o + s], lam = m->sf_lambda[o + s];
                    const double eta = m->sf_eta[o + s];
                    const double b = 1.0 + lam * cost;
                    const double pre = pow(2.0, 1.0 - zeta);
                    const double tp = b > 0.0 ? pow(b, zeta - 1.0) : 0.0, cosp = tp * b;
                    const double ex5 = exp(-(eta * (rij * rij + rik * rik)));
                    const double fc5 = fcij * fcik;
                    double *dj = dsf + ((size_t)s * nn + j) * 3, *dk = dsf + ((size_t)s * nn + kk) * 3;
                    double *di = dsf + (size_t)s * nn * 3;
                    for (int k = 0; k < 3; ++k) {
                        const double dfj5 = -(fcik * dfcij * dij[k]), dfk5 = -(fcij * dfcik * dik[k]);
                        const double dej5 = ex5 * eta * 2.0 * dij[k], dek5 = ex5 * eta * 2.0 * dik[k];
                        const double gj = pre * (cosp * ex5 * dfj5 + cosp * dej5 * fc5 +
                                                 tp * lam * zeta * dcj[k] * ex5 * fc5);
                        const double gk = pre * (cosp * ex5 * dfk5 + cosp * dek5 * fc5 +
                                                 tp * lam * zeta * dck[k] * ex5 * fc5);
                        dj[k] = dj[k] + gj;
                        dk[k] = dk[k] + gk;
                        di[k] = di[k] - (gj + gk);
                    }
                    sf[s] = sf[s] + pre * cosp * ex5 * fc5;
                    continue;
                }
                const double zeta = m->sf_zeta[o + s], lam = m->sf_lambda[o + s];
                const double eta = m->sf_eta[o + s];
                const int iz = (int)zeta;
                const double b = 1.0 + lam * cost;
                double pre, tp;
                if (fabs(zeta - iz) < 1.e-10) {
                    pre = powi(2.0, 1 - iz);
                    tp = powi(b, iz - 1);
                } else {
                    pre = pow(2.0, 1.0 - zeta);
                    tp = pow(b, zeta - 1.0);
                }
                const double cosp = tp * b;
                if (fabs(eta - lasteta) > 1.e-10) {
                    ex = exp(-(eta * (rij * rij + rik * rik + rjk * rjk)));
                    for (int k = 0; k < 3; ++k) {
                        dej[k] = ex * eta * 2.0 * (dij[k] - djk[k]);
                        dek[k] = ex * eta * 2.0 * (dik[k] + djk[k]);
                    }
                    lasteta = eta;
                }
                double *dj = dsf + ((size_t)s * nn + j) * 3, *dk = dsf + ((size_t)s * nn + kk) * 3;
                double *di = dsf + (size_t)s * nn * 3;
                for (int k = 0; k < 3; ++k) {
                    const double gj = pre * (cosp * ex * dfj[k] + cosp * dej[k] * fcpart +
                                             tp * lam * zeta * dcj[k] * ex * fcpart);
                    const double gk = pre * (cosp * ex * dfk[k] + cosp * dek[k] * fcpart +
                                             tp * lam * zeta * dck[k] * ex * fcpart);
                    dj[k] = dj[k] + gj;
                    dk[k] = dk[k] + gk;
                    di[k] = di[k] - (gj + gk);
                }
                sf[s] = sf[s] + pre * cosp * ex * fcpart;
            }
        }
    }
    return 0;
}

/* activate / deactivate, neuralNetworks.f90:33-84 */
static double act_f(char a, double x)
{
    switch (a) {
    case 't': return tanh(x);
    case 'p': return log(1.0 + exp(x));
    case 'r': return x > 0.0 ? x : 0.0;
    default: return x;
    }
}
static double act_d(char a, double x)
{
    switch (a) {
    case 't': { const double t = tanh(x); return 1.0 - t * t; }
    case 'p': return 1.0 / (1.0 + exp(-1.0 * x));
    case 'r': return x > 0.0 ? 1.0 : 0.0;
    default: return 1.0;
    }
}

/* nn_forward + nn_backward(gradOut = 1), neuralNetworks.f90:94-158,220-313.
 * Returns the network output; dx receives d out / d x. */
static double nn_eval(const orc_model *m, int e, const double *x, double *dx, double *work)
{
    const double *W = m->weights[e];
    int nin = m->nsf[e];
    /* work: per layer pre-activation c and input copies */
    double *bufs[ORC_MAXLAYERS + 1];
    double *cs[ORC_MAXLAYERS];
    int sizes[ORC_MAXLAYERS + 1];
    const double *Ws[ORC_MAXLAYERS];
    sizes[0] = nin;
    double *w = work;
    bufs[0] = w;
    memcpy(w, x, sizeof(double) * nin);
    w += nin;
    const double *wp = W;
    for (int l = 0; l < m->nlayers; ++l) {             /* layer_forward :94-114 */
        const int no = m->nodes[l], ni = sizes[l];
        Ws[l] = wp;
        const double *bias = wp + (size_t)no * ni;
        cs[l] = w;
        w += no;
        bufs[l + 1] = w;
        w += no;
        for (int i = 0; i < no; ++i) {
            double acc = 0.0;
            for (int j = 0; j < ni; ++j) acc = acc + wp[i + (size_t)no * j] * bufs[l][j];
            cs[l][i] = acc + bias[i];
            bufs[l + 1][i] = act_f(m->act[l], cs[l][i]);
        }
        sizes[l + 1] = no;
        wp = bias + no;
    }
    const double out = bufs[m->nlayers][0];
    double *g = w;                                      /* layer_backward :116-158 */
    w += 256;
    double *gi = w;
    g[0] = 1.0;
    for (int l = m->nlayers - 1; l >= 0; --l) {
        const int no = m->nodes[l], ni = sizes[l];
        double tmp[256];
        for (int i = 0; i < no; ++i) tmp[i] = g[i] * act_d(m->act[l], cs[l][i]);
        for (int j = 0; j < ni; ++j) {
            double acc = 0.0;
            for (int i = 0; i < no; ++i) acc = acc + tmp[i] * Ws[l][i + (size_t)no * j];
            gi[j] = acc;
        }
        memcpy(g, gi, sizeof(double) * ni);
    }
    memcpy(dx, g, sizeof(double) * nin);
    return out;
}

static int check_model(const orc_model *m)
{
    for (int l = 0; l < m->nlayers; ++l)
        if (m->nodes[l] > 256) return 5;
    for (int e = 0; e < m->nel; ++e)
        if (m->nsf[e] > 256) return 5;
    return 0;
}

typedef struct {
    double *P;
    int *els;
    double *sf, *dsf, *scratch, *work, *dEdG;
    size_t cap_nn;
} atom_ws;

static void ws_reserve(atom_ws *w, int nn, int nsf)
{
    if ((size_t)nn <= w->cap_nn) return;
    w->cap_nn = (size_t)nn * 2;
    free(w->P); free(w->els); free(w->dsf); free(w->scratch);
    w->P = (double *)malloc(sizeof(double) * 3 * w->cap_nn);
    w->els = (int *)malloc(sizeof(int) * w->cap_nn);
    w->dsf = (double *)malloc(sizeof(double) * 3 * w->cap_nn * (nsf > 0 ? nsf : 1) + 8);
    w->scratch = (double *)malloc(sizeof(double) * 6 * w->cap_nn);
}
static void ws_free(atom_ws *w)
{
    free(w->P); free(w->els); free(w->sf); free(w->dsf); free(w->scratch); free(w->work); free(w->dEdG);
}

/* Per-atom work of hdnnpCalcSymfunctions (hdnnps.f90:417-482) and
 * hdnnpCalcNNs (:515-538): slot list, calcAll, centre/scale, NN forward and
 * backward. Leaves the scaled dsf in w->dsf and dE/dG in w->dEdG. */
static int atom_work(const orc_model *m, const int *Z, const nlist *nl, const double *selfpos,
                     int i, int maxsf, atom_ws *w, double *Gout, double *Ei)
{
    const int e = elem_index(m, Z[i]);
    if (e < 0) return 2;
    const int nn = nl[i].n + 1, nsf = m->nsf[e], o = m->sf_off[e];
    ws_reserve(w, nn, maxsf);
    memcpy(w->P, selfpos + 3 * i, 3 * sizeof(double));
    w->els[0] = Z[i];
    for (int n = 1; n < nn; ++n) {
        memcpy(w->P + 3 * n, nl[i].p + 3 * (n - 1), 3 * sizeof(double));
        w->els[n] = Z[nl[i].j[n - 1]];
    }
    int rc = calc_all(m, e, nn, w->P, w->els, w->sf, w->dsf, w->scratch);
    if (rc) return rc;
    for (int s = 0; s < nsf; ++s) {                    /* hdnnps.f90:471-481 */
        if (m->scale_range && m->scalesf) {            /* RuNNer: smin + (smax-smin)(G-shift)/(max-min) */
            const double shift = m->centersf ? m->sf_mean[o + s] : m->sf_min[o + s];
            const double k = (m->scale_max - m->scale_min) / (m->sf_max[o + s] - m->sf_min[o + s]);
            w->sf[s] = m->scale_min + k * (w->sf[s] - shift);
            double *d = w->dsf + (size_t)s * nn * 3;
            for (int q = 0; q < nn * 3; ++q) d[q] = k * d[q];
            continue;
        }
        if (m->centersf) w->sf[s] = w->sf[s] - m->sf_mean[o + s];
        if (m->scalesf) {
            const double den = m->sf_max[o + s] - m->sf_min[o + s];
            w->sf[s] = w->sf[s] / den;
            double *d = w->dsf + (size_t)s * nn * 3;
            for (int q = 0; q < nn * 3; ++q) d[q] = d[q] / den;
        }
    }
    const double out = nn_eval(m, e, w->sf, w->dEdG, w->work);
    *Ei = out + m->atom_energy[e];                       /* :526 */
    if (Gout) memcpy(Gout, w->sf, sizeof(double) * nsf);
    return 0;
}

static int max_nsf(const orc_model *m)
{
    int mx = 1;
    for (int e = 0; e < m->nel; ++e)
        if (m->nsf[e] > mx) mx = m->nsf[e];
    return mx;
}

static void ws_init(atom_ws *w, const orc_model *m)
{
    memset(w, 0, sizeof *w);
    const int mx = max_nsf(m);
    w->sf = (double *)malloc(sizeof(double) * mx);
    w->dEdG = (double *)malloc(sizeof(double) * mx);
    w->work = (double *)malloc(sizeof(double) * (4 * 256 * (ORC_MAXLAYERS + 2)));
}

int orc_evaluate(const orc_model *m, int nat, const int *Z, const double *pos,
                 int periodic, const double *lat, double *energy, double *force,
                 double *dEdLat, double *G, double *dEdG, int ldg, int nthreads)
{
    int err = check_model(m);
    if (err) return err;
    for (int i = 0; i < nat; ++i)
        if (elem_index(m, Z[i]) < 0) return 2;
    double *selfpos = (double *)malloc(sizeof(double) * 3 * (nat ? nat : 1));
    nlist *nl = build_nl(nat, pos, periodic, lat, m->maxrc, selfpos);
    double invlat[9];
    const int want_lat = dEdLat != NULL && periodic;
    if (want_lat) orc_inv3(lat, invlat);               /* hdnnps.f90:509 */
    double *atEns = (double *)calloc(nat ? nat : 1, sizeof(double));
    const int mx = max_nsf(m);
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    const int nt = 1;
    (void)nthreads;
#endif
    double *fbuf = (double *)calloc((size_t)nt * 3 * (nat ? nat : 1), sizeof(double));
    double *lbuf = (double *)calloc((size_t)nt * 9, sizeof(double));
    int status = 0;
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        atom_ws w;
        ws_init(&w, m);
        double *F = fbuf + (size_t)tid * 3 * nat, *T = lbuf + 9 * tid;
#pragma omp for schedule(static)
        for (int i = 0; i < nat; ++i) {                 /* hdnnps.f90:515-554 */
            int rc = atom_work(m, Z, nl, selfpos, i, mx, &w, G ? G + (size_t)i * ldg : NULL, &atEns[i]);
            if (rc) {
#pragma omp atomic write
                status = rc;
                continue;
            }
            const int e = elem_index(m, Z[i]);
            const int nsf = m->nsf[e], nn = nl[i].n + 1;
            if (dEdG) memcpy(dEdG + (size_t)i * ldg, w.dEdG, sizeof(double) * nsf);
            for (int s = 0; s < nsf; ++s) {             /* :533-538 */
                const double *d = w.dsf + (size_t)s * nn * 3;
                for (int n = 0; n < nn; ++n) {
                    const int jj = n == 0 ? i : nl[i].j[n - 1];
                    for (int k = 0; k < 3; ++k) F[3 * jj + k] = F[3 * jj + k] + -(w.dEdG[s] * d[3 * n + k]);
                }
            }
            if (want_lat) {                             /* :539-553 */
                for (int s = 0; s < nsf; ++s) {
                    const double *d = w.dsf + (size_t)s * nn * 3;
                    for (int n = 0; n < nn; ++n) {
                        double relat[3];
                        matvec3(invlat, w.P + 3 * n, relat);
                        for (int x = 0; x < 3; ++x)
                            for (int a = 0; a < 3; ++a)
                                M3(T, a, x) = M3(T, a, x) + w.dEdG[s] * d[3 * n + a] * relat[x];
                    }
                }
            }
        }
        ws_free(&w);
    }
    if (!status) {
        for (int q = 0; q < 3 * nat; ++q) {
            double acc = 0.0;
            for (int t = 0; t < nt; ++t) acc += fbuf[(size_t)t * 3 * nat + q];
            force[q] = acc;
        }
        double E = 0.0;                                   /* :557 */
        for (int i = 0; i < nat; ++i) E = E + atEns[i];
        *energy = E;
        if (dEdLat) {
            for (int q = 0; q < 9; ++q) {
                double acc = 0.0;
                if (want_lat)
                    for (int t = 0; t < nt; ++t) acc += lbuf[9 * t + q];
                dEdLat[q] = acc;
            }
        }
    }
    free(fbuf);
    free(lbuf);
    free(atEns);
    nl_free(nl, nat);
    free(selfpos);
    return status;
}

int orc_atoms(const orc_model *m, int nat, const int *Z, const double *pos, int periodic,
              const double *lat, int nsel, const int *sel, double *Ei, double *G,
              double *dEdG, int ldg)
{
    int err = check_model(m);
    if (err) return err;
    for (int i = 0; i < nat; ++i)
        if (elem_index(m, Z[i]) < 0) return 2;
    double *selfpos = (double *)malloc(sizeof(double) * 3 * (nat ? nat : 1));
    nlist *nl = build_nl(nat, pos, periodic, lat, m->maxrc, selfpos);
    const int mx = max_nsf(m);
    int status = 0;
#pragma omp parallel
    {
        atom_ws w;
        ws_init(&w, m);
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < nsel; ++q) {
            const int i = sel[q];
            int rc = atom_work(m, Z, nl, selfpos, i, mx, &w, G + (size_t)q * ldg, &Ei[q]);
            if (rc) {
#pragma omp atomic write
                status = rc;
                continue;
            }
            const int e = elem_index(m, Z[i]);
            memcpy(dEdG + (size_t)q * ldg, w.dEdG, sizeof(double) * m->nsf[e]);
        }
        ws_free(&w);
    }
    nl_free(nl, nat);
    free(selfpos);
    return status;
}
