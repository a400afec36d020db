/*
 * ORACLE — test infrastructure only.
 *
 * FP64 restatement of omc_matrad's scoring and dose-influence assembly,
 * written from the reference's ucodes/omc_matrad/omc_matrad.c and following
 * it statement by statement:
 *
 *   orc_mr_create          initScore (:856-873): endep, accum_endep and
 *                          accum_endep2 over nvox + 1 regions, zeroed; the
 *                          CSC output of mexFunction starts with jcs[0] = 0
 *                          (:1533-1540)
 *   orc_mr_accum_endep     accumEndep (:897-918): per region e = endep/n,
 *                          accum_endep += e, accum_endep2 += e*e, endep = 0
 *   orc_mr_accumulate      accumulateResults (:920-1012), both iout values:
 *                          Gy conversion 1.602e-10/(rho*V) for iout = 1,
 *                          mean over nbatch, the Q11 variance
 *                          (accum_endep2/(nbatch-1) - mean^2)/nbatch, zero
 *                          for a zero mean, then zero dose and variance
 *                          where rho < 0.044
 *   orc_mr_beamlet_end     the rest of the beamlet loop body (:1607-1702):
 *                          maximum, threshold doseMax*rel, entries strictly
 *                          above it appended in region order (row irl-1),
 *                          jcs[b+1], and the reset of accum_endep only --
 *                          accum_endep2, now the beamlet's variance, is
 *                          carried into the next beamlet's sum of squares
 *
 * It shares no code with the product (ompmc_amd/): the phantom arrives as
 * plain arrays, and the arithmetic and its order are the reference's.  The
 * output arrays grow by doubling; the reference's growth policy
 * (:1627-1664) changes only capacity, never the entries.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_mr {
    int isize, jsize, ksize;
    double *xbounds, *ybounds, *zbounds, *med_densities;
    long gridsize;
    double *endep, *accum_endep, *accum_endep2;
    int nbeamlets;
    int64_t *jcs, *irs;
    double *sr, *sr_var;
    int64_t lin_ix, nzmax;
} orc_mr;

static double *dup(const double *p, long n) {
    double *q = malloc(sizeof(double) * n);
    if (q) memcpy(q, p, sizeof(double) * n);
    return q;
}

void orc_mr_destroy(orc_mr *s) {
    if (!s) return;
    free(s->xbounds); free(s->ybounds); free(s->zbounds); free(s->med_densities);
    free(s->endep); free(s->accum_endep); free(s->accum_endep2);
    free(s->jcs); free(s->irs); free(s->sr); free(s->sr_var);
    free(s);
}

orc_mr *orc_mr_create(int nx, int ny, int nz, const double *xb, const double *yb, const double *zb,
                      const double *rho, int nbeamlets) {
    orc_mr *s = calloc(1, sizeof *s);
    if (!s) return NULL;
    s->isize = nx; s->jsize = ny; s->ksize = nz;
    s->gridsize = (long)nx * ny * nz;
    s->xbounds = dup(xb, nx + 1);
    s->ybounds = dup(yb, ny + 1);
    s->zbounds = dup(zb, nz + 1);
    s->med_densities = dup(rho, s->gridsize);
    s->endep = calloc(s->gridsize + 1, sizeof(double));
    s->accum_endep = calloc(s->gridsize + 1, sizeof(double));
    s->accum_endep2 = calloc(s->gridsize + 1, sizeof(double));
    s->nbeamlets = nbeamlets;
    s->jcs = calloc((size_t)nbeamlets + 1, sizeof(int64_t));
    s->nzmax = 1024;
    s->irs = malloc(sizeof(int64_t) * s->nzmax);
    s->sr = malloc(sizeof(double) * s->nzmax);
    s->sr_var = malloc(sizeof(double) * s->nzmax);
    if (!s->xbounds || !s->ybounds || !s->zbounds || !s->med_densities || !s->endep || !s->accum_endep ||
        !s->accum_endep2 || !s->jcs || !s->irs || !s->sr || !s->sr_var) {
        orc_mr_destroy(s);
        return NULL;
    }
    s->jcs[0] = 0;
    s->lin_ix = 0;
    return s;
}

/* The batch tally (region 0 = outside), filled by the caller as ausgab
 * would (:883-895). */
double *orc_mr_endep(orc_mr *s) { return s->endep; }
double *orc_mr_accum(orc_mr *s, int which) { return which ? s->accum_endep2 : s->accum_endep; }

/* accumEndep (:897-918) */
void orc_mr_accum_endep(orc_mr *s, int nperbatch) {
    for (long irl = 0; irl < s->gridsize + 1; irl++) {
        double edep = s->endep[irl];
        edep /= (double)nperbatch;
        s->accum_endep[irl] += edep;
        s->accum_endep2[irl] += edep * edep;
    }
    memset(s->endep, 0, (s->gridsize + 1) * sizeof(double));
}

/* accumulateResults (:920-1012) */
void orc_mr_accumulate(orc_mr *s, int iout, int nbatch) {
    int imax = s->isize, ijmax = s->isize * s->jsize;
    for (int iz = 0; iz < s->ksize; iz++)
        for (int iy = 0; iy < s->jsize; iy++)
            for (int ix = 0; ix < s->isize; ix++) {
                long irl = 1 + ix + (long)iy * imax + (long)iz * ijmax;
                double endep = s->accum_endep[irl];
                double endep2 = s->accum_endep2[irl];
                double unc_endep, factor;
                if (iout) {
                    double mass = (s->xbounds[ix + 1] - s->xbounds[ix]) * (s->ybounds[iy + 1] - s->ybounds[iy]) *
                                  (s->zbounds[iz + 1] - s->zbounds[iz]);
                    mass *= s->med_densities[irl - 1];
                    factor = 1.602E-10 / (mass);
                } else {
                    factor = 1.0;
                }
                endep *= factor;
                endep2 *= factor * factor;
                endep /= (double)nbatch;
                endep2 /= (double)(nbatch - 1);
                if (endep != 0.0) {
                    unc_endep = endep2 - endep * endep;
                    unc_endep /= nbatch;
                } else {
                    endep = 0.0;
                    unc_endep = 0.0;
                }
                s->accum_endep[irl] = endep;
                s->accum_endep2[irl] = unc_endep;
            }
    for (int iz = 0; iz < s->ksize; iz++)
        for (int iy = 0; iy < s->jsize; iy++)
            for (int ix = 0; ix < s->isize; ix++) {
                long irl = 1 + ix + (long)iy * imax + (long)iz * ijmax;
                if (s->med_densities[irl - 1] < 0.044) {
                    s->accum_endep[irl] = 0.0;
                    s->accum_endep2[irl] = 0.0;
                }
            }
}

static int grow(orc_mr *s, int64_t need) {
    if (need <= s->nzmax) return 0;
    int64_t n = s->nzmax;
    while (n < need) n *= 2;
    int64_t *ir = realloc(s->irs, sizeof(int64_t) * n);
    if (!ir) return -1;
    s->irs = ir;
    double *v = realloc(s->sr, sizeof(double) * n);
    if (!v) return -1;
    s->sr = v;
    v = realloc(s->sr_var, sizeof(double) * n);
    if (!v) return -1;
    s->sr_var = v;
    s->nzmax = n;
    return 0;
}

/* The end of beamlet ibeamlet after accumulateResults (:1607-1702). */
int orc_mr_beamlet_end(orc_mr *s, int ibeamlet, double relDoseThreshold) {
    double doseMax = 0.0;
    for (long irl = 1; irl < s->gridsize + 1; irl++)
        if (s->accum_endep[irl] > doseMax) doseMax = s->accum_endep[irl];
    double thresh = doseMax * relDoseThreshold;
    int64_t j_nnz = 0;
    for (long irl = 1; irl < s->gridsize + 1; irl++)
        if (s->accum_endep[irl] > thresh) j_nnz++;
    if (grow(s, j_nnz + s->lin_ix)) return -1;
    for (long irl = 1; irl < s->gridsize + 1; irl++) {
        if (s->accum_endep[irl] > thresh) {
            s->sr[s->lin_ix] = s->accum_endep[irl];
            s->irs[s->lin_ix] = irl - 1;
            s->sr_var[s->lin_ix] = s->accum_endep2[irl];
            s->lin_ix++;
        }
    }
    s->jcs[ibeamlet + 1] = s->lin_ix;
    memset(s->accum_endep, 0, (s->gridsize + 1) * sizeof(double));
    return 0;
}

int64_t orc_mr_nnz(const orc_mr *s) { return s->lin_ix; }

void orc_mr_copy(const orc_mr *s, int64_t *jc, int64_t *ir, double *val, double *var) {
    memcpy(jc, s->jcs, sizeof(int64_t) * ((size_t)s->nbeamlets + 1));
    memcpy(ir, s->irs, sizeof(int64_t) * (size_t)s->lin_ix);
    memcpy(val, s->sr, sizeof(double) * (size_t)s->lin_ix);
    memcpy(var, s->sr_var, sizeof(double) * (size_t)s->lin_ix);
}

/* The beamlet loop of mexFunction (:1583-1702) over batch tallies given in
 * full: endep[(b * nbatch + k) * (nvox + 1) + irl] is beamlet b's energy in
 * batch k (region 0 = outside), nper[b] its histories per batch.  Returns
 * the scoring state holding the CSC; NULL when memory runs out. */
orc_mr *orc_mr_run(int nx, int ny, int nz, const double *xb, const double *yb, const double *zb, const double *rho,
                   int nbeamlets, int nbatch, const int *nper, const double *endep, double relDoseThreshold) {
    orc_mr *s = orc_mr_create(nx, ny, nz, xb, yb, zb, rho, nbeamlets);
    if (!s) return NULL;
    const long nreg = s->gridsize + 1;
    for (int ibeamlet = 0; ibeamlet < nbeamlets; ibeamlet++) {
        for (int ibatch = 0; ibatch < nbatch; ibatch++) {
            memcpy(s->endep, endep + ((long)ibeamlet * nbatch + ibatch) * nreg, sizeof(double) * nreg);
            orc_mr_accum_endep(s, nper[ibeamlet]);
        }
        orc_mr_accumulate(s, 1, nbatch);
        if (orc_mr_beamlet_end(s, ibeamlet, relDoseThreshold)) {
            orc_mr_destroy(s);
            return NULL;
        }
    }
    return s;
}
