/*
 * ORACLE — test infrastructure only.
 *
 * FP64 CPU restatement of the reference's transport path (reference
 * src/ompmc.c shower() and everything it calls, plus the geometry, source and
 * scoring callbacks of ucodes/omc_dosxyz/omc_dosxyz.c and
 * ucodes/omc_matrad/omc_matrad.c).  Used only by tests/, by the smoke check
 * and by bench.py's cpu_baseline leg, as the checker; the product never links
 * it.  Physics tables come from the product's host layer
 * (ompmc_amd/csrc/omc_host.c), whose FP64 tables are themselves compared bit
 * for bit with tables dumped from the compiled reference.
 *
 * Pinning: every function here is compared against golden vectors produced by
 * the unmodified reference (oracle/ref_harness.c, tests/golden/), and the
 * whole-run dose of a single random-number stream is required to equal the
 * reference's bit for bit.
 */
#ifndef OMC_ORACLE_H
#define OMC_ORACLE_H

#include <stdint.h>
#include "omc_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Geometry: rectilinear voxels, regions 1 + ix + iy*nx + iz*nx*ny, 0 outside. */
typedef struct orc_geom {
    int nx, ny, nz;
    const double *xb, *yb, *zb;
    const double *rho;         /* per voxel (dose conversion) */
} orc_geom;

/* Creates a transport context over shared read-only tables/regions/geometry.
 * rng_mode 0: RANMAR seeded (ixx, jxx); 1: injected list (orc_inject). */
orc *orc_create(const omc_tables *t, const omc_regions *r, const orc_geom *g, int nsplit, int ixx, int jxx);
void orc_destroy(orc *o);
void orc_inject(orc *o, const double *rn, long n);     /* switch to injected random numbers */
long orc_rn_count(const orc *o);                      /* random numbers drawn so far */
double orc_rand(orc *o);                              /* next random number of the context */
int orc_error(const orc *o);                          /* non-zero after a stack overflow or RN underflow */

/* tally access: endep[nreg] (energy*weight per region, region 0 = outside) */
double *orc_endep(orc *o);
/* optional per-call log of ausgab calls (ir, edep, wt, iq), capacity n entries */
void orc_log_tally(orc *o, double *buf, int cap);
int orc_tally_count(const orc *o);
long orc_howfar_calls(const orc *o);
long orc_hownear_calls(const orc *o);

/* stack access: particle as 11 doubles (iq ir e x y z u v w wt dnear) */
void orc_set_particle(orc *o, int i, const double *p);
void orc_get_particle(const orc *o, int i, double *p);
void orc_set_np(orc *o, int np);
int orc_np(const orc *o);

/* physics entry points (names follow the reference) */
void orc_shower(orc *o);
void orc_photon(orc *o);
void orc_electron(orc *o);
void orc_compton(orc *o);
void orc_pair(orc *o, int imed);
void orc_photo(orc *o);
void orc_rayleigh(orc *o, int imed, double eig, double gle, int lgle);
void orc_brems(orc *o);
void orc_moller(orc *o);
void orc_bhabha(orc *o);
void orc_annih(orc *o);
void orc_rannih(orc *o);
double orc_msdist(orc *o, int imed, int iq, double rhof, double de, double tustep, double eke,
                  double *xf, double *yf, double *zf, double *uf, double *vf, double *wf);
void orc_mscat_fresh(orc *o, int imed, int qel, double elke, double beta2, double q1, double lambda,
                     double chia2, double *cost, double *sint);
void orc_sscat(orc *o, int imed, int qel, double chia2, double elke, double beta2, double *cost, double *sint);
double orc_drange(orc *o, int imed, int iq, int lelke, double ekei, double ekef, double elkei, double elkef);
double orc_eloss(orc *o, int imed, int iq, int irl, double rhof, double tustep, double range, double eke,
                 double elke, int lelke);

/* sources */
void orc_init_history_dosxyz(orc *o, const omc_dosxyz_source *s);
typedef struct orc_beamlets {
    int n;                      /* bixels */
    const int *ibeam;           /* 0-based beam of each bixel */
    const double *src;          /* [nbeams*3] */
    const double *corner, *side1, *side2;   /* [n*3] */
    int gaussian;               /* 0 point, 1 gaussian focal spot */
    double sigma;               /* gaussian width (cm) */
} orc_beamlets;
void orc_init_history_matrad(orc *o, const omc_spectrum *spec, const orc_beamlets *b, int ibeamlet);

/* drivers */
/* Serial dosxyz run (reference omc_dosxyz.c:1206-1282): nperbatch*nbatch
 * histories on one RANMAR stream; fills sum/sum2 [nreg] (per-batch Σ, Σ²) and
 * optionally per-history random-number counts. */
int orc_run_dosxyz(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_dosxyz_source *s,
                   int nsplit, int ixx, int jxx, int nperbatch, int nbatch, double *sum, double *sum2,
                   int *rn_per_history, double *ebal);
/* OpenMP dosxyz run as the reference runs it: thread k seeds (ixx, jxx+k),
 * histories scheduled dynamically; used for CPU timing. Returns seconds. */
double orc_run_dosxyz_omp(const omc_tables *t, const omc_regions *r, const orc_geom *g,
                          const omc_dosxyz_source *s, int nsplit, int ixx, int jxx, int nperbatch, int nbatch,
                          int nthreads, double *sum, double *sum2);
/* Serial matRad Dij run (omc_matrad.c:1576-1713) for one beamlet: per-batch
 * means accumulated into sum/sum2 [nreg] (sum2 must hold the carried variance). */
int orc_run_matrad_beamlet(orc *o, const omc_spectrum *spec, const orc_beamlets *b, int ibeamlet,
                           int nperbatch, int nbatch, double *sum, double *sum2);
/* Serial matRad Dij transport of all beamlets in the reference's loop order
 * (beamlet, batch, history) on one RANMAR stream; per-beamlet sums [n*nreg]
 * without the carried variance, draws per history optional. */
int orc_run_matrad(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_spectrum *spec,
                   const orc_beamlets *b, int nsplit, int ixx, int jxx, const int *nper, int nbatch, double *sum,
                   double *sum2, int *rn_per_history, double *batches);
/* OpenMP variant for CPU timing (thread k seeds (ixx, jxx+k)); returns seconds. */
double orc_run_matrad_omp(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_spectrum *spec,
                          const orc_beamlets *b, int nsplit, int ixx, int jxx, const int *nper, int nbatch,
                          int nthreads, double *sum, double *sum2);

#ifdef __cplusplus
}
#endif
#endif
