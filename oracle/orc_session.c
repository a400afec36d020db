/*
 * ORACLE — test infrastructure only.
 *
 * Session wrapper used through ctypes by tests/ and bench.py's cpu_baseline:
 * reads a dosxyz input file the way the reference's main does
 * (ucodes/omc_dosxyz/omc_dosxyz.c:1072-1282), then replays single golden
 * cases with injected random numbers or runs whole dose simulations.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "omc_oracle.h"

void orc_ranmar_stream(int ixx, int jxx, long skip, long n, int *out);

typedef struct orc_session {
    omc_inp inp;
    omc_phantom ph;
    omc_tables t;
    omc_regions r;
    orc_geom g;
    omc_dosxyz_source src;
    int nsplit, ixx, jxx, ncase, nbatch;
} orc_session;

static int get(const omc_inp *inp, const char *key, char *buf) { return omc_inp_get(inp, key, buf, 256); }

orc_session *orc_session_open(const char *stem) {
    orc_session *s = calloc(1, sizeof *s);
    char buf[256], path[256], path2[256], folder[256];
    if (omc_inp_parse(&s->inp, stem)) goto fail;
    if (!get(&s->inp, "phantom file", buf)) { omc_host_set_error(OMC_ERR_ARG, "no 'phantom file'"); goto fail; }
    omc_strip_spaces(path, buf, sizeof path);
    if (omc_read_egsphant(&s->ph, path)) goto fail;
    if (!get(&s->inp, "pegs file", buf)) goto fail;
    omc_strip_spaces(path, buf, sizeof path);
    if (!get(&s->inp, "pgs4form file", buf)) goto fail;
    omc_strip_spaces(path2, buf, sizeof path2);
    if (!get(&s->inp, "data folder", buf)) goto fail;
    omc_strip_spaces(folder, buf, sizeof folder);
    omc_paths p = {path, path2, folder};
    if (omc_tables_build(&s->t, s->ph.nmed, (const char(*)[60])s->ph.med_names, &p)) goto fail;
    if (get(&s->inp, "spectrum file", buf)) {
        omc_strip_spaces(path, buf, sizeof path);
        if (omc_spectrum_read(&s->src.spec, path)) goto fail;
    } else {
        if (!get(&s->inp, "mono energy", buf)) goto fail;
        omc_spectrum_mono(&s->src.spec, atof(buf));
    }
    double col[4] = {0, 0, 0, 0};
    if (!get(&s->inp, "collimator bounds", buf) || sscanf(buf, "%lf %lf %lf %lf", &col[0], &col[1], &col[2], &col[3]) != 4) goto fail;
    if (!get(&s->inp, "charge", buf)) goto fail;
    int charge = atoi(buf);
    if (!get(&s->inp, "ssd", buf)) goto fail;
    double ssd = atof(buf);
    omc_spectrum spec = s->src.spec;
    if (omc_dosxyz_source_init(&s->src, &s->ph, col, ssd, charge)) goto fail;
    s->src.spec = spec;
    if (!get(&s->inp, "global ecut", buf)) goto fail;
    double ecut = atof(buf);
    if (!get(&s->inp, "global pcut", buf)) goto fail;
    double pcut = atof(buf);
    if (omc_regions_build(&s->r, &s->ph, &s->t, ecut, pcut)) goto fail;
    if (!get(&s->inp, "nsplit", buf)) goto fail;
    s->nsplit = atoi(buf);
    if (!get(&s->inp, "rng seeds", buf) || sscanf(buf, "%d %d", &s->ixx, &s->jxx) != 2) goto fail;
    if (get(&s->inp, "ncase", buf)) s->ncase = atoi(buf);
    if (get(&s->inp, "nbatch", buf)) s->nbatch = atoi(buf);
    s->g.nx = s->ph.nx; s->g.ny = s->ph.ny; s->g.nz = s->ph.nz;
    s->g.xb = s->ph.xb; s->g.yb = s->ph.yb; s->g.zb = s->ph.zb; s->g.rho = s->ph.rho;
    return s;
fail:
    fprintf(stderr, "orc_session_open(%s): %s\n", stem, omc_host_last_error());
    free(s);
    return NULL;
}

void orc_session_close(orc_session *s) {
    if (!s) return;
    omc_regions_free(&s->r);
    omc_tables_free(&s->t);
    omc_phantom_free(&s->ph);
    free(s);
}

const char *orc_last_error(void) { return omc_host_last_error(); }
int orc_session_nreg(const orc_session *s) { return s->r.nreg; }
const omc_tables *orc_session_tables(const orc_session *s) { return &s->t; }
const omc_regions *orc_session_regions(const orc_session *s) { return &s->r; }

/* table access by name for comparisons with reference dumps */
const void *orc_session_table(orc_session *s, const char *name, long *n, int *is_int) {
    omc_tables *t = &s->t;
    long nm = t->nmed;
    *is_int = 0;
#define T(nm_, ptr, cnt) if (!strcmp(name, nm_)) { *n = (cnt); return (ptr); }
#define TI(nm_, ptr, cnt) if (!strcmp(name, nm_)) { *n = (cnt); *is_int = 1; return (ptr); }
    T("ge0", t->ge0, nm) T("ge1", t->ge1, nm)
    T("gmfp0", t->gmfp0, nm * OMC_MXGE) T("gmfp1", t->gmfp1, nm * OMC_MXGE)
    T("gbr10", t->gbr10, nm * OMC_MXGE) T("gbr11", t->gbr11, nm * OMC_MXGE)
    T("gbr20", t->gbr20, nm * OMC_MXGE) T("gbr21", t->gbr21, nm * OMC_MXGE)
    T("cohe0", t->cohe0, nm * OMC_MXGE) T("cohe1", t->cohe1, nm * OMC_MXGE)
    T("ray_xgrid", t->ray_xgrid, nm * OMC_MXRAYFF) T("ray_fcum", t->ray_fcum, nm * OMC_MXRAYFF)
    T("ray_b", t->ray_b, nm * OMC_MXRAYFF) T("ray_c", t->ray_c, nm * OMC_MXRAYFF)
    T("ray_pmax0", t->ray_pmax0, nm * OMC_MXGE) T("ray_pmax1", t->ray_pmax1, nm * OMC_MXGE)
    TI("ray_iarr", t->ray_iarr, nm * OMC_RAYCDFSIZE)
    T("dl1", t->dl1, nm * 8) T("dl2", t->dl2, nm * 8) T("dl3", t->dl3, nm * 8)
    T("dl4", t->dl4, nm * 8) T("dl5", t->dl5, nm * 8) T("dl6", t->dl6, nm * 8)
    T("bpar0", t->bpar0, nm) T("bpar1", t->bpar1, nm) T("pair_delcm", t->delcm, nm) T("zbrang", t->zbrang, nm)
#define E(x) T(#x, t->x, nm * OMC_MXEKE)
    E(esig0) E(esig1) E(psig0) E(psig1) E(ededx0) E(ededx1) E(pdedx0) E(pdedx1)
    E(ebr10) E(ebr11) E(pbr10) E(pbr11) E(pbr20) E(pbr21) E(tmxs0) E(tmxs1)
    E(blcce0) E(blcce1) E(etae_ms0) E(etae_ms1) E(etap_ms0) E(etap_ms1)
    E(q1ce_ms0) E(q1ce_ms1) E(q1cp_ms0) E(q1cp_ms1) E(q2ce_ms0) E(q2ce_ms1) E(q2cp_ms0) E(q2cp_ms1) E(e_array)
#undef E
    T("range_ep", t->range_ep, 2 * nm * OMC_MXEKE)
    T("eke0", t->eke0, nm) T("eke1", t->eke1, nm)
    TI("sig_ismonotone", t->sig_ismonotone, 2 * nm)
    T("esig_e", t->esig_e, nm) T("psig_e", t->psig_e, nm)
    T("xcc", t->xcc, nm) T("blcc", t->blcc, nm) T("expeke1", t->expeke1, nm)
    T("spin_rej", t->spin_rej, nm * OMC_NSPIN_MED)
    T("ums", t->ums, OMC_NMS) T("fms", t->fms, OMC_NMS) T("wms", t->wms, OMC_NMS) TI("ims", t->ims, OMC_NMS)
    T("region_rhof", s->r.rhof, s->r.nreg) T("region_pcut", s->r.pcut, s->r.nreg) T("region_ecut", s->r.ecut, s->r.nreg)
    TI("region_med", s->r.med, s->r.nreg)
#undef T
#undef TI
    *n = 0;
    return NULL;
}

void orc_session_scalars(orc_session *s, double *out) {
    /* spin: b2spin_min dbeta2i espml dleneri dqq1i; mscat: dllambi dqmsi */
    out[0] = s->t.b2spin_min; out[1] = s->t.dbeta2i; out[2] = s->t.espml; out[3] = s->t.dleneri; out[4] = s->t.dqq1i;
    out[5] = s->t.dllambi; out[6] = s->t.dqmsi;
}

void orc_session_pegs(orc_session *s, double *f, int *ii) {
    for (int i = 0; i < s->t.nmed; i++) {
        double *q = f + 12 * i; int *qi = ii + 8 * i;
        q[0] = s->t.rho[i]; q[1] = s->t.rlc[i]; q[2] = s->t.ae[i]; q[3] = s->t.ap[i];
        q[4] = s->t.ue[i]; q[5] = s->t.up[i]; q[6] = s->t.te[i]; q[7] = s->t.thmoll[i];
        q[8] = s->t.delcm[i]; q[9] = 0; q[10] = 0; q[11] = 0;
        qi[0] = s->t.ne[i]; qi[1] = s->t.meke[i]; qi[2] = s->t.mge[i]; qi[3] = s->t.msge[i];
        qi[4] = s->t.iunrst[i]; qi[5] = s->t.epstfl[i]; qi[6] = s->t.iaprim[i]; qi[7] = 0;
    }
}

enum { K_COMPTON = 1, K_PAIR, K_PHOTO, K_RAYLEIGH, K_BREMS, K_MOLLER, K_BHABHA, K_ANNIH,
       K_RANNIH, K_MSDIST, K_MSCAT, K_SSCAT, K_DRANGE, K_ELOSS, K_PHOTON, K_ELECTRON };

/* Replays one harness case (oracle/ref_harness.c run_samplers) with the
 * recorded random numbers; writes the same output layout.  Returns the number
 * of outputs, or -1 on an unknown kind; *rn_used gets the number drawn. */
int orc_session_case(orc_session *s, int kind, const double *in, int nin, const double *rn, long nrn, double *out,
                     int cap, long *rn_used) {
    (void)nin;
    orc *o = orc_create(&s->t, &s->r, &s->g, s->nsplit, s->ixx, s->jxx);
    orc_inject(o, rn, nrn);
    double tlog[4 * 4096];
    int k = 0;
    orc_set_np(o, 0);
    switch (kind) {
    case K_COMPTON: case K_PHOTO: case K_BREMS: case K_MOLLER: case K_BHABHA: case K_ANNIH: case K_RANNIH:
    case K_PAIR: case K_RAYLEIGH: case K_PHOTON: case K_ELECTRON:
        orc_set_particle(o, 0, in);
        if (kind == K_PHOTON || kind == K_ELECTRON) orc_log_tally(o, tlog, 4096);
        switch (kind) {
        case K_COMPTON: orc_compton(o); break;
        case K_PHOTO: orc_photo(o); break;
        case K_BREMS: orc_brems(o); break;
        case K_MOLLER: orc_moller(o); break;
        case K_BHABHA: orc_bhabha(o); break;
        case K_ANNIH: orc_annih(o); break;
        case K_RANNIH: orc_rannih(o); break;
        case K_PAIR: orc_pair(o, (int)in[11]); break;
        case K_RAYLEIGH: orc_rayleigh(o, (int)in[11], in[2], in[12], (int)in[13]); break;
        case K_PHOTON: orc_photon(o); break;
        case K_ELECTRON: orc_electron(o); break;
        }
        out[k++] = orc_np(o);
        for (int i = 0; i <= orc_np(o) && i < 16 && k + 11 <= cap; i++) { orc_get_particle(o, i, out + k); k += 11; }
        if (kind == K_PHOTON || kind == K_ELECTRON) {
            int n = orc_tally_count(o);
            out[k++] = n;
            for (int i = 0; i < 4 * n && k < cap; i++) out[k++] = tlog[i];
            out[k++] = (double)orc_howfar_calls(o);
            out[k++] = (double)orc_hownear_calls(o);
        }
        break;
    case K_MSDIST: {
        orc_set_particle(o, 0, in + 6);
        double xf, yf, zf, uf, vf, wf;
        out[0] = orc_msdist(o, (int)in[0], (int)in[1], in[2], in[3], in[4], in[5], &xf, &yf, &zf, &uf, &vf, &wf);
        out[1] = xf; out[2] = yf; out[3] = zf; out[4] = uf; out[5] = vf; out[6] = wf;
        k = 7;
        break;
    }
    case K_MSCAT: {
        double c = 7.0, sn = 7.0;
        orc_mscat_fresh(o, (int)in[0], (int)in[1], in[2], in[3], in[4], in[5], in[6], &c, &sn);
        out[0] = c; out[1] = sn; k = 2;
        break;
    }
    case K_SSCAT: {
        double c, sn;
        orc_sscat(o, (int)in[0], (int)in[1], in[2], in[3], in[4], &c, &sn);
        out[0] = c; out[1] = sn; k = 2;
        break;
    }
    case K_DRANGE:
        out[0] = orc_drange(o, (int)in[0], (int)in[1], (int)in[2], in[3], in[4], in[5], in[6]); k = 1;
        break;
    case K_ELOSS:
        out[0] = orc_eloss(o, (int)in[0], (int)in[1], (int)in[2], in[3], in[4], in[5], in[6], in[7], (int)in[8]); k = 1;
        break;
    default:
        orc_destroy(o);
        return -1;
    }
    *rn_used = orc_rn_count(o);
    if (orc_error(o)) k = -2;
    orc_destroy(o);
    return k;
}

/* Serial dosxyz run from the input file's source and seeds. */
int orc_session_dose(orc_session *s, int nperbatch, int nbatch, double *sum, double *sum2, int *rn_per_history,
                     double *ebal) {
    return orc_run_dosxyz(&s->t, &s->r, &s->g, &s->src, s->nsplit, s->ixx, s->jxx, nperbatch, nbatch, sum, sum2,
                          rn_per_history, ebal);
}

double orc_session_dose_omp(orc_session *s, int nperbatch, int nbatch, int nthreads, double *sum, double *sum2) {
    return orc_run_dosxyz_omp(&s->t, &s->r, &s->g, &s->src, s->nsplit, s->ixx, s->jxx, nperbatch, nbatch, nthreads,
                              sum, sum2);
}

void orc_session_finish(orc_session *s, const double *sum, const double *sum2, int nperbatch, int nbatch, double *dose,
                        double *unc) {
    omc_dosxyz_finish(&s->ph, sum, sum2, nperbatch, nbatch, dose, unc);
}

/* RANMAR stream of this session's seeds: n values as 24-bit integers */
void orc_session_ranmar(orc_session *s, long n, int *out) { orc_ranmar_stream(s->ixx, s->jxx, 0, n, out); }

/* standalone RANMAR for the Marsaglia-Zaman known-answer test */
void orc_ranmar_stream(int ixx, int jxx, long skip, long n, int *out) {
    orc *o = orc_create(NULL, NULL, NULL, 1, ixx, jxx);
    for (long i = 0; i < skip; i++) orc_rand(o);
    for (long i = 0; i < n; i++) out[i] = (int)(orc_rand(o) * 16777216.0);
    orc_destroy(o);
}

/* matRad Dij transport of the session's phantom and tables, serial on one
 * RANMAR stream in the reference's loop order (omc_matrad.c:1576-1600):
 * beamlet by beamlet, batch by batch, nper[b] histories per batch.  For
 * beamlet b, sum/sum2 + b*nreg receive the per-batch sums of endep/nper[b]
 * and of its square (accumEndep, :897-918) without the variance carried from
 * the previous beamlet; rn_per_history (optional) the draws of each history
 * in run order; batches (optional) every batch's raw energy tally,
 * [nbix][nbatch][nreg]. */
int orc_session_matrad(orc_session *s, int spectrum, double energy, const double *cdf1, const double *cdf2, int nbix,
                       const int *ibeam, const double *src, const double *corner, const double *side1,
                       const double *side2, int gaussian, double sigma, const int *nper, int nbatch, double *sum,
                       double *sum2, int *rn_per_history, double *batches) {
    omc_spectrum sp;
    memset(&sp, 0, sizeof sp);
    if (spectrum) {
        sp.spectrum = 1;
        sp.deltak = OMC_INVDIM;
        memcpy(sp.cdfinv1, cdf1, sizeof sp.cdfinv1);
        memcpy(sp.cdfinv2, cdf2, sizeof sp.cdfinv2);
    } else {
        omc_spectrum_mono(&sp, energy);
    }
    orc_beamlets bl = {nbix, ibeam, src, corner, side1, side2, gaussian, sigma};
    return orc_run_matrad(&s->t, &s->r, &s->g, &sp, &bl, s->nsplit, s->ixx, s->jxx, nper, nbatch, sum, sum2,
                          rn_per_history, batches);
}

/* OpenMP matRad transport for CPU timing; returns the transport seconds. */
double orc_session_matrad_omp(orc_session *s, int spectrum, double energy, const double *cdf1, const double *cdf2,
                              int nbix, const int *ibeam, const double *src, const double *corner,
                              const double *side1, const double *side2, int gaussian, double sigma, const int *nper,
                              int nbatch, int nthreads, double *sum, double *sum2) {
    omc_spectrum sp;
    memset(&sp, 0, sizeof sp);
    if (spectrum) {
        sp.spectrum = 1;
        sp.deltak = OMC_INVDIM;
        memcpy(sp.cdfinv1, cdf1, sizeof sp.cdfinv1);
        memcpy(sp.cdfinv2, cdf2, sizeof sp.cdfinv2);
    } else {
        omc_spectrum_mono(&sp, energy);
    }
    orc_beamlets bl = {nbix, ibeam, src, corner, side1, side2, gaussian, sigma};
    return orc_run_matrad_omp(&s->t, &s->r, &s->g, &sp, &bl, s->nsplit, s->ixx, s->jxx, nper, nbatch, nthreads, sum,
                              sum2);
}

/* Samples n primaries of beamlet ib on the session's RANMAR stream with the
 * matRad source (omc_matrad.c:1177-1413); out[11*n]: iq ir e x y z u v w wt
 * dnear per primary. */
int orc_session_matrad_source(orc_session *s, int spectrum, double energy, const double *cdf1, const double *cdf2,
                              int nbix, const int *ibeam, const double *src, const double *corner,
                              const double *side1, const double *side2, int gaussian, double sigma, int ib, int n,
                              double *out) {
    omc_spectrum sp;
    memset(&sp, 0, sizeof sp);
    if (spectrum) {
        sp.spectrum = 1;
        sp.deltak = OMC_INVDIM;
        memcpy(sp.cdfinv1, cdf1, sizeof sp.cdfinv1);
        memcpy(sp.cdfinv2, cdf2, sizeof sp.cdfinv2);
    } else {
        omc_spectrum_mono(&sp, energy);
    }
    orc_beamlets bl = {nbix, ibeam, src, corner, side1, side2, gaussian, sigma};
    orc *o = orc_create(&s->t, &s->r, &s->g, s->nsplit, s->ixx, s->jxx);
    if (!o) return OMC_ERR_NOMEM;
    for (int i = 0; i < n; i++) {
        orc_init_history_matrad(o, &sp, &bl, ib);
        orc_get_particle(o, 0, out + 11 * (long)i);
    }
    orc_destroy(o);
    return OMC_OK;
}

/* The .3ddose writer of the product's host layer (omc_host.c, reference
 * omc_dosxyz.c:801-886) on the session's phantom. */
int orc_session_write_3ddose(orc_session *s, const char *path, const double *dose, const double *unc) {
    return omc_write_3ddose(path, &s->ph, dose, unc);
}
