/*
 * ORACLE — test infrastructure only (see omc_oracle.h).
 *
 * FP64 restatement of the reference transport path.  Each function names the
 * reference lines it follows.  Arithmetic is written in the reference's
 * operation order and with the same libm calls so that, on one random-number
 * stream, results equal the reference's bit for bit; the state lives in a
 * per-thread context instead of thread-private globals.
 */
#include "omc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RM OMC_RM
#define MXGE OMC_MXGE
#define MXEKE OMC_MXEKE
#define MXSTACK OMC_MXSTACK

struct orc {
    const omc_tables *t;
    const omc_regions *r;
    const orc_geom *g;
    int nsplit;
    int np, npold;
    int *iq, *ir;
    double *e, *x, *y, *z, *u, *v, *w, *wt, *dnear;
    /* RANMAR state (reference src/omc_random.c:63-150) */
    int ru[97], rc, rcd, rcm, ri, rj, rbuf[128], rpos;
    int inj;
    const double *rn;
    long nrn, irn;
    long count;
    int err;
    double *endep;
    double *tlog;
    int tcap, tn;
    long nhowfar, nhownear;
    double ensrc;
    /* energy balance (not in the reference): photons that left the phantom,
     * energy dropped without a deposit (quirk Q6) */
    double esc, lost;
};

/* ------------------------------------------------------------------ RNG */
static void ranmar_seed(orc *o, int ixx, int jxx) {
    if (ixx <= 0 || ixx > 31328) ixx = 1802;
    if (jxx <= 0 || jxx > 31328) jxx = 9373;
    int i = (ixx / 177 % 177) + 2, j = (ixx % 177) + 2, k = (jxx / 169 % 178) + 1, l = (jxx % 169);
    for (int ii = 0; ii < 97; ii++) {
        int s = 0, t = 8388608;
        for (int jj = 0; jj < 24; jj++) {
            int m = ((i * j % 179) * k) % 179;
            i = j; j = k; k = m;
            l = (53 * l + 1) % 169;
            if (l * m % 64 >= 32) s += t;
            t /= 2;
        }
        o->ru[ii] = s;
    }
    o->rc = 362436; o->rcd = 7654321; o->rcm = 16777213;
    o->ri = 97; o->rj = 33;
    o->rpos = 128;
}

static void ranmar_fill(orc *o) {
    for (int n = 0; n < 128; n++) {
        int v = o->ru[o->ri - 1] - o->ru[o->rj - 1];
        if (v < 0) v += 16777216;
        o->ru[o->ri - 1] = v;
        o->ri -= 1; o->rj -= 1;
        if (o->ri == 0) o->ri = 97;
        else if (o->rj == 0) o->rj = 97;
        o->rc -= o->rcd;
        if (o->rc < 0) o->rc += o->rcm;
        v -= o->rc;
        if (v < 0) v += 16777216;
        o->rbuf[n] = v;
    }
    o->rpos = 0;
}

static double rnd(orc *o) {
    o->count++;
    if (o->inj) {
        if (o->irn >= o->nrn) { o->err = 2; return 0.5; }
        return o->rn[o->irn++];
    }
    if (o->rpos >= 128) ranmar_fill(o);
    return o->rbuf[o->rpos++] * (1.0 / 16777216.0);
}

/* ------------------------------------------------------------------ context */
orc *orc_create(const omc_tables *t, const omc_regions *r, const orc_geom *g, int nsplit, int ixx, int jxx) {
    orc *o = calloc(1, sizeof *o);
    if (!o) return NULL;
    o->t = t; o->r = r; o->g = g; o->nsplit = nsplit < 1 ? 1 : nsplit;
    o->iq = calloc(MXSTACK, sizeof(int)); o->ir = calloc(MXSTACK, sizeof(int));
    double **a[] = {&o->e, &o->x, &o->y, &o->z, &o->u, &o->v, &o->w, &o->wt, &o->dnear};
    for (int i = 0; i < 9; i++) *a[i] = calloc(MXSTACK, sizeof(double));
    o->endep = calloc(r ? r->nreg : 1, sizeof(double));
    ranmar_seed(o, ixx, jxx);
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    free(o->iq); free(o->ir); free(o->e); free(o->x); free(o->y); free(o->z);
    free(o->u); free(o->v); free(o->w); free(o->wt); free(o->dnear); free(o->endep);
    free(o);
}

double orc_rand(orc *o) { return rnd(o); }

void orc_inject(orc *o, const double *rn, long n) { o->inj = 1; o->rn = rn; o->nrn = n; o->irn = 0; }
long orc_rn_count(const orc *o) { return o->count; }
int orc_error(const orc *o) { return o->err; }
double *orc_endep(orc *o) { return o->endep; }
void orc_log_tally(orc *o, double *buf, int cap) { o->tlog = buf; o->tcap = cap; o->tn = 0; o->nhowfar = 0; o->nhownear = 0; }
int orc_tally_count(const orc *o) { return o->tn; }
long orc_howfar_calls(const orc *o) { return o->nhowfar; }
long orc_hownear_calls(const orc *o) { return o->nhownear; }
void orc_set_np(orc *o, int np) { o->np = np; o->npold = np; }
int orc_np(const orc *o) { return o->np; }

void orc_set_particle(orc *o, int i, const double *p) {
    o->iq[i] = (int)p[0]; o->ir[i] = (int)p[1]; o->e[i] = p[2];
    o->x[i] = p[3]; o->y[i] = p[4]; o->z[i] = p[5];
    o->u[i] = p[6]; o->v[i] = p[7]; o->w[i] = p[8]; o->wt[i] = p[9]; o->dnear[i] = p[10];
}

void orc_get_particle(const orc *o, int i, double *p) {
    p[0] = o->iq[i]; p[1] = o->ir[i]; p[2] = o->e[i];
    p[3] = o->x[i]; p[4] = o->y[i]; p[5] = o->z[i];
    p[6] = o->u[i]; p[7] = o->v[i]; p[8] = o->w[i]; p[9] = o->wt[i]; p[10] = o->dnear[i];
}

/* ------------------------------------------------------------------ user code: geometry and scoring */
/* ausgab (ucodes/omc_dosxyz/omc_dosxyz.c:683-694) */
static void ausgab(orc *o, double edep) {
    int np = o->np, irl = o->ir[np];
    double endep = o->wt[np] * edep;
    o->endep[irl] += endep;
    if (o->tlog && o->tn < o->tcap) {
        double *q = o->tlog + 4 * o->tn++;
        q[0] = irl; q[1] = edep; q[2] = o->wt[np]; q[3] = o->iq[np];
    }
}

/* howfar (omc_dosxyz.c:187-297): z, then x, then y; strict '<' */
static void howfar(orc *o, int *idisc, int *irnew, double *ustep) {
    o->nhowfar++;
    int np = o->np, irl = o->ir[np];
    const orc_geom *g = o->g;
    double dist;
    if (o->ir[np] == 0) { *idisc = 1; return; }
    int ijmax = g->nx * g->ny, imax = g->nx;
    int irx = (irl - 1) % imax;
    int irz = (irl - 1 - irx) / ijmax;
    int iry = ((irl - 1 - irx) - irz * ijmax) / imax;
    if (o->w[np] > 0.0) {
        dist = (g->zb[irz + 1] - o->z[np]) / o->w[np];
        if (dist < *ustep) { *ustep = dist; *irnew = irz != g->nz - 1 ? irl + ijmax : 0; }
    } else if (o->w[np] < 0.0) {
        dist = -(o->z[np] - g->zb[irz]) / o->w[np];
        if (dist < *ustep) { *ustep = dist; *irnew = irz != 0 ? irl - ijmax : 0; }
    }
    if (o->u[np] > 0.0) {
        dist = (g->xb[irx + 1] - o->x[np]) / o->u[np];
        if (dist < *ustep) { *ustep = dist; *irnew = irx != g->nx - 1 ? irl + 1 : 0; }
    } else if (o->u[np] < 0.0) {
        dist = -(o->x[np] - g->xb[irx]) / o->u[np];
        if (dist < *ustep) { *ustep = dist; *irnew = irx != 0 ? irl - 1 : 0; }
    }
    if (o->v[np] > 0.0) {
        dist = (g->yb[iry + 1] - o->y[np]) / o->v[np];
        if (dist < *ustep) { *ustep = dist; *irnew = iry != g->ny - 1 ? irl + imax : 0; }
    } else if (o->v[np] < 0.0) {
        dist = -(o->y[np] - g->yb[iry]) / o->v[np];
        if (dist < *ustep) { *ustep = dist; *irnew = iry != 0 ? irl - imax : 0; }
    }
}

/* hownear (omc_dosxyz.c:299-334) */
static double hownear(orc *o) {
    o->nhownear++;
    int np = o->np, irl = o->ir[np];
    const orc_geom *g = o->g;
    double tperp = 1.0E10;
    if (irl == 0) return 0.0;
    int ijmax = g->nx * g->ny, imax = g->nx;
    int irx = (irl - 1) % imax;
    int irz = (irl - 1 - irx) / ijmax;
    int iry = ((irl - 1 - irx) - irz * ijmax) / imax;
    tperp = fmin(tperp, g->xb[irx + 1] - o->x[np]);
    tperp = fmin(tperp, o->x[np] - g->xb[irx]);
    tperp = fmin(tperp, g->yb[iry + 1] - o->y[np]);
    tperp = fmin(tperp, o->y[np] - g->yb[iry]);
    tperp = fmin(tperp, g->zb[irz + 1] - o->z[np]);
    tperp = fmin(tperp, o->z[np] - g->zb[irz]);
    return tperp;
}

/* ------------------------------------------------------------------ kinematics helpers */
typedef struct { double A, B, C, cosphi, sinphi; } uphi_t;

static void transfer(orc *o, int n, int from) {
    o->x[n] = o->x[from]; o->y[n] = o->y[from]; o->z[n] = o->z[from];
    o->u[n] = o->u[from]; o->v[n] = o->v[from]; o->w[n] = o->w[from];
    o->ir[n] = o->ir[from]; o->wt[n] = o->wt[from]; o->dnear[n] = o->dnear[from];
}

/* box-rejection azimuth (src/ompmc.c:102-122) */
static void azimuth(orc *o, double *c, double *s) {
    double xphi, xphi2, yphi, yphi2, r2;
    do {
        xphi = rnd(o);
        xphi = 2.0 * xphi - 1.0;
        xphi2 = xphi * xphi;
        yphi = rnd(o);
        yphi2 = yphi * yphi;
        r2 = xphi2 + yphi2;
    } while (r2 > 1.0);
    r2 = 1 / r2;
    *c = (xphi2 - yphi2) * r2;
    *s = 2.0 * xphi * yphi * r2;
}

static void rotate(orc *o, int np, const uphi_t *q, double costhe, double sinthe) {
    double sinps2 = q->A * q->A + q->B * q->B;
    if (sinps2 < 1.0E-20) {
        o->u[np] = sinthe * q->cosphi;
        o->v[np] = sinthe * q->sinphi;
        o->w[np] = q->C * costhe;
    } else {
        double sinpsi = sqrt(sinps2);
        double us = sinthe * q->cosphi, vs = sinthe * q->sinphi;
        double sindel = q->B / sinpsi, cosdel = q->A / sinpsi;
        o->u[np] = q->C * cosdel * us - sindel * vs + q->A * costhe;
        o->v[np] = q->C * sindel * us + cosdel * vs + q->B * costhe;
        o->w[np] = -sinpsi * us + q->C * costhe;
    }
}

/* uphi21 (src/ompmc.c:128-165): new azimuth, rotate the particle on top */
static void uphi21(orc *o, uphi_t *q, double costhe, double sinthe) {
    int np = o->np;
    azimuth(o, &q->cosphi, &q->sinphi);
    q->A = o->u[np]; q->B = o->v[np]; q->C = o->w[np];
    rotate(o, np, q, costhe, sinthe);
}

/* uphi32 (src/ompmc.c:167-202): copy np-1 into np, rotate with the saved frame */
static void uphi32(orc *o, const uphi_t *q, double costhe, double sinthe) {
    int np = o->np;
    transfer(o, np, np - 1);
    rotate(o, np, q, costhe, sinthe);
}

static int pwlf_i(int idx, double lv, const double *c1, const double *c0) { return (int)(lv * c1[idx] + c0[idx]); }
static double pwlf(int idx, double lv, const double *c1, const double *c0) { return lv * c1[idx] + c0[idx]; }

/* ------------------------------------------------------------------ photon interactions */
/* Rayleigh (src/ompmc.c:1120-1163); CDF arrays are read without a medium
 * offset and the bin index is always 0, as in the reference (quirk Q1) */
void orc_rayleigh(orc *o, int imed, double eig, double gle, int lgle) {
    const omc_tables *t = o->t;
    int ibin, ib;
    double xv, costhe, csqthe, sinthe, rnno0, rnno1;
    double pmax = pwlf(imed * MXGE + lgle, gle, t->ray_pmax1, t->ray_pmax0);
    double xmax = OMC_HC_INVERSE * eig;
    double dwi = (double)OMC_RAYCDFSIZE - 1.0;
    do {
        rnno1 = rnd(o);
        do {
            rnno0 = rnd(o);
            rnno0 *= pmax;
            ibin = (int)rnno0 * dwi;
            ib = t->ray_iarr[ibin] - 1;
            if ((t->ray_iarr[ibin + 1] - 1) > ib)
                while (rnno0 >= t->ray_fcum[ib + 1]) ib++;
            rnno0 = (rnno0 - t->ray_fcum[ib]) * t->ray_c[ib];
            xv = t->ray_xgrid[ib] * exp(log(1.0 + rnno0) * t->ray_b[ib]);
        } while (xv >= xmax);
        xv /= eig;
        costhe = 1.0 - OMC_TWICE_HC2 * pow(xv, 2.0);
        csqthe = pow(costhe, 2.0);
    } while (2.0 * rnno1 >= (1.0 + csqthe));
    sinthe = sqrt(1.0 - csqthe);
    uphi_t q;
    uphi21(o, &q, costhe, sinthe);
}

static double pair_rej(const omc_tables *t, int imed, double xi, double esedei, double eseder, double tteig) {
    return 2.0 + 3.0 * (esedei + eseder) -
           4.0 * (esedei + eseder + 1.0 - 4.0 * pow((xi - 0.5), 2.0)) *
               (1.0 + 0.25 * log((pow(((1.0 + eseder) * (1.0 + esedei) / (2.0 * tteig)), 2.0)) +
                                 t->zbrang[imed] * pow(xi, 2.0)));
}

/* pair production (src/ompmc.c:1447-1685) */
void orc_pair(orc *o, int imed) {
    const omc_tables *t = o->t;
    int np = o->np;
    double eig = o->e[np];
    double ese1, ese2;
    int iq1, iq2, l, l1;
    o->npold = np;
    if (eig <= 2.1) {
        double r0 = rnd(o), r1 = rnd(o);
        ese2 = RM + 0.5 * r0 * (eig - 2.0 * RM);
        ese1 = eig - ese2;
        if (r1 < 0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    } else {
        double Amax, Bmax, delta, aux;
        const double *d1 = t->dl1 + imed * 8, *d2 = t->dl2 + imed * 8, *d3 = t->dl3 + imed * 8;
        const double *d4 = t->dl4 + imed * 8, *d5 = t->dl5 + imed * 8, *d6 = t->dl6 + imed * 8;
        if (eig < 50.0) {
            l = 4; l1 = l + 1;
            delta = 4.0 * t->delcm[imed] / eig;
            if (delta < 1.0) {
                Amax = d1[l] + delta * (d2[l] + delta * d3[l]);
                Bmax = d1[l1] + delta * (d2[l1] + delta * d3[l1]);
            } else {
                aux = log(delta + d6[l]);
                Amax = d4[l] + d5[l] * aux;
                Bmax = d4[l1] + d5[l1] * aux;
            }
            aux = 1.0 - 2.0 * RM / eig;
            aux = aux * aux;
            aux *= Amax / 3.0;
            aux /= (Bmax + aux);
        } else {
            l = 6;
            Amax = d1[l];
            Bmax = d1[l + 1];
            aux = t->bpar1[imed] * (1.0 - t->bpar0[imed] * RM / eig);
        }
        double eavail = eig - 2.0 * RM, rejf, rejmax, r4;
        do {
            double br, r0 = rnd(o), r1 = rnd(o);
            r4 = rnd(o);
            if (r0 > aux) {
                br = 0.5 * r1;
                rejmax = Bmax;
                l1 = l + 1;
            } else {
                double r2 = rnd(o), r3 = rnd(o);
                br = 0.5 * (1.0 - fmax(fmax(r1, r2), r3));
                rejmax = Amax;
                l1 = l;
            }
            ese2 = br * eavail + RM;
            ese1 = eig - ese2;
            delta = (eig * t->delcm[imed]) / (ese2 * ese1);
            if (delta < 1.0) rejf = d1[l1] + delta * (d2[l1] + delta * d3[l1]);
            else rejf = d4[l1] + d5[l1] * log(delta + d6[l1]);
        } while (r4 * rejmax > rejf);
        ese1 = eig - ese2;
        r4 = rnd(o);
        if (r4 < 0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    }
    o->e[np] = ese1;
    o->e[np + 1] = ese2;
    uphi_t q;
    for (int j = 0; j < 2; j++) {
        double ese = j == 0 ? ese1 : ese2;
        double tteig = eig / RM, ttese = ese / RM;
        double esedei = ttese / (tteig - ttese), eseder = 1.0 / esedei;
        double ximin = 1.0 / (1.0 + pow((M_PI * ttese), 2.0));
        double rejmin = pair_rej(t, imed, ximin, esedei, eseder, tteig);
        double ya = pow((2.0 / tteig), 2.0);
        double xitry = fmax(0.01, fmax(ximin, fmin(0.5, sqrt(ya / t->zbrang[imed]))));
        double galpha = 1.0 + 0.25 * log(ya + t->zbrang[imed] * xitry * xitry);
        double gbeta = 0.5 * t->zbrang[imed] * xitry / (ya + t->zbrang[imed] * xitry * xitry);
        galpha -= gbeta * (xitry - 0.5);
        double ximid = galpha / (3.0 * gbeta);
        if (galpha >= 0.0) ximid = 0.5 - ximid + sqrt(pow(ximid, 2.0) + 0.25);
        else ximid = 0.5 - ximid - sqrt(pow(ximid, 2.0) + 0.25);
        ximid = fmax(0.01, fmax(ximin, fmin(0.5, ximid)));
        double rejmid = pair_rej(t, imed, ximid, esedei, eseder, tteig);
        double rejtop = 1.0 * fmax(rejmin, rejmid);
        double theta, rtest, rejfactor;
        /* accepted when either test passes (quirk Q2) */
        do {
            double xitst = rnd(o);
            double rejtst = pair_rej(t, imed, xitst, esedei, eseder, tteig);
            rtest = rnd(o);
            theta = sqrt(1.0 / xitst - 1.0) / ttese;
            rejfactor = rejtst / rejtop;
        } while ((rtest > rejfactor) && (theta >= M_PI));
        double sinthe = sin(theta), costhe = cos(theta);
        if (j == 0) {
            uphi21(o, &q, costhe, sinthe);
        } else {
            sinthe = -sinthe;
            np += 1;
            o->np = np;
            uphi32(o, &q, costhe, sinthe);
        }
    }
    o->iq[np] = iq2;
    o->iq[np - 1] = iq1;
}

/* Klein-Nishina Compton (src/ompmc.c:1688-1801) */
void orc_compton(orc *o) {
    int np = o->np;
    double eig = o->e[np], ko = o->e[np] / RM;
    double broi = 1.0 + 2.0 * ko, bro = 1.0 / broi;
    int first = 1;
    double sinthe = 0.0, costhe, br, r1, r2, r3, aux, rejf3, temp = 0.0;
    double alph1 = 0.0, alph2 = 0.0, alpha = 0.0, rejmax = 0.0;
    o->npold = np;
    do {
        if (ko > 2.0) {
            if (first) {
                alph1 = log(broi);
                alph2 = ko * (broi + 1.0) * pow(bro, 2.0);
                alpha = alph1 + alph2;
            }
            do {
                r1 = rnd(o);
                r2 = rnd(o);
                if (r1 * alpha < alph1) br = exp(alph1 * r2) * bro;
                else br = sqrt(r2 * pow(broi, 2.0) + (1.0 - r2)) * bro;
                temp = (1.0 - br) / (ko * br);
                sinthe = fmax(0.0, temp * (2.0 - temp));
                aux = 1.0 + pow(br, 2.0);
                rejf3 = aux - br * sinthe;
                r3 = rnd(o);
            } while (r3 * aux > rejf3);
        } else {
            if (first) rejmax = broi + bro;
            do {
                r1 = rnd(o);
                r2 = rnd(o);
                br = bro + (1.0 - bro) * r1;
                temp = (1.0 - br) / (ko * br);
                sinthe = fmax(0.0, temp * (2.0 - temp));
                rejf3 = 1.0 + br * br - br * sinthe;
            } while (r2 * br * rejmax > rejf3);
        }
        first = 0;
    } while ((br < bro) || (br > 1));
    costhe = 1.0 - temp;
    sinthe = sqrt(sinthe);
    double esg = br * eig, ese = eig - esg + RM;
    o->e[np] = esg;
    uphi_t q;
    uphi21(o, &q, costhe, sinthe);
    np += 1;
    o->np = np;
    aux = 1.0 + br * br - 2.0 * br * costhe;
    if (aux > 1.0E-8) {
        costhe = (1.0 - br * costhe) / sqrt(aux);
        sinthe = (1.0 - costhe) * (1.0 + costhe);
        sinthe = sinthe > 0.0 ? -sqrt(sinthe) : 0.0;
    } else {
        costhe = 0.0;
        sinthe = -1.0;
    }
    uphi32(o, &q, costhe, sinthe);
    o->e[np] = ese;
    o->iq[np] = -1;
}

/* photo-electric (src/ompmc.c:1804-1864): no binding energy (quirk Q3) */
void orc_photo(orc *o) {
    int np = o->np;
    o->npold = np;
    o->e[np] += RM;
    o->iq[np] = -1;
    double eelec = o->e[np];
    if (eelec > o->r->ecut[o->ir[np]]) {
        double beta = sqrt((eelec - RM) * (eelec + RM)) / eelec;
        double costhe, sinth2, gamma = eelec / RM;
        double alpha = 0.5 * gamma - 0.5 + 1.0 / gamma;
        double ratio = beta / alpha, rnpht2, xi;
        do {
            double rnpht = rnd(o);
            rnpht = 2.0 * rnpht - 1.0;
            if (ratio <= 0.2) {
                double fkappa = rnpht + 0.5 * ratio * (1.0 - rnpht) * (1.0 + rnpht);
                if (gamma < 100.0) {
                    costhe = (beta + fkappa) / (1.0 + beta * fkappa);
                } else {
                    if (fkappa > 0.0)
                        costhe = 1.0 - (1.0 - fkappa) * (gamma - 3.0) / (2.0 * (1.0 + fkappa) * pow((gamma - 1.0), 3.0));
                    else
                        costhe = (beta + fkappa) / (1.0 + beta * fkappa);
                }
                xi = (1.0 + beta * fkappa) * pow(gamma, 2.0);
            } else {
                xi = pow(gamma, 2.0) * (1.0 + alpha * (sqrt(1.0f + ratio * (2.0 * rnpht + ratio)) - 1.0));
                costhe = (1.0 - 1.0 / xi) / beta;
            }
            sinth2 = fmax((1.0 - costhe) * (1.0 + costhe), 0.0);
            rnpht2 = rnd(o);
        } while (rnpht2 > 0.5 * (1.0 + gamma) * sinth2 * xi / gamma);
        uphi_t q;
        uphi21(o, &q, costhe, sqrt(sinth2));
    }
}

/* photon step with splitting (src/ompmc.c:1867-2139) */
void orc_photon(orc *o) {
    const omc_tables *t = o->t;
    const omc_regions *R = o->r;
    int np = o->np;
    int irl = o->ir[np], irold, irnew;
    int imed = R->med[irl];
    int idisc;
    double rhof, tstep, ustep, edep, eig = o->e[np];
    double dpmfp, dpmfp_old, gmfpr0 = 0.0, gmfp = 0.0, gbr1, gbr2, rnno;
    int nsplit = o->nsplit, i_survive_s, ip;
    double d_eta, eta_prime, a_survive;
    double xsave, ysave, zsave, usave, vsave, wsave, esave, wtsave;
    int irsave;
    if (eig <= R->pcut[irl] || o->wt[np] == 0) {
        edep = eig;
        ausgab(o, edep);
        o->np -= 1;
        return;
    }
    int lgle = 0;
    double gle = log(eig), cohfac = 0.0;
    int ptrans;
start_mfp_loop:
    rnno = rnd(o);
    rnno /= (double)nsplit;
    d_eta = 1.0 / (double)nsplit;
    eta_prime = 1.0 - rnno + d_eta;
    xsave = o->x[np]; ysave = o->y[np]; zsave = o->z[np];
    usave = o->u[np]; vsave = o->v[np]; wsave = o->w[np];
    esave = o->e[np]; wtsave = o->wt[np] / (double)nsplit;
    irsave = o->ir[np];
    np -= 1;
    rnno = rnd(o);
    a_survive = rnno * nsplit;
    i_survive_s = (int)a_survive;
    dpmfp_old = 0.0;
    for (int isplit = 0; isplit < nsplit; isplit++) {
        ptrans = 1;
        eta_prime -= d_eta;
        if (eta_prime <= 0.0) goto end_mfp_loop;
        dpmfp = -log(eta_prime) - dpmfp_old;
        dpmfp_old += dpmfp;
        np += 1;
        o->np = np;
        if (np >= MXSTACK) { o->err = 1; o->np = -1; return; }
        o->x[np] = xsave; o->y[np] = ysave; o->z[np] = zsave;
        o->u[np] = usave; o->v[np] = vsave; o->w[np] = wsave;
        o->e[np] = esave; o->wt[np] = wtsave;
        o->ir[np] = irsave; o->iq[np] = 0;
        irl = o->ir[np];
        irold = irl;
        imed = R->med[irl];
        do {
            if (imed != -1) {
                lgle = pwlf_i(imed, gle, t->ge1, t->ge0) - 1;
                gmfpr0 = pwlf(imed * MXGE + lgle, gle, t->gmfp1, t->gmfp0);
                rhof = R->rhof[irl];
                gmfp = gmfpr0 / rhof;
                cohfac = pwlf(imed * MXGE + lgle, gle, t->cohe1, t->cohe0);
                gmfp *= cohfac;
                tstep = gmfp * dpmfp;
            } else {
                tstep = 1.0E8;
            }
            irnew = irl;
            idisc = 0;
            ustep = tstep;
            howfar(o, &idisc, &irnew, &ustep);
            edep = 0.0;
            o->x[np] += ustep * o->u[np];
            o->y[np] += ustep * o->v[np];
            o->z[np] += ustep * o->w[np];
            if (idisc > 0) {
                o->esc += o->wt[np] * o->e[np];     /* leaves untallied (src/ompmc.c:2008-2017) */
                np -= 1;
                o->np = np;
                if (np < 0) return;
                goto end_mfp_loop;
            }
            if (imed != -1) dpmfp = fmax(0.0, dpmfp - ustep / gmfp);
            if (irnew != irold) {
                o->ir[np] = irnew;
                irl = irnew;
                irold = irnew;
                imed = R->med[irl];
            }
            if (imed != -1 && dpmfp <= OMC_SGMFP) ptrans = 0;
        } while (ptrans);
        xsave = o->x[np]; ysave = o->y[np]; zsave = o->z[np];
        irsave = o->ir[np];
        rnno = rnd(o);
        if (rnno <= 1.0 - cohfac) {
            if (isplit != i_survive_s) {
                np -= 1;
                o->np = np;
                continue;
            } else {
                o->wt[np] *= nsplit;
                orc_rayleigh(o, imed, eig, gle, lgle);
                continue;
            }
        } else {
            rnno = rnd(o);
            gbr1 = pwlf(imed * MXGE + lgle, gle, t->gbr11, t->gbr10);
            if (rnno <= gbr1 && eig > 2.0 * RM) {
                orc_pair(o, imed);
                np = o->np;
            } else {
                gbr2 = pwlf(imed * MXGE + lgle, gle, t->gbr21, t->gbr20);
                if (rnno < gbr2) {
                    orc_compton(o);
                    np = o->np;
                } else {
                    orc_photo(o);
                    np = o->np;
                }
            }
        }
        /* keep one scattered photon of the split set, carrying full weight */
        ip = o->npold;
        do {
            if (o->iq[ip] == 0) {
                if (isplit != i_survive_s) {
                    if (ip < np) {
                        o->e[ip] = o->e[np]; o->iq[ip] = o->iq[np];
                        o->u[ip] = o->u[np]; o->v[ip] = o->v[np];
                        o->w[ip] = o->w[np]; o->wt[ip] = o->wt[np];
                    }
                    np -= 1;
                } else {
                    o->wt[ip] *= nsplit;
                    ip += 1;
                }
            } else {
                ip += 1;
            }
        } while (ip <= np);
        o->np = np;
    }
end_mfp_loop:
    if (np < 0) return;
    if (o->iq[np] == 0) {
        eig = o->e[np];
        irl = o->ir[np];
        imed = R->med[irl];
        if (eig <= R->pcut[irl]) {
            edep = eig;
            ausgab(o, edep);
            np -= 1;
            o->np = np;
            return;
        }
        gle = log(eig);
        goto start_mfp_loop;
    }
}

/* ------------------------------------------------------------------ electron physics */
/* spin rejection (src/ompmc.c:3115-3186) */
typedef struct { int i, j; } spinr_t;
typedef struct { int i, j; double omega2; } mscats_t;

static double spin_rejection(orc *o, int imed, int qel, double elke, double beta2, double q1, double cost,
                             int *spin_index, int is_single, spinr_t *sr) {
    const omc_tables *t = o->t;
    double ai, aj, ak, qq1, xi, rnno;
    int k;
    if (*spin_index) {
        *spin_index = 0;
        if (beta2 >= t->b2spin_min) {
            ai = (beta2 - t->b2spin_min) * t->dbeta2i;
            sr->i = (int)ai;
            ai -= (double)sr->i;
            sr->i += OMC_MXE_SPIN + 1;
        } else if (elke > t->espml) {
            ai = (elke - t->espml) * t->dleneri;
            sr->i = (int)ai;
            ai -= sr->i;
        } else {
            sr->i = 0;
            ai = -1.0f;
        }
        rnno = rnd(o);
        if (rnno < ai) sr->i += 1;
        if (is_single) {
            sr->j = 0;
        } else {
            qq1 = 2.0 * q1;
            qq1 = qq1 / (1.0 + qq1);
            aj = qq1 * t->dqq1i;
            sr->j = (int)aj;
            if (sr->j >= OMC_MXQ_SPIN) {
                sr->j = OMC_MXQ_SPIN;
            } else {
                aj -= (double)sr->j;
                rnno = rnd(o);
                if (rnno < aj) sr->j += 1;
            }
        }
    }
    xi = sqrt(0.5 * (1.0 - cost));
    ak = xi * OMC_MXU_SPIN;
    k = (int)ak;
    ak -= (double)k;
    size_t base = (size_t)imed * OMC_NSPIN_MED + (size_t)qel * (OMC_MXE_SPIN1 + 1) * (OMC_MXQ_SPIN + 1) * (OMC_MXU_SPIN + 1) +
                  (size_t)sr->i * (OMC_MXQ_SPIN + 1) * (OMC_MXU_SPIN + 1) + (size_t)sr->j * (OMC_MXU_SPIN + 1);
    double s0 = t->spin_rej[base + k], s1 = t->spin_rej[base + k + 1];
    return (1.0 - ak) * (s0) + ak * s1;
}

/* single elastic scattering (src/ompmc.c:3188-3217) */
void orc_sscat(orc *o, int imed, int qel, double chia2, double elke, double beta2, double *cost, double *sint) {
    int spin_index = 1;
    spinr_t sr = {0, 0};
    double xi, rejf, rnno;
    do {
        xi = rnd(o);
        xi = 2.0 * chia2 * xi / (1.0 - xi + chia2);
        *cost = 1.0 - xi;
        rejf = spin_rejection(o, imed, qel, elke, beta2, 0, *cost, &spin_index, 1, &sr);
        rnno = rnd(o);
    } while (rnno > rejf);
    *sint = sqrt(xi * (2.0 - xi));
}

/* multiple-scattering angle (src/ompmc.c:3624-3803).  The interpolation
 * inside the q^(2+) surface bin uses du = ums[k] - ums[k] = 0, so the sampled
 * u is the bin's tabulated value, as in the reference. */
static void mscat(orc *o, int imed, int qel, int *spin_index, int *find_index, double elke, double beta2, double q1,
                  double lambda, double chia2, double *cost, double *sint, mscats_t *ms, spinr_t *sr) {
    const omc_tables *t = o->t;
    double xi, rejf, rnno;
    double explambda = exp(-lambda);
    if (lambda <= 13.8) {
        double sprob = rnd(o);
        if (sprob < explambda) { *cost = 1.0; *sint = 0.0; return; }
        double wsum = (1.0 + lambda) * explambda;
        if (sprob < wsum) {
            do {
                xi = rnd(o);
                xi = 2.0 * chia2 * xi / (1.0 - xi + chia2);
                *cost = 1.0 - xi;
                rejf = spin_rejection(o, imed, qel, elke, beta2, q1, *cost, spin_index, 0, sr);
                rnno = rnd(o);
            } while (rnno > rejf);
            *sint = sqrt(xi * (2.0 - xi));
            return;
        }
        if (lambda <= 1) {
            int icount = 0;
            double wprob = explambda, sinz, cosz, phi;
            wsum = explambda;
            *cost = 1.0;
            *sint = 0.0;
            do {
                icount += 1;
                if (icount > 20) break;
                wprob = wprob * lambda / icount;
                wsum = wsum + wprob;
                do {
                    xi = rnd(o);
                    xi = 2.0 * chia2 * xi / (1.0 - xi + chia2);
                    cosz = 1.0 - xi;
                    rejf = spin_rejection(o, imed, qel, elke, beta2, q1, cosz, spin_index, 0, sr);
                    rnno = rnd(o);
                } while (rnno > rejf);
                sinz = xi * (2.0 - xi);
                if (sinz > 1.0E-20) {
                    sinz = sqrt(sinz);
                    xi = rnd(o);
                    phi = xi * 6.2831853;
                    *cost = (*cost) * cosz - *sint * sinz * cos(phi);
                    *sint = sqrt(fmax(0.0, (double)((1.0 - (*cost)) * (1.0 + (*cost)))));
                }
            } while (wsum <= sprob);
            return;
        }
    }
    if (lambda <= OMC_LAMBMAX_MS) {
        double ai, aj, llmbda = log(lambda);
        if (*find_index) {
            ai = llmbda * t->dllambi;
            ms->i = (int)ai;
            ai -= (double)ms->i;
            xi = rnd(o);
            if (xi < ai) ms->i += 1;
            if (q1 < OMC_QMIN_MS) {
                ms->j = 0;
            } else if (q1 < OMC_QMAX_MS) {
                aj = q1 * t->dqmsi;
                ms->j = (int)aj;
                aj -= (double)ms->j;
                xi = rnd(o);
                if (xi < aj) ms->j += 1;
            } else {
                ms->j = OMC_MXQ_MS;
            }
            if (llmbda < 2.2299)
                ms->omega2 = chia2 * (lambda + 4.0) *
                             (1.347006 + llmbda * (0.209364 - llmbda * (0.45525 - llmbda * (0.50142 - 0.081234 * llmbda))));
            else
                ms->omega2 = chia2 * (lambda + 4.0) * (-2.77164 + llmbda * (2.94874 - llmbda * (0.1535754 - llmbda * 0.00552888)));
            *find_index = 0;
        }
        int k;
        double a, ak, u, du, x1;
        do {
            xi = rnd(o);
            ak = xi * OMC_MXU_MS;
            k = ak;
            ak -= k;
            int base = ms->i * (OMC_MXQ_MS + 1) * (OMC_MXU_MS + 1) + (ms->j) * (OMC_MXU_MS + 1);
            if (ak > t->wms[base + k]) k = t->ims[base + k];
            a = t->fms[base + k];
            u = t->ums[base + k];
            du = t->ums[base + k] - u;
            xi = rnd(o);
            if (fabs(a) < 0.2) {
                x1 = 0.5 * (1.0 - xi) * a;
                u += xi * du * (1.0 + x1 * (1.0 - xi * a));
            } else {
                u -= du / a * (1.0 - sqrt(1.0 + xi * a * (2.0 + a)));
            }
            xi = ms->omega2 * u / (1.0 + 0.5 * ms->omega2 - u);
            if (xi > 1.99999) xi = 1.99999;
            *cost = 1.0 - xi;
            rejf = spin_rejection(o, imed, qel, elke, beta2, q1, *cost, spin_index, 0, sr);
            rnno = rnd(o);
        } while (rnno > rejf);
        *sint = sqrt(xi * (2.0 - xi));
    }
}

void orc_mscat_fresh(orc *o, int imed, int qel, double elke, double beta2, double q1, double lambda, double chia2,
                     double *cost, double *sint) {
    int si = 1, fi = 1;
    mscats_t ms = {0, 0, 0.0};
    spinr_t sr = {0, 0};
    mscat(o, imed, qel, &si, &fi, elke, beta2, q1, lambda, chia2, cost, sint, &ms, &sr);
}

/* PRESTA-II style two-substep multiple scattering (src/ompmc.c:3805-3994) */
double orc_msdist(orc *o, int imed, int iq, double rhof, double de, double tustep, double eke, double *x_final,
                  double *y_final, double *z_final, double *u_final, double *v_final, double *w_final) {
    const omc_tables *t = o->t;
    int qel = (1 + iq) / 2;
    mscats_t ms = {0, 0, 0.0};
    spinr_t sr = {0, 0};
    double blccc = t->blcc[imed], xcccc = t->xcc[imed];
    double e = eke - 0.5 * de;
    double tau = e / RM, tau2 = pow(tau, 2.0);
    double epsilon = de / eke, epsilonp = de / e;
    e *= (1.0 - pow(epsilonp, 2.0) * (6.0 + 10.0 * tau + 5.0 * tau2) / (24.0 * tau2 + 72.0 * tau + 48.0));
    double p2 = e * (e + 2.0 * RM);
    double beta2 = p2 / (p2 + pow(RM, 2.0));
    double chia2 = xcccc / (4.0 * p2 * blccc);
    double lambda = 0.5 * tustep * rhof * blccc / beta2;
    double temp2 = 0.166666 * (4.0 + tau * (6.0 + tau * (7.0 + tau * (4.0 + tau)))) *
                   (epsilonp / ((tau + 1.0) * (tau + 2.0))) * (epsilonp / ((tau + 1.0) * (tau + 2.0)));
    lambda *= (1.0 - temp2);
    double elke = log(e);
    int lelke = pwlf_i(imed, elke, t->eke1, t->eke0) - 1;
    if (lelke < 0) {
        lelke = 0;
        elke = (1.0 - t->eke0[imed]) / t->eke1[imed];
    }
    double etap, xi_corr, gamma;
    int ix = MXEKE * imed + lelke;
    if (qel == 0) {
        etap = pwlf(ix, elke, t->etae_ms1, t->etae_ms0);
        xi_corr = pwlf(ix, elke, t->q1ce_ms1, t->q1ce_ms0);
        gamma = pwlf(ix, elke, t->q2ce_ms1, t->q2ce_ms0);
    } else {
        etap = pwlf(ix, elke, t->etap_ms1, t->etap_ms0);
        xi_corr = pwlf(ix, elke, t->q1cp_ms1, t->q1cp_ms0);
        gamma = pwlf(ix, elke, t->q2cp_ms1, t->q2cp_ms0);
    }
    double ms_corr = pwlf(ix, elke, t->blcce1, t->blcce0);
    chia2 *= etap;
    lambda /= (etap * (1.0 + chia2));
    lambda *= ms_corr;
    double chilog = log(1.0 + 1.0 / chia2);
    double q1 = 2.0 * chia2 * (chilog * (1.0 + chia2) - 1.0);
    gamma = 6.0 * chia2 * (1.0 + chia2) * (chilog * (1.0 + 2.0 * chia2) - 2.0) / q1 * gamma;
    double xi = q1 * lambda;
    int find_index = 1, spin_index = 1;
    double w1 = 0.0, sint1 = 0.0, w2 = 0.0, sint2 = 0.0, cphi1, sphi1, cphi2, sphi2;
    mscat(o, imed, qel, &spin_index, &find_index, elke, beta2, xi, lambda, chia2, &w1, &sint1, &ms, &sr);
    azimuth(o, &cphi1, &sphi1);
    mscat(o, imed, qel, &spin_index, &find_index, elke, beta2, xi, lambda, chia2, &w2, &sint2, &ms, &sr);
    azimuth(o, &cphi2, &sphi2);
    double u2 = sint2 * cphi2, v2 = sint2 * sphi2;
    double u2p = w1 * u2 + sint1 * w2;
    double us = u2p * cphi1 - v2 * sphi1;
    double vs = u2p * sphi1 + v2 * cphi1;
    double ws = w1 * w2 - sint1 * u2;
    xi *= 2 * xi_corr;
    double eta = rnd(o);
    double eta1 = 0.5 * (1.0 - eta);
    double delta = 0.9082483 - (0.1020621 - 0.0263747 * gamma) * xi;
    double temp1 = 2.0 + tau;
    double temp = (2.0 + tau * temp1) / ((tau + 1.0) * temp1);
    temp -= (tau + 1.0) / ((tau + 2.0) * (chilog * (1.0 + chia2) - 1.0));
    temp *= epsilonp;
    temp1 = 1.0 - temp;
    delta += 0.40824829 * (epsilon * (tau + 1.0) / ((tau + 2.0) * (chilog * (1.0 + chia2) - 1.0) * (chilog * (1.0 + 2.0 * chia2) - 2.0)) -
                           0.25 * pow(temp, 2.0));
    double b = eta * delta, c = eta * (1.0 - delta);
    double w1v2 = w1 * v2;
    double ut = b * sint1 * cphi1 + c * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
    double vt = b * sint1 * sphi1 + c * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
    double wt = eta1 * (1.0 + temp) + b * w1 + c * w2 + eta1 * ws * temp1;
    double ustep = tustep * sqrt(ut * ut + vt * vt + wt * wt);
    int np = o->np;
    double x0 = o->x[np], y0 = o->y[np], z0 = o->z[np];
    double u0 = o->u[np], v0 = o->v[np], w0 = o->w[np];
    double sint02 = u0 * u0 + v0 * v0;
    if (sint02 > 1.0E-20) {
        double sint0 = sqrt(sint02), sint0i = 1.0 / sint0;
        double cphi0 = sint0i * u0, sphi0 = sint0i * v0;
        u2p = w0 * us + sint0 * ws;
        ws = w0 * ws - sint0 * us;
        us = u2p * cphi0 - vs * sphi0;
        vs = u2p * sphi0 + vs * cphi0;
        u2p = w0 * ut + sint0 * wt;
        wt = w0 * wt - sint0 * ut;
        ut = u2p * cphi0 - vt * sphi0;
        vt = u2p * sphi0 + vt * cphi0;
    } else {
        wt = w0 * wt;
        ws = w0 * ws;
    }
    *x_final = x0 + tustep * ut;
    *y_final = y0 + tustep * vt;
    *z_final = z0 + tustep * wt;
    *u_final = us;
    *v_final = vs;
    *w_final = ws;
    return ustep;
}

/* CSDA path length within one PWLF interval (src/ompmc.c:3997-4032) */
double orc_drange(orc *o, int imed, int iq, int lelke, double ekei, double ekef, double elkei, double elkef) {
    const omc_tables *t = o->t;
    double dedxmid, aux, fedep = 1.0 - ekef / ekei;
    double elktmp = 0.5 * (elkei + elkef + 0.25 * fedep * fedep * (1.0 + fedep * (1.0 + 0.875 * fedep)));
    int ix = MXEKE * imed + lelke;
    if (iq < 0) {
        dedxmid = pwlf(ix, elktmp, t->ededx1, t->ededx0);
        dedxmid = 1.0 / dedxmid;
        aux = t->ededx1[ix] * dedxmid;
    } else {
        dedxmid = pwlf(ix, elktmp, t->pdedx1, t->pdedx0);
        dedxmid = 1.0 / dedxmid;
        aux = t->pdedx1[ix] * dedxmid;
    }
    aux = aux * (1.0 + 2.0 * aux) * fedep * fedep / (6.0 * (2.0 - fedep) * (2.0 - fedep));
    return fedep * ekei * dedxmid * (1.0 + aux);
}

/* sub-threshold energy loss over a path tustep (src/ompmc.c:4034-4126) */
double orc_eloss(orc *o, int imed, int iq, int irl, double rhof, double tustep, double range, double eke, double elke,
                 int lelke) {
    const omc_tables *t = o->t;
    (void)irl;
    double aux, dedxmid, de, fedep, elktmp, eketmp;
    int lelktmp, qel = (1 + iq) / 2;
    const double *rep = t->range_ep + (size_t)qel * t->nmed * MXEKE + (size_t)imed * MXEKE;
    double tuss = range - rep[lelke] / rhof;
    if (tuss >= tustep) {
        int ix = imed * MXEKE + lelke;
        if (iq < 0) {
            dedxmid = pwlf(ix, elke, t->ededx1, t->ededx0);
            aux = t->ededx1[ix] / dedxmid;
        } else {
            dedxmid = pwlf(ix, elke, t->pdedx1, t->pdedx0);
            aux = t->pdedx1[ix] / dedxmid;
        }
        de = dedxmid * tustep * rhof;
        fedep = de / eke;
        de *= (1.0 - 0.5 * fedep * aux * (1.0 - 0.333333 * fedep * (aux - 1.0 - 0.25 * fedep * (2.0 - aux * (4.0 - aux)))));
    } else {
        lelktmp = lelke;
        tuss = (range - tustep) * rhof;
        if (tuss <= 0) {
            de = eke - t->te[imed] * 0.99;
        } else {
            while (tuss < rep[lelktmp]) lelktmp -= 1;
            elktmp = (lelktmp + 2 - t->eke0[imed]) / t->eke1[imed];
            eketmp = t->e_array[imed * MXEKE + lelktmp + 1];
            tuss = (rep[lelktmp + 1] - tuss) / rhof;
            int ix = MXEKE * imed + lelktmp;
            if (iq < 0) {
                dedxmid = pwlf(ix, elktmp, t->ededx1, t->ededx0);
                aux = t->ededx1[ix] / dedxmid;
            } else {
                dedxmid = pwlf(ix, elktmp, t->pdedx1, t->pdedx0);
                aux = t->pdedx1[ix] / dedxmid;
            }
            de = dedxmid * tuss * rhof;
            fedep = de / eketmp;
            de *= (1.0 - 0.5 * fedep * aux * (1.0 - 0.333333 * fedep * (aux - 1.0 - 0.25 * fedep * (2.0 - aux * (4.0 - aux)))));
            de += eke - eketmp;
        }
    }
    return de;
}

/* photon roulette after an electron interaction when splitting is on */
static void roulette_photons(orc *o) {
    if (o->nsplit > 1) {
        for (int ip = o->npold; ip <= o->np; ip++) {
            if (o->iq[ip] == 0) {
                double r = rnd(o);
                if (r * (double)o->nsplit > 1.0) { o->wt[ip] = 0.0; o->e[ip] = 0.0; }
                else o->wt[ip] *= o->nsplit;
            }
        }
    }
}

/* two-photon annihilation at rest (src/ompmc.c:4129-4185) */
void orc_rannih(orc *o) {
    int np = o->np;
    o->npold = np;
    double rnno = rnd(o);
    double costhe = 2.0 * rnno - 1;
    double sinthe = sqrt(fmax(0.0, (1.0 - costhe) * (1.0 + costhe)));
    rnno = rnd(o);          /* drawn and unused, as in the reference */
    (void)rnno;
    double cosphi, sinphi;
    azimuth(o, &cosphi, &sinphi);
    o->e[np] = RM;
    o->iq[np] = 0;
    o->u[np] = sinthe * cosphi;
    o->v[np] = sinthe * sinphi;
    o->w[np] = costhe;
    np += 1;
    o->e[np] = RM;
    o->iq[np] = 0;
    transfer(o, np, np - 1);
    o->u[np] = -1.0 * o->u[np - 1];
    o->v[np] = -1.0 * o->v[np - 1];
    o->w[np] = -1.0 * o->w[np - 1];
    o->np = np;
    roulette_photons(o);
}

/* bremsstrahlung (src/ompmc.c:4188-4374) */
void orc_brems(orc *o) {
    const omc_tables *t = o->t;
    int np = o->np, irl = o->ir[np], imed = o->r->med[irl];
    double eie = o->e[np], phi1, phi2;
    o->npold = np;
    int l = eie < 50.0 ? 1 : 3, l1 = l + 1;
    double ekin = eie - RM, brmin = t->ap[imed] / ekin, waux = -log(brmin);
    double a = o->u[np], b = o->v[np], c = o->w[np];
    double sinpsi = a * a + b * b, sindel = 0.0, cosdel = 0.0;
    if (sinpsi > 1.0E-20) {
        sinpsi = sqrt(sinpsi);
        sindel = b / sinpsi;
        cosdel = a / sinpsi;
    }
    double ztarg = t->zbrang[imed], tteie = eie / RM;
    double beta = sqrt((tteie - 1.0) * (tteie + 1.0)) / tteie;
    double y2max = 2.0 * beta * (1.0 + beta) * tteie * tteie;
    double y2maxi = 1.0 / y2max;
    double z2max = y2max + 1.0;
    double z2maxi = sqrt(z2max);
    double aux, br, delta, r6, r7, rejf, ese, esg;
    const double *d1 = t->dl1 + imed * 8, *d2 = t->dl2 + imed * 8, *d3 = t->dl3 + imed * 8;
    const double *d4 = t->dl4 + imed * 8, *d5 = t->dl5 + imed * 8, *d6 = t->dl6 + imed * 8;
    do {
        r6 = rnd(o);
        r7 = rnd(o);
        br = brmin * exp(r6 * waux);
        esg = ekin * br;
        ese = eie - esg;
        delta = esg / eie / ese * t->delcm[imed];
        aux = ese / eie;
        if (delta < 1.0) {
            phi1 = d1[l - 1] + delta * (d2[l - 1] + delta * d3[l - 1]);
            phi2 = d1[l1 - 1] + delta * (d2[l1 - 1] + delta * d3[l1 - 1]);
        } else {
            phi1 = d4[l - 1] + d5[l - 1] * log(delta + d6[l - 1]);
            phi2 = phi1;
        }
        rejf = (1.0 + pow(aux, 2.0)) * phi1 - 2.0 * aux * phi2 / 3.0;
    } while (r7 >= rejf);
    np += 1;
    o->e[np] = esg;
    o->iq[np] = 0;
    transfer(o, np, np - 1);
    double y2tst = 0.0, ttese = ese / RM, esedei = ttese / tteie;
    double rjarg1 = 1.0 + esedei * esedei, rjarg2 = rjarg1 + 2.0 * esedei, rjarg3;
    double rtest = 1.0, rejtst = 0.0;
    aux = 2.0 * ese * tteie / esg;
    aux = aux * aux;
    double aux1 = aux * ztarg;
    if (aux1 > 10.0) rjarg3 = -log(t->zbrang[imed]) + (1.0 - aux1) / pow(aux1, 2.0);
    else rjarg3 = log(aux / (1.0 + aux1));
    double rejmax = rjarg1 * rjarg3 - rjarg2;
    while (rtest >= rejtst) {
        y2tst = rnd(o);
        rtest = rnd(o);
        double aux3 = z2maxi / (y2tst + (1.0 - y2tst) * z2maxi);
        rtest = rtest * aux3 * rejmax;
        y2tst = pow(aux3, 2.0) - 1.0;
        double y2tst1 = esedei * y2tst / pow(aux3, 4.0);
        double aux4 = 16.0 * y2tst1 - rjarg2;
        double aux5 = rjarg1 - 4.0 * y2tst1;
        if (rtest < aux4 + aux5 * rjarg3) break;
        double aux2 = log(aux / (1.0 + aux1 / pow(aux3, 4.0)));
        rejtst = aux4 + aux5 * aux2;
    }
    double costhe = 1.0 - 2.0 * y2tst * y2maxi;
    double sinthe = sqrt(fmax(0.0, (1.0 - pow(costhe, 2.0))));
    double cphi, sphi;
    azimuth(o, &cphi, &sphi);
    if (sinpsi >= 1.0E-10) {
        double us = sinthe * cphi, vs = sinthe * sphi;
        o->u[np] = c * cosdel * us - sindel * vs + a * costhe;
        o->v[np] = c * sindel * us + cosdel * vs + b * costhe;
        o->w[np] = c * costhe - sinpsi * us;
    } else {
        o->u[np] = sinthe * cphi;
        o->v[np] = sinthe * sphi;
        o->w[np] = c * costhe;
    }
    o->e[np - 1] = ese;
    o->np = np;
    roulette_photons(o);
}

/* Moller (src/ompmc.c:4377-4453); no interaction below 2*TE (quirk Q4) */
void orc_moller(orc *o) {
    const omc_tables *t = o->t;
    int np = o->np, irl = o->ir[np], imed = o->r->med[irl];
    double eie = o->e[np], ekin = eie - RM;
    o->npold = np;
    if (ekin <= 2.0 * t->te[imed]) return;
    double t0 = ekin / RM, e0 = t0 + 1.0, extrae = eie - t->thmoll[imed];
    double g2 = pow(t0, 2.0) / pow(e0, 2.0), g3 = (2.0 * t0 + 1.0) / pow(e0, 2.0);
    double br, gmax = (1.0 + 1.25 * g2), rejf4, r, r27, r28;
    do {
        r27 = rnd(o);
        br = t->te[imed] / (ekin - extrae * r27);
        r = br / (1.0 - br);
        r28 = rnd(o);
        rejf4 = (1.0 + g2 * pow(br, 2.0) + r * (r - g3));
        r28 *= gmax;
    } while (r28 > rejf4);
    double ekse2 = br * ekin, ese1 = eie - ekse2, ese2 = ekse2 + RM;
    o->e[np] = ese1;
    o->e[np + 1] = ese2;
    double h1 = (eie + RM) / ekin;
    double costh = h1 * (ese1 - RM) / (ese1 + RM);
    double sinthe = sqrt(1.0 - costh), costhe = sqrt(costh);
    uphi_t q;
    uphi21(o, &q, costhe, sinthe);
    np += 1;
    o->np = np;
    o->iq[np] = -1;
    costh = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrt(1.0 - costh);
    costhe = sqrt(costh);
    uphi32(o, &q, costhe, sinthe);
}

/* Bhabha (src/ompmc.c:4456-4543) */
void orc_bhabha(orc *o) {
    const omc_tables *t = o->t;
    int np = o->np, irl = o->ir[np], imed = o->r->med[irl];
    double eip = o->e[np], ekin = eip - RM, t0 = ekin / RM, e0 = t0 + 1.0;
    double yy = 1.0 / (t0 + 2.0);
    double beta2 = (pow(e0, 2.0) - 1.0) / pow(e0, 2.0);
    double ep0 = t->te[imed] / ekin, ep0c = 1.0 - ep0, yp = 1.0 - 2.0 * yy;
    o->npold = np;
    double b1, b2, b3, b4;
    b4 = pow(yp, 3.0);
    b3 = b4 + pow(yp, 2.0);
    b2 = yp * (3.0 + pow(yy, 2.0));
    b1 = 2.0 - pow(yy, 2.0);
    double r3, r4, br, rejf2;
    do {
        r3 = rnd(o);
        br = ep0 / (1.0 - ep0c * r3);
        r4 = rnd(o);
        rejf2 = (1.0 - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4))));
    } while (r4 > rejf2);
    if (br < 0.5) {
        o->iq[np + 1] = -1;
    } else {
        o->iq[np] = -1;
        o->iq[np + 1] = 1;
        br = 1.0 - br;
    }
    br = fmax(br, 0.0);
    double ekse2 = br * ekin, ese1 = eip - ekse2, ese2 = ekse2 + RM;
    o->e[np] = ese1;
    o->e[np + 1] = ese2;
    double h1 = (eip + RM) / ekin;
    double costh = fmin(1.0, h1 * (ese1 - RM) / (ese1 + RM));
    double sinthe = sqrt(1.0 - costh), costhe = sqrt(costh);
    uphi_t q;
    uphi21(o, &q, costhe, sinthe);
    np += 1;
    o->np = np;
    costh = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrt(1.0 - costh);
    costhe = sqrt(costh);
    uphi32(o, &q, costhe, sinthe);
}

/* two-photon annihilation in flight (src/ompmc.c:4546-4663) */
void orc_annih(orc *o) {
    int np = o->np;
    double avip = o->e[np] + RM, a = avip / RM, g, tt, p;
    g = a - 1.0;
    tt = g - 1.0;
    p = sqrt(a * tt);
    o->npold = np;
    double pot = p / tt, ep0 = 1.0 / (a + p), wsamp = log((1.0 - ep0) / ep0);
    double aa = o->u[np], bb = o->v[np], cc = o->w[np];
    double sinpsi = pow(aa, 2.0) + pow(bb, 2.0), sindel = 0.0, cosdel = 0.0;
    if (sinpsi > 1.0E-20) {
        sinpsi = sqrt(sinpsi);
        sindel = bb / sinpsi;
        cosdel = aa / sinpsi;
    }
    double ep, rejf, r1, r2;
    do {
        r1 = rnd(o);
        ep = ep0 * exp(r1 * wsamp);
        r2 = rnd(o);
        rejf = 1.0 - pow(ep * a - 1.0, 2.0) / (ep * (pow(a, 2.0) - 2.0));
    } while (r2 > rejf);
    double esg1 = avip * ep;
    o->e[np] = esg1;
    o->iq[np] = 0;
    double costhe = fmin(1.0, (esg1 - RM) * pot / esg1);
    double sinthe = sqrt(1.0 - pow(costhe, 2.0));
    double sphi, cphi, us, vs;
    azimuth(o, &cphi, &sphi);
    if (sinpsi >= 1.0E-10) {
        us = sinthe * cphi;
        vs = sinthe * sphi;
        o->u[np] = cc * cosdel * us - sindel * vs + aa * costhe;
        o->v[np] = cc * sindel * us + cosdel * vs + bb * costhe;
        o->w[np] = cc * costhe - sinpsi * us;
    } else {
        o->u[np] = sinthe * cphi;
        o->v[np] = sinthe * sphi;
        o->w[np] = cc * costhe;
    }
    np += 1;
    double esg2 = avip - esg1;
    o->e[np] = esg2;
    o->iq[np] = 0;
    transfer(o, np, np - 1);
    costhe = fmin(1.0, (esg2 - RM) * pot / esg2);
    sinthe = -sqrt(1.0 - pow(costhe, 2.0));
    if (sinpsi >= 1.0E-10) {
        us = sinthe * cphi;
        vs = sinthe * sphi;
        o->u[np] = cc * cosdel * us - sindel * vs + aa * costhe;
        o->v[np] = cc * sindel * us + cosdel * vs + bb * costhe;
        o->w[np] = cc * costhe - sinpsi * us;
    } else {
        o->u[np] = sinthe * cphi;
        o->v[np] = sinthe * sphi;
        o->w[np] = cc * costhe;
    }
    o->np = np;
    roulette_photons(o);
}

/* deposit the kinetic energy of an electron below cut-off (or E+m of a
 * discarded positron) and annihilate a stopped positron
 * (src/ompmc.c:4683-4705, 5079-5106) */
static void electron_stop(orc *o, int iq, double eie, int discard) {
    int np = o->np;
    double edep = (discard && iq > 0) ? o->e[np] + RM : o->e[np] - RM;
    ausgab(o, edep);
    if (iq > 0 && edep < eie) {
        orc_rannih(o);
        return;
    }
    o->np -= 1;
}

/* condensed-history electron step (src/ompmc.c:4666-5451).  The variables
 * keep the reference's scopes: total_de and total_tstep are re-declared in
 * every ustep iteration (so a step that does not reach the interaction site
 * is followed by a zero-length step that resamples the path to the next
 * interaction), and the function returns to shower() after every discrete
 * interaction, cut-off or discard. */
void orc_electron(orc *o) {
    const omc_tables *t = o->t;
    const omc_regions *R = o->r;
    int nmed = t->nmed;
    int np = o->np, irl = o->ir[np], imed = R->med[irl];
    double rhof = R->rhof[irl], edep = 0.0;
    uphi_t q;
    double eie = o->e[np];
    int iq = o->iq[np], qel = (1 + iq) / 2, medold = imed;
    double rnno;
    if (eie <= R->ecut[irl]) { electron_stop(o, iq, eie, 0); return; }
    double elke = 0.0;
    int lelke = 0;
    double sigratio = 0.0, rfict = 0.0;
    do {
        int compute_tstep = 1;
        double eke = eie - RM, demfp = 0.0, sigf, sig0 = 0.0, dedx0, ustep;
        if (imed != -1) {
            rnno = rnd(o);
            if (rnno == 0.0) rnno = 1.0E-30;
            demfp = fmax(-log(rnno), OMC_EPSEMFP);
            elke = log(eke);
            lelke = pwlf_i(imed, elke, t->eke1, t->eke0) - 1;
            int ix = imed * MXEKE + lelke;
            if (t->sig_ismonotone[qel * nmed + imed]) {
                if (iq < 0) {
                    sig0 = pwlf(ix, elke, t->esig1, t->esig0);
                    dedx0 = pwlf(ix, elke, t->ededx1, t->ededx0);
                    sig0 /= dedx0;
                } else {
                    sig0 = pwlf(ix, elke, t->psig1, t->psig0);
                    dedx0 = pwlf(ix, elke, t->pdedx1, t->pdedx0);
                    sig0 /= dedx0;
                }
            } else {
                sig0 = iq < 0 ? t->esig_e[imed] : t->psig_e[imed];
            }
        }
        do {
            int call_howfar = 0, do_single = 0, called_msdist = 0;
            double tstep, tustep, total_de = 0.0, ekef, ekei, elkei, tuss;
            double total_tstep = 0.0;
            double range = 0.0, p2, beta2, etap, tvstep = 0.0, de = 0.0;
            double x_final = 0, y_final = 0, z_final = 0, u_final = 0, v_final = 0, w_final = 0;
            const double *rep = t->range_ep + (size_t)qel * nmed * MXEKE + (size_t)(imed < 0 ? 0 : imed) * MXEKE;
            (void)total_de;
            if (imed == -1) {
                tstep = 10.0E8;
                ustep = tstep;
                tustep = ustep;
                call_howfar = 1;
            } else {
                rhof = R->rhof[irl];
                if (sig0 <= 0.0) {
                    tstep = 10.0E8;
                    sig0 = 1.0E-15;
                } else {
                    if (compute_tstep) {
                        total_de = demfp / sig0;
                        ekef = eke - total_de;
                        if (ekef <= t->e_array[imed * MXEKE + 0]) {
                            tstep = 10.0E8;
                        } else {
                            double elkef = log(ekef);
                            int lelkef = pwlf_i(imed, elkef, t->eke1, t->eke0) - 1;
                            if (lelkef == lelke) {
                                tstep = orc_drange(o, imed, iq, lelke, eke, ekef, elke, elkef);
                            } else {
                                ekei = t->e_array[imed * MXEKE + lelke];
                                elkei = (lelke + 1 - t->eke0[imed]) / t->eke1[imed];
                                tuss = orc_drange(o, imed, iq, lelke, eke, ekei, elke, elkei);
                                ekei = t->e_array[imed * MXEKE + lelkef + 1];
                                elkei = ((lelkef + 2) - t->eke0[imed]) / t->eke1[imed];
                                tstep = orc_drange(o, imed, iq, lelkef, ekei, ekef, elkei, elkef);
                                tstep += tuss + rep[lelke] - rep[lelkef + 1];
                            }
                        }
                        total_tstep = tstep;
                        compute_tstep = 0;
                    }
                    tstep = total_tstep / rhof;
                }
                int ix = imed * MXEKE + lelke;
                if (iq < 0) dedx0 = pwlf(ix, elke, t->ededx1, t->ededx0);
                else dedx0 = pwlf(ix, elke, t->pdedx1, t->pdedx0);
                double dedx = rhof * dedx0;
                double tmxs = pwlf(ix, elke, t->tmxs1, t->tmxs0);
                tmxs /= rhof;
                ekei = t->e_array[imed * MXEKE + lelke];
                elkei = (lelke + 1 - t->eke0[imed]) / t->eke1[imed];
                range = orc_drange(o, imed, iq, lelke, eke, ekei, elke, elkei);
                range += rep[lelke];
                range /= rhof;
                tustep = fmin(fmin(tstep, tmxs), range);
                double tperp = hownear(o);
                o->dnear[np] = tperp;
                double blccl = rhof * t->blcc[imed], xccl = rhof * t->xcc[imed];
                p2 = eke * (eke + 2.0 * RM);
                beta2 = p2 / (p2 + pow(RM, 2.0));
                if (iq < 0) etap = pwlf(MXEKE * imed + lelke, elke, t->etae_ms1, t->etae_ms0);
                else etap = pwlf(MXEKE * imed + lelke, elke, t->etap_ms1, t->etap_ms0);
                double ms_corr = pwlf(MXEKE * imed + lelke, elke, t->blcce1, t->blcce0);
                blccl = blccl / etap / (1.0 + 0.25 * etap * xccl / blccl / p2) * ms_corr;
                double ssmfp = beta2 / blccl;
                double skindepth = OMC_SKIN_DEPTH_FOR_BCA * ssmfp;
                tustep = fmin(tustep, fmax(tperp, skindepth));
                if ((tustep <= tperp) && (tustep > skindepth)) {
                    call_howfar = 0;
                    do_single = 0;
                    called_msdist = 1;
                    de = orc_eloss(o, imed, iq, irl, rhof, tustep, range, eke, elke, lelke);
                    tvstep = tustep;
                    ustep = orc_msdist(o, imed, iq, rhof, de, tustep, eke, &x_final, &y_final, &z_final, &u_final,
                                       &v_final, &w_final);
                } else {
                    called_msdist = 0;
                    rnno = rnd(o);
                    if (rnno < 1.0E-30) rnno = 1.0E-30;
                    double lambda = (-1.0) * log(1.0 - rnno);
                    double lambda_max = 0.5 * blccl * RM / dedx;
                    lambda_max *= (eke / RM + 1.0) * (eke / RM + 1.0) * (eke / RM + 1.0);
                    if (lambda >= 0.0 && lambda_max > 0.0) {
                        if (lambda < lambda_max) tuss = lambda * ssmfp * (1.0 - 0.5 * lambda / lambda_max);
                        else tuss = 0.5 * lambda * ssmfp;
                        if (tuss < tustep) {
                            tustep = tuss;
                            do_single = 1;
                        } else {
                            do_single = 0;
                        }
                    } else {
                        /* the reference warns and drops the particle (quirk Q6) */
                        o->lost += o->wt[np] * (eie + (iq > 0 ? RM : -RM));
                        o->np -= 1;
                        return;
                    }
                    ustep = tustep;
                    call_howfar = ustep < tperp ? 0 : 1;
                }
            }
            int irold = o->ir[np], irnew = o->ir[np], idisc = 0;
            if (call_howfar) howfar(o, &idisc, &irnew, &ustep);
            if (idisc > 0) { electron_stop(o, iq, eie, 1); return; }
            if (ustep < 0) ustep = 0.0;
            double vstep;
            if (ustep == 0.0f || imed == -1) {
                if (ustep != 0.0f) {
                    vstep = ustep;
                    tvstep = vstep;
                    o->x[np] += o->u[np] * vstep;
                    o->y[np] += o->v[np] * vstep;
                    o->z[np] += o->w[np] * vstep;
                    o->dnear[np] -= vstep;
                }
                if (irnew != irold) {
                    o->ir[np] = irnew;
                    irl = irnew;
                    imed = R->med[irl];
                }
                if (eie <= R->ecut[irl]) { electron_stop(o, iq, eie, 0); return; }
                break;
            }
            vstep = ustep;
            if (call_howfar) {
                tvstep = vstep;
                if (tvstep != tustep) do_single = 0;
                de = orc_eloss(o, imed, iq, irl, rhof, tvstep, range, eke, elke, lelke);
            } else {
                tvstep = tustep;
                if (called_msdist == 0) de = orc_eloss(o, imed, iq, irl, rhof, tvstep, range, eke, elke, lelke);
            }
            edep = de;
            ekef = eke - de;
            double sinthe = 0.0, costhe = 1.0;
            if (called_msdist == 0) {
                if (do_single) {
                    double ekems = fmax(ekef, R->ecut[irl] - RM);
                    p2 = ekems * (ekems + 2.0 * RM);
                    beta2 = p2 / (p2 + pow(RM, 2.0));
                    double chia2 = t->xcc[imed] / (4.0 * t->blcc[imed] * p2);
                    double elkems = log(ekems);
                    int lelkems = pwlf_i(imed, elkems, t->eke1, t->eke0) - 1;
                    if (iq < 0) etap = pwlf(MXEKE * imed + lelkems, elkems, t->etae_ms1, t->etae_ms0);
                    else etap = pwlf(MXEKE * imed + lelkems, elkems, t->etap_ms1, t->etap_ms0);
                    chia2 *= etap;
                    orc_sscat(o, imed, qel, chia2, elkems, beta2, &costhe, &sinthe);
                } else {
                    sinthe = 0.0f;
                    costhe = 1.0f;
                }
            }
            if (called_msdist == 0) {
                x_final = o->x[np] + o->u[np] * vstep;
                y_final = o->y[np] + o->v[np] * vstep;
                z_final = o->z[np] + o->w[np] * vstep;
                if (do_single) uphi21(o, &q, costhe, sinthe);
                u_final = o->u[np];
                v_final = o->v[np];
                w_final = o->w[np];
            }
            ausgab(o, edep);
            o->x[np] = x_final; o->y[np] = y_final; o->z[np] = z_final;
            o->u[np] = u_final; o->v[np] = v_final; o->w[np] = w_final;
            o->dnear[np] -= vstep;
            irold = o->ir[np];
            eie -= edep;
            o->e[np] = eie;
            if (irnew == irl && eie <= R->ecut[irl]) { electron_stop(o, iq, eie, 0); return; }
            medold = imed;
            if (imed != -1) {
                eke = eie - RM;
                elke = log(eke);
                lelke = pwlf_i(imed, elke, t->eke1, t->eke0) - 1;
            }
            if (irnew != irold) {
                o->ir[np] = irnew;
                irl = irnew;
                imed = R->med[irl];
            }
            if (eie <= R->ecut[irl]) { electron_stop(o, iq, eie, 0); return; }
            if (imed != medold) break;
            demfp -= de * sig0;
            total_de -= de;
            total_tstep -= tvstep * rhof;
            if (total_tstep < 1.0E-9) demfp = 0.0;
        } while (demfp >= OMC_EPSEMFP);
        if ((imed != medold) || (ustep == 0.0) || (imed == -1)) continue;
        int ix = imed * MXEKE + lelke;
        if (iq < 0) {
            sigf = pwlf(ix, elke, t->esig1, t->esig0);
            dedx0 = pwlf(ix, elke, t->ededx1, t->ededx0);
            sigf /= dedx0;
        } else {
            sigf = pwlf(ix, elke, t->psig1, t->psig0);
            dedx0 = pwlf(ix, elke, t->pdedx1, t->pdedx0);
            sigf /= dedx0;
        }
        sigratio = sigf / sig0;
        rfict = rnd(o);
    } while (rfict >= sigratio);
    int ix = imed * MXEKE + lelke;
    if (iq < 0) {
        double ebr1 = pwlf(ix, elke, t->ebr11, t->ebr10);
        rnno = rnd(o);
        if (rnno <= ebr1) {
            orc_brems(o);
        } else {
            if (o->e[np] <= t->thmoll[imed]) {
                if (ebr1 <= 0) return;
                orc_brems(o);
            } else {
                orc_moller(o);
            }
        }
    } else {
        double pbr1 = pwlf(ix, elke, t->pbr11, t->pbr10);
        rnno = rnd(o);
        if (rnno < pbr1) {
            orc_brems(o);
        } else {
            double pbr2 = pwlf(ix, elke, t->pbr21, t->pbr20);
            if (rnno < pbr2) orc_bhabha(o);
            else orc_annih(o);
        }
    }
}

void orc_shower(orc *o) {
    while (o->np >= 0 && !o->err) {
        if (o->iq[o->np] == 0) orc_photon(o);
        else orc_electron(o);
    }
}

/* ------------------------------------------------------------------ sources */
/* omc_dosxyz initHistory (ucodes/omc_dosxyz/omc_dosxyz.c:964-1068) */
void orc_init_history_dosxyz(orc *o, const omc_dosxyz_source *s) {
    const orc_geom *g = o->g;
    double r1, r2;
    o->np = 0;
    int np = 0;
    o->iq[np] = s->charge;
    double ein;
    if (s->spec.spectrum) {
        r1 = rnd(o);
        r2 = rnd(o);
        int k = (int)fmin(s->spec.deltak * r1, s->spec.deltak - 1.0);
        ein = s->spec.cdfinv1[k] + r2 * s->spec.cdfinv2[k];
    } else {
        ein = s->spec.energy;
    }
    o->e[np] = o->iq[np] != 0 ? ein + RM : ein;
    o->ensrc += ein;
    double rxyz = 0.0;
    if (s->xsize == 0.0 || s->ysize == 0.0) {
        o->x[np] = s->xinl;
        o->y[np] = s->yinl;
        rxyz = sqrt(pow(s->ssd, 2.0) + pow(o->x[np], 2.0) + pow(o->y[np], 2.0));
        o->w[np] = s->ssd / rxyz;
    } else {
        double fw, r3;
        do {
            r3 = rnd(o);
            o->x[np] = r3 * s->xsize + s->xinl;
            r3 = rnd(o);
            o->y[np] = r3 * s->ysize + s->yinl;
            r3 = rnd(o);
            rxyz = sqrt(s->ssd * s->ssd + o->x[np] * o->x[np] + o->y[np] * o->y[np]);
            o->w[np] = s->ssd / rxyz;
            fw = o->w[np] * o->w[np] * o->w[np];
        } while (r3 >= fw);
    }
    o->z[np] = g->zb[0];
    o->u[np] = o->x[np] / rxyz;
    o->v[np] = o->y[np] / rxyz;
    int ix, iy;
    if (s->xsize == 0.0) {
        ix = s->ixinl;
    } else {
        ix = s->ixinl - 1;
        while ((g->xb[ix + 1] < o->x[np]) && ix < g->nx - 1) ix++;
    }
    if (s->ysize == 0.0) {
        iy = s->iyinl;
    } else {
        iy = s->iyinl - 1;
        while ((g->yb[iy + 1] < o->y[np]) && iy < g->ny - 1) iy++;
    }
    o->ir[np] = 1 + ix + iy * g->nx;
    o->wt[np] = 1.0;
    o->dnear[np] = 0.0;
}

/* Box-Muller pair (src/omc_random.c:215-224) */
static void box_muller(orc *o, double rn[2]) {
    rn[0] = rnd(o);
    rn[1] = rnd(o);
    double R = sqrt(-2.0 * log(rn[0]));
    double th = 2.0 * M_PI * rn[1];
    rn[0] = R * cos(th);
    rn[1] = R * sin(th);
}

/* omc_matrad initHistory (ucodes/omc_matrad/omc_matrad.c:1177-1413).  The
 * source is always a photon spectrum (quirks Q13, Q14); the z clamp below the
 * phantom uses ybounds[0] (quirk Q12). */
void orc_init_history_matrad(orc *o, const omc_spectrum *spec, const orc_beamlets *b, int ib) {
    const orc_geom *g = o->g;
    int ijmax = g->nx * g->ny, imax = g->nx;
    double r1, r2;
    o->np = 0;
    int np = 0;
    o->iq[np] = 0;
    double ein;
    if (spec->spectrum) {
        r1 = rnd(o);
        r2 = rnd(o);
        int k = (int)fmin(spec->deltak * r1, spec->deltak - 1.0);
        ein = spec->cdfinv1[k] + r2 * spec->cdfinv2[k];
    } else {
        ein = spec->energy;
    }
    o->e[np] = ein;
    o->ensrc += ein;
    r1 = rnd(o);
    r2 = rnd(o);
    const double *s1 = b->side1 + 3 * ib, *s2 = b->side2 + 3 * ib, *c = b->corner + 3 * ib;
    double xiso = r1 * s1[0] + r2 * s2[0] + c[0];
    double yiso = r1 * s1[1] + r2 * s2[1] + c[1];
    double ziso = r1 * s1[2] + r2 * s2[2] + c[2];
    int ibeam = b->ibeam[ib];
    double sp[3];
    if (!b->gaussian) {
        sp[0] = b->src[3 * ibeam]; sp[1] = b->src[3 * ibeam + 1]; sp[2] = b->src[3 * ibeam + 2];
    } else {
        double n1 = sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
        double n2 = sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
        double p1[3] = {s1[0] / n1, s1[1] / n1, s1[2] / n1}, p2[3] = {s2[0] / n2, s2[1] / n2, s2[2] / n2};
        double rn[2];
        box_muller(o, rn);
        rn[0] *= b->sigma;
        rn[1] *= b->sigma;
        for (int k = 0; k < 3; k++) sp[k] = b->src[3 * ibeam + k] + rn[0] * p1[k] + rn[1] * p2[k];
    }
    double xd = xiso - sp[0], yd = yiso - sp[1], zd = ziso - sp[2];
    double vnorm = sqrt(xd * xd + yd * yd + zd * zd);
    double u = -(xd) / vnorm, v = -(yd) / vnorm, w = -(zd) / vnorm;
    double ustep = DBL_MAX, dist;
    if (u > 0.0) { dist = (g->xb[g->nx] - xiso) / u; if (dist < ustep) ustep = dist; }
    if (u < 0.0) { dist = -(xiso - g->xb[0]) / u; if (dist < ustep) ustep = dist; }
    if (v > 0.0) { dist = (g->yb[g->ny] - yiso) / v; if (dist < ustep) ustep = dist; }
    if (v < 0.0) { dist = -(yiso - g->yb[0]) / v; if (dist < ustep) ustep = dist; }
    if (w > 0.0) { dist = (g->zb[g->nz] - ziso) / w; if (dist < ustep) ustep = dist; }
    if (w < 0.0) { dist = -(ziso - g->zb[0]) / w; if (dist < ustep) ustep = dist; }
    o->x[np] = xiso + ustep * u;
    o->y[np] = yiso + ustep * v;
    o->z[np] = ziso + ustep * w;
    o->u[np] = -u; o->v[np] = -v; o->w[np] = -w;
    if (o->x[np] < g->xb[0]) o->x[np] = g->xb[0] + 2.0 * DBL_MIN;
    if (o->x[np] > g->xb[g->nx]) o->x[np] = g->xb[g->nx] - 2.0 * DBL_MIN;
    if (o->y[np] < g->yb[0]) o->y[np] = g->yb[0] + 2.0 * DBL_MIN;
    if (o->y[np] > g->yb[g->ny]) o->y[np] = g->yb[g->ny] - 2.0 * DBL_MIN;
    if (o->z[np] < g->zb[0]) o->z[np] = g->yb[0] + 2.0 * DBL_MIN;
    if (o->z[np] > g->zb[g->nz]) o->z[np] = g->zb[g->nz] - 2.0 * DBL_MIN;
    int ix = 0, iy = 0, iz = 0;
    while (ix < g->nx - 1 && g->xb[ix + 1] < o->x[np]) ix++;
    while (iy < g->ny - 1 && g->yb[iy + 1] < o->y[np]) iy++;
    while (iz < g->nz - 1 && g->zb[iz + 1] < o->z[np]) iz++;
    o->ir[np] = 1 + ix + iy * imax + iz * ijmax;
    o->wt[np] = 1.0;
    o->dnear[np] = 0.0;
}

/* ------------------------------------------------------------------ drivers */
/* Energy balance of a run (not in the reference): ebal[0] the primaries'
 * energy (photons E, electrons E - m, positrons E + m), [1] photons that
 * left the phantom, [2] energy dropped without a deposit (quirk Q6), [3]
 * energy deposited in all regions including region 0. */
static double primary_energy(const orc *o) {
    int np = o->np;
    return o->wt[np] * (o->iq[np] == 0 ? o->e[np] : o->e[np] + (o->iq[np] > 0 ? RM : -RM));
}

int orc_run_dosxyz(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_dosxyz_source *s,
                   int nsplit, int ixx, int jxx, int nperbatch, int nbatch, double *sum, double *sum2,
                   int *rn_per_history, double *ebal) {
    orc *o = orc_create(t, r, g, nsplit, ixx, jxx);
    if (!o) return OMC_ERR_NOMEM;
    long h = 0;
    double src = 0.0, dep = 0.0;
    for (int b = 0; b < nbatch; b++) {
        for (int i = 0; i < nperbatch; i++) {
            long before = o->count;
            orc_init_history_dosxyz(o, s);
            src += primary_energy(o);
            orc_shower(o);
            if (rn_per_history) rn_per_history[h] = (int)(o->count - before);
            h++;
            if (o->err) { orc_destroy(o); return OMC_ERR_STACK; }
        }
        for (int i = 0; i < r->nreg; i++) {
            double e = o->endep[i];
            dep += e;
            sum[i] += e;
            sum2[i] += e * e;
        }
        memset(o->endep, 0, sizeof(double) * r->nreg);
    }
    if (ebal) {
        ebal[0] = src;
        ebal[1] = o->esc;
        ebal[2] = o->lost;
        ebal[3] = dep;
    }
    orc_destroy(o);
    return OMC_OK;
}

double orc_run_dosxyz_omp(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_dosxyz_source *s,
                          int nsplit, int ixx, int jxx, int nperbatch, int nbatch, int nthreads, double *sum,
                          double *sum2) {
    double *endep = calloc(r->nreg, sizeof(double));
    double t0 = 0.0, t1 = 0.0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_num_procs();
    orc **ctx = calloc(nthreads, sizeof(orc *));
    for (int k = 0; k < nthreads; k++) ctx[k] = orc_create(t, r, g, nsplit, ixx, jxx + k);
    t0 = omp_get_wtime();
    for (int b = 0; b < nbatch; b++) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int i = 0; i < nperbatch; i++) {
            orc *o = ctx[omp_get_thread_num()];
            orc_init_history_dosxyz(o, s);
            orc_shower(o);
        }
        for (int k = 0; k < nthreads; k++) {
            for (int i = 0; i < r->nreg; i++) endep[i] += ctx[k]->endep[i];
            memset(ctx[k]->endep, 0, sizeof(double) * r->nreg);
        }
        for (int i = 0; i < r->nreg; i++) {
            sum[i] += endep[i];
            sum2[i] += endep[i] * endep[i];
            endep[i] = 0.0;
        }
    }
    t1 = omp_get_wtime();
    for (int k = 0; k < nthreads; k++) orc_destroy(ctx[k]);
    free(ctx);
#else
    (void)nthreads;
#endif
    free(endep);
    return t1 - t0;
}

int orc_run_matrad_beamlet(orc *o, const omc_spectrum *spec, const orc_beamlets *b, int ib, int nperbatch,
                           int nbatch, double *sum, double *sum2) {
    for (int bt = 0; bt < nbatch; bt++) {
        for (int i = 0; i < nperbatch; i++) {
            orc_init_history_matrad(o, spec, b, ib);
            orc_shower(o);
            if (o->err) return OMC_ERR_STACK;
        }
        for (int i = 0; i < o->r->nreg; i++) {
            double e = o->endep[i];
            e /= (double)nperbatch;
            sum[i] += e;
            sum2[i] += e * e;
        }
        memset(o->endep, 0, sizeof(double) * o->r->nreg);
    }
    return OMC_OK;
}

/* Serial matRad Dij transport in the reference's loop order
 * (omc_matrad.c:1576-1600): beamlet by beamlet, batch by batch, nper[b]
 * histories per batch on one RANMAR stream.  sum/sum2 + b*nreg receive
 * beamlet b's per-batch sums of endep/nper[b] and of its square (accumEndep,
 * :897-918); rn_per_history (optional) the draws of each history; batches
 * (optional, [nbix][nbatch][nreg]) each batch's energy tally before it is
 * folded in. */
int orc_run_matrad(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_spectrum *spec,
                   const orc_beamlets *b, int nsplit, int ixx, int jxx, const int *nper, int nbatch, double *sum,
                   double *sum2, int *rn_per_history, double *batches) {
    orc *o = orc_create(t, r, g, nsplit, ixx, jxx);
    if (!o) return OMC_ERR_NOMEM;
    long h = 0;
    int nreg = r->nreg;
    for (int ib = 0; ib < b->n; ib++) {
        double *sb = sum + (long)ib * nreg, *s2b = sum2 + (long)ib * nreg;
        for (int bt = 0; bt < nbatch; bt++) {
            for (int i = 0; i < nper[ib]; i++) {
                long before = o->count;
                orc_init_history_matrad(o, spec, b, ib);
                orc_shower(o);
                if (rn_per_history) rn_per_history[h] = (int)(o->count - before);
                h++;
                if (o->err) { orc_destroy(o); return OMC_ERR_STACK; }
            }
            if (batches) memcpy(batches + ((long)ib * nbatch + bt) * nreg, o->endep, sizeof(double) * nreg);
            for (int i = 0; i < nreg; i++) {
                double e = o->endep[i];
                e /= (double)nper[ib];
                sb[i] += e;
                s2b[i] += e * e;
            }
            memset(o->endep, 0, sizeof(double) * nreg);
        }
    }
    orc_destroy(o);
    return OMC_OK;
}

/* OpenMP matRad Dij transport as the reference runs it (omc_matrad.c:1576-1600):
 * for each beamlet and batch, the histories of the batch are shared by the
 * team (thread k seeds (ixx, jxx+k)), then the batch sums are folded in.
 * Thread-private tallies replace the reference's atomic updates.  Used for
 * CPU timing only; returns the transport seconds. */
double orc_run_matrad_omp(const omc_tables *t, const omc_regions *r, const orc_geom *g, const omc_spectrum *spec,
                          const orc_beamlets *b, int nsplit, int ixx, int jxx, const int *nper, int nbatch,
                          int nthreads, double *sum, double *sum2) {
    double t0 = 0.0, t1 = 0.0;
#ifdef _OPENMP
    int nreg = r->nreg;
    if (nthreads < 1) nthreads = omp_get_num_procs();
    orc **ctx = calloc(nthreads, sizeof(orc *));
    for (int k = 0; k < nthreads; k++) ctx[k] = orc_create(t, r, g, nsplit, ixx, jxx + k);
    t0 = omp_get_wtime();
    for (int ib = 0; ib < b->n; ib++) {
        double *sb = sum + (long)ib * nreg, *s2b = sum2 + (long)ib * nreg;
        for (int bt = 0; bt < nbatch; bt++) {
#pragma omp parallel for schedule(guided) num_threads(nthreads)
            for (int i = 0; i < nper[ib]; i++) {
                orc *o = ctx[omp_get_thread_num()];
                orc_init_history_matrad(o, spec, b, ib);
                orc_shower(o);
            }
#pragma omp parallel for num_threads(nthreads)
            for (int i = 0; i < nreg; i++) {
                double e = 0.0;
                for (int k = 0; k < nthreads; k++) {
                    e += ctx[k]->endep[i];
                    ctx[k]->endep[i] = 0.0;
                }
                e /= (double)nper[ib];
                sb[i] += e;
                s2b[i] += e * e;
            }
        }
    }
    t1 = omp_get_wtime();
    for (int k = 0; k < nthreads; k++) orc_destroy(ctx[k]);
    free(ctx);
#else
    (void)t; (void)r; (void)g; (void)spec; (void)b; (void)nsplit; (void)ixx; (void)jxx; (void)nper; (void)nbatch;
    (void)nthreads; (void)sum; (void)sum2;
#endif
    return t1 - t0;
}
