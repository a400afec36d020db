/*
 * Golden-vector harness around the UNMODIFIED reference sources.
 *
 * Test infrastructure only: this file is compiled together with the reference's
 * src/ompmc.c, src/omc_random.c, src/omc_utilities.c and
 * ucodes/omc_dosxyz/omc_dosxyz.c (its main renamed by -Dmain=) by
 * oracle/Makefile, into oracle/_ref/ref_harness.  Nothing in the product links
 * it.  It uses the link seam of GNU ld's --wrap on the engine's calls to
 * setRandom, ausgab, howfar and hownear: every random number the engine draws
 * and every tally it makes is logged (or, in injection mode, the random numbers
 * come from a caller-supplied list).
 *
 * Modes (argv[1]):
 *   tables   <inp-stem> <out>            dump every physics table, binary
 *   samplers <inp-stem> <out> <ncase>    per-sampler cases: inputs, RNs, outputs
 *   dose     <inp-stem> <out> <nhist>    dosxyz batch loop, serial: per-history
 *                                        RN counts, raw tallies, final dose
 *
 * Output format ("OMCG" records): see write_arr()/case_*() below; the Python
 * reader is tests/golden/golden_io.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ompmc.h"
#include "omc_random.h"
#include "omc_utilities.h"

/* ---- reference symbols from omc_dosxyz.c (non-static there) ---- */
struct Geom {
    int *med_indices; double *med_densities;
    int isize, jsize, ksize;
    double *xbounds, *ybounds, *zbounds;
};
extern struct Geom geometry;
struct Score { double ensrc; double *endep; double *accum_endep; double *accum_endep2; };
extern struct Score score;
void initPhantom(void);
void initSource(void);
void initScore(void);
void initHistory(void);
void accumEndep(void);
void accumulateResults(int iout, int nhist, int nbatch);

/* ---- link seam ---- */
double __real_setRandom(void);
void __real_ausgab(double edep);
void __real_howfar(int *idisc, int *irnew, double *ustep);
double __real_hownear(void);

static int g_inject = 0;          /* 1: return injected values */
static double *g_inj = NULL; static long g_ninj = 0, g_iinj = 0;
static double *g_log = NULL; static long g_nlog = 0, g_caplog = 0;
static long g_nrn_total = 0;
static int g_log_rn = 0;
static int g_log_tally = 0;
static double g_tally[4 * 4096]; static int g_ntally = 0;
static long g_nhowfar = 0, g_nhownear = 0;

double __wrap_setRandom(void) {
    double v;
    if (g_inject) {
        if (g_iinj >= g_ninj) { fprintf(stderr, "harness: injected RNs exhausted\n"); exit(3); }
        v = g_inj[g_iinj++];
    } else {
        v = __real_setRandom();
    }
    g_nrn_total++;
    if (g_log_rn) {
        if (g_nlog == g_caplog) {
            g_caplog = g_caplog ? 2 * g_caplog : 4096;
            g_log = realloc(g_log, g_caplog * sizeof(double));
        }
        g_log[g_nlog++] = v;
    }
    return v;
}

void __wrap_ausgab(double edep) {
    if (g_log_tally && g_ntally < 4096) {
        g_tally[4 * g_ntally + 0] = (double)stack.ir[stack.np];
        g_tally[4 * g_ntally + 1] = edep;
        g_tally[4 * g_ntally + 2] = stack.wt[stack.np];
        g_tally[4 * g_ntally + 3] = (double)stack.iq[stack.np];
        g_ntally++;
    }
    __real_ausgab(edep);
}

void __wrap_howfar(int *idisc, int *irnew, double *ustep) { g_nhowfar++; __real_howfar(idisc, irnew, ustep); }
double __wrap_hownear(void) { g_nhownear++; return __real_hownear(); }

/* ---- output ---- */
static FILE *g_out;

static void write_arr(const char *name, int dtype, long n, const void *p) {
    char nm[32]; memset(nm, 0, 32); strncpy(nm, name, 31);
    fwrite("OMCA", 1, 4, g_out); fwrite(nm, 1, 32, g_out);
    int32_t dt = dtype; int64_t nn = n;
    fwrite(&dt, 4, 1, g_out); fwrite(&nn, 8, 1, g_out);
    fwrite(p, dtype == 0 ? 8 : 4, n, g_out);
}
#define WD(name, ptr, n) write_arr(name, 0, (long)(n), ptr)
#define WI(name, ptr, n) write_arr(name, 1, (long)(n), ptr)

/* deterministic input generator (independent of RANMAR) */
static uint64_t g_s = 0x9E3779B97F4A7C15ull;
static double urand(void) {
    g_s ^= g_s << 13; g_s ^= g_s >> 7; g_s ^= g_s << 17;
    return (double)(g_s >> 11) * (1.0 / 9007199254740992.0);
}
static double logu(double a, double b) { return exp(log(a) + urand() * (log(b) - log(a))); }

static void rand_dir(double *u, double *v, double *w) {
    double ct = 2.0 * urand() - 1.0, ph = 2.0 * M_PI * urand();
    double st = sqrt(1.0 - ct * ct);
    *u = st * cos(ph); *v = st * sin(ph); *w = ct;
    if (urand() < 0.1) { *u = 0.0; *v = 0.0; *w = urand() < 0.5 ? 1.0 : -1.0; }
}

/* particle <-> doubles: iq ir e x y z u v w wt dnear */
static void put_particle(int i, const double *p) {
    stack.iq[i] = (int)p[0]; stack.ir[i] = (int)p[1]; stack.e[i] = p[2];
    stack.x[i] = p[3]; stack.y[i] = p[4]; stack.z[i] = p[5];
    stack.u[i] = p[6]; stack.v[i] = p[7]; stack.w[i] = p[8];
    stack.wt[i] = p[9]; stack.dnear[i] = p[10];
}
static void get_particle(int i, double *p) {
    p[0] = stack.iq[i]; p[1] = stack.ir[i]; p[2] = stack.e[i];
    p[3] = stack.x[i]; p[4] = stack.y[i]; p[5] = stack.z[i];
    p[6] = stack.u[i]; p[7] = stack.v[i]; p[8] = stack.w[i];
    p[9] = stack.wt[i]; p[10] = stack.dnear[i];
}

/* random particle inside voxel ir (1-based) of the loaded phantom */
static void rand_particle(double *p, int iq, double e, int ir) {
    int nx = geometry.isize, ny = geometry.jsize;
    int ix = (ir - 1) % nx, iy = ((ir - 1) / nx) % ny, iz = (ir - 1) / (nx * ny);
    p[0] = iq; p[1] = ir; p[2] = e;
    p[3] = geometry.xbounds[ix] + urand() * (geometry.xbounds[ix + 1] - geometry.xbounds[ix]);
    p[4] = geometry.ybounds[iy] + urand() * (geometry.ybounds[iy + 1] - geometry.ybounds[iy]);
    p[5] = geometry.zbounds[iz] + urand() * (geometry.zbounds[iz + 1] - geometry.zbounds[iz]);
    rand_dir(&p[6], &p[7], &p[8]);
    p[9] = 1.0; p[10] = 0.0;
}

/* one case record: "OMCC", kind, n_in, n_rn, n_out, then doubles */
static void case_begin(void) { g_nlog = 0; g_ntally = 0; g_log_rn = 1; g_log_tally = 1; g_nhowfar = g_nhownear = 0; }
static void case_write(int kind, const double *in, int nin, const double *out, int nout) {
    int32_t h[4] = {kind, nin, (int32_t)g_nlog, nout};
    fwrite("OMCC", 1, 4, g_out); fwrite(h, 4, 4, g_out);
    fwrite(in, 8, nin, g_out); fwrite(g_log, 8, g_nlog, g_out); fwrite(out, 8, nout, g_out);
    g_log_rn = 0; g_log_tally = 0;
}

static void init_all(const char *stem) {
    parseInputFile((char *)stem);
    initPhantom();
    initMediaData();
    initSource();
    initRegions();
    initVrt();
    initScore();
    initRandom();
    initStack();
}

static int nreg(void) { return geometry.isize * geometry.jsize * geometry.ksize; }

static void dump_tables(void) {
    int nmed = media.nmed, nm = nmed;
    WD("ge0", photon_data.ge0, nm); WD("ge1", photon_data.ge1, nm);
    WD("gmfp0", photon_data.gmfp0, nm * MXGE); WD("gmfp1", photon_data.gmfp1, nm * MXGE);
    WD("gbr10", photon_data.gbr10, nm * MXGE); WD("gbr11", photon_data.gbr11, nm * MXGE);
    WD("gbr20", photon_data.gbr20, nm * MXGE); WD("gbr21", photon_data.gbr21, nm * MXGE);
    WD("cohe0", photon_data.cohe0, nm * MXGE); WD("cohe1", photon_data.cohe1, nm * MXGE);
    WD("ray_xgrid", rayleigh_data.xgrid, nm * MXRAYFF); WD("ray_fcum", rayleigh_data.fcum, nm * MXRAYFF);
    WD("ray_b", rayleigh_data.b_array, nm * MXRAYFF); WD("ray_c", rayleigh_data.c_array, nm * MXRAYFF);
    WD("ray_pmax0", rayleigh_data.pmax0, nm * MXGE); WD("ray_pmax1", rayleigh_data.pmax1, nm * MXGE);
    WI("ray_iarr", rayleigh_data.i_array, nm * RAYCDFSIZE);
    WD("dl1", pair_data.dl1, nm * 8); WD("dl2", pair_data.dl2, nm * 8); WD("dl3", pair_data.dl3, nm * 8);
    WD("dl4", pair_data.dl4, nm * 8); WD("dl5", pair_data.dl5, nm * 8); WD("dl6", pair_data.dl6, nm * 8);
    WD("bpar0", pair_data.bpar0, nm); WD("bpar1", pair_data.bpar1, nm);
    WD("pair_delcm", pair_data.delcm, nm); WD("zbrang", pair_data.zbrang, nm);
#define EL(n) WD(#n, electron_data.n, nm * MXEKE)
    EL(esig0); EL(esig1); EL(psig0); EL(psig1); EL(ededx0); EL(ededx1); EL(pdedx0); EL(pdedx1);
    EL(ebr10); EL(ebr11); EL(pbr10); EL(pbr11); EL(pbr20); EL(pbr21); EL(tmxs0); EL(tmxs1);
    EL(blcce0); EL(blcce1); EL(etae_ms0); EL(etae_ms1); EL(etap_ms0); EL(etap_ms1);
    EL(q1ce_ms0); EL(q1ce_ms1); EL(q1cp_ms0); EL(q1cp_ms1); EL(q2ce_ms0); EL(q2ce_ms1);
    EL(q2cp_ms0); EL(q2cp_ms1); EL(e_array);
#undef EL
    WD("range_ep", electron_data.range_ep, 2 * nm * MXEKE);
    WD("eke0", electron_data.eke0, nm); WD("eke1", electron_data.eke1, nm);
    WI("sig_ismonotone", electron_data.sig_ismonotone, 2 * nm);
    WD("esig_e", electron_data.esig_e, nm); WD("psig_e", electron_data.psig_e, nm);
    WD("xcc", electron_data.xcc, nm); WD("blcc", electron_data.blcc, nm);
    WD("expeke1", electron_data.expeke1, nm);
    double sp[5] = {spin_data.b2spin_min, spin_data.dbeta2i, spin_data.espml, spin_data.dleneri, spin_data.dqq1i};
    WD("spin_scalars", sp, 5);
    WD("spin_rej", spin_data.spin_rej, (long)nm * 2 * (MXE_SPIN1 + 1) * (MXQ_SPIN + 1) * (MXU_SPIN + 1));
    long nms = (MXL_MS + 1) * (MXQ_MS + 1) * (MXU_MS + 1);
    WD("ums", mscat_data.ums_array, nms); WD("fms", mscat_data.fms_array, nms);
    WD("wms", mscat_data.wms_array, nms); WI("ims", mscat_data.ims_array, nms);
    double ms[2] = {mscat_data.dllambi, mscat_data.dqmsi};
    WD("mscat_scalars", ms, 2);
    double pg[12 * MXMED]; int pgi[8 * MXMED];
    for (int i = 0; i < nmed; i++) {
        double *q = pg + 12 * i; int *qi = pgi + 8 * i;
        q[0] = pegs_data.rho[i]; q[1] = pegs_data.rlc[i]; q[2] = pegs_data.ae[i]; q[3] = pegs_data.ap[i];
        q[4] = pegs_data.ue[i]; q[5] = pegs_data.up[i]; q[6] = pegs_data.te[i]; q[7] = pegs_data.thmoll[i];
        q[8] = pegs_data.delcm[i]; q[9] = 0; q[10] = 0; q[11] = 0;
        qi[0] = pegs_data.ne[i]; qi[1] = pegs_data.meke[i]; qi[2] = pegs_data.mge[i]; qi[3] = pegs_data.msge[i];
        qi[4] = pegs_data.iunrst[i]; qi[5] = pegs_data.epstfl[i]; qi[6] = pegs_data.iaprim[i]; qi[7] = 0;
    }
    WD("pegs_f", pg, 12 * nmed); WI("pegs_i", pgi, 8 * nmed);
    double el[MXMED * MXELEMENT * 4]; int nel = 0;
    for (int i = 0; i < nmed; i++)
        for (int j = 0; j < pegs_data.ne[i]; j++) {
            el[4 * nel + 0] = pegs_data.elements[i][j].z; el[4 * nel + 1] = pegs_data.elements[i][j].wa;
            el[4 * nel + 2] = pegs_data.elements[i][j].pz; el[4 * nel + 3] = pegs_data.elements[i][j].rhoz;
            nel++;
        }
    WD("elements", el, 4 * nel);
    int n = nreg() + 1;
    WI("region_med", region.med, n); WD("region_rhof", region.rhof, n);
    WD("region_pcut", region.pcut, n); WD("region_ecut", region.ecut, n);
}

enum { K_COMPTON = 1, K_PAIR, K_PHOTO, K_RAYLEIGH, K_BREMS, K_MOLLER, K_BHABHA, K_ANNIH,
       K_RANNIH, K_MSDIST, K_MSCAT, K_SSCAT, K_DRANGE, K_ELOSS, K_PHOTON, K_ELECTRON, K_SHOWER };

/* output: np, then particles 0..np (11 doubles each), then tally log */
static int pack_stack(double *out) {
    int k = 0; out[k++] = stack.np;
    for (int i = 0; i <= stack.np && i < 16; i++) { get_particle(i, out + k); k += 11; }
    return k;
}
static int pack_tally(double *out) {
    int k = 0; out[k++] = g_ntally;
    for (int i = 0; i < 4 * g_ntally; i++) out[k++] = g_tally[i];
    out[k++] = (double)g_nhowfar; out[k++] = (double)g_nhownear;
    return k;
}

static int region_of_med(int imed) {
    for (int i = 1; i <= nreg(); i++) if (region.med[i] == imed) return i;
    return 1;
}

static void run_samplers(int ncase) {
    static double in[64], out[16 * 11 + 4 * 4096 + 16];
    int nmed = media.nmed;
    for (int c = 0; c < ncase; c++) {
        for (int kind = K_COMPTON; kind <= K_ELECTRON; kind++) {
            int imed = c % nmed, ir = region_of_med(imed);
            int k = 0, nin = 0;
            double p[11];
            stack.np = 0; stack.npold = 0;
            switch (kind) {
            case K_COMPTON: case K_PHOTO: {
                double e = kind == K_COMPTON ? logu(0.005, 30.0) : logu(0.001, 2.0);
                rand_particle(p, 0, e, ir); put_particle(0, p);
                memcpy(in, p, sizeof p); nin = 11;
                case_begin();
                if (kind == K_COMPTON) compton(); else photo();
                k = pack_stack(out); break;
            }
            case K_PAIR: {
                double e = c % 5 == 4 ? logu(50.0, 120.0) : logu(1.03, 50.0);
                rand_particle(p, 0, e, ir); put_particle(0, p);
                memcpy(in, p, sizeof p); in[11] = imed; nin = 12;
                case_begin(); pair(imed); k = pack_stack(out); break;
            }
            case K_RAYLEIGH: {
                double e = logu(0.01, 5.0);
                rand_particle(p, 0, e, ir); put_particle(0, p);
                double gle = log(e);
                int lgle = pwlfInterval(imed, gle, photon_data.ge1, photon_data.ge0) - 1;
                memcpy(in, p, sizeof p); in[11] = imed; in[12] = gle; in[13] = lgle; nin = 14;
                case_begin(); rayleigh(imed, e, gle, lgle); k = pack_stack(out); break;
            }
            case K_BREMS: case K_MOLLER: case K_BHABHA: case K_ANNIH: case K_RANNIH: {
                int iq = (kind == K_BHABHA || kind == K_ANNIH || kind == K_RANNIH) ? 1 : -1;
                double ek = kind == K_MOLLER ? logu(2.0 * pegs_data.te[imed] * 1.001, 25.0) :
                            kind == K_BREMS ? (c % 7 == 6 ? logu(50.0, 80.0) : logu(pegs_data.ap[imed] * 1.5, 30.0)) :
                            kind == K_RANNIH ? logu(0.001, 0.3) : logu(0.3, 30.0);
                rand_particle(p, iq, ek + RM, ir); put_particle(0, p);
                memcpy(in, p, sizeof p); nin = 11;
                case_begin();
                if (kind == K_BREMS) brems(); else if (kind == K_MOLLER) moller();
                else if (kind == K_BHABHA) bhabha(); else if (kind == K_ANNIH) annih(); else rannih();
                k = pack_stack(out); break;
            }
            case K_MSDIST: {
                int iq = (c & 1) ? 1 : -1;
                double eke = logu(0.3, 20.0);
                double rhof = region.rhof[ir];
                double tustep = logu(1e-3, 0.5) / rhof;
                double de = eke * urand() * 0.2;
                rand_particle(p, iq, eke + RM, ir); put_particle(0, p);
                in[0] = imed; in[1] = iq; in[2] = rhof; in[3] = de; in[4] = tustep; in[5] = eke;
                memcpy(in + 6, p, sizeof p); nin = 17;
                case_begin();
                double xf, yf, zf, uf, vf, wf;
                double us = msdist(imed, iq, rhof, de, tustep, eke, &xf, &yf, &zf, &uf, &vf, &wf);
                out[0] = us; out[1] = xf; out[2] = yf; out[3] = zf; out[4] = uf; out[5] = vf; out[6] = wf;
                k = 7; break;
            }
            case K_MSCAT: {
                int qel = c & 1;
                double eke = logu(0.3, 20.0), elke = log(eke);
                double p2 = eke * (eke + 2.0 * RM), beta2 = p2 / (p2 + RM * RM);
                double chia2 = electron_data.xcc[imed] / (4.0 * p2 * electron_data.blcc[imed]);
                int branch = c % 4;
                double lambda = branch == 0 ? logu(0.05, 1.0) : branch == 1 ? logu(1.0, 13.8) : logu(13.8, 5e4);
                double q1 = 2.0 * chia2 * (log(1.0 + 1.0 / chia2) * (1.0 + chia2) - 1.0) * lambda;
                if (branch == 3) q1 = logu(1e-4, 0.7);
                in[0] = imed; in[1] = qel; in[2] = elke; in[3] = beta2; in[4] = q1; in[5] = lambda; in[6] = chia2;
                nin = 7;
                case_begin();
                int spin_index = 1, find_index = 1; struct Mscats ms; struct Spinr sr;
                double cost = 7.0, sint = 7.0;
                mscat(imed, qel, &spin_index, &find_index, elke, beta2, q1, lambda, chia2, &cost, &sint, &ms, &sr);
                out[0] = cost; out[1] = sint; k = 2; break;
            }
            case K_SSCAT: {
                int qel = c & 1;
                double eke = logu(0.2, 20.0), elke = log(eke);
                double p2 = eke * (eke + 2.0 * RM), beta2 = p2 / (p2 + RM * RM);
                double chia2 = electron_data.xcc[imed] / (4.0 * p2 * electron_data.blcc[imed]) * logu(0.8, 1.3);
                in[0] = imed; in[1] = qel; in[2] = chia2; in[3] = elke; in[4] = beta2; nin = 5;
                case_begin();
                double cost, sint;
                sscat(imed, qel, chia2, elke, beta2, &cost, &sint);
                out[0] = cost; out[1] = sint; k = 2; break;
            }
            case K_DRANGE: case K_ELOSS: {
                int iq = (c & 1) ? 1 : -1;
                double eke = logu(0.25, 40.0), elke = log(eke);
                int lelke = pwlfInterval(imed, elke, electron_data.eke1, electron_data.eke0) - 1;
                if (kind == K_DRANGE) {
                    double ekef = eke * (1.0 - urand() * 0.3);
                    int lf = pwlfInterval(imed, log(ekef), electron_data.eke1, electron_data.eke0) - 1;
                    if (lf != lelke) ekef = electron_data.e_array[imed * MXEKE + lelke] * (1.0 + 1e-9);
                    if (ekef > eke) ekef = eke;
                    double elkef = log(ekef);
                    in[0] = imed; in[1] = iq; in[2] = lelke; in[3] = eke; in[4] = ekef; in[5] = elke; in[6] = elkef; nin = 7;
                    case_begin();
                    out[0] = computeDrange(imed, iq, lelke, eke, ekef, elke, elkef); k = 1;
                } else {
                    double rhof = region.rhof[ir];
                    int qel = (1 + iq) / 2;
                    double ekei = electron_data.e_array[imed * MXEKE + lelke];
                    double elkei = (lelke + 1 - electron_data.eke0[imed]) / electron_data.eke1[imed];
                    double range = computeDrange(imed, iq, lelke, eke, ekei, elke, elkei);
                    range += electron_data.range_ep[qel * nmed * MXEKE + imed * MXEKE + lelke];
                    range /= rhof;
                    double tustep = range * (c % 3 == 0 ? 1.2 : urand());
                    in[0] = imed; in[1] = iq; in[2] = ir; in[3] = rhof; in[4] = tustep; in[5] = range;
                    in[6] = eke; in[7] = elke; in[8] = lelke; nin = 9;
                    case_begin();
                    out[0] = computeEloss(imed, iq, ir, rhof, tustep, range, eke, elke, lelke); k = 1;
                }
                break;
            }
            case K_PHOTON: case K_ELECTRON: {
                int iq = kind == K_PHOTON ? 0 : ((c & 1) ? 1 : -1);
                int irr = 1 + (int)(urand() * nreg()); if (irr > nreg()) irr = nreg();
                double ek = kind == K_PHOTON ? logu(0.005, 15.0) : logu(0.05, 15.0);
                rand_particle(p, iq, iq ? ek + RM : ek, irr); put_particle(0, p);
                memcpy(in, p, sizeof p); nin = 11;
                case_begin();
                if (kind == K_PHOTON) photon(); else electron();
                k = pack_stack(out);
                k += pack_tally(out + k);
                break;
            }
            }
            case_write(kind, in, nin, out, k);
        }
    }
}

static void run_dose(int nhist, int nbatch) {
    int nper = nhist / nbatch;
    int *rncount = malloc(sizeof(int) * nhist);
    int h = 0;
    g_log_rn = 0; g_log_tally = 0;
    for (int b = 0; b < nbatch; b++) {
        for (int i = 0; i < nper; i++) {
            long before = g_nrn_total;
            initHistory();
            shower();
            rncount[h++] = (int)(g_nrn_total - before);
        }
        if (b == 0) {
            /* raw tallies after the first batch, before accumulation */
            WD("endep_batch0", score.endep, nreg() + 1);
        }
        accumEndep();
    }
    WI("rncount", rncount, h);
    WD("accum_endep", score.accum_endep, nreg() + 1);
    WD("accum_endep2", score.accum_endep2, nreg() + 1);
    double ens = score.ensrc; WD("ensrc", &ens, 1);
    accumulateResults(1, nper, nbatch);
    WD("dose", score.accum_endep + 1, nreg());
    WD("dose_unc", score.accum_endep2 + 1, nreg());
    free(rncount);
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s tables|samplers|dose <inp-stem> <out> [n] [nbatch]\n", argv[0]);
        return 2;
    }
    init_all(argv[2]);
    g_out = fopen(argv[3], "wb");
    if (!g_out) { perror("open out"); return 2; }
    fwrite("OMCG", 1, 4, g_out);
    if (!strcmp(argv[1], "tables")) dump_tables();
    else if (!strcmp(argv[1], "samplers")) run_samplers(argc > 4 ? atoi(argv[4]) : 50);
    else if (!strcmp(argv[1], "dose")) run_dose(argc > 4 ? atoi(argv[4]) : 1000, argc > 5 ? atoi(argv[5]) : 10);
    else { fprintf(stderr, "unknown mode\n"); return 2; }
    fclose(g_out);
    return 0;
}
