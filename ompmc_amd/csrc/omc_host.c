/*
 * Host-side C layer: input formats, physics table builders, regions, source
 * set-up, scoring post-processing and the RANMAR generator.
 *
 * The table builders reproduce the reference's arithmetic operation by
 * operation (same libm calls, same evaluation order), because the FP64 tables
 * are required to equal the reference's bit for bit; each builder cites the
 * reference lines it follows.  Control flow, storage and error handling are
 * this project's own (return codes instead of exit()).
 */
#include "omc_host.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define RM OMC_RM
#define MXGE OMC_MXGE
#define MXEKE OMC_MXEKE
#define MXRAYFF OMC_MXRAYFF
#define RAYCDFSIZE OMC_RAYCDFSIZE

/* ------------------------------------------------------------------ errors */
static __thread char g_err[512];

const char *omc_host_last_error(void) { return g_err; }

int omc_host_set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

/* ------------------------------------------------------------------ .inp */
static int line_blank(const char *s) {
    for (; *s; ++s)
        if (!isspace((unsigned char)*s)) return 0;
    return 1;
}

void omc_strip_spaces(char *dst, const char *src, size_t cap) {
    size_t n = 0;
    for (; *src && n + 1 < cap; ++src)
        if (!isspace((unsigned char)*src)) dst[n++] = *src;
    dst[n] = '\0';
}

/* reference src/omc_utilities.c:63-102 */
int omc_inp_parse(omc_inp *inp, const char *stem) {
    char path[1024];
    snprintf(path, sizeof path, "%s.inp", stem);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    memset(inp, 0, sizeof *inp);
    char buf[OMC_INP_LEN];
    int n = 0;
    while (fgets(buf, sizeof buf, fp)) {
        if (strchr(buf, '#') || line_blank(buf)) continue;
        if (n >= OMC_INP_MAX) {
            fclose(fp);
            return omc_host_set_error(OMC_ERR_LIMIT, "more than %d keys in %s", OMC_INP_MAX, path);
        }
        /* key: up to the first '=', CR or LF; value: the rest up to CR/LF */
        size_t kl = strcspn(buf, "=\r\n");
        memcpy(inp->key[n], buf, kl);
        inp->key[n][kl] = '\0';
        const char *v = buf + kl;
        if (*v == '=') v++;
        size_t vl = strcspn(v, "\r\n");
        memcpy(inp->value[n], v, vl);
        inp->value[n][vl] = '\0';
        n++;
    }
    fclose(fp);
    inp->count = n;
    return OMC_OK;
}

/* reference src/omc_utilities.c:105-123: the item count is kept as count-1
 * and a zero count means "nothing parsed", so one-key files read nothing. */
int omc_inp_get(const omc_inp *inp, const char *key, char *dest, size_t cap) {
    int last = inp->count - 1;
    if (last <= 0) return 0;
    for (int i = 0; i <= last; i++) {
        if (strstr(inp->key[i], key)) {
            snprintf(dest, cap, "%s", inp->value[i]);
            return 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ egsphant */
void omc_phantom_free(omc_phantom *ph) {
    if (ph->owns) {
        free(ph->xb); free(ph->yb); free(ph->zb); free(ph->med); free(ph->rho);
    }
    memset(ph, 0, sizeof *ph);
}

/* reference ucodes/omc_dosxyz/omc_dosxyz.c:84-152 */
int omc_read_egsphant(omc_phantom *ph, const char *path) {
    memset(ph, 0, sizeof *ph);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    char buf[4096];
    int rc = OMC_OK;
    if (!fgets(buf, sizeof buf, fp)) { rc = OMC_ERR_FORMAT; goto fail; }
    ph->nmed = atoi(buf);
    if (ph->nmed < 1 || ph->nmed > OMC_MXMED) {
        fclose(fp);
        return omc_host_set_error(OMC_ERR_LIMIT, "%s: %d media (1..%d supported)", path, ph->nmed, OMC_MXMED);
    }
    for (int i = 0; i < ph->nmed; i++) {
        if (!fgets(buf, sizeof buf, fp)) { rc = OMC_ERR_FORMAT; goto fail; }
        omc_strip_spaces(ph->med_names[i], buf, 60);
    }
    if (!fgets(buf, sizeof buf, fp)) { rc = OMC_ERR_FORMAT; goto fail; }   /* estepe line */
    if (!fgets(buf, sizeof buf, fp) || sscanf(buf, "%d %d %d", &ph->nx, &ph->ny, &ph->nz) != 3 ||
        ph->nx < 1 || ph->ny < 1 || ph->nz < 1) { rc = OMC_ERR_FORMAT; goto fail; }
    ph->owns = 1;
    long nvox = (long)ph->nx * ph->ny * ph->nz;
    ph->xb = malloc(sizeof(double) * (ph->nx + 1));
    ph->yb = malloc(sizeof(double) * (ph->ny + 1));
    ph->zb = malloc(sizeof(double) * (ph->nz + 1));
    ph->med = malloc(sizeof(int) * nvox);
    ph->rho = malloc(sizeof(double) * nvox);
    if (!ph->xb || !ph->yb || !ph->zb || !ph->med || !ph->rho) { rc = OMC_ERR_NOMEM; goto fail; }
    for (int i = 0; i <= ph->nx; i++) if (fscanf(fp, "%lf", &ph->xb[i]) != 1) { rc = OMC_ERR_FORMAT; goto fail; }
    for (int i = 0; i <= ph->ny; i++) if (fscanf(fp, "%lf", &ph->yb[i]) != 1) { rc = OMC_ERR_FORMAT; goto fail; }
    for (int i = 0; i <= ph->nz; i++) if (fscanf(fp, "%lf", &ph->zb[i]) != 1) { rc = OMC_ERR_FORMAT; goto fail; }
    if (!fgets(buf, sizeof buf, fp)) { rc = OMC_ERR_FORMAT; goto fail; }     /* rest of the line */
    /* media: exactly nx single-digit characters per row, a newline, a blank
     * line after each slice */
    for (int k = 0; k < ph->nz; k++) {
        for (int j = 0; j < ph->ny; j++) {
            for (int i = 0; i < ph->nx; i++) {
                int c = fgetc(fp);
                if (c == EOF) { rc = OMC_ERR_FORMAT; goto fail; }
                ph->med[i + (long)j * ph->nx + (long)k * ph->nx * ph->ny] = c - '0';
            }
            if (!fgets(buf, sizeof buf, fp)) { rc = OMC_ERR_FORMAT; goto fail; }
        }
        if (!fgets(buf, sizeof buf, fp)) { rc = OMC_ERR_FORMAT; goto fail; }
    }
    for (long v = 0; v < nvox; v++) {
        if (fscanf(fp, "%lf", &ph->rho[v]) != 1) { rc = OMC_ERR_FORMAT; goto fail; }
    }
    for (long v = 0; v < nvox; v++) {
        if (ph->med[v] < 0 || ph->med[v] > ph->nmed) {
            fclose(fp);
            int code = omc_host_set_error(OMC_ERR_FORMAT, "%s: voxel %ld has medium %d (nmed %d)", path, v,
                                          ph->med[v], ph->nmed);
            omc_phantom_free(ph);
            return code;
        }
    }
    fclose(fp);
    return OMC_OK;
fail:
    fclose(fp);
    omc_phantom_free(ph);
    return omc_host_set_error(rc, "%s: malformed egsphant file", path);
}

/* ------------------------------------------------------------------ PEGS4 */
/* Looks for "KEY=" (blank-insensitive key) in a comma-separated line and
 * returns a pointer to the value text, or NULL. */
static const char *pegs_field(const char *line, const char *key) {
    const char *p = line;
    size_t kl = strlen(key);
    while (*p) {
        while (*p == ' ' || *p == ',') p++;
        const char *start = p;
        while (*p && *p != '=' && *p != ',') p++;
        if (*p == '=') {
            /* compare the trimmed name */
            const char *e = p;
            while (e > start && e[-1] == ' ') e--;
            if ((size_t)(e - start) == kl && strncmp(start, key, kl) == 0) return p + 1;
            p++;
            while (*p && *p != ',') p++;
        }
        if (*p == ',') p++;
    }
    return NULL;
}

/* reference src/ompmc.c:5507-5961 (media located by name, values scaled to cm) */
static int read_pegs(omc_tables *t, const char *pegs_file) {
    FILE *fp = fopen(pegs_file, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", pegs_file);
    int found[OMC_MXMED] = {0}, nfound = 0;
    char line[512];
    while (nfound < t->nmed && fgets(line, sizeof line, fp)) {
        if (strncmp(line, " MEDIUM=", 8) != 0) continue;
        char name[25];
        int j = 0;
        for (int i = 0; i < 24 && line[8 + i] && line[8 + i] != ' ' && line[8 + i] != '\n'; i++) name[j++] = line[8 + i];
        name[j] = '\0';
        int imed = -1;
        for (int i = 0; i < t->nmed; i++)
            if (strncmp(name, t->med_names[i], 20) == 0 && strlen(t->med_names[i]) == strlen(name)) { imed = i; break; }
        if (imed < 0 || found[imed]) continue;
        t->ne[imed] = 0;
        if (!fgets(line, sizeof line, fp)) break;
        const char *v;
        if ((v = pegs_field(line, "RHO"))) t->rho[imed] = strtod(v, NULL);
        if ((v = pegs_field(line, "NE"))) t->ne[imed] = atoi(v);
        if ((v = pegs_field(line, "IUNRST"))) t->iunrst[imed] = atoi(v);
        if ((v = pegs_field(line, "EPSTFL"))) t->epstfl[imed] = atoi(v);
        if ((v = pegs_field(line, "IAPRIM"))) t->iaprim[imed] = atoi(v);
        if (t->ne[imed] > OMC_MXELEMENT) {
            fclose(fp);
            return omc_host_set_error(OMC_ERR_LIMIT, "medium %s has %d elements", name, t->ne[imed]);
        }
        int ok = 1;
        for (int m = 0; m < t->ne[imed]; m++) {
            omc_element el;
            memset(&el, 0, sizeof el);
            if (!fgets(line, sizeof line, fp)) { ok = 0; break; }
            if ((v = pegs_field(line, "ASYM"))) { el.symbol[0] = v[0]; el.symbol[1] = v[1]; el.symbol[2] = 0; }
            if ((v = pegs_field(line, "Z"))) el.z = strtod(v, NULL);
            if ((v = pegs_field(line, "A"))) el.wa = strtod(v, NULL);
            if ((v = pegs_field(line, "PZ"))) el.pz = strtod(v, NULL);
            if ((v = pegs_field(line, "RHOZ"))) el.rhoz = strtod(v, NULL);
            t->elements[imed][m] = el;
        }
        if (!ok) break;
        if (!fgets(line, sizeof line, fp) ||
            sscanf(line, "%lf %lf %lf %lf %lf", &t->rlc[imed], &t->ae[imed], &t->ap[imed], &t->ue[imed],
                   &t->up[imed]) != 5)
            continue;
        t->te[imed] = t->ae[imed] - RM;
        t->thmoll[imed] = (t->te[imed]) * 2 + RM;
        if (!fgets(line, sizeof line, fp) ||
            sscanf(line, "%d %d %d %d %d %d %d", &t->msge[imed], &t->mge[imed], &t->mseke[imed], &t->meke[imed],
                   &t->mleke[imed], &t->mcmfp[imed], &t->mrange[imed]) != 7)
            continue;
        if (t->meke[imed] > MXEKE) {
            fclose(fp);
            return omc_host_set_error(OMC_ERR_LIMIT, "medium %s has MEKE %d > %d", name, t->meke[imed], MXEKE);
        }
        for (int i = 0; i < 7; i++)
            if (!fgets(line, sizeof line, fp)) { ok = 0; break; }
        double d1, d2, d3, d4, d5;
        if (!ok || sscanf(line, "%lf %lf %lf %lf %lf", &d1, &d2, &d3, &d4, &d5) != 5) continue;
        double dl6, a1, a2, b1, b2, p1, p2, xr0, teff0;
        if (fscanf(fp, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &dl6, &t->delcm[imed], &a1, &a2,
                   &b1, &b2, &p1, &p2, &xr0, &teff0, &t->blcc[imed], &t->xcc[imed], &t->eke0[imed],
                   &t->eke1[imed]) != 14) {
            fclose(fp);
            return omc_host_set_error(OMC_ERR_FORMAT, "%s: medium %s header truncated", pegs_file, name);
        }
        int n = t->meke[imed];
        size_t o = (size_t)imed * MXEKE;
        for (int k = 0; k < n; k++) {
            if (fscanf(fp, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf",
                       &t->esig0[o + k], &t->esig1[o + k], &t->psig0[o + k], &t->psig1[o + k], &t->ededx0[o + k],
                       &t->ededx1[o + k], &t->pdedx0[o + k], &t->pdedx1[o + k], &t->ebr10[o + k], &t->ebr11[o + k],
                       &t->pbr10[o + k], &t->pbr11[o + k], &t->pbr20[o + k], &t->pbr21[o + k], &t->tmxs0[o + k],
                       &t->tmxs1[o + k]) != 16) {
                fclose(fp);
                return omc_host_set_error(OMC_ERR_FORMAT, "%s: medium %s table truncated", pegs_file, name);
            }
        }
        /* lengths in cm (reference src/ompmc.c:5903-5915) */
        double dfacti = 1.0 / (t->rlc[imed]);
        t->blcc[imed] *= dfacti;
        for (int k = 0; k < n; k++) {
            t->esig0[o + k] *= dfacti; t->psig0[o + k] *= dfacti;
            t->ededx0[o + k] *= dfacti; t->pdedx0[o + k] *= dfacti;
            t->pdedx1[o + k] *= dfacti; t->esig1[o + k] *= dfacti;
            t->psig1[o + k] *= dfacti; t->ededx1[o + k] *= dfacti;
        }
        t->xcc[imed] *= sqrt(dfacti);
        found[imed] = 1;
        nfound++;
    }
    fclose(fp);
    if (nfound < t->nmed) {
        char miss[512] = "";
        for (int i = 0; i < t->nmed; i++)
            if (!found[i]) { strncat(miss, " ", sizeof miss - strlen(miss) - 1); strncat(miss, t->med_names[i], sizeof miss - strlen(miss) - 1); }
        return omc_host_set_error(OMC_ERR_MEDIUM, "media not found or unreadable in %s:%s", pegs_file, miss);
    }
    return OMC_OK;
}

/* ------------------------------------------------------------------ photon */
typedef struct xsec { int n[OMC_MXELEMENT]; double *d0[OMC_MXELEMENT], *d1[OMC_MXELEMENT]; } xsec;

static void xsec_free(xsec *x) {
    for (int i = 0; i < OMC_MXELEMENT; i++) { free(x->d0[i]); free(x->d1[i]); }
}

/* reference src/ompmc.c:217-279: the first 50 elements of an XCOM file */
static int read_xsec(xsec *x, const char *folder, const char *name) {
    char path[1024];
    snprintf(path, sizeof path, "%s%s", folder, name);
    memset(x, 0, sizeof *x);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    for (int i = 0; i < OMC_MXELEMENT; i++) {
        unsigned n;
        if (fscanf(fp, "%u", &n) != 1) goto bad;
        x->n[i] = (int)n;
        x->d0[i] = malloc(sizeof(double) * (n + 1));
        x->d1[i] = malloc(sizeof(double) * (n + 1));
        for (unsigned j = 0; j < n; j++)
            if (fscanf(fp, "%lf %lf", &x->d0[i][j], &x->d1[i][j]) != 2) goto bad;
    }
    fclose(fp);
    return OMC_OK;
bad:
    fclose(fp);
    xsec_free(x);
    return omc_host_set_error(OMC_ERR_FORMAT, "Could not read the data file %s", path);
}

/* reference src/ompmc.c:281-346 (heap sort that also permutes 1-based indices) */
static void heap_sort_idx(int n, double *a, int *idx) {
    for (int i = 0; i < n; i++) idx[i] = i + 1;
    if (n < 2) return;
    int l = n / 2 + 1, ir = n;
    for (;;) {
        double va; int vi;
        if (l > 1) {
            l--;
            va = a[l - 1]; vi = l;
        } else {
            va = a[ir - 1]; vi = idx[ir - 1];
            a[ir - 1] = a[0]; idx[ir - 1] = idx[0];
            if (--ir == 0) { a[0] = va; idx[0] = vi; return; }
        }
        int i = l, j = 2 * l;
        while (j <= ir) {
            if (j < ir && a[j - 1] < a[j]) j++;
            if (va < a[j - 1]) {
                a[i - 1] = a[j - 1]; idx[i - 1] = idx[j - 1];
                i = j; j = 2 * j;
            } else {
                j = ir + 1;
            }
        }
        a[i - 1] = va; idx[i - 1] = vi;
    }
}

/* reference src/ompmc.c:348-450: element cross sections interpolated
 * log-log on the MXGE grid and summed with weights pz (flag 1 pair, flag 2
 * triplet: threshold factor (1-eth/E)^3 removed before and restored after). */
static void mix_xsec(double *res, int flag, int ne, const xsec *x, const double *zs, const double *pzs,
                     double ge0, double ge1) {
    for (int j = 0; j < MXGE; j++) res[j] = 0.0;
    for (int i = 0; i < ne; i++) {
        int z = (int)(zs[i] + 0.5) - 1;
        int n = x->n[z];
        double eth = 0.0;
        double *d0 = malloc(sizeof(double) * (n + 1)), *d1 = malloc(sizeof(double) * (n + 1));
        if (flag == 0) {
            for (int j = 0; j < n; j++) { d0[j] = x->d0[z][j]; d1[j] = x->d1[z][j]; }
        } else {
            for (int j = 0; j < n; j++) { d0[j + 1] = x->d0[z][j]; d1[j + 1] = x->d1[z][j]; }
            eth = flag == 1 ? 2.0 * RM : 4.0 * RM;
            n++;
            for (int j = 1; j < n; j++) d1[j] -= 3.0 * log(1.0 - eth / exp(d0[j]));
            d0[0] = (double)log(eth);
            d1[0] = (double)d1[1];
        }
        for (int j = 0; j < MXGE; j++) {
            double gle = ((double)(j + 1) - ge0) / ge1;
            double e = exp(gle);
            double sig = 0.0;
            if (gle < d0[0] || gle >= d0[n - 1]) {
                if (flag != 0) sig = gle < d0[0] ? 0.0 : exp(d1[n - 1]);
            } else {
                int k = 0;
                while (k < n - 1 && !(gle >= d0[k] && gle < d0[k + 1])) k++;
                double p = (gle - d0[k]) / (d0[k + 1] - d0[k]);
                sig = exp(p * d1[k + 1] + (1.0 - p) * d1[k]);
            }
            if (flag != 0 && e > eth) sig *= (1.0 - eth / e) * (1.0 - eth / e) * (1.0 - eth / e);
            res[j] += pzs[i] * sig;
        }
        free(d0); free(d1);
    }
}

/* Klein-Nishina total cross section per electron (reference src/ompmc.c:452-470) */
static double kn_total(double e) {
    double con = 0.1274783851;
    double ko = e / RM;
    if (ko < 0.01) return 8.0 * con / 3.0 * (1.0 - ko * (2.0 - ko * (5.2 - 13.3 * ko))) / RM;
    double c1 = 1.0 / (ko * ko);
    double c2 = 1.0 - 2.0 * (1.0 + ko) * c1;
    double c3 = (1.0 + 2.0 * ko) * c1;
    double eps2 = 1.0, eps1 = 1.0 / (1.0 + 2.0 * ko);
    return (c1 * (1.0 / eps1 - 1.0 / eps2) + c2 * log(eps2 / eps1) + eps2 * (c3 + 0.5 * eps2) -
            eps1 * (c3 + 0.5 * eps1)) / e * con;
}

/* PWLF coefficients between consecutive grid points: c1 = slope*ge1,
 * c0 = value - c1*gle (reference src/ompmc.c:594-626) */
static void pwlf_store(double *c0, double *c1, int idx, double v, double vold, double ge1, double gle) {
    c1[idx] = (v - vold) * ge1;
    c0[idx] = v - c1[idx] * gle;
}

/* reference src/ompmc.c:472-666 */
static int build_photon(omc_tables *t, const char *folder) {
    xsec xp, xr, xpair, xtrip;
    int rc;
    if ((rc = read_xsec(&xp, folder, "xcom_photo.data"))) return rc;
    if ((rc = read_xsec(&xr, folder, "xcom_rayleigh.data"))) { xsec_free(&xp); return rc; }
    if ((rc = read_xsec(&xpair, folder, "xcom_pair.data"))) { xsec_free(&xp); xsec_free(&xr); return rc; }
    if ((rc = read_xsec(&xtrip, folder, "xcom_triplet.data"))) {
        xsec_free(&xp); xsec_free(&xr); xsec_free(&xpair); return rc;
    }
    double *sp = malloc(sizeof(double) * MXGE * 4);
    double *sr = sp + MXGE, *spa = sp + 2 * MXGE, *st = sp + 3 * MXGE;
    for (int i = 0; i < t->nmed && rc == OMC_OK; i++) {
        int ne = t->ne[i];
        for (int j = 0; j < ne; j++)
            if (t->elements[i][j].z > OMC_MXELEMENT + 0.5 || t->elements[i][j].z < 0.5) {
                rc = omc_host_set_error(OMC_ERR_LIMIT, "medium %s: element Z=%g outside 1..%d", t->med_names[i],
                                        t->elements[i][j].z, OMC_MXELEMENT);
            }
        if (rc) break;
        t->ge1[i] = (double)(MXGE - 1) / log(t->up[i] / t->ap[i]);
        t->ge0[i] = 1.0 - t->ge1[i] * log(t->ap[i]);
        double sumA = 0.0, sumZ = 0.0;
        double zs[OMC_MXELEMENT], pzs[OMC_MXELEMENT];
        int srt[OMC_MXELEMENT];
        for (int j = 0; j < ne; j++) {
            zs[j] = t->elements[i][j].z;
            sumA += t->elements[i][j].pz * t->elements[i][j].wa;
            sumZ += t->elements[i][j].pz * t->elements[i][j].z;
        }
        double con2 = t->rho[i] / (sumA * 1.6605655);
        heap_sort_idx(ne, zs, srt);
        for (int j = 0; j < ne; j++) pzs[j] = t->elements[i][srt[j] - 1].pz;
        mix_xsec(sp, 0, ne, &xp, zs, pzs, t->ge0[i], t->ge1[i]);
        mix_xsec(sr, 0, ne, &xr, zs, pzs, t->ge0[i], t->ge1[i]);
        mix_xsec(spa, 1, ne, &xpair, zs, pzs, t->ge0[i], t->ge1[i]);
        mix_xsec(st, 2, ne, &xtrip, zs, pzs, t->ge0[i], t->ge1[i]);
        double gle = 0.0, gmfp = 0.0, gbr1 = 0.0, gbr2 = 0.0, cohe = 0.0;
        double gmfp_o = 0.0, gbr1_o = 0.0, gbr2_o = 0.0, cohe_o = 0.0;
        for (int j = 0; j < MXGE; j++) {
            gle = ((double)(j + 1) - t->ge0[i]) / t->ge1[i];
            double e = exp(gle);
            double sig_kn = sumZ * kn_total(e);
            double sig_p = spa[j] + st[j];
            double sigma = sig_kn + sig_p + sp[j];
            gmfp = 1.0 / (sigma * con2);
            gbr1 = sig_p / sigma;
            gbr2 = gbr1 + sig_kn / sigma;
            cohe = sigma / (sr[j] + sigma);
            if (j > 0) {
                int idx = i * MXGE + (j - 1);
                pwlf_store(t->gmfp0, t->gmfp1, idx, gmfp, gmfp_o, t->ge1[i], gle);
                pwlf_store(t->gbr10, t->gbr11, idx, gbr1, gbr1_o, t->ge1[i], gle);
                pwlf_store(t->gbr20, t->gbr21, idx, gbr2, gbr2_o, t->ge1[i], gle);
                pwlf_store(t->cohe0, t->cohe1, idx, cohe, cohe_o, t->ge1[i], gle);
            }
            gmfp_o = gmfp; gbr1_o = gbr1; gbr2_o = gbr2; cohe_o = cohe;
        }
        int idx = i * MXGE + MXGE - 1;
        t->gmfp1[idx] = t->gmfp1[idx - 1]; t->gmfp0[idx] = gmfp - t->gmfp1[idx] * gle;
        t->gbr11[idx] = t->gbr11[idx - 1]; t->gbr10[idx] = gbr1 - t->gbr11[idx] * gle;
        t->gbr21[idx] = t->gbr21[idx - 1]; t->gbr20[idx] = gbr2 - t->gbr21[idx] * gle;
        t->cohe1[idx] = t->cohe1[idx - 1]; t->cohe0[idx] = cohe - t->cohe1[idx] * gle;
    }
    free(sp);
    xsec_free(&xp); xsec_free(&xr); xsec_free(&xpair); xsec_free(&xtrip);
    return rc;
}

/* ------------------------------------------------------------------ Rayleigh */
/* reference src/ompmc.c:758-819 */
static int read_ff(double *xval, double *aff, const char *path) {
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    for (int i = 0; i < MXRAYFF; i++)
        if (fscanf(fp, "%lf", &xval[i]) != 1) goto bad;
    for (int i = 0; i < OMC_MXELEMENT; i++)
        for (int j = 0; j < MXRAYFF; j++)
            if (fscanf(fp, "%lf", &aff[i * MXRAYFF + j]) != 1) goto bad;
    fclose(fp);
    return OMC_OK;
bad:
    fclose(fp);
    return omc_host_set_error(OMC_ERR_FORMAT, "Could not read atomic form factors file %s", path);
}

/* reference src/ompmc.c:821-1025 (EGSnrc prepare_rayleigh_data) */
static int build_rayleigh(omc_tables *t, const char *pgs4form) {
    double xval[MXRAYFF];
    double *aff = malloc(sizeof(double) * OMC_MXELEMENT * MXRAYFF);
    int rc = read_ff(xval, aff, pgs4form);
    if (rc) { free(aff); return rc; }
    double ff[MXRAYFF], pe[MXGE];
    for (int i = 0; i < t->nmed; i++) {
        double *xg = t->ray_xgrid + i * MXRAYFF, *fc = t->ray_fcum + i * MXRAYFF;
        double *ba = t->ray_b + i * MXRAYFF, *ca = t->ray_c + i * MXRAYFF;
        int *ia = t->ray_iarr + i * RAYCDFSIZE;
        for (int j = 0; j < MXRAYFF; j++) {
            double s = 0.0;
            xg[j] = xval[j];
            for (int k = 0; k < t->ne[i]; k++) {
                int z = (int)t->elements[i][k].z - 1;
                s += t->elements[i][k].pz * pow(aff[z * MXRAYFF + j], 2);
            }
            ff[j] = sqrt(s);
        }
        if (xg[0] < 1.0E-6) xg[0] = 0.0001;
        double emin = exp((1.0 - t->ge0[i]) / t->ge1[i]);
        double emax = exp((MXGE - t->ge0[i]) / t->ge1[i]);
        /* a form factor of exactly +0.0 becomes the smallest subnormal, so
         * that its logarithm stays finite */
        for (int j = 0; j < MXRAYFF; j++) {
            uint64_t bits;
            memcpy(&bits, &ff[j], 8);
            if (bits == 0) { bits = 1; memcpy(&ff[j], &bits, 8); }
        }
        double sum0 = 0.0;
        fc[0] = 0.0;
        ba[MXRAYFF - 1] = 0.0;   /* never assigned by the reference; unused */
        for (int j = 0; j < MXRAYFF - 1; j++) {
            double b = log(ff[j + 1] / ff[j]) / log(xg[j + 1] / xg[j]);
            ba[j] = b;
            double x1 = xg[j], x2 = xg[j + 1];
            double px1 = pow(x1, 2.0 * b), px2 = pow(x2, 2.0 * b);
            sum0 += pow(ff[j], 2) * (pow(x2, 2) * px2 - pow(x1, 2) * px1) / ((1.0 + b) * px1);
            fc[j + 1] = sum0;
        }
        double dle = log(emax / emin) / ((double)MXGE - 1.0);
        int idx = 1;
        for (int j = 1; j <= MXGE; j++) {
            double e = emin * exp(dle * ((double)j - 1.0));
            double xmax = 20.607544 * 2.0 * e / RM;
            int k;
            for (k = 1; k <= MXRAYFF - 1; k++)
                if (xmax >= xg[k - 1] && xmax < xg[k]) break;
            idx = k;
            double b = ba[idx - 1];
            double x1 = xg[idx - 1], x2 = xmax;
            double px1 = pow(x1, 2.0 * b), px2 = pow(x2, 2.0 * b);
            pe[j - 1] = fc[idx - 1] + pow(ff[idx - 1], 2) * (pow(x2, 2) * px2 - pow(x1, 2) * px1) / ((1.0 + b) * px1);
        }
        ia[RAYCDFSIZE - 1] = idx;
        double anorm = 1.0 / sqrt(pe[MXGE - 1]);
        double anorm1 = 1.005 / pe[MXGE - 1];
        double anorm2 = 1.0 / pe[MXGE - 1];
        for (int j = 0; j < MXGE; j++) {
            pe[j] *= anorm1;
            if (pe[j] > 1.0) pe[j] = 1.0;
        }
        for (int j = 0; j < MXRAYFF; j++) {
            ff[j] *= anorm;
            fc[j] *= anorm2;
            ca[j] = (1.0 + ba[j]) / pow(xg[j] * ff[j], 2);
        }
        /* uniform cumulative bins */
        double dw = 1.0 / ((double)RAYCDFSIZE - 1.0);
        double xold = xg[0];
        int ibin = 1;
        double b = ba[0];
        double px1 = pow(xg[0], 2.0 * b);
        ia[0] = 1;
        for (int j = 2; j <= RAYCDFSIZE - 1; j++) {
            double w = dw;
            for (;;) {
                double x1 = xold, x2 = xg[ibin];
                double tt = pow(x1, 2) * pow(x1, 2.0 * b);
                double px2 = pow(x2, 2.0 * b);
                double aux = pow(ff[ibin - 1], 2) * (pow(x2, 2) * px2 - tt) / ((1.0 + b) * px1);
                if (aux > w) {
                    xold = exp(log(tt + w * (1.0 + b) * px1 / pow(ff[ibin - 1], 2)) / (2.0 + 2.0 * b));
                    ia[j - 1] = ibin;
                    break;
                }
                w -= aux;
                xold = x2;
                ibin++;
                b = ba[ibin - 1];
                px1 = pow(xold, 2.0 * b);
            }
        }
        for (int j = 0; j < MXRAYFF; j++) ba[j] = 0.5 / (1.0 + ba[j]);
        for (int j = 0; j < MXGE - 1; j++) {
            double gle = ((j + 1) - t->ge0[i]) / t->ge1[i];
            t->ray_pmax1[i * MXGE + j] = (pe[j + 1] - pe[j]) * t->ge1[i];
            t->ray_pmax0[i * MXGE + j] = pe[j] - t->ray_pmax1[i * MXGE + j] * gle;
        }
        t->ray_pmax0[i * MXGE + MXGE - 1] = t->ray_pmax0[i * MXGE + MXGE - 2];
        t->ray_pmax1[i * MXGE + MXGE - 1] = t->ray_pmax1[i * MXGE + MXGE - 2];
    }
    free(aff);
    return OMC_OK;
}

/* ------------------------------------------------------------------ pair / brems */
/* Coulomb correction (reference src/ompmc.c:1166-1181) */
static double coulomb_fc(double zi) {
    double a = OMC_FSC * zi;
    double fc = 1 + pow(a, 2);
    fc = 1.0 / fc;
    fc = fc + 0.20206 - 0.0369 * pow(a, 2);
    fc = fc + 0.0083 * pow(a, 4);
    fc = fc - 0.002 * pow(a, 6);
    fc = fc * pow(a, 2);
    return fc;
}

/* atomic-electron screening term (reference src/ompmc.c:1183-1205) */
static double screening_xi(double zi, double fc) {
    if (zi == 4) return 5.924 / (4.710 - fc);
    if (zi == 3) return 5.805 / (4.740 - fc);
    if (zi == 2) return 5.621 / (4.790 - fc);
    if (zi == 1) return 6.144 / (5.310 - fc);
    return log(1194.0 * pow(zi, -2.0 / 3.0)) / log(184.15 * pow(zi, -1.0 / 3.0)) - fc;
}

/* reference src/ompmc.c:1207-1336 (EGSnrc fix_brems + INITIALIZE-PAIR-ANGLE) */
static void build_pair(omc_tables *t) {
    for (int m = 0; m < t->nmed; m++) {
        double Zt = 0.0, Zb = 0.0, Zf = 0.0;
        for (int i = 0; i < t->ne[m]; i++) {
            double zi = t->elements[m][i].z, pi = t->elements[m][i].pz;
            double fc = coulomb_fc(zi);
            double xi = screening_xi(zi, fc);
            double aux = pi * zi * (zi + xi);
            Zt = Zt + aux;
            Zb = Zb - aux * log(zi) / 3.0;
            Zf = Zf + aux * fc;
        }
        double Zv = (Zb - Zf) / Zt, Zg = Zb / Zt;
        double fmax1 = 2.0 * (20.863 + 4.0 * Zg) - 2.0 * (20.029 + 4.0 * Zg) / 3.0;
        double fmax2 = 2.0 * (20.863 + 4.0 * Zv) - 2.0 * (20.029 + 4.0 * Zv) / 3.0;
        double *l1 = t->dl1 + 8 * m, *l2 = t->dl2 + 8 * m, *l3 = t->dl3 + 8 * m;
        double *l4 = t->dl4 + 8 * m, *l5 = t->dl5 + 8 * m, *l6 = t->dl6 + 8 * m;
        /* brems screening rejection coefficients: (A,B) x (unscreened, Coulomb corrected) */
        double zz[2] = {Zg, Zv}, fm[2] = {fmax1, fmax2};
        for (int c = 0; c < 2; c++) {
            double Z = zz[c], f = fm[c];
            l1[2 * c] = (20.863 + 4.0 * Z) / f; l2[2 * c] = -3.242 / f; l3[2 * c] = 0.625 / f;
            l4[2 * c] = (21.12 + 4.0 * Z) / f; l5[2 * c] = -4.184 / f; l6[2 * c] = 0.952;
            l1[2 * c + 1] = (20.029 + 4.0 * Z) / f; l2[2 * c + 1] = -1.93 / f; l3[2 * c + 1] = -0.086 / f;
            l4[2 * c + 1] = (21.12 + 4.0 * Z) / f; l5[2 * c + 1] = -4.184 / f; l6[2 * c + 1] = 0.952;
        }
        /* pair production rejection coefficients */
        for (int c = 0; c < 2; c++) {
            double Z = zz[c];
            int a = 4 + 2 * c, b = a + 1;
            l1[a] = (3.0 * (20.863 + 4.0 * Z) - (20.029 + 4.0 * Z));
            l2[a] = (3.0 * (-3.242) - (-1.930));
            l3[a] = (3.0 * (0.625) - (-0.086));
            l4[a] = (2.0 * 21.12 + 8.0 * Z);
            l5[a] = (2.0 * (-4.184));
            l6[a] = 0.952;
            l1[b] = (3.0 * (20.863 + 4.0 * Z) + (20.029 + 4.0 * Z));
            l2[b] = (3.0 * (-3.242) + (-1.930));
            l3[b] = (3.0 * 0.625 + (-0.086));
            l4[b] = (4.0 * 21.12 + 16.0 * Z);
            l5[b] = (4.0 * (-4.184));
            l6[b] = 0.952;
        }
        t->bpar1[m] = l1[6] / (3.0 * l1[7] + l1[6]);
        t->bpar0[m] = 12.0 * l1[7] / (3.0 * l1[7] + l1[6]);
        double zb = 0.0, pzn = 0.0;
        for (int i = 0; i < t->ne[m]; i++) {
            zb += (double)(t->elements[m][i].pz) * (t->elements[m][i].z) * ((t->elements[m][i].z) + 1.0f);
            pzn += t->elements[m][i].pz;
        }
        t->zbrang[m] = (8.116224E-05) * pow(zb / pzn, 1.0 / 3.0);
    }
}

/* ------------------------------------------------------------------ spin */
/* natural cubic spline set-up and evaluation (reference src/ompmc.c:3034-3113) */
static void spline_setup(const double *x, const double *f, double *a, double *b, double *c, double *d, int n) {
    double s = 0, r;
    int m1 = 2, m2 = n - 1, m, mr;
    for (m = 1; m <= m2; m++) {
        d[m - 1] = x[m] - x[m - 1];
        r = (f[m] - f[m - 1]) / d[m - 1];
        c[m - 1] = r - s;
        s = r;
    }
    s = 0; r = 0;
    c[0] = 0; c[n - 1] = 0;
    for (m = m1; m <= m2; m++) {
        c[m - 1] = c[m - 1] + r * c[m - 2];
        b[m - 1] = 2 * (x[m - 2] - x[m]) - r * s;
        s = d[m - 1];
        r = s / b[m - 1];
    }
    mr = m2;
    for (m = m1; m <= m2; m++) {
        c[mr - 1] = (d[mr - 1] * c[mr] - c[mr - 1]) / b[mr - 1];
        mr = mr - 1;
    }
    for (m = 1; m <= m2; m++) {
        s = d[m - 1];
        r = c[m] - c[m - 1];
        d[m - 1] = r / s;
        c[m - 1] = 3 * c[m - 1];
        b[m - 1] = (f[m] - f[m - 1]) / s - (c[m - 1] + r) * s;
        a[m - 1] = f[m - 1];
    }
}

static double spline_eval(double s, const double *x, const double *a, const double *b, const double *c,
                          const double *d, int n) {
    int lo, hi, dir, m;
    if (x[0] > x[n - 1]) { dir = 1; lo = n; hi = 0; }
    else { dir = 0; lo = 0; hi = n; }
    if (s >= x[hi + dir - 1]) m = hi + 2 * dir - 1;
    else if (s <= x[lo - dir]) m = lo - 2 * dir + 1;
    else {
        int ml = lo, mu = hi;
        while (abs(mu - ml) > 1) {
            int mav = (ml + mu) / 2;
            if (s < x[mav - 1]) mu = mav; else ml = mav;
        }
        m = mu + dir - 1;
    }
    double q = s - x[m - 1];
    return a[m - 1] + q * (b[m - 1] + q * (c[m - 1] + q * d[m - 1]));
}

/* PWLF of a function sampled on the electron grid: interval i-1 gets the
 * line through (eil(i-1), f(i-1)) and (eil(i), f(i)) expressed with f(i). */
static void e_pwlf(double *c0, double *c1, int o, int i, double si2, double si1, double eke1, double eil) {
    c1[o + i - 1] = (si2 - si1) * eke1;
    c0[o + i - 1] = si2 - c1[o + i - 1] * eil;
}

#define NSPIN1 (OMC_MXE_SPIN1 + 1)

/* Screening-table lookup.  At the top of the beta^2 grid the reference
 * interpolates with the entry one past the end of its 2x32 array
 * (src/ompmc.c:2712-2715, 2740-2743).  For electrons that is the first
 * positron entry, which is kept; for positrons it is uninitialised heap
 * memory (a stale value of a freed buffer in the reference's own runs), which
 * cannot be reproduced, so the last positron entry is used instead.  Only
 * positrons above ~45 MeV are affected. */
static double eta_at(const double *eta, int idx) { return eta[idx < 2 * NSPIN1 ? idx : 2 * NSPIN1 - 1]; }
#define SPIN_IDX(imed, iq, i, j, k) \
    ((size_t)(imed) * OMC_NSPIN_MED + (size_t)(iq) * NSPIN1 * (OMC_MXQ_SPIN + 1) * (OMC_MXU_SPIN + 1) + \
     (size_t)(i) * (OMC_MXQ_SPIN + 1) * (OMC_MXU_SPIN + 1) + (size_t)(j) * (OMC_MXU_SPIN + 1) + (k))

/* reference src/ompmc.c:2394-2965.  The file's header values are float32 and
 * enter some expressions as float arithmetic, which is kept. */
static int build_spin(omc_tables *t, const char *folder) {
    char path[1024];
    snprintf(path, sizeof path, "%sspinms.data", folder);
    FILE *fp = fopen(path, "rb");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    fseek(fp, 0, SEEK_END);
    long len = ftell(fp);
    rewind(fp);
    unsigned char *raw = malloc(len > 0 ? len : 1);
    if (fread(raw, 1, len, fp) != (size_t)len) { fclose(fp); free(raw); return omc_host_set_error(OMC_ERR_IO, "short read %s", path); }
    fclose(fp);
    const int nener = OMC_MXE_SPIN;
    float espin_min, espin_max, b2spin_min, b2spin_max;
    if (len < 1104) { free(raw); return omc_host_set_error(OMC_ERR_FORMAT, "%s too short", path); }
    memcpy(&espin_min, raw + 36, 4); memcpy(&espin_max, raw + 40, 4);
    memcpy(&b2spin_min, raw + 44, 4); memcpy(&b2spin_max, raw + 48, 4);
    t->b2spin_min = (double)b2spin_min;
    double earray[NSPIN1];
    double dloge = log(espin_max / espin_min) / (double)nener;
    double eloge = log(espin_min);
    earray[0] = espin_min;
    for (int i = 1; i <= nener; i++) { eloge += dloge; earray[i] = exp(eloge); }
    double dbeta2 = (b2spin_max - b2spin_min) / nener;
    double beta2 = b2spin_min;
    earray[nener + 1] = espin_max;
    for (int i = nener + 2; i <= 2 * nener + 1; i++) {
        beta2 += dbeta2;
        earray[i] = beta2 < 0.999 ? RM * 1000.0 * (1.0 / sqrt(1.0 - beta2) - 1) : 50585.1;
    }
    espin_min /= 1000.0;
    espin_max /= 1000.0;
    double dlener = log(espin_max / espin_min) / OMC_MXE_SPIN;
    t->dleneri = 1.0 / dlener;
    t->espml = log(espin_min);
    dbeta2 = (b2spin_max - b2spin_min) / OMC_MXE_SPIN;
    t->dbeta2i = 1.0 / dbeta2;
    double dqq1 = 0.5 / OMC_MXQ_SPIN;
    t->dqq1i = 1.0 / dqq1;

    double eta_a[2 * NSPIN1], c_a[2 * NSPIN1], g_a[2 * NSPIN1];
    double elarray[NSPIN1], farray[NSPIN1], af[NSPIN1], bf[NSPIN1], cf[NSPIN1], df[NSPIN1];
    for (int imed = 0; imed < t->nmed; imed++) {
        double sum_Z2 = 0.0, sum_A = 0.0, sum_pz = 0.0, sum_Z = 0.0;
        memset(eta_a, 0, sizeof eta_a); memset(c_a, 0, sizeof c_a); memset(g_a, 0, sizeof g_a);
        for (int ie = 0; ie < t->ne[imed]; ie++) {
            double z = t->elements[imed][ie].z;
            int iz = (int)(z + 0.5);
            double pz = t->elements[imed][ie].pz;
            double tmp = z * (z + 1.0) * pz;
            sum_Z2 += tmp; sum_Z += pz * z;
            sum_A += pz * t->elements[imed][ie].wa; sum_pz += pz;
            double z23 = pow(z, 2.0 / 3.0);
            for (int iq = 0; iq < 2; iq++) {
                for (int i = 0; i <= OMC_MXE_SPIN1; i++) {
                    long rec = (long)(iz - 1) * 4 * (nener + 1) + 2L * iq * (nener + 1) + i + 1; /* 0-based */
                    if ((rec + 1) * 1104 > len) {
                        free(raw);
                        return omc_host_set_error(OMC_ERR_FORMAT, "%s has no record for Z=%d", path, iz);
                    }
                    const unsigned char *r = raw + rec * 1104;
                    float f4[20];
                    memcpy(f4, r, sizeof f4);
                    double dum1 = f4[0], dum2 = f4[1], dum3 = f4[2], aux_o = f4[3];
                    eta_a[iq * NSPIN1 + i] += tmp * log(z23 * aux_o);
                    double tau = earray[i] / (1000.0 * RM);
                    double eta = z23 / ((137.03604 * 0.88534138) * (137.03604 * 0.88534138)) * aux_o / 4 / tau / (tau + 2);
                    c_a[iq * NSPIN1 + i] += tmp * (log(1.0 + 1.0 / eta) - 1.0 / (1.0 + eta)) * dum1 * dum3;
                    g_a[iq * NSPIN1 + i] += tmp * dum2;
                    for (int j = 0; j <= OMC_MXQ_SPIN; j++) {
                        double fmaxj = f4[4 + j];
                        for (int k = 0; k <= OMC_MXU_SPIN; k++) {
                            int16_t q16;
                            memcpy(&q16, r + 80 + 2 * ((OMC_MXU_SPIN + 1) * j + k), 2);
                            int ii2 = q16;
                            if (ii2 < 0) ii2 += 65536;
                            double v = ii2;
                            v = v * fmaxj / 65535;
                            t->spin_rej[SPIN_IDX(imed, iq, i, j, k)] += tmp * v;
                        }
                    }
                }
            }
        }
        /* scale each rejection row to a maximum of one */
        for (int iq = 0; iq < 2; iq++)
            for (int i = 0; i <= OMC_MXE_SPIN1; i++)
                for (int j = 0; j <= OMC_MXQ_SPIN; j++) {
                    double flmax = 0.0;
                    for (int k = 0; k <= OMC_MXU_SPIN; k++)
                        if (flmax < t->spin_rej[SPIN_IDX(imed, iq, i, j, k)]) flmax = t->spin_rej[SPIN_IDX(imed, iq, i, j, k)];
                    for (int k = 0; k <= OMC_MXU_SPIN; k++)
                        t->spin_rej[SPIN_IDX(imed, iq, i, j, k)] = t->spin_rej[SPIN_IDX(imed, iq, i, j, k)] / flmax;
                }
        for (int i = 0; i <= OMC_MXE_SPIN1; i++) {
            double tau = (earray[i] / RM) * 0.001;
            for (int iq = 0; iq < 2; iq++) {
                double aux_o = exp(eta_a[iq * NSPIN1 + i] / sum_Z2) / (pow(137.03604 * 0.88534138, 2.0));
                eta_a[iq * NSPIN1 + i] = 0.26112447 * aux_o * (t->blcc[imed]) / (t->xcc[imed]);
                double eta = aux_o / 4.0 / tau / (tau + 2);
                double gamma = 3.0 * (1.0 + eta) * (log(1.0 + 1.0 / eta) * (1.0 + 2.0 * eta) - 2.0) /
                               (log(1.0 + 1.0 / eta) * (1.0 + eta) - 1.0);
                g_a[iq * NSPIN1 + i] = g_a[iq * NSPIN1 + i] / sum_Z2 / gamma;
                c_a[iq * NSPIN1 + i] = c_a[iq * NSPIN1 + i] / sum_Z2 / (log(1.0 + 1.0 / eta) - 1.0 / (1.0 + eta));
            }
        }
        /* screening-parameter correction on the electron grid */
        int o = imed * MXEKE;
        double eke0 = t->eke0[imed], eke1 = t->eke1[imed];
        double eil = (1.0 - eke0) / eke1;
        double e = exp(eil);
        double si1e, si1p, si2e, si2p, aae;
        int je = 0;
        if (e <= espin_min) {
            si1e = eta_a[0]; si1p = eta_a[NSPIN1];
        } else {
            if (e <= espin_max) {
                aae = (eil - t->espml) * t->dleneri;
                je = (int)aae; aae = aae - je;
            } else {
                double tau = e / RM;
                double b2 = tau * (tau + 2.0) / pow(tau + 1.0, 2.0);
                aae = (b2 - t->b2spin_min) * t->dbeta2i;
                je = (int)aae; aae = aae - je;
                je = je + OMC_MXE_SPIN + 1;
            }
            si1e = (1 - aae) * eta_at(eta_a, je) + aae * eta_at(eta_a, je + 1);
            si1p = (1 - aae) * eta_at(eta_a, NSPIN1 + je) + aae * eta_at(eta_a, NSPIN1 + je + 1);
        }
        int neke = t->meke[imed];
        for (int i = 1; i < neke; i++) {
            eil = (i + 1.0 - eke0) / eke1;
            e = exp(eil);
            if (e <= espin_min) {
                si2e = eta_a[0]; si2p = eta_a[NSPIN1];
            } else {
                if (e <= espin_max) {
                    aae = (eil - t->espml) * t->dleneri;
                    je = (int)aae; aae = aae - je;
                } else {
                    double tau = e / RM;
                    double b2 = tau * (tau + 2.0) / ((tau + 1.0) * (tau + 1.0));
                    aae = (b2 - t->b2spin_min) * t->dbeta2i;
                    je = (int)aae; aae = aae - je;
                    je = je + OMC_MXE_SPIN + 1;
                }
                si2e = (1.0 - aae) * eta_at(eta_a, je) + aae * eta_at(eta_a, je + 1);
                si2p = (1.0 - aae) * eta_at(eta_a, NSPIN1 + je) + aae * eta_at(eta_a, NSPIN1 + je + 1);
            }
            e_pwlf(t->etae_ms0, t->etae_ms1, o, i, si2e, si1e, eke1, eil);
            e_pwlf(t->etap_ms0, t->etap_ms1, o, i, si2p, si1p, eke1, eil);
            si1e = si2e; si1p = si2p;
        }
        t->etae_ms1[o + neke - 1] = t->etae_ms1[o + neke - 2];
        t->etae_ms0[o + neke - 1] = t->etae_ms0[o + neke - 2];
        t->etap_ms1[o + neke - 1] = t->etap_ms1[o + neke - 2];
        t->etap_ms0[o + neke - 1] = t->etap_ms0[o + neke - 2];

        /* first- and second-moment spin corrections: spline over the spin
         * energy grid (the beta^2 node at index MXE_SPIN+1 is skipped) */
        for (int i = 0; i <= OMC_MXE_SPIN; i++) elarray[i] = log(earray[i] / 1000.0);
        for (int i = OMC_MXE_SPIN + 1; i <= OMC_MXE_SPIN1 - 1; i++) elarray[i] = log(earray[i + 1] / 1000.0);
        int ndata = NSPIN1;
        elarray[ndata - 1] = t->ue[imed] > 1.0E5 ? log(t->ue[imed]) : log(1.0E5);
        double *dst0[4] = {t->q1ce_ms0, t->q1cp_ms0, t->q2ce_ms0, t->q2cp_ms0};
        double *dst1[4] = {t->q1ce_ms1, t->q1cp_ms1, t->q2ce_ms1, t->q2cp_ms1};
        const double *src[4] = {c_a, c_a + NSPIN1, g_a, g_a + NSPIN1};
        for (int s = 0; s < 4; s++) {
            for (int i = 0; i <= OMC_MXE_SPIN; i++) farray[i] = src[s][i];
            for (int i = OMC_MXE_SPIN + 1; i <= OMC_MXE_SPIN1 - 1; i++) farray[i] = src[s][i + 1];
            farray[ndata - 1] = 1.0;
            spline_setup(elarray, farray, af, bf, cf, df, ndata);
            eil = (1.0 - eke0) / eke1;
            si1e = spline_eval(eil, elarray, af, bf, cf, df, ndata);
            for (int i = 1; i <= neke - 1; i++) {
                eil = (i + 1 - eke0) / eke1;
                si2e = spline_eval(eil, elarray, af, bf, cf, df, ndata);
                e_pwlf(dst0[s], dst1[s], o, i, si2e, si1e, eke1, eil);
                si1e = si2e;
            }
            dst1[s][o + neke - 1] = dst1[s][o + neke - 2];
            /* the reference copies the slope into the q1ce intercept of the
             * last interval (src/ompmc.c:2802-2803); kept for parity */
            dst0[s][o + neke - 1] = s == 0 ? dst1[s][o + neke - 2] : dst0[s][o + neke - 2];
        }
        /* subtract the scattering power already in discrete Moller/Bhabha events */
        double tauc = t->te[imed] / RM;
        si1e = 1.0;
        for (int i = 1; i <= neke - 1; i++) {
            eil = ((double)(i + 1) - eke0) / eke1;
            e = exp(eil);
            int leil = i;
            double tau = e / RM;
            if (tau > 2.0 * tauc) {
                double sig = t->esig1[o + leil] * eil + t->esig0[o + leil];
                double dedx = t->ededx1[o + leil] * eil + t->ededx0[o + leil];
                sig /= dedx;
                if (sig > 1.0E-6) {
                    double etap = t->etae_ms1[o + leil] * eil + t->etae_ms0[o + leil];
                    double eta = 0.25 * etap * (t->xcc[imed]) / (t->blcc[imed]) / tau / (tau + 2);
                    double g_r = (1.0 + 2.0 * eta) * log(1.0 + 1.0 / eta) - 2.0;
                    double g_m = log(0.5 * tau / tauc) +
                                 (1.0 + ((tau + 2.0) / (tau + 1.0)) * ((tau + 2.0) / (tau + 1.0))) *
                                     log(2.0 * (tau - tauc + 2.0) / (tau + 4.0)) -
                                 0.25 * (tau + 2.0) * (tau + 2.0 + 2.0 * (2.0 * tau + 1.0) / ((tau + 1.0) * (tau + 1.0))) *
                                     log((tau + 4.0) * (tau - tauc) / tau / (tau - tauc + 2.0)) +
                                 0.5 * (tau - 2.0 * tauc) * (tau + 2.0) * (1.0 / (tau - tauc) - 1.0 / ((tau + 1.0) * (tau + 1.0)));
                    g_m = g_m < g_r ? g_m / g_r : 1.0;
                    si2e = 1.0 - g_m * sum_Z / sum_Z2;
                } else {
                    si2e = 1.0;
                }
            } else {
                si2e = 1.0;
            }
            e_pwlf(t->blcce0, t->blcce1, o, i, si2e, si1e, eke1, eil);
            si1e = si2e;
        }
        t->blcce1[o + neke - 1] = t->blcce1[o + neke - 2];
        t->blcce0[o + neke - 1] = t->blcce0[o + neke - 2];
        (void)sum_A; (void)sum_pz;
    }
    free(raw);
    return OMC_OK;
}

/* ------------------------------------------------------------------ multiple scattering */
/* reference src/ompmc.c:3219-3298 */
static int read_msnew(omc_tables *t, const char *folder) {
    char path[1024];
    snprintf(path, sizeof path, "%smsnew.data", folder);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    for (int i = 0; i <= OMC_MXL_MS; i++)
        for (int j = 0; j <= OMC_MXQ_MS; j++) {
            int base = i * (OMC_MXQ_MS + 1) * (OMC_MXU_MS + 1) + j * (OMC_MXU_MS + 1);
            int ok = 1;
            for (int k = 0; k <= OMC_MXU_MS; k++) ok &= fscanf(fp, "%lf", &t->ums[base + k]) == 1;
            for (int k = 0; k <= OMC_MXU_MS; k++) ok &= fscanf(fp, "%lf", &t->fms[base + k]) == 1;
            for (int k = 0; k < OMC_MXU_MS; k++) ok &= fscanf(fp, "%lf", &t->wms[base + k]) == 1;
            for (int k = 0; k < OMC_MXU_MS; k++) ok &= fscanf(fp, "%d", &t->ims[base + k]) == 1;
            if (!ok) { fclose(fp); return omc_host_set_error(OMC_ERR_FORMAT, "%s truncated", path); }
            for (int k = 0; k < OMC_MXU_MS; k++) {
                t->fms[base + k] = t->fms[base + k + 1] / t->fms[base + k] - 1.0;
                t->ims[base + k] = t->ims[base + k] - 1;
            }
            t->fms[base + OMC_MXU_MS] = t->fms[base + OMC_MXU_MS - 1];
            /* the reference leaves the last alias entry unset; it is never read */
            t->wms[base + OMC_MXU_MS] = 0.0;
            t->ims[base + OMC_MXU_MS] = 0;
        }
    fclose(fp);
    double llammin = log(OMC_LAMBMIN_MS), llammax = log(OMC_LAMBMAX_MS);
    double dllamb = (llammax - llammin) / OMC_MXL_MS;
    t->dllambi = 1.0 / dllamb;
    double dqms = OMC_QMAX_MS / OMC_MXQ_MS;
    t->dqmsi = 1.0 / dqms;
    return OMC_OK;
}

/* reference src/ompmc.c:3300-3529 */
static int build_mscat(omc_tables *t, const char *folder) {
    int rc = read_msnew(t, folder);
    if (rc) return rc;
    for (int m = 0; m < t->nmed; m++) {
        t->blcc[m] = 1.16699413758864573 * t->blcc[m];
        t->xcc[m] = pow(t->xcc[m], 2.0);
    }
    if ((rc = build_spin(t, folder))) return rc;
    int nmed = t->nmed;
    for (int m = 0; m < nmed; m++) {
        double sigee = 1.0E-15, sigep = 1.0E-15, sige_o = -1.0, sigp_o = -1.0;
        int emono = 1, pmono = 1;
        int o = m * MXEKE;
        for (int i = 1; i <= t->meke[m]; i++) {
            double ei = exp(((double)i - t->eke0[m]) / t->eke1[m]);
            double eil = log(ei);
            int l = i - 1;
            double dedx = t->ededx1[o + l] * eil + t->ededx0[o + l];
            double sig = t->esig1[o + l] * eil + t->esig0[o + l];
            sig /= dedx;
            if (sig > sigee) sigee = sig;
            if (sig < sige_o) emono = 0;
            sige_o = sig;
            dedx = t->pdedx1[o + l] * eil + t->pdedx0[o + l];
            sig = t->psig1[o + l] * eil + t->psig0[o + l];
            sig /= dedx;
            if (sig > sigep) sigep = sig;
            if (sig < sigp_o) pmono = 0;
            sigp_o = sig;
        }
        t->sig_ismonotone[m] = emono;
        t->sig_ismonotone[nmed + m] = pmono;
        t->esig_e[m] = sigee;
        t->psig_e[m] = sigep;
    }
    for (int m = 0; m < nmed; m++) {
        int o = m * MXEKE;
        double *re = t->range_ep + o, *rp = t->range_ep + (size_t)nmed * MXEKE + o;
        double eke0 = t->eke0[m], eke1 = t->eke1[m];
        double ei = exp((1.0 - eke0) / eke1);
        t->e_array[o] = ei;
        t->expeke1[m] = exp(1.0 / eke1) - 1.0;
        re[0] = 0.0; rp[0] = 0.0;
        int neke = t->meke[m];
        /* CSDA range on the grid: log-interpolated dE/dx at the midpoint plus
         * the second-order series term */
        for (int i = 1; i <= neke - 1; i++) {
            double eip1 = exp(((double)(i + 1) - eke0) / eke1);
            t->e_array[o + i] = eip1;
            double eke = 0.5 * (eip1 + ei);
            double elke = log(eke);
            int lelke = (int)(eke1 * elke + eke0) - 1;
            double dedx = t->pdedx1[o + lelke] * elke + t->pdedx0[o + lelke];
            double aux = t->pdedx1[o + i - 1] / dedx;
            rp[i] = rp[i - 1] + (eip1 - ei) / dedx * (1.0 + aux * (1.0 + 2.0 * aux) * pow((eip1 - ei) / eke, 2.0) / 24.0);
            dedx = t->ededx1[o + lelke] * elke + t->ededx0[o + lelke];
            aux = t->ededx1[o + i - 1] / dedx;
            re[i] = re[i - 1] + (eip1 - ei) / dedx * (1.0 + aux * (1.0 + 2.0 * aux) * pow(((eip1 - ei) / eke), 2.0) / 24.0);
            ei = eip1;
        }
        /* maximum step for multiple scattering (tmxs), PEGS values replaced */
        double eil = (1.0 - eke0) / eke1;
        ei = exp(eil);
        int leil = 1;
        double p2 = ei * (ei + 2.0 * RM);
        double beta2 = p2 / (p2 + pow(RM, 2.0));
        double chi_a2 = t->xcc[m] / (4.0 * p2 * t->blcc[m]);
        double dedx0 = t->ededx1[o + leil] * eil + t->ededx0[o + leil];
        double estepx = 2.0 * p2 * beta2 * dedx0 / ei / t->xcc[m] / (log(1.0 + 1.0 / chi_a2) * (1.0 + chi_a2) - 1.0);
        estepx *= OMC_XIMAX;
        if (estepx > OMC_ESTEPE) estepx = OMC_ESTEPE;
        double si = estepx * ei / dedx0;
        for (int i = 1; i <= neke - 1; i++) {
            double elke = ((double)(i + 1) - eke0) / eke1;
            double eke = exp(elke);
            int lelke = i;
            p2 = eke * (eke + 2.0 * RM);
            beta2 = p2 / (p2 + pow(RM, 2.0));
            chi_a2 = t->xcc[m] / (4.0 * p2 * t->blcc[m]);
            double dedx = t->ededx1[o + lelke] * elke + t->ededx0[o + lelke];
            estepx = 2.0 * p2 * beta2 * dedx / eke / (t->xcc[m]) / (log(1.0 + 1.0 / chi_a2) * (1.0 + chi_a2) - 1.0);
            estepx = estepx * OMC_XIMAX;
            if (estepx > OMC_ESTEPE) estepx = OMC_ESTEPE;
            double ekef = (1.0 - estepx) * eke;
            double sip1;
            int lelkef;
            if (ekef <= t->e_array[o]) {
                sip1 = (t->e_array[o] - ekef) / dedx0;
                ekef = t->e_array[o];
                lelkef = 0;
            } else {
                double elkef = log(ekef);
                lelkef = eke1 * elkef + eke0 - 1;
                int leip1l = lelkef + 1;
                double eip1l = ((double)(leip1l + 1) - eke0) / eke1;
                double eip1 = t->e_array[o + leip1l];
                double aux = (eip1 - ekef) / eip1;
                double elktmp = 0.5 * (elkef + eip1l + 0.25 * aux * aux * (1.0 + aux * (1.0 + 0.875 * aux)));
                double ektmp = 0.5 * (ekef + eip1);
                int lelktmp = lelkef;
                dedx = t->ededx1[o + lelktmp] * elktmp + t->ededx0[o + lelktmp];
                aux = t->ededx1[o + lelktmp] / dedx;
                sip1 = (eip1 - ekef) / dedx * (1.0 + aux * (1.0 + 2.0 * aux) * (pow(((eip1 - ekef) / ektmp), 2.0) / 24.0));
            }
            sip1 += re[lelke] - re[lelkef + 1];
            t->tmxs1[o + i - 1] = (sip1 - si) * eke1;
            t->tmxs0[o + i - 1] = sip1 - t->tmxs1[o + i - 1] * elke;
            si = sip1;
        }
        t->tmxs0[o + neke - 1] = t->tmxs0[o + neke - 2];
        t->tmxs1[o + neke - 1] = t->tmxs1[o + neke - 2];
    }
    return OMC_OK;
}

/* ------------------------------------------------------------------ driver */
void omc_tables_free(omc_tables *t) {
    double **dp[] = {&t->gmfp0, &t->gmfp1, &t->gbr10, &t->gbr11, &t->gbr20, &t->gbr21, &t->cohe0, &t->cohe1,
                     &t->ray_xgrid, &t->ray_fcum, &t->ray_b, &t->ray_c, &t->ray_pmax0, &t->ray_pmax1,
                     &t->esig0, &t->esig1, &t->psig0, &t->psig1, &t->ededx0, &t->ededx1, &t->pdedx0, &t->pdedx1,
                     &t->ebr10, &t->ebr11, &t->pbr10, &t->pbr11, &t->pbr20, &t->pbr21, &t->tmxs0, &t->tmxs1,
                     &t->blcce0, &t->blcce1, &t->etae_ms0, &t->etae_ms1, &t->etap_ms0, &t->etap_ms1,
                     &t->q1ce_ms0, &t->q1ce_ms1, &t->q1cp_ms0, &t->q1cp_ms1, &t->q2ce_ms0, &t->q2ce_ms1,
                     &t->q2cp_ms0, &t->q2cp_ms1, &t->e_array, &t->range_ep, &t->spin_rej, &t->ums, &t->fms,
                     &t->wms};
    for (size_t i = 0; i < sizeof dp / sizeof dp[0]; i++) { free(*dp[i]); *dp[i] = NULL; }
    free(t->ray_iarr); t->ray_iarr = NULL;
    free(t->ims); t->ims = NULL;
}

static int alloc_tables(omc_tables *t) {
    size_t nm = (size_t)t->nmed;
    double **ge[] = {&t->gmfp0, &t->gmfp1, &t->gbr10, &t->gbr11, &t->gbr20, &t->gbr21, &t->cohe0, &t->cohe1,
                     &t->ray_pmax0, &t->ray_pmax1};
    for (size_t i = 0; i < sizeof ge / sizeof ge[0]; i++) *ge[i] = calloc(nm * MXGE, sizeof(double));
    double **rf[] = {&t->ray_xgrid, &t->ray_fcum, &t->ray_b, &t->ray_c};
    for (size_t i = 0; i < 4; i++) *rf[i] = calloc(nm * MXRAYFF, sizeof(double));
    t->ray_iarr = calloc(nm * RAYCDFSIZE, sizeof(int));
    double **ek[] = {&t->esig0, &t->esig1, &t->psig0, &t->psig1, &t->ededx0, &t->ededx1, &t->pdedx0, &t->pdedx1,
                     &t->ebr10, &t->ebr11, &t->pbr10, &t->pbr11, &t->pbr20, &t->pbr21, &t->tmxs0, &t->tmxs1,
                     &t->blcce0, &t->blcce1, &t->etae_ms0, &t->etae_ms1, &t->etap_ms0, &t->etap_ms1,
                     &t->q1ce_ms0, &t->q1ce_ms1, &t->q1cp_ms0, &t->q1cp_ms1, &t->q2ce_ms0, &t->q2ce_ms1,
                     &t->q2cp_ms0, &t->q2cp_ms1, &t->e_array};
    for (size_t i = 0; i < sizeof ek / sizeof ek[0]; i++) *ek[i] = calloc(nm * MXEKE, sizeof(double));
    t->range_ep = calloc(2 * nm * MXEKE, sizeof(double));
    /* one zero entry past the end: the rejection lookup interpolates with
     * the next entry, which for cos(theta) = -1 on the last row of the last
     * medium lies one past the table (reference quirk Q17) */
    t->spin_rej = calloc(nm * OMC_NSPIN_MED + 1, sizeof(double));
    t->ums = calloc(OMC_NMS, sizeof(double));
    t->fms = calloc(OMC_NMS, sizeof(double));
    t->wms = calloc(OMC_NMS, sizeof(double));
    t->ims = calloc(OMC_NMS, sizeof(int));
    if (!t->gmfp0 || !t->spin_rej || !t->ims || !t->range_ep || !t->e_array || !t->ray_iarr)
        return omc_host_set_error(OMC_ERR_NOMEM, "table allocation failed");
    return OMC_OK;
}

int omc_tables_build(omc_tables *t, int nmed, const char names[][60], const omc_paths *p) {
    memset(t, 0, sizeof *t);
    if (nmed < 1 || nmed > OMC_MXMED) return omc_host_set_error(OMC_ERR_LIMIT, "nmed %d outside 1..%d", nmed, OMC_MXMED);
    if (!p || !p->pegs_file || !p->pgs4form_file || !p->data_folder)
        return omc_host_set_error(OMC_ERR_ARG, "missing data paths");
    t->nmed = nmed;
    for (int i = 0; i < nmed; i++) snprintf(t->med_names[i], 60, "%s", names[i]);
    int rc = alloc_tables(t);
    char folder[1024];
    size_t fl = strlen(p->data_folder);
    snprintf(folder, sizeof folder, "%s%s", p->data_folder, (fl && p->data_folder[fl - 1] == '/') ? "" : "/");
    if (!rc) rc = read_pegs(t, p->pegs_file);
    if (!rc) rc = build_photon(t, folder);
    if (!rc) rc = build_rayleigh(t, p->pgs4form_file);
    if (!rc) build_pair(t);
    if (!rc) rc = build_mscat(t, folder);
    if (rc) omc_tables_free(t);
    return rc;
}

/* ------------------------------------------------------------------ regions */
/* reference ucodes/omc_dosxyz/omc_dosxyz.c:890-962.  Where the PEGS AE lies
 * above the global ECUT the reference leaves the region's ECUT unset while
 * announcing that the PEGS value is used; the PEGS value is used here. */
int omc_regions_build(omc_regions *r, const omc_phantom *ph, const omc_tables *t, double ecut, double pcut) {
    long nvox = (long)ph->nx * ph->ny * ph->nz;
    r->nreg = (int)(nvox + 1);
    r->med = malloc(sizeof(int) * r->nreg);
    r->rhof = malloc(sizeof(double) * r->nreg);
    r->pcut = malloc(sizeof(double) * r->nreg);
    r->ecut = malloc(sizeof(double) * r->nreg);
    if (!r->med || !r->rhof || !r->pcut || !r->ecut) { omc_regions_free(r); return omc_host_set_error(OMC_ERR_NOMEM, "regions"); }
    r->med[0] = -1; r->rhof[0] = 0.0; r->pcut[0] = 0.0; r->ecut[0] = 0.0;
    for (long i = 1; i <= nvox; i++) {
        int imed = ph->med[i - 1] - 1;
        r->med[i] = imed;
        if (imed == -1) {
            r->rhof[i] = 0.0; r->pcut[i] = 0.0; r->ecut[i] = 0.0;
            continue;
        }
        if (imed >= t->nmed) { omc_regions_free(r); return omc_host_set_error(OMC_ERR_FORMAT, "voxel medium %d", imed + 1); }
        r->rhof[i] = ph->rho[i - 1] == 0.0F ? 1.0 : ph->rho[i - 1] / t->rho[imed];
        r->pcut[i] = t->ap[imed] <= pcut ? pcut : t->ap[imed];
        r->ecut[i] = t->ae[imed] <= ecut ? ecut : t->ae[imed];
    }
    return OMC_OK;
}

void omc_regions_free(omc_regions *r) {
    free(r->med); free(r->rhof); free(r->pcut); free(r->ecut);
    memset(r, 0, sizeof *r);
}

/* ------------------------------------------------------------------ source */
void omc_spectrum_mono(omc_spectrum *s, double energy) {
    memset(s, 0, sizeof *s);
    s->spectrum = 0;
    s->energy = energy;
}

/* reference ucodes/omc_dosxyz/omc_dosxyz.c:396-510 */
int omc_spectrum_read(omc_spectrum *s, const char *path) {
    memset(s, 0, sizeof *s);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    char buf[512];
    int nens = 0, imode = -1;
    double enmin = 0.0;
    if (!fgets(buf, sizeof buf, fp) || !fgets(buf, sizeof buf, fp) || sscanf(buf, "%d %lf %d", &nens, &enmin, &imode) != 3) {
        fclose(fp);
        return omc_host_set_error(OMC_ERR_FORMAT, "%s: malformed spectrum header", path);
    }
    if (nens > OMC_MXEBIN || nens < 1) {
        fclose(fp);
        return omc_host_set_error(OMC_ERR_LIMIT, "%s: %d energy bins (1..%d supported)", path, nens, OMC_MXEBIN);
    }
    double ens[OMC_MXEBIN], pdf[OMC_MXEBIN], cdf[OMC_MXEBIN];
    for (int i = 0; i < nens; i++) {
        if (!fgets(buf, sizeof buf, fp) || sscanf(buf, "%lf %lf", &ens[i], &pdf[i]) != 2) {
            fclose(fp);
            return omc_host_set_error(OMC_ERR_FORMAT, "%s: bin %d unreadable", path, i);
        }
    }
    fclose(fp);
    if (imode == 1) {
        pdf[0] *= (ens[0] - enmin);
        for (int i = 1; i < nens; i++) pdf[i] *= (ens[i] - ens[i - 1]);
    } else if (imode != 0) {
        return omc_host_set_error(OMC_ERR_FORMAT, "%s: invalid mode number %d", path, imode);
    }
    cdf[0] = pdf[0];
    for (int i = 1; i < nens; i++) cdf[i] = cdf[i - 1] + pdf[i];
    double fnorm = 1.0 / cdf[nens - 1];
    s->spectrum = 1;
    s->deltak = OMC_INVDIM;
    double gridsz = 1.0f / s->deltak;
    for (int i = 0; i < nens; i++) cdf[i] *= fnorm;
    for (int k = 0; k < s->deltak; k++) {
        double ak = (double)k * gridsz;
        int i;
        for (i = 0; i < nens; i++)
            if (ak <= cdf[i]) break;
        if (i >= nens) i = nens - 1;   /* unreachable for a normalised CDF */
        s->cdfinv1[k] = i != 0 ? ens[i - 1] : enmin;
        s->cdfinv2[k] = ens[i] - s->cdfinv1[k];
    }
    return OMC_OK;
}

/* reference ucodes/omc_dosxyz/omc_dosxyz.c:521-617 */
int omc_dosxyz_source_init(omc_dosxyz_source *s, const omc_phantom *ph, const double col[4], double ssd, int charge) {
    if (charge < -1 || charge > 1) return omc_host_set_error(OMC_ERR_ARG, "Particle kind not recognized.");
    if (ssd < 0) return omc_host_set_error(OMC_ERR_ARG, "SSD must be greater than zero.");
    s->charge = charge;
    s->ssd = ssd;
    s->xinl = col[0]; s->xinu = col[1]; s->yinl = col[2]; s->yinu = col[3];
    const double *xb = ph->xb, *yb = ph->yb;
    if (s->xinl < xb[0]) s->xinl = xb[0];
    if (s->xinu <= s->xinl) s->xinu = s->xinl;
    if (s->xinu > xb[ph->nx]) s->xinu = xb[ph->nx];
    if (s->xinl > xb[ph->nx]) s->xinl = xb[ph->nx];
    s->ixinl = 0;
    while (s->ixinl < ph->nx - 1 && xb[s->ixinl] <= s->xinl && xb[s->ixinl + 1] < s->xinl) s->ixinl++;
    if (s->yinl < yb[0]) s->yinl = yb[0];
    if (s->yinu <= s->yinl) s->yinu = s->yinl;
    if (s->yinu > yb[ph->ny]) s->yinu = yb[ph->ny];
    if (s->yinl > yb[ph->ny]) s->yinl = yb[ph->ny];
    s->iyinl = 0;
    while (s->iyinl < ph->ny - 1 && yb[s->iyinl] <= s->yinl && yb[s->iyinl + 1] < s->yinl) s->iyinl++;
    s->xsize = s->xinu - s->xinl;
    s->ysize = s->yinu - s->yinl;
    return OMC_OK;
}

/* ------------------------------------------------------------------ scoring */
/* reference ucodes/omc_dosxyz/omc_dosxyz.c:719-799 (iout = 1) */
void omc_dosxyz_finish(const omc_phantom *ph, const double *sum, const double *sum2, int nperbatch, int nbatch,
                       double *dose, double *unc) {
    int nx = ph->nx, ny = ph->ny, nz = ph->nz;
    double inc = (double)nperbatch;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                long irl = 1 + ix + (long)iy * nx + (long)iz * nx * ny;
                double e = sum[irl], e2 = sum2[irl], u;
                e /= (double)nbatch;
                e2 /= (double)nbatch;
                if (e != 0.0) {
                    u = e2 - e * e;
                    u /= (double)(nbatch - 1);
                    u = sqrt(u) / e;
                } else {
                    e = 0.0;
                    u = 0.9999999;
                }
                double mass = (ph->xb[ix + 1] - ph->xb[ix]) * (ph->yb[iy + 1] - ph->yb[iy]) * (ph->zb[iz + 1] - ph->zb[iz]);
                mass *= ph->rho[irl - 1];
                e *= 1.602E-10 / (mass * inc);
                if (ph->rho[irl - 1] < 0.044) { e = 0.0; u = 0.9999999; }
                dose[irl - 1] = e;
                unc[irl - 1] = u;
            }
}

int omc_write_3ddose(const char *path, const omc_phantom *ph, const double *dose, const double *unc) {
    FILE *fp = fopen(path, "w");
    if (!fp) return omc_host_set_error(OMC_ERR_IO, "Unable to open file: %s", path);
    long nvox = (long)ph->nx * ph->ny * ph->nz;
    fprintf(fp, "%5d%5d%5d\n", ph->nx, ph->ny, ph->nz);
    for (int i = 0; i <= ph->nx; i++) fprintf(fp, "%f ", ph->xb[i]);
    fprintf(fp, "\n");
    for (int i = 0; i <= ph->ny; i++) fprintf(fp, "%f ", ph->yb[i]);
    fprintf(fp, "\n");
    for (int i = 0; i <= ph->nz; i++) fprintf(fp, "%f ", ph->zb[i]);
    fprintf(fp, "\n");
    for (long v = 0; v < nvox; v++) fprintf(fp, "%e ", dose[v]);
    fprintf(fp, "\n");
    for (long v = 0; v < nvox; v++) fprintf(fp, "%f ", unc[v]);
    fprintf(fp, "\n");
    if (fclose(fp) != 0) return omc_host_set_error(OMC_ERR_IO, "write failed: %s", path);
    return OMC_OK;
}

/* ------------------------------------------------------------------ RANMAR */
/* reference src/omc_random.c:63-150 */
void omc_ranmar_init(omc_ranmar *r, int ixx, int jxx) {
    if (ixx <= 0 || ixx > 31328) ixx = 1802;
    if (jxx <= 0 || jxx > 31328) jxx = 9373;
    int i = (ixx / 177 % 177) + 2, j = (ixx % 177) + 2;
    int k = (jxx / 169 % 178) + 1, l = (jxx % 169);
    for (int ii = 0; ii < 97; ii++) {
        int s = 0, tt = 8388608;
        for (int jj = 0; jj < 24; jj++) {
            int m = ((i * j % 179) * k) % 179;
            i = j; j = k; k = m;
            l = (53 * l + 1) % 169;
            if (l * m % 64 >= 32) s += tt;
            tt /= 2;
        }
        r->u[ii] = s;
    }
    r->c = 362436; r->cd = 7654321; r->cm = 16777213;
    r->i97 = 97; r->j97 = 33;
    r->pos = 128;
}

static void ranmar_refill(omc_ranmar *r) {
    for (int n = 0; n < 128; n++) {
        int v = r->u[r->i97 - 1] - r->u[r->j97 - 1];
        if (v < 0) v += 16777216;
        r->u[r->i97 - 1] = v;
        r->i97 -= 1; r->j97 -= 1;
        if (r->i97 == 0) r->i97 = 97;
        else if (r->j97 == 0) r->j97 = 97;
        r->c -= r->cd;
        if (r->c < 0) r->c += r->cm;
        v -= r->c;
        if (v < 0) v += 16777216;
        r->buf[n] = v;
    }
    r->pos = 0;
}

int omc_ranmar_next24(omc_ranmar *r) {
    if (r->pos >= 128) ranmar_refill(r);
    return r->buf[r->pos++];
}

double omc_ranmar_next(omc_ranmar *r) { return omc_ranmar_next24(r) * (1.0 / 16777216.0); }
