/*
 * Host-side C layer of the transport engine: input formats and the physics
 * table builders.  Everything here runs once per process (tables) or once per
 * request (phantom, source); the transport itself runs on the GPU.
 *
 * Table layouts are flat FP64 arrays indexed [imed*N + i] exactly as the
 * reference builds them (src/ompmc.h:107-435), so the FP64 tables can be
 * compared bit for bit with the reference's; the GPU upload re-packs them
 * (interleaved coefficient pairs, FP32 or FP64).
 */
#ifndef OMC_HOST_H
#define OMC_HOST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limits and constants of the reference (src/ompmc.h, src/omc_random.h). */
#define OMC_RM 0.5109989461
#define OMC_MXSTACK 10000
#define OMC_MXGE 2000
#define OMC_SGMFP 1.0E-05
#define OMC_MXRAYFF 100
#define OMC_RAYCDFSIZE 100
#define OMC_HC_INVERSE 80.65506856998
#define OMC_TWICE_HC2 0.000307444456
#define OMC_FSC 0.00729735255664
#define OMC_XIMAX 0.5
#define OMC_ESTEPE 0.25
#define OMC_EPSEMFP 1.0E-5
#define OMC_SKIN_DEPTH_FOR_BCA 3
#define OMC_MXE_SPIN 15
#define OMC_MXE_SPIN1 (2 * OMC_MXE_SPIN + 1)
#define OMC_MXQ_SPIN 15
#define OMC_MXU_SPIN 31
#define OMC_NSPIN_MED (2 * (OMC_MXE_SPIN1 + 1) * (OMC_MXQ_SPIN + 1) * (OMC_MXU_SPIN + 1))
#define OMC_MXL_MS 63
#define OMC_MXQ_MS 7
#define OMC_MXU_MS 31
#define OMC_NMS ((OMC_MXL_MS + 1) * (OMC_MXQ_MS + 1) * (OMC_MXU_MS + 1))
#define OMC_LAMBMIN_MS 1.0
#define OMC_LAMBMAX_MS 1.0E5
#define OMC_QMIN_MS 1.0E-3
#define OMC_QMAX_MS 0.5
#define OMC_MXMED 9
#define OMC_MXELEMENT 50
#define OMC_MXEKE 500
#define OMC_MXEBIN 200
#define OMC_INVDIM 1000
#define OMC_PI 3.14159265358979323846

/* Error codes returned by every entry point (0 = success); the same values
 * as include/omc_gpu.h. */
#ifndef OMC_OK
#define OMC_OK 0
#define OMC_ERR_IO 1        /* file missing or unreadable */
#define OMC_ERR_FORMAT 2    /* malformed file content */
#define OMC_ERR_MEDIUM 3    /* medium not found in the PEGS4 file */
#define OMC_ERR_LIMIT 4     /* a compile-time limit of the reference exceeded */
#define OMC_ERR_ARG 5       /* invalid argument */
#define OMC_ERR_DEVICE 6    /* HIP runtime failure */
#define OMC_ERR_NOMEM 7     /* host or device allocation failed */
#define OMC_ERR_BUDGET 8    /* request exceeds the HBM admission budget */
#define OMC_ERR_STACK 9     /* particle stack overflow */
#define OMC_ERR_CANCELLED 10
#endif

/* Last error message of the calling thread. */
const char *omc_host_last_error(void);
int omc_host_set_error(int code, const char *fmt, ...);

/* ---------------------------------------------------------------- .inp */
/* key = value input file (reference src/omc_utilities.c:63-161).  Lines that
 * contain '#' anywhere or are blank are dropped; lookups match the key by
 * substring and return the first match; values keep their leading blank. */
#define OMC_INP_MAX 80
#define OMC_INP_LEN 256
typedef struct omc_inp {
    int count;
    char key[OMC_INP_MAX][OMC_INP_LEN];
    char value[OMC_INP_MAX][OMC_INP_LEN];
} omc_inp;

/* Parses "<stem>.inp". */
int omc_inp_parse(omc_inp *inp, const char *stem);
/* Returns 1 and copies the value when found, 0 otherwise.  Like the reference,
 * a file holding a single key yields no lookups (src/omc_utilities.c:91,111). */
int omc_inp_get(const omc_inp *inp, const char *key, char *dest, size_t cap);
/* Copies `src` without any whitespace (reference removeSpaces). */
void omc_strip_spaces(char *dst, const char *src, size_t cap);

/* ---------------------------------------------------------------- phantom */
typedef struct omc_phantom {
    int nmed;
    char med_names[OMC_MXMED][60];
    int nx, ny, nz;
    double *xb, *yb, *zb;        /* nx+1, ny+1, nz+1 boundaries (cm) */
    int *med;                    /* 1-based medium per voxel, x fastest; 0 = vacuum */
    double *rho;                 /* g/cm^3 per voxel */
    int owns;                    /* 1 if the arrays belong to this struct */
} omc_phantom;

/* egsphant reader (reference ucodes/omc_dosxyz/omc_dosxyz.c:62-175). */
int omc_read_egsphant(omc_phantom *ph, const char *path);
void omc_phantom_free(omc_phantom *ph);

/* ---------------------------------------------------------------- tables */
typedef struct omc_element { double z, wa, pz, rhoz; char symbol[3]; } omc_element;

typedef struct omc_tables {
    int nmed;
    char med_names[OMC_MXMED][60];
    /* PEGS4 per-medium data (reference struct Pegs, src/ompmc.h:376-405) */
    int ne[OMC_MXMED], meke[OMC_MXMED], mge[OMC_MXMED], msge[OMC_MXMED];
    int mseke[OMC_MXMED], mleke[OMC_MXMED], mcmfp[OMC_MXMED], mrange[OMC_MXMED];
    int iunrst[OMC_MXMED], epstfl[OMC_MXMED], iaprim[OMC_MXMED];
    double rho[OMC_MXMED], rlc[OMC_MXMED], ae[OMC_MXMED], ap[OMC_MXMED];
    double ue[OMC_MXMED], up[OMC_MXMED], te[OMC_MXMED], thmoll[OMC_MXMED];
    double delcm[OMC_MXMED];
    omc_element elements[OMC_MXMED][OMC_MXELEMENT];

    /* photon PWLF tables [nmed*MXGE] */
    double ge0[OMC_MXMED], ge1[OMC_MXMED];
    double *gmfp0, *gmfp1, *gbr10, *gbr11, *gbr20, *gbr21, *cohe0, *cohe1;
    /* Rayleigh */
    double *ray_xgrid, *ray_fcum, *ray_b, *ray_c;   /* [nmed*MXRAYFF] */
    double *ray_pmax0, *ray_pmax1;                  /* [nmed*MXGE] */
    int *ray_iarr;                                  /* [nmed*RAYCDFSIZE] */
    /* pair / brems parameters */
    double dl1[OMC_MXMED * 8], dl2[OMC_MXMED * 8], dl3[OMC_MXMED * 8];
    double dl4[OMC_MXMED * 8], dl5[OMC_MXMED * 8], dl6[OMC_MXMED * 8];
    double bpar0[OMC_MXMED], bpar1[OMC_MXMED], zbrang[OMC_MXMED];
    /* electron PWLF tables [nmed*MXEKE] */
    double *esig0, *esig1, *psig0, *psig1, *ededx0, *ededx1, *pdedx0, *pdedx1;
    double *ebr10, *ebr11, *pbr10, *pbr11, *pbr20, *pbr21, *tmxs0, *tmxs1;
    double *blcce0, *blcce1, *etae_ms0, *etae_ms1, *etap_ms0, *etap_ms1;
    double *q1ce_ms0, *q1ce_ms1, *q1cp_ms0, *q1cp_ms1, *q2ce_ms0, *q2ce_ms1;
    double *q2cp_ms0, *q2cp_ms1, *e_array;
    double *range_ep;                               /* [2*nmed*MXEKE] */
    double eke0[OMC_MXMED], eke1[OMC_MXMED];
    int sig_ismonotone[2 * OMC_MXMED];
    double esig_e[OMC_MXMED], psig_e[OMC_MXMED];
    double xcc[OMC_MXMED], blcc[OMC_MXMED], expeke1[OMC_MXMED];
    /* spin rejection surfaces [nmed][2][32][16][32] and interpolation scalars */
    double b2spin_min, dbeta2i, espml, dleneri, dqq1i;
    double *spin_rej;
    /* screened Rutherford multiple scattering (msnew) */
    double *ums, *fms, *wms;
    int *ims;
    double dllambi, dqmsi;
} omc_tables;

typedef struct omc_paths {
    const char *pegs_file;      /* .pegs4dat */
    const char *pgs4form_file;  /* pgs4form.dat */
    const char *data_folder;    /* holds xcom_*.data, msnew.data, spinms.data; trailing '/' optional */
} omc_paths;

/* Builds every table for the named media (reference initMediaData,
 * src/ompmc.c:5468-5505). */
int omc_tables_build(omc_tables *t, int nmed, const char names[][60], const omc_paths *paths);
void omc_tables_free(omc_tables *t);

/* ---------------------------------------------------------------- regions */
/* Per-region transport parameters (reference initRegions,
 * ucodes/omc_dosxyz/omc_dosxyz.c:890-962).  Region 0 is outside the phantom. */
typedef struct omc_regions {
    int nreg;           /* nvox + 1 */
    int *med;           /* 0-based medium, -1 vacuum */
    double *rhof, *pcut, *ecut;
} omc_regions;

int omc_regions_build(omc_regions *r, const omc_phantom *ph, const omc_tables *t,
                      double ecut, double pcut);
void omc_regions_free(omc_regions *r);

/* ---------------------------------------------------------------- source */
/* Spectrum inverse CDF (reference omc_dosxyz.c:383-510; omc_matrad.c:627-769). */
typedef struct omc_spectrum {
    int spectrum;           /* 0 mono, 1 spectrum */
    double energy;          /* mono energy (MeV) */
    double deltak;          /* INVDIM */
    double cdfinv1[OMC_INVDIM], cdfinv2[OMC_INVDIM];
} omc_spectrum;

int omc_spectrum_read(omc_spectrum *s, const char *path);
void omc_spectrum_mono(omc_spectrum *s, double energy);

/* Collimated point source of omc_dosxyz (omc_dosxyz.c:521-632). */
typedef struct omc_dosxyz_source {
    omc_spectrum spec;
    int charge;
    double ssd;
    double xinl, xinu, yinl, yinu, xsize, ysize;
    int ixinl, iyinl;
} omc_dosxyz_source;

int omc_dosxyz_source_init(omc_dosxyz_source *s, const omc_phantom *ph, const double collimator[4],
                           double ssd, int charge);

/* ---------------------------------------------------------------- scoring */
/* dosxyz dose and relative batch uncertainty (omc_dosxyz.c:719-799):
 * sum/sum2 are the per-batch accumulators [nreg]; writes dose/unc [nvox]. */
void omc_dosxyz_finish(const omc_phantom *ph, const double *sum, const double *sum2,
                       int nperbatch, int nbatch, double *dose, double *unc);
/* .3ddose writer (omc_dosxyz.c:841-879). */
int omc_write_3ddose(const char *path, const omc_phantom *ph, const double *dose, const double *unc);

/* ---------------------------------------------------------------- RANMAR */
/* Marsaglia-Zaman RANMAR as the reference implements it
 * (src/omc_random.c:63-192): 24-bit integers, refilled 128 at a time. */
typedef struct omc_ranmar {
    int u[97];
    int c, cd, cm, i97, j97;
    int buf[128];
    int pos;
} omc_ranmar;

void omc_ranmar_init(omc_ranmar *r, int ixx, int jxx);
/* next value as the 24-bit integer k (the random number is k * 2^-24) */
int omc_ranmar_next24(omc_ranmar *r);
double omc_ranmar_next(omc_ranmar *r);

#ifdef __cplusplus
}
#endif
#endif
