/*
 * FP64 CPU restatement of Radtrans' CIRSRADG_WAVE (correlated-k thermal
 * emission with analytic Jacobian), used as TEST INFRASTRUCTURE only: the
 * parity tests, the smoke check and the CPU-baseline leg of the benchmark call
 * it as the checker.  The product library never links or calls it.
 *
 * Each function cites the reference file:line it restates.  All REAL
 * quantities are double, as in the reference built with REAL promoted to
 * 8 bytes.
 */
#ifndef RADG_ORACLE_H
#define RADG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One gas's k-table (the content of a .kta file, read_khead.f:38-109). */
typedef struct orc_ktable {
  int idgas, isogas;
  int npoint, np, nt, ng;
  double vmin, delv;
  const double *lnp;   /* [np]  ln(P / atm), as read_khead.f:80-85 stores it */
  const double *tk;    /* [nt]  K */
  const double *g_ord; /* [ng] */
  const double *delg;  /* [ng] */
  const double *k;     /* [npoint][np][nt][ng], k * 1e20 */
} orc_ktable;

/* The /CIATABLE/ block (ciacom.f:9-18): kcia(9,25,1501) == [1501][25][9];
 * with npara > 0 also /FPTABLE/ (the para-H2 tables fpread_all.f reads):
 * kh2h2/kh2he(npara,25,1501) == [1501][25][npara], kcia5(5,25,1501). */
typedef struct orc_cia {
  double dnu;
  const double *temps; /* [25] */
  const double *kcia;  /* [1501][25][9] */
  int npara;
  double dnu2;
  const double *temps2, *fp, *kh2h2, *kh2he, *kcia5;
  /* /E3TABLE/ (e3interp.f): 1001 transmissions and E3 values (NULL: none;
   * imod 18 needs it) */
  const double *e3tran, *e3val;
  /* N2's near-infrared CIA (n2n2cia.f, n2h2cia.f) as those functions return
   * it at their knots, cm-1 amagat-2: N2-N2 on wn1[n1], N2-H2 on wn2[n2]
   * (n1 = n2 = 0: none; then N2 is refused at 3995-4995 cm-1) */
  int n1, n2;
  const double *wn1, *k1, *wn2, *k2;
  /* /SOLARDAT/ stellar spectrum (iread = 999 when nsol > 0): swave[nsol],
   * srad[nsol] (get_solar_wave.f) */
  int nsol;
  const double *swave, *srad;
} orc_cia;

/* The argument surface of cirsradg_wave (cirsradg_wave.f:1-5), compact C
 * layout: per-layer arrays [L], per-(layer,gas) arrays [L][ngas],
 * per-(path,layer) arrays [npath][L]. */
typedef struct orc_req {
  int nlayer, npath, ngas, nsw, nem;
  int inormal, iray, ispace, itype, flagh2p, amform;
  double dist, tsurf, radius1, radextra;
  const double *press, *temp, *delh, *hfp;
  const double *pp, *amount;
  const int *idgas, *isogas;
  const int *nlayin, *imod;
  const int *layinc;             /* 1-based layer numbers */
  const double *scale, *emtemp;
  const int *isw;                /* 1-based parameter ids */
  const double *vem, *emissivity;
  /* dust: cont(k,j) particles/cm2 as [L][ncont]; the /DUST/ block's
   * wavenumbers vsec[nsec] and extinction cross-sections xsec(1,k,l) as
   * [ncont][nsec] (cirsradg_wave.f:225,312). ncont = 0: no dust. */
  int ncont, nsec;
  const double *cont, *vsec, *xsec;
  /* /NLTE_FLAGS/ (cirsradg_wave.f:328): 0 = LTE */
  int inlte;
  double nlte_k1, nlte_a;
  /* /SOLARDAT/ iform (no solar file, iread != 999): 1 and 3 scale imod 3
   * radiances to the planet's total flux (cirsradg_wave.f:912-925) */
  int iform;
} orc_req;

/* Status codes (the reference STOPs instead). */
enum {
  ORC_OK = 0,
  ORC_ERR_LIMITS = 1,      /* check_limits.f:10-78 */
  ORC_ERR_UNSUPPORTED = 2, /* branch outside this restatement */
  ORC_ERR_EMISSIVITY = 3,  /* interpem.f:17-22 */
  ORC_ERR_ALLOC = 4
};

/* Jacobian recurrence: 0 = the reference's O(L^2) triangular recurrence
 * (cirsradg_wave.f:954-978); 1 = the O(L) adjoint form (same quantity). */
int orc_cirsradg_wave(const orc_ktable *tabs, int ntab, const orc_cia *cia,
                      const orc_req *r, double vwave, int jac_mode,
                      double *output, double *gtsurf, double *doutputdq);

/* cirsrtfg_wave.f:361-448 wavenumber loop: outputs [nwave][npath],
 * [nwave][npath] and [nwave][npath][nlayer][nparam].  nthreads>1 splits the
 * wavenumbers over OpenMP threads (the calls are independent). */
int orc_spectrum(const orc_ktable *tabs, int ntab, const orc_cia *cia,
                 const orc_req *r, const double *vwave, int nwave,
                 int jac_mode, int nthreads,
                 double *output, double *gtsurf, double *doutputdq);

/* Pieces exposed for unit tests. */
int orc_locate(const double *xx, int n, double x);                 /* locate.f */
void orc_sort2g(int n, double *ra, int *ib);                        /* sort2g.f */
void orc_rankg(int nparam, const double *gw, int ng, int nloop,
               double *weight, double *cont, double *grad, int grad_ld,
               double *k_g, double *dkdq, int dkdq_ld);             /* rankg.f */
void orc_noverlapg(const double *delg, int ng, int ngas,
                   const double *amount, const double *k_gn,
                   const double *dkgndt, double *k_g, double *dkdq); /* noverlapg.f */
double orc_planck(int ispace, double v, double t);                  /* planck_wave.f */
double orc_planckg(int ispace, double v, double t);                 /* planckg_wave.f */
void orc_ciaread(const orc_cia *cia, double vv, double temp,
                 double *alpha, double *dalphadt);                  /* ciaread.f */
void orc_nciacon(const orc_cia *cia, double v0, double p, double t,
                 int inormal, int ngas, const int *idgas, const int *isogas,
                 const double *amount, const double *pp, double xlen,
                 double *absorb, int *iabs, double *dabsorb);       /* nciacon.f */
double orc_nlte_t(int inlte, double k1, double a, double t, double p, double v); /* calc_nlte_t.f */
double orc_interp(const double *x, const double *y, int n, double xin); /* interp.f */
double orc_verint(const double *x, const double *y, int n, double xin); /* verint.f */
/* e3interp.f on a 1001-point table: E3 at transmission x, dE3/dtr in *de3 */
double orc_e3(const double *trans, const double *e3, double x, double *de3);
double orc_rayleigh(int iray, double v, double p, double t, double fheh2,
                    double fch4h2, double fnh3);                    /* rayleigh*.f */

#ifdef __cplusplus
}
#endif
#endif
