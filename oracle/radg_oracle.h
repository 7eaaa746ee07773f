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

/* One gas's k-table (the content of a .kta file, read_khead.f:38-109).
 * nt < 0: one temperature grid per pressure, t2[np][|nt|] (tk unused);
 * delv <= 0: an irregular wavenumber grid vcen[npoint] (vmin unused). */
typedef struct orc_ktable {
  int idgas, isogas;
  int npoint, np, nt, ng;
  double vmin, delv;
  const double *lnp;   /* [np]  ln(P / atm), as read_khead.f:80-85 stores it */
  const double *tk;    /* [nt]  K */
  const double *g_ord; /* [ng] */
  const double *delg;  /* [ng] */
  const double *k;     /* [npoint][np][|nt|][ng], k * 1e20 */
  const double *t2;    /* [np][|nt|] K when nt < 0 */
  const double *vcen;  /* [npoint] when delv <= 0 */
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
  /* /GASCONDEF/ (gascom.f) and the tables ngascon.f reads.
   *   ih2o = 1: the CKD H2O continuum on nh2o points h2o_v0 + i*h2o_dv
   *     (/CH2OS0/ self at 296 K, /CH2OS1/ self at 260 K, /CH2OF0/ foreign
   *     at 296 K) with /COMCON/'s c2, ts, ps;
   *   ich4 != 0: Karkoschka's CH4 table (ch4cont.f: wavelength micron,
   *     coefficient as karkoschka.dat holds it);
   *   inh3 != 0: Lutz & Owen's NH3 table (nh3cont.f, lutzowen.dat);
   *   io3, iuvscat: not restated (ORC_ERR_UNSUPPORTED where they act).
   * hmin_static: test knob.  hmin_bf.f and hmin_ff.f add to a local
   * `summ` they never initialise; compiled with static locals (flang
   * -fno-automatic, ifort's default) it carries over from call to call,
   * reset only by a call outside the function's wavelength range.  0 (the
   * restatement): each call starts from 0.  1: the carry-over is emulated
   * within one cirsradg_wave call, starting from 0 (a caller that resets
   * both functions before each call sees this). */
  int ih2o, ich4, io3, inh3, iuvscat;
  int nh2o;
  double h2o_v0, h2o_dv, h2o_c2, h2o_ts, h2o_ps;
  const double *h2o_s296, *h2o_s260, *h2o_f296;
  int nch4, nnh3;
  const double *ch4_wl, *ch4_k, *nh3_wl, *nh3_k;
  int hmin_static;
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
  /* /LBLTABLE/ ILBL: 0 = k-tables; 2 = the tables are line-by-line tables
   * (one g-ordinate, the .lta content of read_klblhead.f, all on one
   * wavenumber and P/T grid) looked up by get_klblg and summed by
   * noverlapg1 (cirsradg_wave.f:336-339,502-506,735-741).  1 is refused. */
  int ilbl;
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
void orc_noverlapg1(int ngas, const double *amount, const double *k_gn,
                    const double *dkgndt, double *k_g, double *dkdq); /* noverlapg1.f */
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
/* H- opacity with a fresh `summ` per call: free-free at (v0 cm-1, temp K),
 * cm4 dyne-1 (hmin_ff.f), and bound-free cross-section at v0 (hmin_bf.f) */
double orc_hmin_ff(double v0, double temp);
double orc_hmin_bf(double v0);
/* cirsconv.f:12-246 on one spectrum y[nwave] at vwave -> yout[nconv] at
 * vconv: fwhm > 0 the boxcar of width fwhm (or, with nfwhm > 0, of the width
 * verint.f gives from vfwhm/xfwhm at each vconv) integrated over cspline.f /
 * csplint.f's natural spline by a 20-point trapezoid and divided by fwhm;
 * fwhm = 0 the identity; fwhm < 0 the nfil filter functions of the .fil file
 * (filter k: centre fil_vcen[k], fil_n[k] points fil_v/fil_f, concatenated).
 * Returns ORC_ERR_LIMITS for nwave < 2 with fwhm != 0 and for a conv point
 * no filter matches. */
int orc_cirsconv(double fwhm, int nwave, const double *vwave, const double *y, int nconv,
                 const double *vconv, double *yout, int nfwhm, const double *vfwhm,
                 const double *xfwhm, int nfil, const int *fil_n, const double *fil_vcen,
                 const double *fil_v, const double *fil_f);
/* ch4cont.f / nh3cont.f on the registered tables, in 1e20 cm2/molecule */
double orc_ch4cont(const orc_cia *cia, double nu);
double orc_nh3cont(const orc_cia *cia, double nu);

#ifdef __cplusplus
}
#endif
#endif
