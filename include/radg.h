/*
 * radg: MI355X (gfx950) implementation of Radtrans' CIRSRADG_WAVE --
 * correlated-k thermal emission with its analytic Jacobian.
 *
 * C ABI.  Every entry point takes plain pointers and sizes, returns an int
 * status (0 = success, see RADG_E*), and never aborts the host: the
 * reference STOPs the process on the same conditions.  radg_last_error()
 * returns the message of the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   cirsradg_wave_           radtran/cirsradg/cirsradg_wave.f:1-5 (types :167-178)
 *   radg_spectrum_r4/_r8     the wavenumber loop radtran/cirsradg/cirsrtfg_wave.f:361-448
 *   radg_read_klist          /INTERPK/ filled by radtran/cirsrad/read_klist.f:30-368
 *                            from .kta headers (radtran/cirsrad/read_khead.f:18-120)
 *   radg_register_ktables    the same registry from arrays already in memory
 *   radg_read_klbllist       /INTERPKLBL/ filled by radtran/cirsrad/read_klbllist.f:30-198
 *                            from .lta headers (radtran/cirsrad/read_klblhead.f:19-75)
 *   radg_register_lbltables  the same registry from arrays already in memory
 *   radg_set_ilbl            /LBLTABLE/ ILBL (radtran/cirsradg/cirsradg_wave.f:317,336-339,
 *                            502-506,705-741: get_klblg.f and noverlapg1.f instead of
 *                            get_kg.f and noverlapg.f)
 *   radg_register_cia        /CIATABLE/ filled by radtran/radtran/ciaread.f:63-103
 *   radg_read_cia            the same from the table file ciaread.f:77-89 reads
 *   radg_register_parah2     /FPTABLE/ filled by radtran/radtran/fpread_all.f:72-143
 *   radg_register_e3         /E3TABLE/ filled by radtran/radtran/e3interp.f:30-51
 *   radg_register_solar      /SOLARDAT/ filled for radtran/cirsrad/get_solar_wave.f:1-46
 *   radg_register_n2nir      the tables inside radtran/radtran/n2n2cia.f:7-47 and
 *                            n2h2cia.f:3-80 (nciacon.f:287-296, nparacon_all.f:312-320)
 *   radg_register_gascon     /GASCONDEF/ (radtran/includes/gascom.f) with the tables
 *                            radtran/radtran/ngascon.f:86-92 (BLOCK DATA), ch4cont.f:35-43
 *                            and nh3cont.f:35-44 read
 *   radg_register_dust       /DUST/ vsec, xsec, nsec, ncont (radtran/cirsradg/cirsradg_wave.f:
 *                            225,312) that the caller fills from its .xsc file
 *   radg_set_nlte            /NLTE_FLAGS/ inlte_flag, nlte_k1, nlte_A (cirsradg_wave.f:328)
 *   radg_set_iform           /SOLARDAT/ iform, iread (cirsradg_wave.f:307-309)
 *   radg_set_mode            (none: the reference's precision is fixed at compile time)
 *   radg_batch_*, radg_server_*  (none: device-resident batches and the serving loop
 *                            that replace nemesisPY/nemesis.py:4586-4664's process pool)
 *   radg_cirsconv            radtran/cirsrad/cirsconv.f:12-246 as the caller's tail applies
 *                            it (radtran/cirsradg/cirsrtfg_wave.f:468-516), many spectra at once
 *   radg_batch_set_statemaps / _read_mapped  the caller's tail of the wavenumber loop,
 *                            map2pro + map2xvec (radtran/cirsradg/cirsrtfg_wave.f:403-425,
 *                            map2pro.f:49-84, map2xvec.f:41-54)
 *   radg_oe_*                one Levenberg-Marquardt proposal and the cost function of
 *                            nemesis/coreret.f:554-1068 (calc_gain_matrix.f, calcnextxn.f,
 *                            calc_phiret.f, dinvertm.f) for many retrievals at once
 */
#ifndef RADG_H
#define RADG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Array limits, radtran/includes/arrdef.f:14-23.  They fix the leading
 * dimensions of every array of the Fortran argument surface. */
#define RADG_MAXLAY 400
#define RADG_MAXPAT 200
#define RADG_MAXGAS 30
#define RADG_MAXCON 10
#define RADG_MAXSEC 16000
#define RADG_MAXG 51
#define RADG_MAXK 30
#define RADG_NPARMAX (RADG_MAXGAS + 2 + RADG_MAXCON)
#define RADG_E3N 1001 /* e3interp.f's mtab */
/* The GPU overlap sorts NG*NG keys in one wavefront's registers. */
#define RADG_GPU_MAXG 32

/* Status codes. */
#define RADG_OK 0
#define RADG_ELIMITS 1      /* check_limits.f:10-78 and the other size checks */
#define RADG_EUNSUPPORTED 2 /* a branch this implementation does not provide */
#define RADG_ENOTABLES 3    /* no k-tables or CIA table registered */
#define RADG_EINPUT 4       /* invalid argument value (NaN, non-positive P, ...) */
#define RADG_EHIP 5         /* a HIP runtime call failed */
#define RADG_EBUDGET 6      /* a batch does not fit the device-memory budget */
#define RADG_EIO 7          /* a .kta file could not be read */
#define RADG_ECANCELLED 8   /* a serving request was cancelled before admission */
#define RADG_ETIMEOUT 9     /* radg_server_wait timed out */

/* Arithmetic forms. */
#define RADG_MODE_FP64 0   /* every quantity in FP64 */
#define RADG_MODE_MIXED 1  /* FP32 interpolation/exp/overlap; FP64 Planck, gdist and
                              transmission product; compensated g-sums */
#define RADG_MODE_BF16 2   /* the mixed form reading k-tables stored as 16-bit log10(k)
                              codes: log10(k) = code + offset, one FP32 offset per
                              (wavenumber point, g-ordinate) of a table, the mean of log10(k)
                              over its (P, T) values; the code is BF16, so results carry a
                              rounding of log10(k) up to |log10(k) - offset| * 2^-8 */
#define RADG_MODE_F16 3    /* the same with IEEE half-precision codes (11 significant
                              bits): rounding up to |log10(k) - offset| * 2^-11 */

int radg_set_mode(int mode);
int radg_get_mode(void);
const char *radg_last_error(void);

/* ---------------------------------------------------------------- tables */

/* Register the k-tables of one list (replaces any earlier registration).
 * All tables share one P/T/g grid, as read_klist.f:117-160 requires.
 *   press_atm[np], g_ord[ng], delg[ng]; nt > 0: temp[nt], one temperature
 *   grid; nt < 0 (the .kta's NT < 0): temp[np][|nt|], a grid per pressure
 *   (read_khead.f:86-92, get_kg.f:145-173)
 *   per table t: idgas[t], isogas[t], npoint[t], vmin[t], delv[t] and
 *   k[t] -> npoint[t]*np*|nt|*ng values, k*1e20, g fastest then T then P
 *   (the .kta record order, calc_fnktablec.f:593-609);
 *   delv[t] <= 0: an irregular wavenumber grid vcen[t] -> npoint[t]
 *   increasing values (read_khead.f:93-100), looked up as read_klist.f:
 *   292-330 does (DELV < 0 the nearest point; DELV = 0 the point at or below,
 *   merged with the next).  vcen may be NULL when every delv > 0.  Regular
 *   and irregular tables do not mix in one list. */
int radg_register_ktables(int ntab, const int *idgas, const int *isogas,
                          const int *npoint, const double *vmin, const double *delv,
                          int np, int nt, int ng, const double *press_atm,
                          const double *temp, const double *g_ord, const double *delg,
                          const double *const *k, const double *const *vcen);

/* Read the .kta files named in a .kls list (one path per line) and register
 * them (read_klist.f:30-110, read_khead.f:18-120).  recl is the record
 * length in bytes (4 for the REAL(4) build, 8 for REAL(8)). */
int radg_read_klist(const char *klist_path, int recl);

/* /INTERPKLBL/: line-by-line tables (one g-ordinate of weight 1,
 * cirsradg_wave.f:336-339), all on the wavenumbers vmin + i*delv, i < npoint
 * (delv > 0) and one P/T grid (np pressures in atm; nt > 0 temperatures, or
 * nt < 0: |nt| temperatures per pressure, temp[np][|nt|]).  k[t] holds table
 * t's npoint*np*|nt| values of k * 1e20, temperature fastest, then pressure
 * (the .lta record order, calc_lbltable.f:321-368).  A gas takes the last
 * table of its (idgas, isogas) (read_klbllist.f:190-198); gases without one
 * have k = 0.  Replaces any line-by-line tables registered before; the
 * k-tables are kept. */
int radg_register_lbltables(int ntab, const int *idgas, const int *isogas, int npoint,
                            double vmin, double delv, int np, int nt, const double *press_atm,
                            const double *temp, const double *const *k);

/* Read the .lta files named in an .lls list (one path per line, the
 * extension made .lta as file.f does) and register them
 * (read_klbllist.f:30-198, read_klblhead.f:19-75): the tables must share
 * NP, NT, DELV, NPOINT, VMIN and IREC0, with P/T grid ends within 1e-4
 * (else RADG_EINPUT, where the reference stops); the last file's P/T grid
 * is used.  recl: record length in bytes (4 or 8). */
int radg_read_klbllist(const char *list_path, int recl);

/* /LBLTABLE/ ILBL for the batches made afterwards (and every entry point
 * that makes one): 0 runs get_kg + noverlapg on the k-tables, 2 runs
 * get_klblg + noverlapg1 on the line-by-line tables (a wavenumber must lie
 * within (vmin - delv - 0.02 delv, vmin + npoint delv - 0.02 delv), the
 * slices get_klblg.f:182-185 can reach; else RADG_EINPUT).  ILBL = 1 is
 * RADG_EUNSUPPORTED: noverlapg1.f:64-67 stops unless the k-tables have one
 * g-ordinate.  The T column of the line-by-line Jacobian carries
 * sum(amount_i dk_i/dT) through noverlapg1's branches; the reference's own
 * is undefined there (noverlapg1.f:108 stores it into grad). */
int radg_set_ilbl(int ilbl);
int radg_get_ilbl(void);

/* Copy the 16-bit form (mode RADG_MODE_BF16 or RADG_MODE_F16) of registered
 * table itab (0-based) to the host: codes -> npoint*np*nt*ng values in the
 * order of radg_register_ktables, offsets -> npoint*ng FP32 values, with
 * log10(k) = code + offsets[point][g] (k <= 0: code 0xFF80 in BF16, 0xFC00 in
 * half).  The form is made on the device on first use.  No counterpart in
 * the reference (its tables are REAL(4)/REAL(8) only): an inspection hook
 * for callers that check or archive the encoded tables. */
int radg_ktable_codes(int mode, int itab, unsigned short *codes, float *offsets);

/* /CIATABLE/: temps[25] (K), kcia[1501][25][9] in C order, which is the
 * Fortran kcia(9,25,1501) (ciacom.f:9); wavenumber grid (i-1)*dnu. */
int radg_register_cia(double dnu, const double *temps, const double *kcia);

/* /CIATABLE/ from the table file ciaread.f:77-89 reads (the name and dnu of
 * the run's .cia file): an unformatted sequential file of the 25 REAL*8
 * temperatures and kcia(9,25,1501) in REAL*4 (as the raddata tables hold it) or
 * REAL*8.  Registers it as radg_register_cia does. */
int radg_read_cia(const char *path, double dnu);

/* /DUST/ for the fixed-argument entry points (cirsradg_wave_,
 * radg_spectrum_r4/_r8), whose cont(maxcon,maxlay) argument then holds
 * ncont particle columns per layer: wavenumbers vsec[nsec] (ascending) and
 * extinction cross-sections xsec[ncont][nsec] (cm2 per particle; the
 * reference's xsec(1,k,l)).  ncont = 0 clears it.  nsec <= 16000,
 * ncont <= 10 (check_limits.f:24-33).  Batched requests carry their own. */
int radg_register_dust(int nsec, int ncont, const double *vsec, const double *xsec);

/* /FPTABLE/, the para-H2 CIA that fpread_all.f reads (ciacom.f): npara
 * para fractions fp[npara] (uniform, the first 0.25 as nparacon_all.f
 * assumes), temps[25], kh2h2/kh2he[1501][25][npara] (kh2h2(npara,25,1501)),
 * kcia5[1501][25][5] (H2-N2, N2-CH4, N2-N2, CH4-CH4, H2-CH4) on the grid
 * (i-1)*dnu.  Used by requests with flagh2p = 1; npara = 0 clears it. */
int radg_register_parah2(double dnu, const double *temps, int npara, const double *fp,
                         const double *kh2h2, const double *kh2he, const double *kcia5);

/* /E3TABLE/ for disc-integrated emission (imod 18): E3 of the optical depth
 * -ln(trans) at the n = 1001 transmissions trans[n] = (i-1)/1000, as
 * e3tab.dat holds it (e3interp.f:33-44).  The gradient is formed from the
 * table as e3interp.f:47-51 does.  Needed by requests with an imod 18 path;
 * n = 0 clears it. */
int radg_register_e3(int n, const double *trans, const double *e3);

/* N2's near-infrared CIA: N2-N2 (n2n2cia.f) and N2-H2 (n2h2cia.f) as those
 * functions return them (cm-1 amagat-2) at their knots wn1[n1], wn2[n2]
 * (increasing; the reference's are 4500:5:4825 and 3995:5:4995 cm-1, so a
 * caller fills k1[i] = n2n2cia(wn1[i])).  Linear between knots (verint.f),
 * zero outside.  Needed by requests with N2 at 3995-4995 cm-1; n1 = n2 = 0
 * clears them. */
int radg_register_n2nir(int n1, const double *wn1, const double *k1, int n2, const double *wn2,
                        const double *k2);

/* /NLTE_FLAGS/ (cirsradg_wave.f:328) for the fixed-argument entry points:
 * the emission temperature of every layer of an imod 3 or 18 path becomes
 * calc_nlte_t.f's (flag 0 = LTE; 1 = collisional with k1 in [1e-20, 1e-3]
 * and A; 2 = pressure form).  Batched requests carry their own. */
int radg_set_nlte(int flag, double k1, double a);

/* /SOLARDAT/ iform and iread for the fixed-argument entry points (which take
 * radius1 and radextra as arguments).  iread = 999 reads the spectrum
 * radg_register_solar registered (refused without one; while it is set the
 * spectrum cannot be cleared); any other iread leaves the spectrum unused by
 * these entry points, as cirsradg_wave.f:866,921 test iread.  Batched
 * requests treat a registered spectrum as iread = 999. */
int radg_set_iform(int iform, int iread);

/* /SOLARDAT/ swave, srad, nspt: the stellar spectrum get_solar_wave.f reads
 * (the reference's iread = 999), wave[npt] monotonic in the calculation's
 * units, rad[npt] its total power.  With it, iform 1 divides imod 3 and 18
 * paths by the spectrum at each wavenumber (cirsradg_wave.f:921-925,
 * 1144-1149) and iform 4 multiplies imod 0 transmissions by its flux at the
 * request's dist AU (:866-869; get_solar_wave.f:19-46: interp.f inside the
 * range, zero outside, over 4 pi (dist AU)^2 for dist > 0).  Batches use the
 * spectrum registered when they are created; npt = 0 clears it. */
int radg_register_solar(int npt, const double *wave, const double *rad);

/* /GASCONDEF/ (radtran/includes/gascom.f) and the tables ngascon.f reads
 * (radtran/radtran/ngascon.f:86-92,144-244, ch4cont.f, nh3cont.f), for the
 * gas continua cirsradg_wave.f:590-611 adds per layer and gas.
 *   ih2o = 1: the CKD H2O continuum of every gas with id 1, on nh2o >= 2
 *     points h2o_v0 + i*h2o_dv cm-1: h2o_s296 (self, 296 K; /CH2OS0/),
 *     h2o_s260 (self, 260 K; /CH2OS1/), h2o_f296 (foreign, 296 K;
 *     /CH2OF0/), with /COMCON/'s c2, ts, ps;
 *   ich4 != 0: CH4 (id 6, isotope 0/1) from Karkoschka's table, nch4
 *     wavelengths ch4_wl (micron, non-decreasing) and coefficients ch4_k as
 *     karkoschka.dat holds them (ch4cont.f:44-57);
 *   inh3 != 0: NH3 (id 11, isotope 0/1) from Lutz & Owen's table
 *     (lutzowen.dat, nh3cont.f:46-70), likewise;
 *   ih2o = 2, io3 != 0 and iuvscat bits are accepted here but requests
 *     whose gases they would act on are refused with RADG_EUNSUPPORTED
 *     (ngascon.f:215-221 leaves the absorption unassigned; the ozone and UV
 *     cross-section files are not in the reference tree).
 * The H- continua need no flag and no table: bound-free for id 91 isotope 0
 * (hmin_bf.f) and, with e- (id 80) present, free-free for id 48 isotope 0
 * (hmin_ff.f), whose electron state (electron.f) requires every 80 and 91
 * to precede every such 48 in the gas list.  Batches use the definition
 * registered when they are created; NULL clears it (every flag 0). */
typedef struct radg_gascon {
  int ih2o, ich4, io3, inh3, iuvscat;
  int nh2o;
  double h2o_v0, h2o_dv, h2o_c2, h2o_ts, h2o_ps;
  const double *h2o_s296, *h2o_s260, *h2o_f296;
  int nch4, nnh3;
  const double *ch4_wl, *ch4_k, *nh3_wl, *nh3_k;
} radg_gascon;
int radg_register_gascon(const radg_gascon *g);

/* ------------------------------------------------ Fortran argument surface */

/* Drop-in for CALL cirsradg_wave(...) from a REAL(4) host: same 37
 * arguments, same fixed leading dimensions (arrdef.f), same outputs.  The
 * hidden COMMON inputs come from radg_register_ktables/_cia; ncont = 0.
 * iwave is accepted and ignored: the table slice is found from vwave. */
void cirsradg_wave_(const float *dist, const int *inormal, const int *iray,
                    const float *delh, const int *nlayer, const int *npath,
                    const int *ngas, const float *press, const float *temp,
                    const float *pp, const float *amount, const int *iwave,
                    const int *ispace, const int *amform, const float *vwave,
                    const int *nlayin, const int *layinc, const float *cont,
                    const float *scale, const int *imod, const int *idgas,
                    const int *isogas, const float *emtemp, const int *itype,
                    const int *nem, const float *vem, const float *emissivity,
                    const float *tsurf, float *gtsurf, const float *radius1,
                    const int *flagh2p, const float *hfp, const float *radextra,
                    const int *nsw, const int *isw, float *output, float *doutputdq);

/* Status of the last cirsradg_wave_ call (the Fortran routine has none). */
int radg_cirsradg_status(void);

/* The whole wavenumber loop in one call.  Arguments as cirsradg_wave_ (fixed
 * leading dimensions; cont(maxcon,maxlay) is read for the registered dust
 * types; radius1 and radextra scale imod 3 and 18 paths when radg_set_iform
 * chose iform 1 or 3), plus nwave and vwave[nwave].  Outputs, C order:
 *   output[nwave][npath], gtsurf[nwave][npath],
 *   doutputdq[nwave][npath][nlayer][ngas+1+ncont]   (params: gas amounts
 *   1..ngas, then temperature, then the registered dust types; per molecule
 *   cm-2, per K and per particle cm-2; rows past nlayin and params not in
 *   isw are zero). */
int radg_spectrum_r4(const float *dist, const int *inormal, const int *iray,
                     const float *delh, const int *nlayer, const int *npath,
                     const int *ngas, const float *press, const float *temp,
                     const float *pp, const float *amount, const int *ispace,
                     const int *nlayin, const int *layinc, const float *cont,
                     const float *scale, const int *imod, const int *idgas,
                     const int *isogas,
                     const float *emtemp, const int *itype, const int *nem,
                     const float *vem, const float *emissivity, const float *tsurf,
                     const float *radius1, const float *radextra,
                     const int *flagh2p, const float *hfp, const int *nsw, const int *isw,
                     int nwave, const float *vwave, float *output, float *gtsurf,
                     float *doutputdq);
int radg_spectrum_r8(const double *dist, const int *inormal, const int *iray,
                     const double *delh, const int *nlayer, const int *npath,
                     const int *ngas, const double *press, const double *temp,
                     const double *pp, const double *amount, const int *ispace,
                     const int *nlayin, const int *layinc, const double *cont,
                     const double *scale, const int *imod, const int *idgas,
                     const int *isogas,
                     const double *emtemp, const int *itype, const int *nem,
                     const double *vem, const double *emissivity, const double *tsurf,
                     const double *radius1, const double *radextra,
                     const int *flagh2p, const double *hfp, const int *nsw, const int *isw,
                     int nwave, const double *vwave, double *output, double *gtsurf,
                     double *doutputdq);

/* --------------------------------------------- device-resident batches */

/* One spectrum request in compact C layout (0-based axes):
 *   press/temp/delh[L]; pp/amount[L][ngas]; idgas/isogas[ngas];
 *   nlayin/imod[npath]; layinc (1-based layer numbers), scale, emtemp
 *   [npath][L]; isw[nsw] (1-based parameter ids); vwave[nwave];
 *   vem/emissivity[nem]; path_weight[npath] (angle quadrature weights used
 *   for the weighted spectrum; may be NULL).  imod 0, 1, 3, 18 (disc
 *   integrated, cirsradg_wave.f:1116-1218: needs radg_register_e3 and
 *   tsurf > 0; its surface term carries no Jacobian, as in the reference)
 *   and the SCR paths 13, 14 (cells: the first layer's 1 - e^-tau,
 *   (1 + e^-tau)/2) and 8 (layinc(1) a cell times layinc(2) an atmosphere,
 *   both earlier paths; its Jacobian rows are nlayin(1) on path 1's layers,
 *   :1037-1102, 1251-1252, map2pro.f:52-57).
 * Dust (cirsradg_wave.f:550-577), ncont = 0 for none: cont[L][ncont]
 *   particles/cm2, cross-section grid vsec[nsec] (ascending cm-1) and
 *   extinction xsec[ncont][nsec] (cm2 per particle; nsec = 1: constant).
 *   The Jacobian then has nparam = ngas + 1 + ncont columns, the dust ones
 *   last (d/d(particles/cm2)). */
typedef struct radg_request {
  int nlayer, npath, ngas, nsw, nwave, nem;
  int inormal, iray, ispace, itype, flagh2p;
  double tsurf;
  const double *press, *temp, *delh, *pp, *amount;
  const int *idgas, *isogas, *nlayin, *imod, *layinc;
  const double *scale, *emtemp;
  const int *isw;
  const double *vwave, *vem, *emissivity, *path_weight;
  int ncont, nsec;
  const double *cont, *vsec, *xsec;
  int inlte;                 /* /NLTE_FLAGS/ of calc_nlte_t.f: 0 LTE, 1, 2 */
  double nlte_k1, nlte_a;
  int iform;                 /* /SOLARDAT/ iform: 1 and 3 give imod 3 paths the */
  double radius1, radextra;  /* planet's total flux, 4 pi^2 ((radius1 + radextra)
                                km in cm)^2, times 1e-18 for 3 (cirsradg_wave.f:912-925);
                                imod 18 paths 4 pi (...)^2 (:1134-1149) */
  const double *hfp;         /* [L] para-H2 fractions when flagh2p = 1 (nparacon_all.f;
                                < 0: the table's last fraction); fpara is then the last
                                Jacobian parameter */
  double dist;               /* AU from the star: iform 4 scales imod 0 by the stellar flux
                                there, with a stellar spectrum registered */
} radg_request;

typedef struct radg_batch radg_batch;

/* Pack and upload nreq requests; outputs stay on the device.  Fails with
 * RADG_EBUDGET if the batch needs more than budget_bytes of device memory
 * (0 = no limit). */
int radg_batch_create(int nreq, const radg_request *reqs, int mode,
                      size_t budget_bytes, radg_batch **out);
/* Device bytes a batch of these requests would allocate (for admission). */
int radg_batch_bytes(int nreq, const radg_request *reqs, int mode, size_t *bytes);
/* Launch every kernel of the batch on a HIP stream (NULL = default stream). */
int radg_batch_launch(radg_batch *b, void *hip_stream);
/* Also queue, behind each launch's kernels, copies of the outputs to pinned
 * host memory: bit 0 the angle-weighted arrays, bit 1 the per-path arrays
 * (0 = none).  radg_batch_read then serves from that copy. */
int radg_batch_set_fetch(radg_batch *b, int what);
/* Wait for the batch's stream. */
int radg_batch_sync(radg_batch *b);
/* Copy request r's results to host memory (FP64 regardless of mode).  Any
 * pointer may be NULL.  Shapes as radg_spectrum_r8; the weighted arrays are
 * wout[nwave], wgts[nwave], wjac[nwave][nlayer][nparam] (sum over paths of
 * path_weight * result). */
int radg_batch_read(radg_batch *b, int r, double *output, double *gtsurf,
                    double *doutputdq, double *wout, double *wgts, double *wjac);
/* Kernel timings of the last launch in milliseconds (HIP events on the
 * batch's stream): [0] layer/overlap kernel, [1] emission kernel,
 * [2] path quadrature. */
int radg_batch_timings(radg_batch *b, float *ms3);
/* Sum of the kernel timings over the last (at most 128) launches. */
int radg_batch_timing_stats(radg_batch *b, int *nlaunch, float *sum_ms3);
/* Algorithmic bytes of the batch: alg_bytes3[0] = k-table nodes touched
 * (distinct (P,T) nodes bracketing the layers x NG x bytes per value, per
 * wavenumber and gas), [1] = outputs written, [2] = layer inputs and CIA
 * nodes; plus the number of spectra and of layer-kernel work items. */
int radg_batch_info(radg_batch *b, double *alg_bytes3, int *nspectra, int *nitems_a);
int radg_batch_destroy(radg_batch *b);

/* cirsconv.f (radtran/cirsrad/cirsconv.f:12-246), the convolution the caller's
 * tail applies to every path's radiance, surface gradient and state-vector
 * Jacobian column (radtran/cirsradg/cirsrtfg_wave.f:468-516), for ncol
 * spectra at once: y[nwave][ncol] (device, FP64; wavenumber-major, as a
 * batch's wout and kk) at the increasing wavenumbers vwave[nwave] ->
 * yout[nconv][ncol] (device) at vconv[nconv].
 *   fwhm > 0: the boxcar mean of the natural cubic spline through the
 *     spectrum (cspline.f, csplint.f) over [vconv - w/2, vconv + w/2], a
 *     20-point trapezoid divided by fwhm, with w = fwhm or, with nfwhm > 0
 *     (FWHMEXIST), verint.f of the table vfwhm/xfwhm at vconv;
 *   fwhm = 0: yout(i) = y(i) (nconv <= nwave);
 *   fwhm < 0: the nfil filter functions of the runname.fil file -- filter k
 *     centred on fil_vcen[k] (matching a vconv within 1e-5) with fil_n[k]
 *     points, its wavenumbers and values concatenated in fil_v / fil_f.
 * The map is formed once per call on the host in FP64 and applied in one
 * launch; the call returns when yout is written. */
typedef struct radg_conv {
  double fwhm;
  int nfwhm;
  const double *vfwhm, *xfwhm;
  int nfil;
  const int *fil_n;
  const double *fil_vcen, *fil_v, *fil_f;
} radg_conv;
int radg_cirsconv(int nwave, const double *vwave, int nconv, const double *vconv,
                  const radg_conv *conv, int ncol, const double *y, double *yout, void *hip_stream);

/* The forward model and analytic Jacobian of one request with NO device
 * allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The kernels are those of
 * radg_batch_create + radg_batch_launch; the packed inputs, intermediates and
 * per-path outputs lie in the library's file-scope device array of
 * RADG_NOALLOC_BYTES per GPU, held until the process ends.  Outputs are the
 * caller's device arrays (FP64): wout[nwave], wgts[nwave] and
 * wjac[nwave][nlayer][nparam], the angle-weighted radiance, surface gradient
 * and Jacobian radg_batch_read returns (gas columns per molecule).  The
 * kernels are queued on hip_stream (NULL: the null stream); the call returns
 * once they are queued.  A request that needs more scratch than
 * RADG_NOALLOC_BYTES is refused with RADG_ELIMITS (split its wavenumbers);
 * the table form of mode must exist (radg_set_mode(mode) after registering
 * the tables makes it).  Calls on one GPU share the array: calls that overlap
 * in time on one GPU (other streams, other threads) race. */
#define RADG_NOALLOC_BYTES ((size_t)128 << 20)
int radg_spectrum_noalloc(const radg_request *req, int mode, double *wout, double *wgts,
                          double *wjac, void *hip_stream);
/* ------------------------------------- layer -> profile -> state maps */
/* Per request: the matrices cirsrtfg_wave.f:403-425 applies after each
 * cirsradg_wave call (0-based axes, C order):
 *   dam[nlayer][npro]  d(layer amount)/d(level vmr), the /LGRAD/ DAM of
 *                      radtran/includes/laygrad.f (used for every gas)
 *   dte[nlayer][npro]  d(layer T)/d(level T), /LGRAD/ DTE
 *   dco[nlayer][npro]  d(layer dust)/d(level dust), /LGRAD/ DCO (used for every
 *                      dust type, map2pro.f:72-76); may be NULL without dust
 *   dfp[nlayer][npro]  d(layer para fraction)/d(level), /LGRAD/ DFP
 *                      (map2pro.f:77-80); may be NULL unless flagh2p = 1
 *   xmap[nv][nparam][npro]  d(profile quantity)/d(state element), map2xvec.f's
 *                      xmap(maxv, maxgas+2+maxcon, maxpro) for the request's
 *                      nparam = ngas + 1 + ncont parameters
 * nv = 0: no map for that request. */
typedef struct radg_statemap {
  int nv, npro;
  const double *dam, *dte, *xmap;
  const double *dco;
  const double *dfp;
} radg_statemap;
/* Attach maps[nreq] to a batch (copied to the device).  Every later launch
 * then also forms, on the device and in FP64,
 *   kk[nwave][nv] = sum_p path_weight[p] doutdx[.][p][.]
 * and, when per_path is non-zero or the paths of a request do not share
 * their layers, doutdx[nwave][npath][nv] -- the reference's gradients per
 * path, doutdx = map2xvec(map2pro(doutputdq)) (cirsrtfg_wave.f:436-444). */
int radg_batch_set_statemaps(radg_batch *b, const radg_statemap *maps, int per_path);
/* Copy request r's mapped Jacobians to the host (either pointer may be NULL). */
int radg_batch_read_mapped(radg_batch *b, int r, double *doutdx, double *kk);
/* Device addresses of request r's angle-weighted radiance wout[nwave] and
 * mapped Jacobian kk[nwave][nv] (NULL without a map), valid until the batch
 * is destroyed, for consumers that stay on the device (radg_oe_*). */
int radg_batch_device_outputs(radg_batch *b, int r, const double **wout, const double **kk);

/* ------------------------------------------------ optimal estimation */
/* nret independent retrievals of nx state elements from ny measurements,
 * everything in FP64.  Constants are copied to the device at creation:
 *   y[nret][ny] measurements, se[nret][ny] their (diagonal) variances
 *   (coreret.f:255-285), xa[nret][nx] a-priori states, sa the a-priori
 *   covariances, [nret][nx][nx] or, with sa_shared, one [nx][nx] for all.
 * Sa is inverted on the device by Cholesky (dinvertm JMOD=2 in
 * coreret.f:225-250).  Needs nx <= 4096; state elements not retrieved are
 * the caller's business (coreret.f:613-619 resets fixed ones). */
typedef struct radg_oe radg_oe;
int radg_oe_create(int nret, int nx, int ny, const double *y, const double *se,
                   const double *xa, const double *sa, int sa_shared, radg_oe **out);
/* calc_phiret.f for every retrieval at states xn[nret][nx] (host) with
 * forward radiances yn_dev[r] -> [ny] (device): chisq = (yn - y)^T Se^-1
 * (yn - y) and phi = chisq + (xn - xa)^T Sa^-1 (xn - xa).  chisq may be NULL. */
int radg_oe_cost(radg_oe *h, const double *xn, const double *const *yn_dev, double *phi,
                 double *chisq);
/* One damped step per retrieval (coreret.f:609-632): from xn (host), the
 * forward radiances yn_dev[r] and Jacobians kk_dev[r] -> [ny][nx] (device),
 *   x_out = xa + dd (y - yn) - aa (xa - xn)          (calcnextxn.f)
 * with dd = (Sa^-1 + K^T Se^-1 K)^-1 K^T Se^-1 and aa = dd K
 * (calc_gain_matrix.f), computed as xa + H^-1 K^T Se^-1 [(y - yn) + K (xn - xa)]
 * by one Cholesky factorisation and solve of H = Sa^-1 + K^T Se^-1 K, then
 *   xn1 = xn + (x_out - xn) / (1 + lambda[r])   -> xn1[nret][nx] (host). */
int radg_oe_propose(radg_oe *h, const double *xn, const double *const *yn_dev,
                    const double *const *kk_dev, const double *lambda, double *xn1);
int radg_oe_destroy(radg_oe *h);

/* ------------------------------------------------------- serving loop */

/* A persistent loop on a worker thread: requests are deep-copied, validated
 * and sized at submit, queued, and admitted in arrival order into device
 * batches while the in-flight batches' memory stays within budget_bytes
 * (0 = unlimited) and a stream slot is free (max_inflight streams, each
 * batch at most max_batch requests).  Requests that arrive while others run
 * are admitted as slots and memory free up; their kernels overlap the running
 * batches.  Results stay with the ticket until radg_server_wait collects them:
 * always the angle-weighted arrays, and the per-path arrays too when
 * per_path_outputs is non-zero. */
typedef struct radg_server radg_server;
int radg_server_create(int mode, size_t budget_bytes, int max_batch, int max_inflight,
                       int per_path_outputs, radg_server **out);
/* Returns the validation status of the request (its ticket then reports the
 * same status); *ticket is always set. */
int radg_server_submit(radg_server *s, const radg_request *req, long long *ticket);
/* Cancel a request still waiting for admission (the reference's .abo file,
 * nemesis/coreret.f:1090-1100); its status becomes RADG_ECANCELLED. */
int radg_server_cancel(radg_server *s, long long ticket);
/* Wait up to timeout_s seconds (< 0: forever) for a ticket; on completion the
 * ticket is released, *status holds the request's status and *latency_s the
 * time from submit to results on the host.  Outputs as radg_batch_read. */
int radg_server_wait(radg_server *s, long long ticket, double timeout_s, double *output,
                     double *gtsurf, double *doutputdq, double *wout, double *wgts, double *wjac,
                     int *status, double *latency_s);
/* steps (batches launched), requests admitted, requests completed, largest
 * batch, peak in-flight bytes, queue length, then worker microseconds spent
 * packing/uploading/launching, retiring results, and in kernels. */
int radg_server_stats(radg_server *s, long long *stats9);
int radg_server_destroy(radg_server *s);

#ifdef __cplusplus
}
#endif
#endif
