/*
 * FP64 CPU restatement of CIRSRADG_WAVE -- TEST INFRASTRUCTURE (the oracle).
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
 * file, as the checker.  It restates, in plain C and in the reference's own
 * loop order, the routines on the path:
 *
 *   locate      frecipes/locate.f:1-16
 *   sort2g      radtran/cirsradg/sort2g.f:1-52
 *   rankg       radtran/cirsradg/rankg.f:1-156
 *   rankk       radtran/cirsradg/rankk.f:17-120
 *   noverlapg   radtran/cirsradg/noverlapg.f:1-202
 *   get_kg      radtran/cirsradg/get_kg.f:1-484 (+ read_klist.f:260-330 lookup)
 *   pindex      radtran/ciatable/pindex.f:1-32
 *   ciaread     radtran/radtran/ciaread.f:104-196
 *   nciacon     radtran/radtran/nciacon.f:1-310
 *   nparacon_all radtran/radtran/nparacon_all.f:1-338 (+ fpread_all.f:168-257)
 *   co2cia      radtran/radtran/co2cia.f
 *   ngascon     radtran/radtran/ngascon.f:136-267 (CKD H2O, ch4cont.f, nh3cont.f,
 *               hmin_bf.f, hmin_ff.f; the /ELECTRON/ state of cirsradg_wave.f:590-611)
 *   rayleigh*   radtran/scatter/rayleighj.f, rayleighv.f, rayleigha.f,
 *               rayleighls.f (the iray branch, cirsradg_wave.f:663-691)
 *   planck      radtran/radtran/planck_wave.f:1-42, planckg_wave.f:1-40
 *   calc_nlte_t radtran/radtran/calc_nlte_t.f:1-89 (+ invplanck_wave.f)
 *   dpexp       radtran/radtran/dpexp.f:1-14
 *   interpem    radtran/cirsrad/interpem.f:1-37
 *   interp      radtran/rtm_util/interp.f:1-26 (dust cross-sections)
 *   cirsconv    radtran/cirsrad/cirsconv.f:12-246 (+ rtm_util/cspline.f, csplint.f)
 *   cirsradg    radtran/cirsradg/cirsradg_wave.f:333-1290 (imod 0, 1, 3)
 *
 * Everything that the reference declares REAL is double here, matching the
 * reference compiled with REAL promoted to 8 bytes.
 */
#include "radg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXLAY 400
#define MAXPAT 200
#define MAXGAS 30
#define MAXG 51
#define MAXK 30
#define MAXCON 10
#define MAXSEC 16000
#define CIA_NUMWN 1501
#define CIA_NUMT 25
#define CIA_NPAIRS 9

/* constdef.f:7-11 */
#define AMAGAT 2.68675e19
#define KBOLTZMANN 1.37947e-23
#define PI_REF 3.1415927

/* ---------------------------------------------------------------- locate */
/* frecipes/locate.f:1-16.  xx is 0-based here; the return value is the
 * reference's 1-based JL in 0..n. */
int orc_locate(const double *xx, int n, double x) {
  int jl = 0, ju = n + 1;
  int ascnd = xx[n - 1] > xx[0];
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if (ascnd == (x > xx[jm - 1]))
      jl = jm;
    else
      ju = jm;
  }
  return jl;
}

/* ---------------------------------------------------------------- sort2g */
/* sort2g.f:1-52: heapsort of ra(1:n) carrying ib.  ra/ib are 1-based views
 * (element 0 unused). */
void orc_sort2g(int n, double *ra, int *ib) {
  int l, ir, i, j, irb;
  double rra;
  if (n < 2) return;
  l = n / 2 + 1;
  ir = n;
  for (;;) {
    if (l > 1) {
      l = l - 1;
      rra = ra[l];
      irb = ib[l];
    } else {
      rra = ra[ir];
      irb = ib[ir];
      ra[ir] = ra[1];
      ib[ir] = ib[1];
      ir = ir - 1;
      if (ir == 1) {
        ra[1] = rra;
        ib[1] = irb;
        return;
      }
    }
    i = l;
    j = l + l;
    while (j <= ir) {
      if (j < ir) {
        if (ra[j] < ra[j + 1]) j = j + 1;
      }
      if (rra < ra[j]) {
        ra[i] = ra[j];
        ib[i] = ib[j];
        i = j;
        j = j + j;
      } else {
        j = ir + 1;
      }
    }
    ra[i] = rra;
    ib[i] = irb;
  }
}

/* ----------------------------------------------------------------- rankg */
/* rankg.f:1-156.  weight/cont are 1-based (index 0 unused), length nloop;
 * grad is 1-based in the element index with leading dimension grad_ld:
 * grad[iparam*grad_ld + I], iparam 0-based.  k_g[ig] 1-based (size ng+2),
 * dkdq[iparam*dkdq_ld + ig] 1-based in ig. */
void orc_rankg(int nparam, const double *gw, int ng, int nloop,
               double *weight, double *cont, double *grad, int grad_ld,
               double *k_g, double *dkdq, int dkdq_ld) {
  double g_ord[MAXG + 2];
  double *gdist = (double *)malloc(sizeof(double) * (nloop + 1));
  double *tmp = (double *)malloc(sizeof(double) * (nloop + 1));
  int *ico = (int *)malloc(sizeof(int) * (nloop + 1));
  int i, ig, ip;
  double sum, frac;

  g_ord[1] = 0.0;
  for (i = 1; i <= ng; i++) g_ord[i + 1] = g_ord[i] + gw[i - 1];
  if (g_ord[ng + 1] < 1.0) g_ord[ng + 1] = 1.0;
  /* g_ord(ng+2) is never initialised by the reference; it is only read if an
   * element lies wholly beyond the last boundary, which cannot happen when
   * the total weight is (sum gw)^2 <= g_ord(ng+1). */
  g_ord[ng + 2] = g_ord[ng + 1];

  for (i = 1; i <= nloop; i++) ico[i] = i;
  orc_sort2g(nloop, cont, ico);

  for (i = 1; i <= nloop; i++) tmp[i] = weight[ico[i]];
  for (i = 1; i <= nloop; i++) weight[i] = tmp[i];
  for (ip = 0; ip < nparam; ip++) {
    double *gp = grad + (size_t)ip * grad_ld;
    for (i = 1; i <= nloop; i++) tmp[i] = gp[ico[i]];
    for (i = 1; i <= nloop; i++) gp[i] = tmp[i];
  }

  gdist[0] = 0.0;
  gdist[1] = weight[1];
  for (i = 2; i <= nloop; i++) gdist[i] = weight[i] + gdist[i - 1];

  for (i = 1; i <= ng + 1; i++) {
    k_g[i] = 0.0;
    for (ip = 0; ip < nparam; ip++) dkdq[ip * dkdq_ld + i] = 0.0;
  }

  ig = 1;
  sum = 0.0;
  for (i = 1; i <= nloop; i++) {
    if (gdist[i] < g_ord[ig + 1]) {
      k_g[ig] = k_g[ig] + cont[i] * weight[i];
      for (ip = 0; ip < nparam; ip++)
        dkdq[ip * dkdq_ld + ig] += grad[(size_t)ip * grad_ld + i] * weight[i];
      sum = sum + weight[i];
    } else {
      frac = (g_ord[ig + 1] - gdist[i - 1]) / (gdist[i] - gdist[i - 1]);
      k_g[ig] = k_g[ig] + frac * cont[i] * weight[i];
      for (ip = 0; ip < nparam; ip++)
        dkdq[ip * dkdq_ld + ig] += frac * grad[(size_t)ip * grad_ld + i] * weight[i];
      sum = sum + frac * weight[i];
      k_g[ig] = k_g[ig] / sum;
      for (ip = 0; ip < nparam; ip++) dkdq[ip * dkdq_ld + ig] /= sum;
      ig = ig + 1;
      sum = (1.0 - frac) * weight[i];
      if (ig <= ng + 1) {
        k_g[ig] = (1.0 - frac) * cont[i] * weight[i];
        for (ip = 0; ip < nparam; ip++)
          dkdq[ip * dkdq_ld + ig] = (1.0 - frac) * grad[(size_t)ip * grad_ld + i] * weight[i];
      }
      if (ig > ng) {
        /* The remaining elements would index g_ord beyond ng+1. */
        break;
      }
    }
  }
  if (ig == ng) {
    k_g[ig] = k_g[ig] / sum;
    for (ip = 0; ip < nparam; ip++) dkdq[ip * dkdq_ld + ig] /= sum;
  }
  free(gdist);
  free(tmp);
  free(ico);
}

/* ----------------------------------------------------------------- rankk */
/* rankk.f:17-120 (non-coincident wavenumber).  1-based arrays. */
static void orc_rankk(const double *gw, int ng, double *cont, double *dcont,
                      double *weight, int nloop, double *k_g, double *dkdt) {
  double g_ord[MAXG + 2], gdist[2 * MAXG + 2], tmp[2 * MAXG + 2];
  int ico[2 * MAXG + 2];
  int i, ig;
  double sum, frac;
  g_ord[1] = 0.0;
  for (i = 1; i <= ng; i++) g_ord[i + 1] = g_ord[i] + gw[i - 1];
  if (g_ord[ng + 1] < 1.0) g_ord[ng + 1] = 1.0;
  for (i = 1; i <= nloop; i++) ico[i] = i;
  orc_sort2g(nloop, cont, ico);
  for (i = 1; i <= nloop; i++) tmp[i] = weight[ico[i]];
  for (i = 1; i <= nloop; i++) weight[i] = tmp[i];
  for (i = 1; i <= nloop; i++) tmp[i] = dcont[ico[i]];
  for (i = 1; i <= nloop; i++) dcont[i] = tmp[i];
  gdist[0] = 0.0;
  gdist[1] = weight[1];
  for (i = 2; i <= nloop; i++) gdist[i] = weight[i] + gdist[i - 1];
  for (i = 1; i <= ng + 1; i++) {
    k_g[i] = 0.0;
    dkdt[i] = 0.0;
  }
  ig = 1;
  sum = 0.0;
  for (i = 1; i <= nloop; i++) {
    if (ig <= ng && gdist[i] <= g_ord[ig + 1]) {
      k_g[ig] += cont[i] * weight[i];
      dkdt[ig] += dcont[i] * weight[i];
      sum += weight[i];
    } else {
      if (ig > ng) break; /* reference writes past ng only */
      frac = (g_ord[ig + 1] - gdist[i - 1]) / (gdist[i] - gdist[i - 1]);
      k_g[ig] += frac * cont[i] * weight[i];
      dkdt[ig] += frac * dcont[i] * weight[i];
      sum += frac * weight[i];
      k_g[ig] /= sum;
      dkdt[ig] /= sum;
      ig = ig + 1;
      if (ig <= ng) {
        sum = (1.0 - frac) * weight[i];
        k_g[ig] = (1.0 - frac) * cont[i] * weight[i];
        dkdt[ig] = (1.0 - frac) * dcont[i] * weight[i];
      }
    }
  }
  if (ig == ng) {
    k_g[ig] /= sum;
    dkdt[ig] /= sum;
  }
}

/* ------------------------------------------------------------- noverlapg */
/* noverlapg.f:1-202.  k_gn/dkgndt are [gas][ng]; k_g is [ng]; dkdq is
 * [param][ng] with ngas+1 params (the last one is T). */
void orc_noverlapg(const double *delg, int ng, int ngas, const double *amount,
                   const double *k_gn, const double *dkgndt, double *k_g_out,
                   double *dkdq_out) {
  const int nl = ng * ng;
  const int ld = ng + 2; /* 1-based rows of k_g / dkdq */
  double k_g[MAXG + 2], dkdq[(MAXGAS + 1) * (MAXG + 2)];
  double k1[MAXG], k2[MAXG], d1[MAXG], d2[MAXG];
  double *weight = (double *)malloc(sizeof(double) * (nl + 1));
  double *contri = (double *)malloc(sizeof(double) * (nl + 1));
  double *grad = (double *)malloc(sizeof(double) * (size_t)(nl + 1) * (ngas + 1));
  int igas, i, j, jg, nloop;
  double a1, a2;

  memset(k_g, 0, sizeof(k_g));
  memset(dkdq, 0, sizeof(dkdq));

  if (ngas == 1) {
    /* The reference leaves k_g/dkdq unwritten for one gas (the DO loop
     * noverlapg.f:64 is empty) and cirsradg_wave then reads stale values.
     * Defined here as the single scaled distribution. */
    for (i = 1; i <= ng; i++) {
      k_g[i] = k_gn[i - 1] * amount[0];
      dkdq[0 * ld + i] = k_gn[i - 1];
      dkdq[1 * ld + i] = dkgndt[i - 1] * amount[0];
    }
  }

  for (igas = 1; igas <= ngas - 1; igas++) {
    /* igas, jgas: 1-based gas numbers; dkdq column p (1-based) at (p-1)*ld */
    if (igas == 1) {
      a1 = amount[0];
      a2 = amount[1];
      for (i = 0; i < ng; i++) {
        k1[i] = k_gn[0 * ng + i];
        k2[i] = k_gn[1 * ng + i];
        d1[i] = dkgndt[0 * ng + i];
        d2[i] = dkgndt[1 * ng + i];
      }
      if (k1[ng - 1] == 0.0) {
        for (i = 1; i <= ng; i++) {
          k_g[i] = k2[i - 1] * a2;
          dkdq[(igas - 1) * ld + i] = 0.0;
          dkdq[(igas)*ld + i] = k2[i - 1];
          dkdq[(igas + 1) * ld + i] = d2[i - 1] * a2;
        }
        continue;
      }
      if (k2[ng - 1] == 0.0) {
        for (i = 1; i <= ng; i++) {
          k_g[i] = k1[i - 1] * a1;
          dkdq[(igas - 1) * ld + i] = k1[i - 1];
          dkdq[(igas)*ld + i] = 0.0;
          dkdq[(igas + 1) * ld + i] = d1[i - 1] * a1;
        }
        continue;
      }
      nloop = 0;
      for (i = 0; i < ng; i++) {
        for (j = 0; j < ng; j++) {
          nloop++;
          weight[nloop] = delg[i] * delg[j];
          contri[nloop] = k1[i] * a1 + k2[j] * a2;
          grad[(size_t)(igas - 1) * (nl + 1) + nloop] = k1[i];
          grad[(size_t)(igas) * (nl + 1) + nloop] = k2[j];
          grad[(size_t)(igas + 1) * (nl + 1) + nloop] = d1[i] * a1 + d2[j] * a2;
        }
      }
    } else {
      a2 = amount[igas];
      for (i = 0; i < ng; i++) {
        k1[i] = k_g[i + 1];
        k2[i] = k_gn[igas * ng + i];
        d1[i] = dkdq[(igas)*ld + i + 1]; /* dK/dT of the running sum */
        d2[i] = dkgndt[igas * ng + i];
      }
      if (k1[ng - 1] == 0.0) {
        for (i = 1; i <= ng; i++) {
          k_g[i] = k2[i - 1] * a2;
          for (jg = 1; jg <= igas; jg++) dkdq[(jg - 1) * ld + i] = 0.0;
          dkdq[(igas)*ld + i] = k2[i - 1];
          dkdq[(igas + 1) * ld + i] = d2[i - 1] * a2;
        }
        continue;
      }
      if (k2[ng - 1] == 0.0) {
        for (i = 1; i <= ng; i++) {
          k_g[i] = k1[i - 1];
          dkdq[(igas)*ld + i] = 0.0;
          dkdq[(igas + 1) * ld + i] = d1[i - 1];
        }
        continue;
      }
      nloop = 0;
      for (i = 0; i < ng; i++) {
        for (j = 0; j < ng; j++) {
          nloop++;
          weight[nloop] = delg[i] * delg[j];
          contri[nloop] = k1[i] + k2[j] * a2;
          for (jg = 1; jg <= igas; jg++)
            grad[(size_t)(jg - 1) * (nl + 1) + nloop] = dkdq[(jg - 1) * ld + i + 1];
          grad[(size_t)(igas) * (nl + 1) + nloop] = k2[j];
          grad[(size_t)(igas + 1) * (nl + 1) + nloop] = d1[i] + d2[j] * a2;
        }
      }
    }
    orc_rankg(igas + 2, delg, ng, nloop, weight, contri, grad, nl + 1, k_g, dkdq, ld);
  }

  for (i = 0; i < ng; i++) {
    k_g_out[i] = k_g[i + 1];
    for (jg = 0; jg < ngas + 1; jg++) dkdq_out[jg * ng + i] = dkdq[jg * ld + i + 1];
  }
  free(weight);
  free(contri);
  free(grad);
}

/* ------------------------------------------------------------ noverlapg1 */
/* noverlapg1.f:1-160, the sum that replaces random overlap for line-by-line
 * tables (NG = 1, else :64-67 stop): k = the amount-weighted sum of the
 * gases' k, taken in gas order with the routine's branches for a zero term
 * (:76-108 first pair, :116-146 later gases).  k_gn/dkgndt are [gas]; dkdq
 * is [ngas + 1], the gas columns then T.
 *
 * k and the gas columns are the reference's: each gas's own k, except that
 * a zero running sum before gas m >= 3 zeroes the columns of the gases
 * before it (:123-128).  The T column follows the same branches with the
 * running sum's derivative carried from gas to gas (amount_i dk_i/dT added
 * for every term the sum takes).  The reference's own T column is undefined
 * or wrong: its first pair stores the sum into grad, not dkdq (:108); later
 * gases read dk1dT(1)/dk2dT(1) of the first pair (:129,137,146) where
 * :119-120 set dk1dT(2)/dk2dT(2); one gas never enters the loop, so k_g and
 * dkdq keep what an earlier call left. */
void orc_noverlapg1(int ngas, const double *amount, const double *k_gn, const double *dkgndt,
                    double *k_g, double *dkdq) {
  double s, d;
  int m, j;
  if (ngas == 1) {
    k_g[0] = k_gn[0] * amount[0];
    dkdq[0] = k_gn[0];
    dkdq[1] = dkgndt[0] * amount[0];
    return;
  }
  {
    const double a1 = amount[0], a2 = amount[1], k1 = k_gn[0], k2 = k_gn[1];
    if (k1 == 0.0) {
      s = k2 * a2;
      dkdq[0] = 0.0;
      dkdq[1] = k2;
      d = dkgndt[1] * a2;
    } else if (k2 == 0.0) {
      s = k1 * a1;
      dkdq[0] = k1;
      dkdq[1] = 0.0;
      d = dkgndt[0] * a1;
    } else {
      s = k1 * a1 + k2 * a2;
      dkdq[0] = k1;
      dkdq[1] = k2;
      d = dkgndt[0] * a1 + dkgndt[1] * a2;
    }
  }
  for (m = 2; m < ngas; m++) {
    const double km = k_gn[m], am = amount[m];
    if (s == 0.0) {
      s = km * am;
      for (j = 0; j < m; j++) dkdq[j] = 0.0;
      dkdq[m] = km;
      d = dkgndt[m] * am;
    } else if (km == 0.0) {
      dkdq[m] = 0.0;
    } else {
      s = s + km * am;
      dkdq[m] = km;
      d = d + dkgndt[m] * am;
    }
  }
  k_g[0] = s;
  dkdq[ngas] = d;
}

/* ---------------------------------------------------------------- planck */
/* planck_wave.f:1-42 (FP64 constants, as in the REAL(8) build). */
double orc_planck(int ispace, double v, double t) {
  const double c1 = 1.1911e-12, c2 = 1.439;
  double y, a;
  if (ispace == 0) {
    y = v;
    a = c1 * (y * y * y);
  } else {
    y = 1e4 / v;
    a = c1 * (y * y * y * y * y) / 1e4;
  }
  if (y == 0.0) return 0.0;
  return a / (exp(c2 * y / t) - 1.0);
}

/* planckg_wave.f:1-40 */
double orc_planckg(int ispace, double v, double t) {
  const double c1 = 1.1911e-12, c2 = 1.439;
  double y, a, tmp, bot, top;
  if (ispace == 0) {
    y = v;
    a = c1 * c2 * (y * y * y * y) / (t * t);
  } else {
    y = 1e4 / v;
    a = c1 * c2 * (y * y * y * y * y * y) * 1e-4 / (t * t);
  }
  tmp = c2 * y / t;
  bot = (exp(tmp) - 1.0) * (exp(tmp) - 1.0);
  top = exp(tmp) * a;
  return top / bot;
}

/* ----------------------------------------------------------------- pindex */
/* pindex.f:1-32; x1 is clamped in place (the reference's side effect). */
static int orc_pindex(const double *x, int nx, double *x1) {
  int i, j = 0;
  if (*x1 > x[nx - 1]) *x1 = x[nx - 1];
  if (*x1 < x[0]) *x1 = x[0];
  for (i = 1; i <= nx - 1; i++) {
    if (*x1 >= x[i - 1])
      j = i;
    else
      break;
  }
  return j;
}

/* ---------------------------------------------------------------- ciaread */
/* ciaread.f:104-196.  alpha/dalphadt are [9]. */
void orc_ciaread(const orc_cia *cia, double vv, double temp, double *alpha,
                 double *dalphadt) {
  double vvk1_lo = 0.0, vvk1_hi = (CIA_NUMWN - 1) * cia->dnu;
  double temp1;
  int cv, ct, i;
  for (i = 0; i < CIA_NPAIRS; i++) {
    alpha[i] = 0.0;
    dalphadt[i] = 0.0;
  }
  if (vv > vvk1_hi || vv < vvk1_lo) return;
  /* ciaread.f:133-157: the second test overwrites the first, so only the
   * lower clamp survives here; pindex then clamps both ends in place. */
  if (temp > cia->temps[CIA_NUMT - 1]) temp1 = cia->temps[CIA_NUMT - 1];
  else temp1 = temp;
  if (temp < cia->temps[0]) temp1 = cia->temps[0];
  else temp1 = temp;
  {
    /* vvk1(i) = (i-1)*dnu1 (ciaread.f:96-98); pindex over that grid. */
    double vtmp = vv;
    cv = 0;
    for (i = 1; i <= CIA_NUMWN - 1; i++) {
      if (vtmp >= (i - 1) * cia->dnu) cv = i;
      else break;
    }
  }
  ct = orc_pindex(cia->temps, CIA_NUMT, &temp1);
  for (i = 0; i < CIA_NPAIRS; i++) {
    /* kcia(pair, T, wn) == kcia[(wn*25 + T)*9 + pair], 0-based */
    const double *kc = cia->kcia;
    double abs11 = kc[((size_t)(cv - 1) * CIA_NUMT + (ct - 1)) * CIA_NPAIRS + i];
    double abs12 = kc[((size_t)(cv)*CIA_NUMT + (ct - 1)) * CIA_NPAIRS + i];
    double abs22 = kc[((size_t)(cv)*CIA_NUMT + (ct)) * CIA_NPAIRS + i];
    double abs21 = kc[((size_t)(cv - 1) * CIA_NUMT + (ct)) * CIA_NPAIRS + i];
    double v_lo = (cv - 1) * cia->dnu, v_hi = cv * cia->dnu;
    double t = (vv - v_lo) / (v_hi - v_lo);
    double u = (temp1 - cia->temps[ct - 1]) / (cia->temps[ct] - cia->temps[ct - 1]);
    double dudt = 1.0 / (cia->temps[ct] - cia->temps[ct - 1]);
    alpha[i] = (1.0 - t) * (1.0 - u) * abs11 + t * (1.0 - u) * abs12 + t * u * abs22 +
               (1.0 - t) * u * abs21;
    dalphadt[i] = (-(1.0 - t) * abs11 - t * abs12 + t * abs22 + (1.0 - t) * abs21) * dudt;
  }
}

/* verint.f:24-48: linear interpolation, the interval above the first grid
 * point past xin (ascending or descending grids), the end intervals
 * extrapolated; a zero-width interval or a NaN result gives y(i). */
double orc_verint(const double *x, const double *y, int n, double xin) {
  int i;
  double yout;
  if (x[0] < x[n - 1]) {
    for (i = 1; i <= n; i++)
      if (x[i - 1] > xin) break;
  } else {
    for (i = 1; i <= n; i++)
      if (x[i - 1] < xin) break;
  }
  if (i > n) i = n;
  if (i == 1) i = 2;
  yout = y[i - 2] + (y[i - 1] - y[i - 2]) * (xin - x[i - 2]) / (x[i - 1] - x[i - 2]);
  if (x[i - 2] == x[i - 1]) yout = y[i - 1];
  if (isnan(yout)) yout = y[i - 1];
  return yout;
}

/* n2n2cia.f and n2h2cia.f on the tables in memory: zero off the knots'
 * range, verint inside (the tables hold the functions' values, the
 * reference's 1e-5 km-1 -> cm-1 factor included). */
static void orc_n2nir(const orc_cia *cia, double v, double *an2n2, double *an2h2) {
  *an2n2 = *an2h2 = 0.0;
  if (cia->n1 > 1 && !(v < cia->wn1[0]) && !(v > cia->wn1[cia->n1 - 1]))
    *an2n2 = orc_verint(cia->wn1, cia->k1, cia->n1, v);
  if (cia->n2 > 1 && !(v < cia->wn2[0] || v > cia->wn2[cia->n2 - 1]))
    *an2h2 = orc_verint(cia->wn2, cia->k2, cia->n2, v);
}

/* nciacon.f:287-296 and nparacon_all.f:312-320: the N2-N2 and N2-H2
 * near-infrared terms, added to sum and DABSORB before the path factor */
static void orc_add_n2nir(const orc_cia *cia, double v, double qh2, double qn2, double *sum,
                          double *dabsorb) {
  double an2n2, an2h2;
  orc_n2nir(cia, v, &an2n2, &an2h2);
  *sum = *sum + an2n2 * qn2 * qn2;
  dabsorb[2] += 2 * qn2 * an2n2;
  *sum = *sum + an2h2 * qn2 * qh2;
  dabsorb[2] += qh2 * an2h2;
  dabsorb[0] += qn2 * an2h2;
}

/* get_solar_wave.f:19-46: interp.f on the stellar spectrum, zero outside
 * its range, over 4 pi (dist AU)^2 for dist > 0 (constdef.f's AU and pi) */
static double orc_solar(const orc_cia *cia, double x, double dist) {
  const int n = cia->nsol;
  double y = orc_interp(cia->swave, cia->srad, n, x);
  if (cia->swave[0] < cia->swave[n - 1]) {
    if (x < cia->swave[0] || x > cia->swave[n - 1]) y = 0.0;
  } else if (x > cia->swave[0] || x < cia->swave[n - 1]) {
    y = 0.0;
  }
  if (dist > 0.0) {
    const double au = 1.49598e13, area = 4 * PI_REF * (dist * au) * (dist * au);
    return y / area;
  }
  return y;
}

/* co2cia.f */
static double orc_co2cia(double v0) {
  double xl = 1e4 / v0, a = 0.0;
  if (xl >= 2.15 && xl <= 2.55) a = 4e-8;
  if (xl >= 1.7 && xl <= 1.76) a = 6e-9;
  if (xl >= 1.25 && xl < 1.35) a = 1.5e-9;
  if (xl >= 1.125 && xl <= 1.225) a = 0.5 * (0.31 + 0.79) * 1e-9;
  if (xl >= 1.06 && xl < 1.125) a = 0.5 * (0.29 + 0.67) * 1e-9;
  return a;
}

/* ----------------------------------------------------------------- dpexp */
/* dpexp.f:1-14: exp with its argument clamped to [-500, 88.7] (imod 0 and 1
 * transmissions, cirsradg_wave.f:861,888) */
static double orc_dpexp(double x) {
  double y = x > -500.0 ? x : -500.0;
  if (y >= 88.7) y = 88.7;
  return exp(y);
}

/* ----------------------------------------------------------- calc_nlte_t */
/* calc_nlte_t.f:59-85: the temperature whose Planck radiance is f times
 * the LTE one.  cirsradg_wave.f:933-935 passes its wavenumber index iwave
 * (>= 1) as the Planck form, so both planck_wave and invplanck_wave take
 * their iwave != 0 (wavelength) branch whatever ispace is; that is kept. */
double orc_nlte_t(int inlte, double k1, double a, double t, double p, double v) {
  const double kboltz = 1.380658e-23, c1 = 1.1911e-12, c2 = 1.439;
  double f, b, y;
  if (inlte == 1) {
    const double ntot = (p * 101325. / (kboltz * t)) * 1.0e-6;
    f = (k1 * ntot) / (a + k1 * ntot);
  } else if (inlte == 2) {
    f = p * 1.0e6 / (1.0 + p * 1.0e6);
  } else {
    return t;
  }
  /* planck_wave.f with iwave != 0 at T, then invplanck_wave.f of B*f */
  y = 1e4 / v;
  b = c1 * (y * y * y * y * y) / 1e4 / (exp(c2 * y / t) - 1.);
  b = b * f;
  if (b == 0.0 || v == 0.0) return 0.0;
  return (c2 * 1.0e4 / v) / log((c1 * 1.0e16) / (b * (v * v * v * v * v)) + 1.0);
}

/* ---------------------------------------------------------------- interp */
/* rtm_util/interp.f:1-26: linear interpolation (and extrapolation from the
 * end intervals) on a monotonic grid; x, y 0-based here. */
double orc_interp(const double *x, const double *y, int n, double xin) {
  int i;  /* 1-based */
  for (i = 1; i <= n; i++) {
    if (x[0] < x[n - 1] ? x[i - 1] > xin : x[i - 1] < xin) break;
  }
  if (i > n) i = n;
  if (i == 1) i = 2;
  return y[i - 2] + (y[i - 1] - y[i - 2]) * (xin - x[i - 2]) / (x[i - 1] - x[i - 2]);
}

/* --------------------------------------------------------------- cirsconv */
/* cspline.f:29-72 with natural ends (yp1, ypn > 0.99e30): second
 * derivatives y2[n] of the cubic spline through (x, y); u is scratch [n]. */
static void orc_cspline(const double *x, const double *y, int n, double *y2, double *u) {
  int i, k;
  y2[0] = 0.0;
  u[0] = 0.0;
  for (i = 1; i < n - 1; i++) {
    const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    const double p = sig * y2[i - 1] + 2.;
    y2[i] = (sig - 1.) / p;
    if (x[i - 1] == x[i] || x[i] == x[i + 1]) {
      u[i] = u[i - 1];
      continue;
    }
    u[i] = (6. * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                (x[i + 1] - x[i - 1]) -
            sig * u[i - 1]) /
           p;
  }
  y2[n - 1] = (0.0 - 0.0 * u[n - 2]) / (0.0 * y2[n - 2] + 1.);
  for (k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

/* csplint.f:10-40: the spline at x by bisection for its interval; results
 * within 1e-37 of zero become 1e-37 (:34-38) */
static double orc_csplint(const double *xa, const double *ya, const double *y2a, int n, double x) {
  int klo = 1, khi = n;
  double h, a, b, y1;
  while (khi - klo > 1) {
    const int k = (khi + klo) / 2;
    if (xa[k - 1] > x) khi = k;
    else klo = k;
  }
  h = xa[khi - 1] - xa[klo - 1];
  a = (xa[khi - 1] - x) / h;
  b = (x - xa[klo - 1]) / h;
  y1 = a * ya[klo - 1] + b * ya[khi - 1] +
       ((a * a * a - a) * y2a[klo - 1] + (b * b * b - b) * y2a[khi - 1]) * (h * h) / 6.;
  return fabs(y1) > 1.00e-37 ? y1 : 1.00e-37;
}

int orc_cirsconv(double fwhm, int nwave, const double *vwave, const double *y, int nconv,
                 const double *vconv, double *yout, int nfwhm, const double *vfwhm,
                 const double *xfwhm, int nfil, const int *fil_n, const double *fil_vcen,
                 const double *fil_v, const double *fil_f) {
  int i, j, k;
  if (fwhm == 0.0) {  /* :142-146 */
    for (i = 0; i < nconv; i++) yout[i] = y[i];
    return ORC_OK;
  }
  if (nwave < 2) return ORC_ERR_LIMITS;  /* :70-77 sets yout(1) = 1 */
  if (fwhm > 0.0) {   /* :79-140 */
    const int nstep = 20;
    double *xc = (double *)malloc(sizeof(double) * (nwave + 2) * 4);
    double *yc = xc + (nwave + 2), *y2 = yc + (nwave + 2), *u = y2 + (nwave + 2);
    int nc = nwave, nc1 = 0;
    if (!xc) return ORC_ERR_ALLOC;
    for (i = 0; i < nwave; i++) {
      xc[i] = vwave[i];
      yc[i] = y[i];
    }
    if (vconv[nconv - 1] > vwave[nwave - 1] - fwhm / 2.) {
      xc[nc] = vconv[nconv - 1] + fwhm;
      yc[nc] = y[nwave - 1] + ((y[nwave - 1] - y[nwave - 2]) / (vwave[nwave - 1] - vwave[nwave - 2]) *
                               (xc[nc] - vwave[nwave - 1]));
      nc++;
    }
    if (vconv[0] < vwave[0] + fwhm / 2.) {
      for (i = nc; i >= 1; i--) {
        xc[i] = xc[i - 1];
        yc[i] = yc[i - 1];
      }
      xc[0] = vconv[0] - fwhm;
      yc[0] = y[0] + ((y[1] - y[0]) / (vwave[1] - vwave[0]) * (xc[0] - vwave[0]));
      nc++;
    }
    for (i = 0; i < nc; i++)  /* NaN points are dropped (:101-108) */
      if (!isnan(yc[i])) {
        xc[nc1] = xc[i];
        yc[nc1] = yc[i];
        nc1++;
      }
    nc = nc1;
    orc_cspline(xc, yc, nc, y2, u);
    for (i = 0; i < nconv; i++) {
      double yf = fwhm, x1, delx, sum = 0.0, yold = 0.0;
      if (nfwhm > 0) yf = orc_verint(vfwhm, xfwhm, nfwhm, vconv[i]);
      x1 = vconv[i] - yf / 2.;
      delx = ((vconv[i] + yf / 2.) - x1) / (double)(nstep - 1);
      for (j = 1; j <= nstep; j++) {
        const double xi = x1 + (j - 1) * delx;
        const double yi = orc_csplint(xc, yc, y2, nc, xi);
        if (j == 1) sum = 0.;
        else sum = sum + (yi + yold) * (delx / 2.);
        yold = yi;
      }
      yout[i] = sum / fwhm;
    }
    free(xc);
    return ORC_OK;
  }
  /* fwhm < 0, :148-243: filter functions, normalised over the wavenumber
   * intervals (xnorm) */
  {
    int off = 0, done = 0;
    char *seen = (char *)calloc((size_t)nconv, 1);
    if (!seen) return ORC_ERR_ALLOC;
    for (k = 0; k < nfil; k++) {
      const double *fv = fil_v + off, *ff = fil_f + off;
      const int ns = fil_n[k];
      for (i = 0; i < nconv; i++) {
        double sum = 0.0, sumf = 0.0;
        if (!(fabs(fil_vcen[k] - vconv[i]) < 0.00001)) continue;
        for (j = 0; j < nwave - 1; j++) {
          double yy = 0.0;
          if (vwave[j] >= fv[0] && vwave[j] <= fv[ns - 1]) yy = orc_interp(fv, ff, ns, vwave[j]);
          if (yy != 0) {
            const double delv = vwave[j + 1] - vwave[j];
            sum = sum + yy * y[j] * delv;
            sumf = sumf + yy * delv;
          }
        }
        yout[i] = sum / sumf;
        if (!seen[i]) done++;
        seen[i] = 1;
      }
      off += ns;
    }
    free(seen);
    return done == nconv ? ORC_OK : ORC_ERR_LIMITS;
  }
}

/* ---------------------------------------------------------------- Rayleigh */
/* Rayleigh cross-section per molecule (cm2) at wavenumber v (cm-1), with the
 * reference's pi literal (PI_REF: rayleigha.f:21, rayleighj.f:22 and
 * constdef.f:11 all write 3.1415927).
 *   iray 1  rayleighj.f:1-52   H2/He mixture (fH2 = 0.864), no anisotropy
 *   iray 2  rayleighv.f:1-29   C / lambda^4, C = 8.8e-28
 *   iray 3  rayleigha.f:1-52   Earth air with the (P, T) refractivity
 *                              correction and faniso(0.035) (rayleigha.f:55-66)
 *   iray 4  rayleighls.f:1-71  H2, He, CH4, NH3 mixture from the ratios
 *                              He/H2, CH4/H2 and the NH3 fraction */
static double orc_faniso(double delta) { return (6.0 + 3.0 * delta) / (6.0 - 7.0 * delta); }

double orc_rayleigh(int iray, double v, double p, double t, double fheh2, double fch4h2,
                    double fnh3) {
  const double pi = PI_REF;
  if (iray == 1) {
    const double ah2 = 13.58e-5, bh2 = 7.52e-3, ahe = 3.48e-5, bhe = 2.30e-3, fh2 = 0.864;
    const double k = 1.37971e-23, p0 = 1.013e5, t0 = 273.15;
    double lambda = 1e-2 / v;           /* m */
    double x = 1.0 / (lambda * 1e6);    /* micron^-1 */
    double nh2 = ah2 * (1.0 + bh2 * x * x);
    double nhe = ahe * (1.0 + bhe * x * x);
    double nair = fh2 * nh2 + (1 - fh2) * nhe;
    double temp = 32 * (pi * pi * pi) * nair * nair;
    double n0 = p0 / (k * t0);
    x = n0 * lambda * lambda;
    return temp * 1e4 * orc_faniso(0.0) / (3 * (x * x));
  }
  if (iray == 2) {
    double lambda = 10000 / v;          /* micron */
    lambda = lambda * lambda * lambda * lambda;
    return 8.8e-28 / lambda;
  }
  if (iray == 3) {
    const double c1 = 8342.12, c2 = 2406030.0, c3 = 15997.0, delta = 0.035, k = 1.37971e-23;
    double lambda = 1e4 / v;            /* micron */
    double nu2 = 1.0 / (lambda * lambda);
    double m1 = c1 + c2 / (130.0 - nu2) + c3 / (38.9 - nu2);
    double p1, t1, xcorr, ns, x, temp;
    m1 = m1 * 1e-8;
    p1 = p * 1.013e5;                   /* Pa */
    t1 = t - 273.15;                    /* C */
    xcorr = p1 * (1 + p1 * (61.3 - t1) * 1e-10) / (96095.4 * (1 + 0.003661 * t1));
    m1 = m1 * xcorr;
    ns = p1 / (k * t);
    lambda = lambda * 1e-6;             /* m */
    x = ns * lambda * lambda;
    temp = 8 * (pi * pi * pi) * ((m1 * (m1 + 2.0)) * (m1 * (m1 + 2.0)));
    return temp * 1e4 * orc_faniso(delta) / (3 * (x * x));
  }
  {
    /* rayleighls.f: A, B, D for H2, He, CH4, NH3 */
    static const double a[4] = {13.58e-5, 3.48e-5, 37.0e-5, 37.0e-5};
    static const double b[4] = {7.52e-3, 2.3e-3, 12.0e-3, 12.0e-3};
    static const double d[4] = {0.0221, 0.025, .0922, .0922};
    double comp[4], h2overtot, loschpm3, wl, xc1 = 0.0, sumwt = 0.0, fact;
    int j;
    h2overtot = (1.0 - fnh3) / (1 + fheh2 + fch4h2);
    comp[0] = h2overtot;
    comp[1] = fheh2 * h2overtot;
    comp[2] = fch4h2 * h2overtot;
    comp[3] = fnh3;
    loschpm3 = 2.687e19 * 1e-12;
    wl = 1e4 / v;
    for (j = 0; j < 4; j++) {
      double nr = 1.0 + a[j] * (1.0 + b[j] / (wl * wl));
      xc1 = xc1 + (nr * nr - 1.0) * (nr * nr - 1.0) * comp[j] * (6 + 3 * d[j]) / (6 - 7 * d[j]);
      sumwt = sumwt + comp[j];
    }
    fact = 8.0 * (pi * pi * pi) / (3.0 * (wl * wl * wl * wl) * (loschpm3 * loschpm3));
    return fact * 1e-8 * xc1 / sumwt;
  }
}

/* --------------------------------------------------------------- ngascon */
/* H- free-free, hmin_ff.f:44-63: the sum over Bell & Berrington's table 3a
 * (wavelength > 0.3645 micron) or 3b (0.1823-0.3645 micron) terms; *inr = 0
 * outside both ranges, where the reference sets summ to 0 (:61-62). */
static double orc_hmin_ff_terms(double v0, double temp, int *inr) {
  static const double a1[6] = {0.0, 2483.346, -3449.889, 2200.040, -696.271, 88.283};
  static const double b1[6] = {0.0, 285.827, -1158.382, 2427.719, -1841.400, 444.517};
  static const double c1[6] = {0.0, -2054.291, 8746.523, -13651.105, 8624.970, -1863.864};
  static const double d1[6] = {0.0, 2827.776, -11485.632, 16755.524, -10051.530, 2095.288};
  static const double e1[6] = {0.0, -1341.537, 5303.609, -7510.494, 4400.067, -901.788};
  static const double f1[6] = {0.0, 208.952, -812.939, 1132.738, -655.020, 132.985};
  static const double a2[4] = {518.1021, 473.2636, -482.2089, 115.5291};
  static const double b2[4] = {-734.8666, 1443.4137, -737.1616, 169.6374};
  static const double c2[4] = {1021.1775, -1977.3395, 1096.8827, -245.649};
  static const double d2[4] = {-479.0721, 922.3575, -521.1341, 114.243};
  static const double e2[4] = {93.1373, -178.9275, 101.7963, -21.9972};
  static const double f2[4] = {-6.4285, 12.3600, -7.0571, 1.5097};
  const double wv = 1e4 / v0, th = 5040.0 / temp;
  const double *A, *B, *Cc, *D, *E, *F;
  double s = 0.0;
  int n, k;
  if (wv > 0.3645) {
    A = a1; B = b1; Cc = c1; D = d1; E = e1; F = f1; n = 6;
  } else if (wv < 0.3645 && wv > 0.1823) {
    A = a2; B = b2; Cc = c2; D = d2; E = e2; F = f2; n = 4;
  } else {
    *inr = 0;
    return 0.0;
  }
  *inr = 1;
  for (k = 0; k < n; k++)
    s = s + pow(th, ((k + 1) + 1.0) / 2.0) *
                (wv * wv * A[k] + B[k] + Cc[k] / wv + D[k] / (wv * wv) + E[k] / (wv * wv * wv) +
                 F[k] / ((wv * wv) * (wv * wv)));
  return s;
}

/* hmin_ff.f:65-67 */
static double orc_hmin_ff_value(double summ) { return summ < 0.0 ? 0.0 : 1e-29 * summ; }

/* H- bound-free, hmin_bf.f:23-33: the sum of the six terms below the
 * photo-detachment threshold lambda0 = 1.6419 micron (*inr = 0 above it,
 * where summ is set to 0), and sig = 1e-18 wv^3 (1/wv - 1/lambda0)^1.5
 * (:34). */
static double orc_hmin_bf_terms(double v0, int *inr, double *sig) {
  static const double coeff[6] = {152.519, 49.534, -118.858, 92.536, -34.194, 4.982};
  const double lambda0 = 1.6419, wv = 1e4 / v0;
  double s = 0.0;
  int i;
  if (!(wv < lambda0)) {
    *inr = 0;
    *sig = 0.0;
    return 0.0;
  }
  *inr = 1;
  for (i = 0; i < 6; i++) s = s + coeff[i] * pow((1.0 / wv) - (1.0 / lambda0), ((i + 1) - 1.0) / 2.0);
  *sig = 1e-18 * pow(wv, 3.0) * pow((1.0 / wv) - (1.0 / lambda0), 3.0 / 2.0);
  return s;
}

double orc_hmin_ff(double v0, double temp) {
  int inr;
  return orc_hmin_ff_value(orc_hmin_ff_terms(v0, temp, &inr));
}

double orc_hmin_bf(double v0) {
  int inr;
  double sig, s = orc_hmin_bf_terms(v0, &inr, &sig);
  return inr ? sig * s : 0.0;
}

/* ch4cont.f:44-57 and nh3cont.f:46-70: the table's coefficient at
 * lambda = 1e4/nu micron (10 micron for nu <= 1000), 0 off the table,
 * linear between the points locate.f brackets.  At the table's last point
 * locate returns n, and the reference reads one element past the wavelength
 * column; the last interval is used here. */
static double orc_tabcont(const double *wl, const double *kt, int n, double nu) {
  double lambda = 10.0, fr;
  int k;
  if (n < 2) return 0.0;
  if (nu > 1000.0) lambda = 1e4 / nu;
  if (lambda < wl[0] || lambda > wl[n - 1]) return 0.0;
  k = orc_locate(wl, n, lambda);
  if (k >= n) k = n - 1;
  if (k < 1) k = 1;
  fr = (lambda - wl[k - 1]) / (wl[k] - wl[k - 1]);
  return kt[k - 1] + fr * (kt[k] - kt[k - 1]);
}

double orc_ch4cont(const orc_cia *cia, double nu) {
  return orc_tabcont(cia->ch4_wl, cia->ch4_k, cia->nch4, nu) / 26850.0;
}

double orc_nh3cont(const orc_cia *cia, double nu) {
  return 1000 * orc_tabcont(cia->nh3_wl, cia->nh3_k, cia->nnh3, nu) / 26850.0;
}

/* CKD 2.1 self-continuum modification factors at 700, 710, ... 1200 cm-1
 * (ngascon.f:103-113) */
static const double CKD_XFAC[51] = {
    1.00000, 1.01792, 1.03767, 1.05749, 1.07730, 1.09708, 1.10489, 1.11268, 1.12047,
    1.12822, 1.13597, 1.14367, 1.15135, 1.15904, 1.16669, 1.17431, 1.18786, 1.20134,
    1.21479, 1.22821, 1.24158, 1.26580, 1.28991, 1.28295, 1.27600, 1.26896, 1.25550,
    1.24213, 1.22879, 1.21560, 1.20230, 1.18162, 1.16112, 1.14063, 1.12016, 1.10195,
    1.09207, 1.08622, 1.08105, 1.07765, 1.07398, 1.06620, 1.05791, 1.04905, 1.03976,
    1.02981, 1.00985, 1.00000, 1.00000, 1.00000, 1.00000};

/* The CKD H2O continuum, ngascon.f:144-213 */
static double orc_h2ocont(const orc_cia *c, double ff, double amount, double ppress,
                          double press, double temp) {
  const double alpha2 = 200.0 * 200.0, alphs2 = 120.0 * 120.0, betas = 5.0e-6, v0s = 1310.0,
               factrs = 0.15;
  const double hwsqf = 330. * 330., betaf = 8.e-11, v0f = 1130., factrf = 0.97, v0f2 = 1900.,
               hwsqf2 = 150. * 150., beta2 = 3.e-06;
  const double vt1 = c->h2o_v0 + c->h2o_dv * (c->nh2o - 1);
  double gr, csff60, csff96, cfff96, tfac, csfft, sfac, vs2, s1, s2, vf2, vf6, vf4, fscal;
  int mk;
  if (!(ff >= c->h2o_v0 && ff < vt1)) return 0.0;
  mk = 1 + (int)((ff - c->h2o_v0) / c->h2o_dv);
  if (mk == c->nh2o) mk = c->nh2o - 1;
  gr = (ff - (c->h2o_v0 + c->h2o_dv * (mk - 1))) / c->h2o_dv;
  csff60 = c->h2o_s260[mk - 1] + gr * (c->h2o_s260[mk] - c->h2o_s260[mk - 1]);
  csff96 = c->h2o_s296[mk - 1] + gr * (c->h2o_s296[mk] - c->h2o_s296[mk - 1]);
  cfff96 = c->h2o_f296[mk - 1] + gr * (c->h2o_f296[mk] - c->h2o_f296[mk - 1]);
  tfac = (temp - 296.0) / (260.0 - 296.0);
  csfft = csff96 * pow(csff60 / csff96, tfac);
  if (ff >= 700.0 && ff < 1200.0) {
    const int ic = (int)(0.1 * (ff - 700.0));
    const double gc = (ff - (700.0 + ic * 10.0)) / 10.0;
    sfac = CKD_XFAC[ic] + gc * (CKD_XFAC[ic + 1] - CKD_XFAC[ic]);
  } else {
    sfac = 1.0;
  }
  vs2 = (ff - v0s) * (ff - v0s);
  s1 = (1.0 - 0.2333 * (alpha2 / ((ff - 1050.0) * (ff - 1050.0) + alpha2)));
  s2 = (1.0 - factrs * (alphs2 / (vs2 + (betas * (vs2 * vs2)) + alphs2)));
  csfft = sfac * s1 * s2 * csfft;
  vf2 = (ff - v0f) * (ff - v0f);
  vf6 = vf2 * vf2 * vf2;
  fscal = (1. - factrf * (hwsqf / (vf2 + (betaf * vf6) + hwsqf)));
  vf2 = (ff - v0f2) * (ff - v0f2);
  vf4 = vf2 * vf2;
  fscal = fscal * (1. - 0.6 * (hwsqf2 / (vf2 + beta2 * vf4 + hwsqf2)));
  cfff96 = cfff96 * fscal;
  {
    const double c11 = 0.5 * c->h2o_c2 / temp;
    const double a1 = ff * amount * tanh(c11 * ff);
    const double a2 = c->h2o_ts / (temp * c->h2o_ps);
    const double a3 = 1.0e-20 * (ppress * csfft + (press - ppress) * cfff96);
    return a1 * a2 * a3;
  }
}

/* /ELECTRON/ (electron.f) as cirsradg_wave.f:590-599 sets it, and the
 * running `summ` of the H- functions (see orc_cia.hmin_static) */
typedef struct orc_hstate {
  double pelec, amounthmin, ppresshmin;
  int eflag;
  double sum_ff, sum_bf;
} orc_hstate;

/* ngascon.f:136-267 for one gas: absorption (a tau) of gas (id, iso) at
 * wavenumber v0.  Returns ORC_ERR_UNSUPPORTED for the branches not
 * restated: IH2O = 2 (the reference leaves ABSORB unassigned, :215-221),
 * ozone (IO3, :238-244: its cross-section file is not in the reference
 * tree) and the UV cross-sections (iuvscat, :274-318). */
static int orc_ngascon(const orc_cia *c, double v0, int id, int iso, double amount,
                       double ppress, double press, double temp, orc_hstate *hs,
                       double *absorb) {
  const int iso01 = iso == 0 || iso == 1;
  *absorb = 0.0;
  if (c->iuvscat > 0) {
    static const int uvid[10] = {11, 28, 26, 27, 30, 6, 5, 33, 1, 34};
    int b;
    for (b = 0; b < 10; b++)
      if (id == uvid[b] && (c->iuvscat & (1 << b))) return ORC_ERR_UNSUPPORTED;
  }
  if (id == 1 && c->ih2o == 1) {
    if (c->nh2o < 2) return ORC_ERR_UNSUPPORTED;
    *absorb = orc_h2ocont(c, v0, amount, ppress, press, temp);
  } else if (id == 1 && iso01 && c->ih2o == 2) {
    return ORC_ERR_UNSUPPORTED;
  } else if (id == 6 && iso01 && c->ich4 != 0) {
    if (c->nch4 < 2) return ORC_ERR_UNSUPPORTED;
    *absorb = amount * 1e-20 * orc_ch4cont(c, v0);
  } else if (id == 11 && iso01 && c->inh3 != 0) {
    if (c->nnh3 < 2) return ORC_ERR_UNSUPPORTED;
    *absorb = amount * 1e-20 * orc_nh3cont(c, v0);
  } else if (id == 3 && iso01 && c->io3 != 0) {
    return ORC_ERR_UNSUPPORTED;
  } else if (id == 48 && iso == 0 && hs->eflag > 0) {
    /* free-free: needs e- (80) and H- (91) set earlier in this layer */
    int inr;
    const double t = orc_hmin_ff_terms(v0, temp, &inr);
    const double nh = (ppress * 1.013e6 * hs->pelec) / hs->ppresshmin;
    if (c->hmin_static) hs->sum_ff = inr ? hs->sum_ff + t : 0.0;
    else hs->sum_ff = t;
    *absorb = hs->amounthmin * orc_hmin_ff_value(hs->sum_ff) * nh;
  } else if (id == 91 && iso == 0) {
    int inr;
    double sig;
    const double t = orc_hmin_bf_terms(v0, &inr, &sig);
    if (c->hmin_static) hs->sum_bf = inr ? hs->sum_bf + t : 0.0;
    else hs->sum_bf = t;
    *absorb = amount * (inr ? sig * hs->sum_bf : 0.0);
  }
  return ORC_OK;
}

/* The gas order the electron state needs: with e- (80) present, H- free-free
 * (48, isotope 0) reads the e- partial pressure and the H- (91) amount and
 * partial pressure cirsradg_wave.f:590-599 stored for earlier gases, so
 * every 80 and 91 must come before every such 48, and a 91 must be present;
 * otherwise the values would be a previous layer's or call's. */
static int orc_hmin_order_ok(int ngas, const int *idgas, const int *isogas) {
  int k, has80 = 0, has91 = 0, first48 = -1, last8091 = -1;
  for (k = 0; k < ngas; k++) {
    if (idgas[k] == 80) has80 = 1;
    if (idgas[k] == 91) has91 = 1;
    if (idgas[k] == 80 || idgas[k] == 91) last8091 = k;
    if (idgas[k] == 48 && isogas[k] == 0 && first48 < 0) first48 = k;
  }
  if (!has80 || first48 < 0) return 1;
  return has91 && last8091 < first48;
}

/* nciacon.f:65-180 (and nparacon_all.f's copy of it): the mole fractions
 * q[5] = H2, He, N2, CH4, CO2 with their 1-based gas numbers iabs[5], the
 * zeroed dabsorb[7] and the path factor tau = height * amagat^2. */
static void orc_cia_setup(double p, double t, int ngas, const int *idgas, const int *isogas,
                          const double *amount, const double *pp, double xlen, double *q,
                          int *iabs, double *dabsorb, double *tau_out) {
  const double P0 = 1.0, T0 = 273.15;
  const double MODBOLTZA = 10.0 * KBOLTZMANN / 1.013;
  double qh2 = 0, qhe = 0, qn2 = 0, qch4 = 0, qco2 = 0;
  double xp1, height = 0.0, tau, totam = 0.0, amag1;
  int i;
  for (i = 0; i < 5; i++) {
    iabs[i] = 0;
    dabsorb[i] = 0.0;
  }
  dabsorb[5] = 0.0;
  dabsorb[6] = 0.0;
  for (i = 0; i < ngas; i++) {
    if (idgas[i] == 39 && (isogas[i] == 0 || isogas[i] == 1)) {
      qh2 = pp[i] / p;
      iabs[0] = i + 1;
    }
    if (idgas[i] == 40) {
      qhe = pp[i] / p;
      iabs[1] = i + 1;
    }
    if (idgas[i] == 22) {
      qn2 = pp[i] / p;
      iabs[2] = i + 1;
    }
    if (idgas[i] == 6 && (isogas[i] == 0 || isogas[i] == 1)) {
      qch4 = pp[i] / p;
      iabs[3] = i + 1;
    }
    if (idgas[i] == 2) {
      qco2 = pp[i] / p;
      iabs[4] = i + 1;
    }
  }
  if (xlen == 0.0) {
    xp1 = 0.0;
    for (i = 0; i < ngas; i++) {
      if (pp[i] > xp1) {
        height = amount[i] * MODBOLTZA * t / pp[i];
        xp1 = pp[i];
      }
    }
    tau = ((p / P0) * (p / P0)) * ((T0 / t) * (T0 / t)) * height;
  } else {
    height = xlen * 1e5;
    xp1 = -1.0;
    for (i = 0; i < ngas; i++) {
      if (pp[i] > xp1 && pp[i] > 0.0) {
        totam = amount[i] * p / pp[i];
        xp1 = pp[i];
      }
    }
    amag1 = (totam / height) / AMAGAT;
    tau = height * amag1 * amag1;
  }
  q[0] = qh2;
  q[1] = qhe;
  q[2] = qn2;
  q[3] = qch4;
  q[4] = qco2;
  *tau_out = tau;
}

/* ---------------------------------------------------------------- nciacon */
/* nciacon.f:1-310.  iabs[5] are 1-based gas numbers (0 = absent),
 * dabsorb[7]. */
void orc_nciacon(const orc_cia *cia, double v0, double p, double t, int inormal,
                 int ngas, const int *idgas, const int *isogas,
                 const double *amount, const double *pp, double xlen,
                 double *absorb, int *iabs, double *dabsorb) {
  double q[5], tau, sum, aco2;
  double alpha[CIA_NPAIRS], dalpha[CIA_NPAIRS];
  int i;
  orc_cia_setup(p, t, ngas, idgas, isogas, amount, pp, xlen, q, iabs, dabsorb, &tau);
  const double qh2 = q[0], qhe = q[1], qn2 = q[2], qch4 = q[3], qco2 = q[4];
  sum = 0.0;
  orc_ciaread(cia, v0, t, alpha, dalpha);
  if (inormal == 0) {
    sum = sum + (alpha[0] * qh2 * qh2) + (alpha[1] * qh2 * qhe);
    dabsorb[0] += 2 * qh2 * alpha[0] + qhe * alpha[1];
    dabsorb[1] += qh2 * alpha[1];
    dabsorb[5] += qh2 * qh2 * dalpha[0] + qh2 * qhe * dalpha[1];
  } else {
    sum = sum + (alpha[2] * qh2 * qh2) + (alpha[3] * qh2 * qhe);
    dabsorb[0] += 2 * qh2 * alpha[2] + qhe * alpha[3];
    dabsorb[1] += qh2 * alpha[3];
    dabsorb[5] += qh2 * qh2 * dalpha[2] + qh2 * qhe * dalpha[3];
  }
  sum = sum + (alpha[4] * qh2 * qn2);
  dabsorb[0] += qn2 * alpha[4];
  dabsorb[2] += qh2 * alpha[4];
  dabsorb[5] += qh2 * qn2 * dalpha[4];
  sum = sum + (alpha[5] * qn2 * qch4);
  dabsorb[2] += qch4 * alpha[5];
  dabsorb[3] += qn2 * alpha[5];
  dabsorb[5] += qn2 * qch4 * dalpha[5];
  sum = sum + (alpha[6] * qn2 * qn2);
  dabsorb[2] += 2 * qn2 * alpha[6];
  dabsorb[5] += qn2 * qn2 * dalpha[6];
  sum = sum + (alpha[7] * qch4 * qch4);
  dabsorb[3] += 2 * qch4 * alpha[7];
  dabsorb[5] += qch4 * qch4 * dalpha[7];
  sum = sum + (alpha[8] * qh2 * qch4);
  dabsorb[0] += qch4 * alpha[8];
  dabsorb[3] += qh2 * alpha[8];
  dabsorb[5] += qh2 * qch4 * dalpha[8];
  aco2 = orc_co2cia(v0);
  sum = sum + aco2 * qco2 * qco2;
  dabsorb[4] += 2 * qco2 * aco2;
  orc_add_n2nir(cia, v0, qh2, qn2, &sum, dabsorb);
  *absorb = sum * tau;
  for (i = 0; i < 7; i++) dabsorb[i] *= tau;
}

/* ------------------------------------------------------------- fpread_all */
/* fpread_all.f:168-257 on the /FPTABLE/ already in memory: bilinear (v, T)
 * interpolation of every para fraction's H2-H2 and H2-He coefficients and of
 * the five other pairs, and their T-derivatives.  The temperature is
 * clamped by pindex in place as in ciaread.  Off the wavenumber grid the
 * reference returns without writing its outputs (stale values); here they
 * are 0. */
static void orc_fpread_all(const orc_cia *cia, double vv, double temp, double *absh2,
                           double *abshe, double *dah2dt, double *dahedt, double *deltaf,
                           double *abs5, double *dabs5) {
  const int np = cia->npara;
  const double d2 = cia->dnu2;
  double temp1 = temp, t, u, dudt;
  int cv = 0, ct, i;
  *deltaf = cia->fp[1] - cia->fp[0];
  for (i = 0; i < np; i++) absh2[i] = abshe[i] = dah2dt[i] = dahedt[i] = 0.0;
  for (i = 0; i < 5; i++) abs5[i] = dabs5[i] = 0.0;
  if (vv > (CIA_NUMWN - 1) * d2 || vv < 0.0) return;
  for (i = 1; i <= CIA_NUMWN - 1; i++) {  /* pindex on vvk2(i) = (i-1)*dnu2 */
    if (vv >= (i - 1) * d2) cv = i;
    else break;
  }
  ct = orc_pindex(cia->temps2, CIA_NUMT, &temp1);
  t = (vv - (cv - 1) * d2) / (cv * d2 - (cv - 1) * d2);
  u = (temp1 - cia->temps2[ct - 1]) / (cia->temps2[ct] - cia->temps2[ct - 1]);
  dudt = 1.0 / (cia->temps2[ct] - cia->temps2[ct - 1]);
  /* kh2h2(i,ct,cv) == kh2h2[cv-1][ct-1][i] */
#define KP(a, c, tt, i) (a)[((size_t)((c) - 1) * CIA_NUMT + ((tt) - 1)) * np + (i)]
  for (i = 0; i < np; i++) {
    const double a11 = KP(cia->kh2h2, cv, ct, i), a12 = KP(cia->kh2h2, cv + 1, ct, i);
    const double a22 = KP(cia->kh2h2, cv + 1, ct + 1, i), a21 = KP(cia->kh2h2, cv, ct + 1, i);
    const double b11 = KP(cia->kh2he, cv, ct, i), b12 = KP(cia->kh2he, cv + 1, ct, i);
    const double b22 = KP(cia->kh2he, cv + 1, ct + 1, i), b21 = KP(cia->kh2he, cv, ct + 1, i);
    absh2[i] = (1.0 - t) * (1.0 - u) * a11 + t * (1.0 - u) * a12 + t * u * a22 + (1.0 - t) * u * a21;
    abshe[i] = (1.0 - t) * (1.0 - u) * b11 + t * (1.0 - u) * b12 + t * u * b22 + (1.0 - t) * u * b21;
    dah2dt[i] = (-(1.0 - t) * a11 - t * a12 + t * a22 + (1.0 - t) * a21) * dudt;
    dahedt[i] = (-(1.0 - t) * b11 - t * b12 + t * b22 + (1.0 - t) * b21) * dudt;
  }
#undef KP
  for (i = 0; i < 5; i++) {
    const double *k5 = cia->kcia5;
    const double a11 = k5[((size_t)(cv - 1) * CIA_NUMT + (ct - 1)) * 5 + i];
    const double a12 = k5[((size_t)cv * CIA_NUMT + (ct - 1)) * 5 + i];
    const double a22 = k5[((size_t)cv * CIA_NUMT + ct) * 5 + i];
    const double a21 = k5[((size_t)(cv - 1) * CIA_NUMT + ct) * 5 + i];
    abs5[i] = (1.0 - t) * (1.0 - u) * a11 + t * (1.0 - u) * a12 + t * u * a22 + (1.0 - t) * u * a21;
    dabs5[i] = (-(1.0 - t) * a11 - t * a12 + t * a22 + (1.0 - t) * a21) * dudt;
  }
}

/* ----------------------------------------------------------- nparacon_all */
/* nparacon_all.f:1-338: CIA with a para-H2 fraction fpara (fpara < 0: the
 * table's last fraction, no fpara derivative).  The fraction index follows
 * the reference, x1 = 1 + (fpara - 0.25)/deltaf, j = INT(x1), clamped to
 * [1, npara-2] with f = 0 below and f = 1 above, so the last interval is
 * never interpolated into.  dabsorb[6] is d(tau)/d(fpara). */
static void orc_nparacon_all(const orc_cia *cia, double vwave, double p, double t, int ngas,
                             const int *idgas, const int *isogas, const double *amount,
                             const double *pp, double fpara, double xlen, double *absorb,
                             int *iabs, double *dabsorb) {
  double q[5], tau, sum, aco2, deltaf, f;
  double h2a[64], hea[64], dh2[64], dhe[64], a5[5], da5[5];
  double ah2h2, ah2he, dh2dt, dhedt, dh2dfp, dhedfp, x1;
  const int kpara = cia->npara;
  int i, j;
  orc_cia_setup(p, t, ngas, idgas, isogas, amount, pp, xlen, q, iabs, dabsorb, &tau);
  const double qh2 = q[0], qhe = q[1], qn2 = q[2], qch4 = q[3], qco2 = q[4];
  orc_fpread_all(cia, vwave, t, h2a, hea, dh2, dhe, &deltaf, a5, da5);
  if (fpara < 0.0) {
    ah2h2 = h2a[kpara - 1];
    ah2he = hea[kpara - 1];
    dh2dt = dh2[kpara - 1];
    dhedt = dhe[kpara - 1];
    dh2dfp = 0.0;
    dhedfp = 0.0;
  } else {
    x1 = 1.0 + (fpara - 0.25) / deltaf;
    j = (int)x1;
    f = x1 - j;
    if (j < 1) {
      j = 1;
      f = 0.0;
    } else if (j > kpara - 2) {
      j = kpara - 2;
      f = 1.0;
    }
    /* 1-based j -> h2a[j-1], h2a[j] */
    ah2h2 = (1.0 - f) * h2a[j - 1] + f * h2a[j];
    ah2he = (1.0 - f) * hea[j - 1] + f * hea[j];
    dh2dt = (1.0 - f) * dh2[j - 1] + f * dh2[j];
    dhedt = (1.0 - f) * dhe[j - 1] + f * dhe[j];
    dh2dfp = (h2a[j] - h2a[j - 1]) / deltaf;
    dhedfp = (hea[j] - hea[j - 1]) / deltaf;
  }
  sum = ah2h2 * qh2 * qh2 + ah2he * qh2 * qhe;
  dabsorb[0] = (2 * qh2 * ah2h2 + qhe * ah2he);
  dabsorb[1] = qh2 * ah2he;
  dabsorb[5] = (dh2dt * qh2 * qh2 + dhedt * qh2 * qhe);
  dabsorb[6] = (dh2dfp * qh2 * qh2 + dhedfp * qh2 * qhe);
  /* H2-N2, N2-CH4, N2-N2, CH4-CH4, H2-CH4 (nparacon_all.f's alpha(1..5)) */
  sum = sum + (a5[0] * qh2 * qn2);
  dabsorb[0] += qn2 * a5[0];
  dabsorb[2] += qh2 * a5[0];
  dabsorb[5] += qh2 * qn2 * da5[0];
  sum = sum + (a5[1] * qn2 * qch4);
  dabsorb[2] += qch4 * a5[1];
  dabsorb[3] += qn2 * a5[1];
  dabsorb[5] += qn2 * qch4 * da5[1];
  sum = sum + (a5[2] * qn2 * qn2);
  dabsorb[2] += 2 * qn2 * a5[2];
  dabsorb[5] += qn2 * qn2 * da5[2];
  sum = sum + (a5[3] * qch4 * qch4);
  dabsorb[3] += 2 * qch4 * a5[3];
  dabsorb[5] += qch4 * qch4 * da5[3];
  sum = sum + (a5[4] * qh2 * qch4);
  dabsorb[0] += qch4 * a5[4];
  dabsorb[3] += qh2 * a5[4];
  dabsorb[5] += qh2 * qch4 * da5[4];
  aco2 = orc_co2cia(vwave);
  sum = sum + aco2 * qco2 * qco2;
  dabsorb[4] += 2 * qco2 * aco2;
  orc_add_n2nir(cia, vwave, qh2, qn2, &sum, dabsorb);
  *absorb = sum * tau;
  for (i = 0; i < 7; i++) dabsorb[i] *= tau;
}

/* --------------------------------------------------------------- e3interp */
/* e3interp.f:14-79 on the table in memory: E3 and dE3/dtr at transmission x,
 * with egrad from the table's differences as e3interp.f:47-51 forms it. */
static void orc_e3interp(const orc_cia *cia, double x, double *e3, double *de3) {
  const int mtab = 1001;
  const double dtrans = 1.0 / (double)(mtab - 1);
  const double *tt = cia->e3tran, *te = cia->e3val;
  int j;
  double f, g0, g1;
  if (x < 0.0 || isnan(x)) {
    *e3 = 0.0;
    *de3 = 0.0;
    return;
  }
  j = 1 + (int)(x / dtrans);
  if (j >= mtab) j = mtab - 1;
  f = (x - tt[j - 1]) / dtrans;
  if (f > 1.0) f = 1.0;
#define EGRAD(i) ((i) == 1 ? (te[1] - te[0]) / dtrans                                      \
                  : (i) == mtab ? (te[mtab - 1] - te[mtab - 2]) / dtrans                   \
                                : 0.5 * (te[(i)] - te[(i) - 2]) / dtrans)
  g0 = EGRAD(j);
  g1 = EGRAD(j + 1);
#undef EGRAD
  *e3 = (1.0 - f) * te[j - 1] + f * te[j];
  *de3 = (1.0 - f) * g0 + f * g1;
}

double orc_e3(const double *trans, const double *e3, double x, double *de3) {
  orc_cia c;
  double v;
  memset(&c, 0, sizeof c);
  c.e3tran = trans;
  c.e3val = e3;
  orc_e3interp(&c, x, &v, de3);
  return v;
}

/* --------------------------------------------------------------- interpem */
static int orc_interpem(int nem, const double *vem, const double *em, double x,
                        double *out) {
  int i;
  double v = 0.0;
  if (nem < 1 || x < vem[0] || x > vem[nem - 1]) return ORC_ERR_EMISSIVITY;
  for (i = 0; i < nem - 1; i++) {
    if (x >= vem[i] && x < vem[i + 1]) {
      double f = (x - vem[i]) / (vem[i + 1] - vem[i]);
      v = (1 - f) * em[i] + f * em[i + 1];
    }
  }
  if (x == vem[nem - 1]) v = em[nem - 1];
  *out = v;
  return ORC_OK;
}

/* ----------------------------------------------------------------- get_kg */
/* Table for gas igas at vwave: the first registered table with the gas's
 * (id, iso) whose grid covers vwave (read_klist.f:260-290, findloc.f). */
/* read_klist.f:260-330: the first regular table of the gas covering the
 * wavenumber, or the gas's irregular table (delv <= 0), which covers every
 * wavenumber (the reference lets the last such table win) */
static const orc_ktable *find_table(const orc_ktable *tabs, int ntab, int id,
                                    int iso, double vwave) {
  int t;
  const orc_ktable *irr = NULL, *reg = NULL;
  for (t = 0; t < ntab; t++) {
    const orc_ktable *kt = &tabs[t];
    double xmax = kt->vmin + (kt->npoint - 1) * kt->delv;
    double tol = (kt->delv >= 0.0001) ? 0.02 * kt->delv : 0.25 * kt->delv;
    if (kt->idgas != id || kt->isogas != iso) continue;
    if (kt->delv <= 0.0) {
      irr = kt;
      continue;
    }
    /* grid ends widened by get_kg's coincidence tolerance (get_kg.f:199-206) */
    if (!reg && vwave >= kt->vmin - tol && vwave <= xmax + tol) reg = kt;
  }
  return irr ? irr : reg;
}

/* read_klist.f:292-330: the slice of an irregular table for wavenumber v.
 * delv < 0: the point nearest v in per cent (the first of equals); delv = 0:
 * the nearest point at or below v, with the lower slice's weight frac.
 * Returns -1 when no point is at or below v (the reference stops). */
static int orc_irregular_slice(const orc_ktable *kt, double v, double *frac) {
  double best = 1e10;
  int ipo, jpo = -1;
  *frac = 1.0;
  for (ipo = 0; ipo < kt->npoint; ipo++) {
    const double dx = 100 * fabs((v - kt->vcen[ipo]) / kt->vcen[ipo]);
    if (kt->delv == 0.0) {
      if (dx < best && kt->vcen[ipo] <= v) {
        jpo = ipo;
        *frac = ipo < kt->npoint - 1
                    ? 1.0 - (v - kt->vcen[ipo]) / (kt->vcen[ipo + 1] - kt->vcen[ipo])
                    : 1.;
        best = dx;
      }
    } else if (dx < best) {
      jpo = ipo;
      best = dx;
    }
  }
  return jpo;
}

/* get_kg.f:1-484 for one gas at all layers.  kout/dkoutdt are [L][ng].
 * lbl: get_klblg.f:1-312 instead (a line-by-line table, ng = 1): the slice
 * 1 + INT((v + 0.02 delv - vmin) / delv) (:182-185), which must exist
 * (:186-189 print a diagnostic and read on past the table), always
 * coincident (no merge), and the same zero-slice test (:200-222), clamps
 * and KLOG rule (:253-261) as get_kg.  A NaN, where get_klblg stops
 * (:284-301), takes get_kg's 1e-37; finite tables give none. */
static int orc_get_kg_gas(const orc_ktable *kt, int nlayer, const double *press,
                          const double *temp, double vwave, double *kout,
                          double *dkoutdt, int lbl) {
  const int np = kt->np, nt = kt->nt < 0 ? -kt->nt : kt->nt, ng = kt->ng, t2 = kt->nt < 0;
  const size_t ntab = (size_t)np * nt * ng;
  const double pmin = kt->lnp[0], pmax = kt->lnp[np - 1];
  const double tmin = t2 ? 0.0 : kt->tk[0], tmax = t2 ? 0.0 : kt->tk[nt - 1];
  double eps = 0.0, tmp = 0.0, frac, frack = 1.0;
  long n1;
  int coinc, il, i;
  const double *table, *table2;
  double ktest = 0.0;

  if (lbl) {
    const double x = (vwave + 0.02 * kt->delv - kt->vmin) / kt->delv;
    if (!(kt->delv > 0.0) || !(x > -1.0 && x < (double)kt->npoint)) return ORC_ERR_LIMITS;
    n1 = (long)x; /* INT truncates toward zero */
  } else if (kt->delv > 0.0) {
    if (kt->delv >= 0.0001) eps = 0.02 * kt->delv;
    else eps = 0.25 * kt->delv;
    n1 = (long)((vwave + eps - kt->vmin) / kt->delv); /* IDINT truncates */
    if (n1 < 0) n1 = 0;
    if (n1 > kt->npoint - 1) n1 = kt->npoint - 1;
  } else {
    n1 = orc_irregular_slice(kt, vwave, &frack);
    if (n1 < 0) return ORC_ERR_LIMITS;
  }
  table = kt->k + (size_t)n1 * ntab;

  for (i = 0; i < (int)ntab; i++) {
    if (table[i] > 0.0) {
      ktest = table[i];
      break;
    }
  }
  if (ktest == 0.0) {
    memset(kout, 0, sizeof(double) * nlayer * ng);
    memset(dkoutdt, 0, sizeof(double) * nlayer * ng);
    return ORC_OK;
  }
  if (lbl) {
    coinc = 1;
  } else if (kt->delv > 0.0) {
    tmp = kt->vmin + kt->delv * n1;
    frac = fabs(vwave - tmp);
    coinc = frac < eps;
  } else {
    coinc = kt->delv < 0.0;  /* get_kg.f:267-281 */
  }
  table2 = (n1 + 1 < kt->npoint) ? table + ntab : table;

  for (il = 0; il < nlayer; il++) {
    double p1 = log(press[il]), t1, v, u, dudt, u2, dudt2;
    int cp, ct, ct2, i1, i2, i3, i4;
    double kg[MAXG + 2], dkdt[MAXG + 2];
    if (p1 < pmin) p1 = pmin;
    if (p1 > pmax) p1 = pmax;
    cp = orc_locate(kt->lnp, np, p1);
    if (cp < 1) cp = 1;
    if (cp >= np) cp = np - 1;
    v = (p1 - kt->lnp[cp - 1]) / (kt->lnp[cp] - kt->lnp[cp - 1]);
    if (!t2) {
      t1 = temp[il];
      if (t1 < tmin) t1 = tmin;
      if (t1 > tmax) t1 = tmax;
      ct = orc_locate(kt->tk, nt, t1);
      if (ct < 1) ct = 1;
      if (ct >= nt) ct = nt - 1;
      u = (t1 - kt->tk[ct - 1]) / (kt->tk[ct] - kt->tk[ct - 1]);
      dudt = 1.0 / (kt->tk[ct] - kt->tk[ct - 1]);
      ct2 = ct;
      u2 = u;
      dudt2 = dudt;
    } else {
      /* get_kg.f:145-173: each bracketing pressure's own temperature grid */
      const double *tn = kt->t2 + (size_t)(cp - 1) * nt;
      t1 = temp[il];
      if (t1 < tn[0]) t1 = tn[0];
      if (t1 > tn[nt - 1]) t1 = tn[nt - 1];
      ct = orc_locate(tn, nt, t1);
      if (ct < 1) ct = 1;
      if (ct >= nt) ct = nt - 1;
      u = (t1 - tn[ct - 1]) / (tn[ct] - tn[ct - 1]);
      dudt = 1.0 / (tn[ct] - tn[ct - 1]);
      tn = kt->t2 + (size_t)cp * nt;
      t1 = temp[il];
      if (t1 < tn[0]) t1 = tn[0];
      if (t1 > tn[nt - 1]) t1 = tn[nt - 1];
      ct2 = orc_locate(tn, nt, t1);
      if (ct2 < 1) ct2 = 1;
      if (ct2 >= nt) ct2 = nt - 1;
      u2 = (t1 - tn[ct2 - 1]) / (tn[ct2] - tn[ct2 - 1]);
      dudt2 = 1.0 / (tn[ct2] - tn[ct2 - 1]);
    }
    /* IOFF, get_kg.f:140-143,168-171 (0-based offsets) */
    i1 = (nt * ng * (cp - 1)) + (ng * (ct - 1));
    i2 = (nt * ng * cp) + (ng * (ct2 - 1));
    i3 = (nt * ng * cp) + (ng * ct2);
    i4 = (nt * ng * (cp - 1)) + (ng * ct);

    if (coinc) {
      int klog = !(table[i1] <= 0.0 || table[i2] <= 0.0 || table[i3] <= 0.0 ||
                   table[i4] <= 0.0);
      for (i = 0; i < ng; i++) {
        double y1 = table[i1 + i], y2 = table[i2 + i], y3 = table[i3 + i], y4 = table[i4 + i];
        double x, dxdt;
        if (klog) {
          y1 = log(y1);
          y2 = log(y2);
          y3 = log(y3);
          y4 = log(y4);
        }
        if (!t2) {
          x = (1.0 - v) * (1.0 - u) * y1 + v * (1.0 - u) * y2 + v * u * y3 + (1.0 - v) * u * y4;
          dxdt = (-(1.0 - v) * y1 - v * y2 + v * y3 + (1 - v) * y4) * dudt;
        } else {  /* get_kg.f:352-357 */
          x = (1.0 - v) * ((1.0 - u) * y1 + u * y4) + v * ((1.0 - u2) * y2 + u2 * y3);
          dxdt = (1.0 - v) * (-y1 + y4) * dudt + v * (-y2 + y3) * dudt2;
        }
        if (klog) {
          kg[i + 1] = exp(x);
          dkdt[i + 1] = exp(x) * dxdt;
        } else {
          kg[i + 1] = x;
          dkdt[i + 1] = dxdt;
        }
      }
    } else {
      double w[2];
      double cont[2 * MAXG + 2], dcont[2 * MAXG + 2], fac[2 * MAXG + 2];
      int loop, count = 0;
      if (kt->delv > 0.0) {  /* get_kg.f:380-385 */
        double weight2 = (vwave - tmp) / kt->delv;
        if (weight2 > 1.0) weight2 = 1.0;
        if (weight2 < 0.0) weight2 = 0.0;
        w[0] = 1.0 - weight2;
        w[1] = weight2;
      } else {  /* :386-389 */
        w[0] = frack;
        w[1] = 1.0 - frack;
      }
      for (loop = 0; loop < 2; loop++) {
        const double *tb = loop == 0 ? table : table2;
        for (i = 0; i < ng; i++) {
          double y1 = tb[i1 + i], y2 = tb[i2 + i], y3 = tb[i3 + i], y4 = tb[i4 + i];
          double x, dxdt;
          if (!t2) {
            x = (1.0 - v) * (1.0 - u) * y1 + v * (1.0 - u) * y2 + v * u * y3 + (1.0 - v) * u * y4;
            dxdt = (-(1.0 - v) * y1 - v * y2 + v * y3 + (1 - v) * y4) * dudt;
          } else {
            x = (1.0 - v) * ((1.0 - u) * y1 + u * y4) + v * ((1.0 - u2) * y2 + u2 * y3);
            dxdt = (1.0 - v) * (-y1 + y4) * dudt + v * (-y2 + y3) * dudt2;
          }
          count++;
          cont[count] = x;
          dcont[count] = dxdt;
          fac[count] = kt->delg[i] * w[loop];
        }
      }
      orc_rankk(kt->delg, ng, cont, dcont, fac, count, kg, dkdt);
    }
    for (i = 0; i < ng; i++) {
      kout[il * ng + i] = isnan(kg[i + 1]) ? 1e-37 : kg[i + 1];
      dkoutdt[il * ng + i] = isnan(dkdt[i + 1]) ? 1e-37 : dkdt[i + 1];
    }
  }
  return ORC_OK;
}

/* ------------------------------------------------------------ cirsradg */
int orc_cirsradg_wave(const orc_ktable *tabs, int ntab, const orc_cia *cia,
                      const orc_req *r, double vwave, int jac_mode,
                      double *output, double *gtsurf, double *doutputdq) {
  const int L = r->nlayer, ngas = r->ngas, npath = r->npath;
  int ng = 0, nparam, ip, j, k, k1, l, ig, rc = ORC_OK;
  int jh2 = -1, jhe = -1, jch4 = -1, jnh3 = -1;
  orc_hstate hs = {0.0, 0.0, 0.0, 0, 0.0, 0.0};
  const double *delg = NULL;
  const orc_ktable *gtab[MAXGAS];
  double esurf = 0.0, vv = vwave;
  double *kout = NULL, *dkoutdt = NULL, *utotl = NULL, *taucon = NULL,
         *dtaucondq = NULL, *kl_g = NULL, *dkdql = NULL, *bb = NULL, *dbdt = NULL,
         *corkout = NULL, *tempgtsurf = NULL, *dcoutdq = NULL, *tautmp = NULL,
         *dtautmpdq = NULL, *taus = NULL, *dtaudq = NULL, *dtolddq = NULL,
         *dtrdq = NULL, *trv = NULL, *wv = NULL, *tlv = NULL;

  /* check_limits.f:10-78 */
  if (L > MAXLAY || L < 1 || npath > MAXPAT || npath < 1 || ngas > MAXGAS || ngas < 1)
    return ORC_ERR_LIMITS;
  for (ip = 0; ip < npath; ip++) {
    if (r->nlayin[ip] > 2 * MAXPAT || r->nlayin[ip] < 1) return ORC_ERR_LIMITS;
    for (j = 0; j < r->nlayin[ip]; j++)
      if (r->layinc[ip * L + j] > L || r->layinc[ip * L + j] < 1) return ORC_ERR_LIMITS;
    if (r->imod[ip] != 0 && r->imod[ip] != 1 && r->imod[ip] != 3 && r->imod[ip] != 8 &&
        r->imod[ip] != 13 && r->imod[ip] != 14 && r->imod[ip] != 18)
      return ORC_ERR_UNSUPPORTED;
    /* imod 8 (cirsradg_wave.f:1037-1082): the product of path layinc(1), a
     * cell (imod 13 or 14, else the reference stops), and path layinc(2),
     * both computed before it in the same g-ordinate; its Jacobian rows are
     * summed over nlayin(1) (:1251-1252), which the atmosphere path must
     * cover.  Anything else would read values the reference never wrote. */
    if (r->imod[ip] == 8) {
      int i1, i2;
      if (r->nlayin[ip] < 2) return ORC_ERR_LIMITS;
      i1 = r->layinc[ip * L] - 1;
      i2 = r->layinc[ip * L + 1] - 1;
      if (i1 < 0 || i2 < 0 || i1 >= ip || i2 >= ip) return ORC_ERR_UNSUPPORTED;
      if (r->imod[i1] != 13 && r->imod[i1] != 14) return ORC_ERR_LIMITS;
      if (r->imod[i2] == 8 || r->nlayin[0] > r->nlayin[i2]) return ORC_ERR_UNSUPPORTED;
    }
    /* imod 18 needs /E3TABLE/ and a ground temperature (its ground term is
     * always added, and planckg at tsurf <= 0 is not a number) */
    if (r->imod[ip] == 18 && (!cia->e3tran || !cia->e3val || !(r->tsurf > 0.0)))
      return ORC_ERR_UNSUPPORTED;
  }
  if (!orc_hmin_order_ok(ngas, r->idgas, r->isogas)) return ORC_ERR_UNSUPPORTED;
  if (r->ncont < 0 || r->ncont > MAXCON || (r->ncont > 0 && (r->nsec < 1 || r->nsec > MAXSEC)))
    return ORC_ERR_LIMITS;                                        /* check_limits.f:24-33 */
  /* cirsradg_wave.f:439-453: the wavenumber must lie on the dust grid */
  if (r->ncont > 0 && r->nsec > 1 &&
      (vwave < r->vsec[0] || vwave > r->vsec[r->nsec - 1]))
    return ORC_ERR_LIMITS;
  if (r->flagh2p == 1 && (cia->npara < 3 || cia->npara > 64 || !r->hfp))
    return ORC_ERR_UNSUPPORTED;                                   /* no /FPTABLE/ */
  /* calc_nlte_t.f:63-67 stops for k1 outside [1e-20, 1e-3] */
  if (r->inlte == 1 && (r->nlte_k1 < 1.0e-20 || r->nlte_k1 > 1.0e-3)) return ORC_ERR_LIMITS;
  /* N2's near-infrared CIA (n2n2cia.f, n2h2cia.f: tabulated, non-zero only
   * for 3995 <= v <= 4995 cm-1) needs its tables when N2 is present there */
  for (k = 0; k < ngas; k++) {
    const double vn = r->ispace == 1 ? 1e4 / vwave : vwave;
    if (r->idgas[k] == 22 && vn >= 3995.0 && vn <= 4995.0 && (cia->n1 < 2 || cia->n2 < 2))
      return ORC_ERR_UNSUPPORTED;
  }

  /* line-by-line tables (/LBLTABLE/ ILBL = 2, cirsradg_wave.f:336-339):
   * one g-ordinate of weight 1; each gas takes the last table of its (id,
   * iso) (read_klbllist.f:190-198), on the P/T grid of the last table read
   * (read_klblhead.f fills the one /INTERPKLBL/ grid per file).  ILBL = 1
   * reaches noverlapg1 with the k-tables' NG, which stops unless NG = 1. */
  if (r->ilbl == 1 || r->ilbl < 0 || r->ilbl > 2) return ORC_ERR_UNSUPPORTED;
  if (r->ilbl == 2) {
    static const double one[1] = {1.0};
    if (ntab < 1) return ORC_ERR_LIMITS;
    for (k = 0; k < ngas; k++) {
      int t;
      gtab[k] = NULL;
      for (t = 0; t < ntab; t++)
        if (tabs[t].idgas == r->idgas[k] && tabs[t].isogas == r->isogas[k]) gtab[k] = &tabs[t];
    }
    ng = 1;
    delg = one;
  }
  /* g-grid: all tables share it (read_klist.f:117-130). */
  for (k = 0; r->ilbl != 2 && k < ngas; k++) {
    gtab[k] = find_table(tabs, ntab, r->idgas[k], r->isogas[k], vwave);
    if (gtab[k]) {
      if (ng == 0) {
        ng = gtab[k]->ng;
        delg = gtab[k]->delg;
      } else if (gtab[k]->ng != ng) {
        return ORC_ERR_LIMITS;
      }
    }
  }
  if (ng == 0 && ntab > 0 && r->ilbl != 2) {
    ng = tabs[0].ng;
    delg = tabs[0].delg;
  }
  if (ng < 1 || ng > MAXG) return ORC_ERR_LIMITS;

  nparam = ngas + 1 + r->ncont + (r->flagh2p == 1);  /* cirsradg_wave.f:375-376 */
  /* cirsradg_wave.f:393-414: the last H2, He, CH4, NH3 (isotope 0 or 1) */
  for (k = 0; k < ngas; k++) {
    if (r->isogas[k] != 0 && r->isogas[k] != 1) continue;
    if (r->idgas[k] == 39) jh2 = k;
    if (r->idgas[k] == 40) jhe = k;
    if (r->idgas[k] == 6) jch4 = k;
    if (r->idgas[k] == 11) jnh3 = k;
  }
  rc = orc_interpem(r->nem, r->vem, r->emissivity, vwave, &esurf);
  if (rc != ORC_OK && r->tsurf > 0.0) return rc;
  rc = ORC_OK;
  if (r->ispace == 1) vv = 1e4 / vwave;

#define ALLOC(p, n)                                   \
  do {                                                \
    p = (double *)calloc((size_t)(n), sizeof(double)); \
    if (!p) { rc = ORC_ERR_ALLOC; goto done; }        \
  } while (0)
  ALLOC(kout, (size_t)ngas * L * ng);
  ALLOC(dkoutdt, (size_t)ngas * L * ng);
  ALLOC(utotl, L);
  ALLOC(taucon, L);
  ALLOC(dtaucondq, (size_t)L * nparam);
  ALLOC(kl_g, (size_t)L * ng);
  ALLOC(dkdql, (size_t)L * ng * (ngas + 1));
  ALLOC(bb, (size_t)npath * L);
  ALLOC(dbdt, (size_t)npath * L);
  ALLOC(corkout, (size_t)npath * ng);
  ALLOC(tempgtsurf, (size_t)npath * ng);
  ALLOC(dcoutdq, (size_t)npath * ng * L * nparam);
  ALLOC(tautmp, L);
  ALLOC(dtautmpdq, (size_t)L * nparam);
  ALLOC(taus, L);
  ALLOC(dtaudq, (size_t)L * nparam);
  ALLOC(dtolddq, (size_t)L * nparam);
  ALLOC(dtrdq, (size_t)L * nparam);
  ALLOC(trv, L + 1);
  ALLOC(wv, L + 1);
  ALLOC(tlv, L + 1);

  /* cirsradg_wave.f:418-426 */
  for (j = 0; j < L; j++) {
    double u = 0.0;
    for (k = 0; k < ngas; k++) u += r->amount[j * ngas + k];
    utotl[j] = u * 1e-20;
  }

  /* get_kg, cirsradg_wave.f:502-503 */
  for (k = 0; k < ngas; k++) {
    if (gtab[k] && r->ilbl == 2) {
      /* get_klblg, cirsradg_wave.f:504-505, on the shared P/T grid */
      orc_ktable kx = *gtab[k];
      kx.np = tabs[ntab - 1].np;
      kx.nt = tabs[ntab - 1].nt;
      kx.lnp = tabs[ntab - 1].lnp;
      kx.tk = tabs[ntab - 1].tk;
      kx.t2 = tabs[ntab - 1].t2;
      rc = orc_get_kg_gas(&kx, L, r->press, r->temp, vwave, kout + (size_t)k * L * ng,
                          dkoutdt + (size_t)k * L * ng, 1);
      if (rc != ORC_OK) goto done;
    } else if (gtab[k]) {
      rc = orc_get_kg_gas(gtab[k], L, r->press, r->temp, vwave, kout + (size_t)k * L * ng,
                          dkoutdt + (size_t)k * L * ng, 0);
      if (rc != ORC_OK) goto done;
    }
  }

  /* per-layer continuum and overlap, cirsradg_wave.f:510-760 */
  for (j = 0; j < L; j++) {
    double aam[MAXGAS], ppp[MAXGAS], absorb, dabsorb[7], amo[MAXGAS];
    double k_gn[MAXGAS * MAXG], dkgndt[MAXGAS * MAXG];
    double k_g[MAXG], dkdq[(MAXGAS + 1) * MAXG];
    int iabs[5];
    taucon[j] = 0.0;
    /* dust, cirsradg_wave.f:550-577: extinction cross-section at vwave
     * (constant for nsec = 1, else interp.f on the /DUST/ grid) times the
     * particles/cm2; d(tau)/d(cont) is the cross-section. */
    for (k = 0; k < r->ncont; k++) {
      const double *xk = r->xsec + (size_t)k * r->nsec;
      const double tau = r->nsec == 1 ? xk[0] : orc_interp(r->vsec, xk, r->nsec, vwave);
      taucon[j] += tau * r->cont[(size_t)j * r->ncont + k];
      dtaucondq[j * nparam + ngas + 1 + k] = tau;
    }
    /* gas continua, cirsradg_wave.f:590-611: the electron state first, then
     * ngascon; d(tau)/d(amount) = absorb/amount, with no T derivative */
    for (k = 0; k < ngas; k++) {
      const double am = r->amount[j * ngas + k], pk = r->pp[j * ngas + k];
      double ab;
      if (r->idgas[k] == 80) {
        hs.pelec = pk;
        hs.eflag = 1;
      }
      if (r->idgas[k] == 91) {
        hs.amounthmin = am;
        hs.ppresshmin = pk;
      }
      rc = orc_ngascon(cia, vv, r->idgas[k], r->isogas[k], am, pk, r->press[j], r->temp[j], &hs,
                       &ab);
      if (rc != ORC_OK) goto done;
      taucon[j] += ab;
      if (ab > 0 && am > 0.0) dtaucondq[j * nparam + k] = ab / am;
    }
    for (k = 0; k < ngas; k++) {
      aam[k] = r->amount[j * ngas + k];
      ppp[k] = r->pp[j * ngas + k];
    }
    if (r->flagh2p == 1)  /* cirsradg_wave.f:633-637 */
      orc_nparacon_all(cia, vv, r->press[j], r->temp[j], ngas, r->idgas, r->isogas, aam, ppp,
                       r->hfp[j], r->delh[j], &absorb, iabs, dabsorb);
    else
      orc_nciacon(cia, vv, r->press[j], r->temp[j], r->inormal, ngas, r->idgas, r->isogas,
                  aam, ppp, r->delh[j], &absorb, iabs, dabsorb);
    taucon[j] += absorb;
    for (l = 0; l < 5; l++) {
      int kk = iabs[l];
      if (kk > 0) dtaucondq[j * nparam + kk - 1] += dabsorb[l] / (utotl[j] * 1.0e20);
    }
    dtaucondq[j * nparam + ngas] += dabsorb[5];
    /* para-H2 fraction, cirsradg_wave.f:658-661 */
    if (r->flagh2p == 1) dtaucondq[j * nparam + nparam - 1] += dabsorb[6];

    /* Rayleigh, cirsradg_wave.f:663-691: an extinction utotl * xray in the
     * continuum and xray * 1e-20 in every gas's d(tau)/d(amount); its P, T
     * and composition derivatives are neglected there and here. */
    if (r->iray > 0 && r->itype != 0) {
      double xray;
      if (r->iray <= 3) {
        xray = orc_rayleigh(r->iray, vv, r->press[j], r->temp[j], 0.0, 0.0, 0.0) * 1e20;
      } else {
        /* FRAC = pp / press (:418-426) of the gases picked at :393-414; a
         * gas that is absent leaves the reference's FH2/FHE/FCH4/FNH3 unset,
         * and the restatement takes 0 for it. */
        const double fh2 = jh2 >= 0 ? r->pp[j * ngas + jh2] / r->press[j] : 0.0;
        const double fhe = jhe >= 0 ? r->pp[j * ngas + jhe] / r->press[j] : 0.0;
        const double fch4 = jch4 >= 0 ? r->pp[j * ngas + jch4] / r->press[j] : 0.0;
        const double fnh3 = jnh3 >= 0 ? r->pp[j * ngas + jnh3] / r->press[j] : 0.0;
        const double fheh2 = fh2 > 0.0 ? fhe / fh2 : 0.0;
        const double fch4h2 = fh2 > 0.0 ? fch4 / fh2 : 0.0;
        xray = orc_rayleigh(4, vv, r->press[j], r->temp[j], fheh2, fch4h2, fnh3) * 1e20;
      }
      taucon[j] += utotl[j] * xray;
      for (k = 0; k < ngas; k++) dtaucondq[j * nparam + k] += xray * 1.0e-20;
    }

    for (k = 0; k < ngas; k++) {
      amo[k] = r->amount[j * ngas + k] * 1.0e-20;
      for (l = 0; l < ng; l++) {
        if (!gtab[k]) {
          k_gn[k * ng + l] = 0.0;
          dkgndt[k * ng + l] = 0.0;
        } else {
          k_gn[k * ng + l] = kout[((size_t)k * L + j) * ng + l];
          dkgndt[k * ng + l] = dkoutdt[((size_t)k * L + j) * ng + l];
        }
      }
    }
    if (r->ilbl == 2) orc_noverlapg1(ngas, amo, k_gn, dkgndt, k_g, dkdq);  /* :739-740 */
    else orc_noverlapg(delg, ng, ngas, amo, k_gn, dkgndt, k_g, dkdq);
    for (l = 0; l < ng; l++) {
      kl_g[j * ng + l] = k_g[l];
      for (k = 0; k < ngas; k++)
        dkdql[((size_t)j * ng + l) * (ngas + 1) + k] = dkdq[k * ng + l] * 1.0e-20;
      dkdql[((size_t)j * ng + l) * (ngas + 1) + ngas] = dkdq[ngas * ng + l];
    }
  }

  /* g loop, cirsradg_wave.f:776-1232 */
  for (ig = 0; ig < ng; ig++) {
    for (j = 0; j < L; j++) {
      tautmp[j] = taucon[j] + kl_g[j * ng + ig];
      for (k = 0; k < nparam; k++) dtautmpdq[j * nparam + k] = dtaucondq[j * nparam + k];
      for (k = 0; k < ngas + 1; k++)
        dtautmpdq[j * nparam + k] += dkdql[((size_t)j * ng + ig) * (ngas + 1) + k];
    }
    for (ip = 0; ip < npath; ip++) {
      const int nlays = r->nlayin[ip];
      double *dco = dcoutdq + ((size_t)ip * ng + ig) * L * nparam;
      double co = 0.0;
      tempgtsurf[ip * ng + ig] = 0.0;
      for (j = 0; j < nlays; j++)
        for (k = 0; k < nparam; k++) dco[j * nparam + k] = 0.0;
      for (j = 0; j < nlays; j++) {
        int li = r->layinc[ip * L + j] - 1;
        double sc = r->scale[ip * L + j];
        taus[j] = tautmp[li] * sc;
        for (k = 0; k < nparam; k++) dtaudq[j * nparam + k] = dtautmpdq[li * nparam + k] * sc;
      }
      if (r->imod[ip] == 13 || r->imod[ip] == 14) {
        /* SCR sideband and wideband cell transmissions, cirsradg_wave.f:1084-1102:
         * the first layer only, no Jacobian */
        co = r->imod[ip] == 13 ? 1.0 - exp(-taus[0]) : 0.5 * (1.0 + exp(-taus[0]));
      } else if (r->imod[ip] == 8) {
        /* cirsradg_wave.f:1037-1082: cell times atmosphere, the atmosphere's
         * Jacobian rows and surface gradient times the cell */
        const int i1 = r->layinc[ip * L] - 1, i2 = r->layinc[ip * L + 1] - 1;
        const double c1 = corkout[i1 * ng + ig];
        const double *d2 = dcoutdq + ((size_t)i2 * ng + ig) * L * nparam;
        co = c1 * corkout[i2 * ng + ig];
        for (j = 0; j < r->nlayin[i2]; j++)
          for (k1 = 0; k1 < r->nsw; k1++) {
            k = r->isw[k1] - 1;
            dco[j * nparam + k] = c1 * d2[j * nparam + k];
          }
        tempgtsurf[ip * ng + ig] = tempgtsurf[i2 * ng + ig] * c1;
      } else if (r->imod[ip] == 0 || r->imod[ip] == 1) {
        /* cirsradg_wave.f:854-895; imod 0 with a stellar spectrum and iform 4
         * is times the stellar flux at dist (:862-880) */
        double tmp = 0.0;
        for (j = 0; j < nlays; j++) tmp += taus[j];
        if (r->imod[ip] == 0) {
          const double xf0 = (cia->nsol > 0 && r->iform == 4) ? orc_solar(cia, vwave, r->dist) : 1.0;
          co = orc_dpexp(-tmp);
          for (j = 0; j < nlays; j++)
            for (k1 = 0; k1 < r->nsw; k1++) {
              k = r->isw[k1] - 1;
              dco[j * nparam + k] = -tmp * co * dtaudq[j * nparam + k] * xf0;
            }
          co = co * xf0;
        } else {
          co = 1.0 - orc_dpexp(-tmp);
          for (j = 0; j < nlays; j++)
            for (k1 = 0; k1 < r->nsw; k1++) {
              k = r->isw[k1] - 1;
              dco[j * nparam + k] = tmp * (1 - co) * dtaudq[j * nparam + k];
            }
        }
      } else if (r->imod[ip] == 18) {
        /* disc-integrated emission, cirsradg_wave.f:1116-1211: the layer
         * weights are differences of E3(tr) (e3interp.f), xfac = 4 pi
         * ((radius1 + radextra) 1e5)^2 for iform 1 and 3 (no extra pi, unlike
         * imod 3), times 2 pi; the ground term is always added and carries
         * no Jacobian (:1196-1205). */
        double xfac = 1.0, trold = 1.0, e3old = 0.5, e3new, e3g, tlayer, tr, bsurf;
        if (r->iform == 1 || r->iform == 3)
          xfac = xfac * 4. * PI_REF * ((r->radius1 + r->radextra) * 1e5) *
                 ((r->radius1 + r->radextra) * 1e5);
        /* with a stellar spectrum, iform 1 divides by its total power
         * (cirsradg_wave.f:921-925, 1144-1149: xdist = -1) */
        if (cia->nsol > 0 && r->iform == 1) xfac = xfac / orc_solar(cia, vwave, -1.0);
        if (r->iform == 3) xfac = xfac * 1e-18;
        if (ig == 0) {
          for (j = 0; j < nlays; j++) {
            const double tn = orc_nlte_t(r->inlte, r->nlte_k1, r->nlte_a, r->emtemp[ip * L + j],
                                         r->press[r->layinc[ip * L + j] - 1], vwave);
            bb[ip * L + j] = orc_planck(r->ispace, vwave, tn);
            dbdt[ip * L + j] = orc_planckg(r->ispace, vwave, tn);
          }
        }
        bsurf = orc_planck(r->ispace, vwave, r->tsurf);
        if (jac_mode == 0) {
          /* de3dq(j1) = e3grad(tr_j) dtr_j/dq, dtrdq as in imod 3 */
          for (k1 = 0; k1 < r->nsw; k1++) {
            k = r->isw[k1] - 1;
            for (j = 0; j < nlays; j++) dtolddq[j * nparam + k] = dtrdq[j * nparam + k] = 0.0;
          }
          {
            double *de3old = (double *)calloc((size_t)nlays * nparam, sizeof(double));
            double *de3 = (double *)calloc((size_t)nlays * nparam, sizeof(double));
            if (!de3old || !de3) {
              free(de3old);
              free(de3);
              rc = ORC_ERR_ALLOC;
              goto done;
            }
            for (j = 0; j < nlays; j++) {
              const double bj = bb[ip * L + j];
              tlayer = exp(-taus[j]);
              tr = trold * tlayer;
              orc_e3interp(cia, tr, &e3new, &e3g);
              co = co + xfac * 2.0 * PI_REF * (e3old - e3new) * bj;
              for (k1 = 0; k1 < r->nsw; k1++) {
                int j1;
                k = r->isw[k1] - 1;
                for (j1 = 0; j1 <= j; j1++) {
                  if (j1 < j)
                    dtrdq[j1 * nparam + k] = dtolddq[j1 * nparam + k] * tlayer;
                  else
                    dtrdq[j1 * nparam + k] = -dtaudq[j1 * nparam + k] * tr;
                  de3[j1 * nparam + k] = e3g * dtrdq[j1 * nparam + k];
                  dco[j1 * nparam + k] +=
                      bj * xfac * 2.0 * PI_REF * (de3old[j1 * nparam + k] - de3[j1 * nparam + k]);
                }
                if (k == ngas)
                  dco[j * nparam + k] += xfac * 2.0 * PI_REF * (e3old - e3new) * dbdt[ip * L + j];
              }
              e3old = e3new;
              trold = tr;
              for (int j1 = 0; j1 <= j; j1++)
                for (k1 = 0; k1 < r->nsw; k1++) {
                  k = r->isw[k1] - 1;
                  dtolddq[j1 * nparam + k] = dtrdq[j1 * nparam + k];
                  de3old[j1 * nparam + k] = de3[j1 * nparam + k];
                }
            }
            free(de3old);
            free(de3);
          }
        } else {
          /* O(L): dR/dtau_m = xfac 2 pi sum_{j>=m} (B_j - B_{j+1}) g_j tr_j,
           * B_{N+1} = 0, g_j = e3grad(tr_j) */
          double s = 0.0;
          trv[0] = 1.0;
          for (j = 0; j < nlays; j++) {
            trv[j + 1] = trv[j] * exp(-taus[j]);
            orc_e3interp(cia, trv[j + 1], &e3new, &e3g);
            tlv[j] = e3new;  /* E3(tr_j) */
            wv[j] = e3g;     /* e3grad(tr_j) */
            co = co + xfac * 2.0 * PI_REF * (e3old - e3new) * bb[ip * L + j];
            e3old = e3new;
          }
          for (j = nlays - 1; j >= 0; j--) {
            const double bnext = j + 1 < nlays ? bb[ip * L + j + 1] : 0.0;
            const double e3prev = j > 0 ? tlv[j - 1] : 0.5;
            s += (bb[ip * L + j] - bnext) * wv[j] * trv[j + 1];
            for (k1 = 0; k1 < r->nsw; k1++) {
              k = r->isw[k1] - 1;
              dco[j * nparam + k] = xfac * 2.0 * PI_REF * s * dtaudq[j * nparam + k];
              if (k == ngas)
                dco[j * nparam + k] += xfac * 2.0 * PI_REF * (e3prev - tlv[j]) * dbdt[ip * L + j];
            }
          }
        }
        co = co + xfac * 2.0 * PI_REF * e3old * (esurf * bsurf);
        tempgtsurf[ip * ng + ig] =
            xfac * 2.0 * PI_REF * e3old * esurf * orc_planckg(r->ispace, vwave, r->tsurf);
      } else {
        /* imod = 3, cirsradg_wave.f:896-1035.  xfac, :912-925 with no solar
         * file (iread != 999): the planet's total flux for iform 1 and 3,
         * with a further 1e-18 for iform 3 */
        double xfac = 1.0;
        if (r->iform == 1 || r->iform == 3)
          xfac = xfac * PI_REF * 4. * PI_REF * ((r->radius1 + r->radextra) * 1e5) *
                 ((r->radius1 + r->radextra) * 1e5);
        /* with a stellar spectrum, iform 1 divides by its total power
         * (cirsradg_wave.f:921-925, 1144-1149: xdist = -1) */
        if (cia->nsol > 0 && r->iform == 1) xfac = xfac / orc_solar(cia, vwave, -1.0);
        if (r->iform == 3) xfac = xfac * 1e-18;
        double trold = 1.0, tr, tlayer, radground, gradground;
        int jm = (int)(0.5 * nlays);
        double p1, p2;
        if (ig == 0) {
          /* cirsradg_wave.f:931-944: Planck at calc_nlte_t's temperature */
          for (j = 0; j < nlays; j++) {
            const double tn = orc_nlte_t(r->inlte, r->nlte_k1, r->nlte_a, r->emtemp[ip * L + j],
                                         r->press[r->layinc[ip * L + j] - 1], vwave);
            bb[ip * L + j] = orc_planck(r->ispace, vwave, tn);
            dbdt[ip * L + j] = orc_planckg(r->ispace, vwave, tn);
          }
        }
        if (jac_mode == 0) {
          for (j = 0; j < nlays; j++)
            for (k = 0; k < nparam; k++) {
              dtolddq[j * nparam + k] = 0.0;
              dtrdq[j * nparam + k] = 0.0;
            }
          for (j = 0; j < nlays; j++) {
            double bj = bb[ip * L + j];
            tlayer = exp(-taus[j]);
            tr = trold * tlayer;
            co = co + xfac * (trold - tr) * bj;
            for (k1 = 0; k1 < r->nsw; k1++) {
              int j1;
              k = r->isw[k1] - 1;
              for (j1 = 0; j1 <= j; j1++) {
                if (j1 < j)
                  dtrdq[j1 * nparam + k] = dtolddq[j1 * nparam + k] * tlayer;
                else
                  dtrdq[j1 * nparam + k] = -dtaudq[j1 * nparam + k] * (tlayer * trold);
                dco[j1 * nparam + k] +=
                    xfac * bj * (dtolddq[j1 * nparam + k] - dtrdq[j1 * nparam + k]);
              }
              if (k == ngas) dco[j * nparam + k] += xfac * (trold - tr) * dbdt[ip * L + j];
            }
            trold = tr;
            for (int j1 = 0; j1 <= j; j1++)
              for (k1 = 0; k1 < r->nsw; k1++) {
                k = r->isw[k1] - 1;
                dtolddq[j1 * nparam + k] = dtrdq[j1 * nparam + k];
              }
          }
        } else {
          /* O(L) form: R = xfac*W_0 with W_{j-1} = B_j (1 - t_j) + t_j W_j and
           * dR/dtau_j = xfac * tr_j * (B_j - W_j) (W_N = ground term). */
          trv[0] = 1.0;
          for (j = 0; j < nlays; j++) {
            tlv[j] = exp(-taus[j]);
            trv[j + 1] = trv[j] * tlv[j];
            co = co + xfac * (trv[j] - trv[j + 1]) * bb[ip * L + j];
          }
          trold = trv[nlays];
        }
        /* cirsradg_wave.f:992-1024: limb test and surface term */
        p1 = r->press[r->layinc[ip * L + (jm > 0 ? jm - 1 : 0)] - 1];
        p2 = r->press[r->layinc[ip * L + nlays - 1] - 1];
        radground = 0.0;
        if (p2 > p1) {
          if (r->tsurf <= 0.0) {
            radground = bb[ip * L + nlays - 1];
            gradground = dbdt[ip * L + nlays - 1];
          } else {
            radground = esurf * orc_planck(r->ispace, vwave, r->tsurf);
            gradground = esurf * orc_planckg(r->ispace, vwave, r->tsurf);
          }
          co = co + xfac * trold * radground;
          tempgtsurf[ip * ng + ig] = xfac * trold * gradground;
          if (jac_mode == 0) {
            for (j = 0; j < nlays; j++)
              for (k1 = 0; k1 < r->nsw; k1++) {
                k = r->isw[k1] - 1;
                dco[j * nparam + k] += xfac * radground * dtolddq[j * nparam + k];
              }
          }
        }
        if (jac_mode != 0) {
          wv[nlays] = radground;
          for (j = nlays - 1; j >= 0; j--) {
            /* wv[j] = upwelling radiance entering layer j from below */
            if (j < nlays - 1)
              wv[j] = bb[ip * L + j + 1] * (1.0 - tlv[j + 1]) + tlv[j + 1] * wv[j + 1];
            else
              wv[j] = radground;
          }
          for (j = 0; j < nlays; j++) {
            double dr = xfac * trv[j + 1] * (bb[ip * L + j] - wv[j]);
            for (k1 = 0; k1 < r->nsw; k1++) {
              k = r->isw[k1] - 1;
              dco[j * nparam + k] = dr * dtaudq[j * nparam + k];
              if (k == ngas) dco[j * nparam + k] += xfac * (trv[j] - trv[j + 1]) * dbdt[ip * L + j];
            }
          }
        }
      }
      corkout[ip * ng + ig] = co;
    }
  }

  /* g-quadrature, cirsradg_wave.f:1240-1286 */
  for (ip = 0; ip < npath; ip++) {
    const int nlay1 = r->imod[ip] == 8 ? r->nlayin[0] : r->nlayin[ip];
    output[ip] = 0.0;
    gtsurf[ip] = 0.0;
    for (j = 0; j < L; j++)
      for (k = 0; k < nparam; k++) doutputdq[((size_t)ip * L + j) * nparam + k] = 0.0;
    for (ig = 0; ig < ng; ig++) {
      const double *dco = dcoutdq + ((size_t)ip * ng + ig) * L * nparam;
      output[ip] += corkout[ip * ng + ig] * delg[ig];
      gtsurf[ip] += tempgtsurf[ip * ng + ig] * delg[ig];
      for (j = 0; j < nlay1; j++)
        for (k1 = 0; k1 < r->nsw; k1++) {
          k = r->isw[k1] - 1;
          doutputdq[((size_t)ip * L + j) * nparam + k] += dco[j * nparam + k] * delg[ig];
        }
    }
  }

done:
  free(kout);
  free(dkoutdt);
  free(utotl);
  free(taucon);
  free(dtaucondq);
  free(kl_g);
  free(dkdql);
  free(bb);
  free(dbdt);
  free(corkout);
  free(tempgtsurf);
  free(dcoutdq);
  free(tautmp);
  free(dtautmpdq);
  free(taus);
  free(dtaudq);
  free(dtolddq);
  free(dtrdq);
  free(trv);
  free(wv);
  free(tlv);
  return rc;
#undef ALLOC
}

int orc_spectrum(const orc_ktable *tabs, int ntab, const orc_cia *cia,
                 const orc_req *r, const double *vwave, int nwave, int jac_mode,
                 int nthreads, double *output, double *gtsurf, double *doutputdq) {
  int iw, rc = ORC_OK;
  const size_t per = (size_t)r->npath * r->nlayer * (r->ngas + 1 + r->ncont + (r->flagh2p == 1));
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(max : rc)
#endif
  for (iw = 0; iw < nwave; iw++) {
    int rci = orc_cirsradg_wave(tabs, ntab, cia, r, vwave[iw], jac_mode,
                                output + (size_t)iw * r->npath, gtsurf + (size_t)iw * r->npath,
                                doutputdq + (size_t)iw * per);
    if (rci > rc) rc = rci;
  }
  (void)nthreads;
  return rc;
}
