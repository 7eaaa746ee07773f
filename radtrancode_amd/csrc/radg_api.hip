// Host side of the radg C ABI: table/CIA registries, request validation and
// packing, device batches and the Fortran-surface entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <map>
#include <mutex>
#include <set>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/radg.h"
#include "radg_dev.h"
#include "radg_internal.h"

// launchers (radg_kernels.hip)
size_t radg_lds_layer_bytes(int ng, int ngas, int E, int fp64);
int radg_layer_waves(int ng, int ngas, int E, int fp64);
int radg_layer_stride(int ng, int ngas, int E, int fp64);
size_t radg_lds_emission_bytes(int RL, int nkrow, int fp64);
int radg_layer_E(int ng);
int radg_emission_RL(int maxlays);
hipError_t radg_launch_layer_lds(const KernelArgs &a, int mode, int E, size_t lds, hipStream_t st);
size_t radg_lds_fused_bytes(int mode, int ng, int E, int RL, int maxgas, int maxnsw, long long maxblock,
                            int *ws);
hipError_t radg_launch_fused(const KernelArgs &a, int mode, int E, int RL, size_t lds, hipStream_t st);
hipError_t radg_convert_f32(const double *src, float *dst, long long n, hipStream_t st);
hipError_t radg_slice_offsets(const double *src, float *off, long long nslice, int ntab, int ng,
                              hipStream_t st);
hipError_t radg_convert_code16(int bf, const double *src, const float *off, unsigned short *dst,
                               long long n, int ntab, int ng, hipStream_t st);
hipError_t radg_launch_emission(const KernelArgs &a, int fp64, int RL, bool disc, size_t lds,
                                hipStream_t st);
hipError_t radg_launch_quadrature(const KernelArgs &a, int fp64, hipStream_t st);
hipError_t radg_launch_gemm(const GemmProblem *probs_dev, int nprob, int maxM, int maxN,
                            int a_fp32, hipStream_t st);
hipError_t radg_launch_pathsum(const void *ps_dev, int nps, long long maxn, hipStream_t st);
size_t radg_pathsum_bytes();
void radg_pathsum_fill(void *dst, const double *in, const double *pw, double *out, int nwave,
                       int npath, int n);

namespace {

// ------------------------------------------------------------- errors
thread_local std::string g_err;
thread_local int g_cirs_status = 0;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess)                                                      \
      return fail(RADG_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                   \
  } while (0)

int g_mode = RADG_MODE_FP64;
std::mutex g_mu;

// constdef.f:7-11
constexpr double AMAGAT = 2.68675e19;
constexpr double KBOLTZMANN = 1.37947e-23;

// ------------------------------------------------------------- tables
struct Table {
  int idgas, isogas, npoint;
  double vmin, delv;
  std::vector<double> vcen;   // irregular grid (delv <= 0, read_khead.f:93-100)
  long long off;              // element offset in the pool
  std::vector<uint8_t> zero;  // per slice: every value <= 0 (get_kg.f:224-250)
};

struct TableSet {
  int np = 0, nt = 0, ng = 0;  // nt: temperatures per pressure
  bool t2 = false;             // a temperature grid per pressure (NT < 0)
  bool lbl = false;            // line-by-line tables (ILBL = 2): ng = 1, one wavenumber grid
  std::vector<double> lnp, tk, g_ord, delg, G;  // tk: [nt], or [np][nt] with t2
  std::vector<Table> tabs;
  long long total = 0;
  double *d_k64 = nullptr;
  float *d_k32 = nullptr;            // made on first use by the mixed form
  unsigned short *d_k16 = nullptr;   // made on first use by the BF16 form
  unsigned short *d_kh = nullptr;    // made on first use by the half form
  float *d_koff = nullptr;           // per (slice, ordinate) offsets of both 16-bit forms
  std::mutex form_mu;
  double *d_G = nullptr, *d_delg64 = nullptr;
  float *d_delg32 = nullptr;
  // the layer kernel's LDS header for E = hdr_E slots per lane, as T = float
  // (d_hdr[0]) and T = double (d_hdr[1])
  char *d_hdr[2] = {nullptr, nullptr};
  int hdr_E = 0;
  ~TableSet() {
    for (char *h : d_hdr)
      if (h) (void)hipFree(h);
    if (d_k64) (void)hipFree(d_k64);
    if (d_k32) (void)hipFree(d_k32);
    if (d_k16) (void)hipFree(d_k16);
    if (d_kh) (void)hipFree(d_kh);
    if (d_koff) (void)hipFree(d_koff);
    if (d_G) (void)hipFree(d_G);
    if (d_delg64) (void)hipFree(d_delg64);
    if (d_delg32) (void)hipFree(d_delg32);
  }
};

struct CIA {
  double dnu = 0;
  double *d_kcia = nullptr, *d_temps = nullptr;
  ~CIA() {
    if (d_kcia) (void)hipFree(d_kcia);
    if (d_temps) (void)hipFree(d_temps);
  }
};

// /FPTABLE/ (ciacom.f), the para-H2 CIA that fpread_all.f reads
struct ParaCIA {
  double dnu = 0;
  int npara = 0;
  double *d_kh2h2 = nullptr, *d_kh2he = nullptr, *d_kcia5 = nullptr, *d_temps = nullptr,
         *d_fp = nullptr;
  ~ParaCIA() {
    for (double *p : {d_kh2h2, d_kh2he, d_kcia5, d_temps, d_fp})
      if (p) (void)hipFree(p);
  }
};

// /E3TABLE/ (e3interp.f:14-51): transmission, E3 and its gradient, 1001
// points each, for imod 18
struct E3Tab {
  double *d_tab = nullptr;  // [3][RADG_E3N]
  ~E3Tab() {
    if (d_tab) (void)hipFree(d_tab);
  }
};

// N2's near-infrared CIA (n2n2cia.f, n2h2cia.f): each function's values
// (cm-1 amagat-2) at its knots
struct N2Nir {
  std::vector<double> wn1, k1, wn2, k2;  // N2-N2, N2-H2
};

// /SOLARDAT/ swave, srad, nspt (get_solar_wave.f): a registered spectrum is
// the reference's iread = 999
struct Solar {
  std::vector<double> wave, rad;
};

// /GASCONDEF/ and the tables ngascon.f reads (radg_register_gascon)
struct GasCon {
  int ih2o = 0, ich4 = 0, io3 = 0, inh3 = 0, iuvscat = 0;
  double h2o_v0 = 0, h2o_dv = 0, h2o_c2 = 0, h2o_ts = 0, h2o_ps = 0;
  std::vector<double> s296, s260, f296;  // CKD H2O
  std::vector<double> ch4_wl, ch4_k, nh3_wl, nh3_k;
};

// The tables beside the k-tables and CIA that some requests need
struct AuxTables {
  std::shared_ptr<ParaCIA> para;
  std::shared_ptr<E3Tab> e3;
  std::shared_ptr<N2Nir> n2nir;
  std::shared_ptr<Solar> solar;
  std::shared_ptr<GasCon> gascon;
};

std::shared_ptr<TableSet> g_tables;
// /LBLTABLE/ ILBL (cirsradg_wave.f:317): 2 runs every batch made afterwards
// on the line-by-line tables g_lbl (get_klblg.f, noverlapg1.f)
std::shared_ptr<TableSet> g_lbl;
std::atomic<int> g_ilbl{0};
std::shared_ptr<CIA> g_cia;
std::shared_ptr<ParaCIA> g_para;
std::shared_ptr<E3Tab> g_e3;
std::shared_ptr<N2Nir> g_n2nir;

std::shared_ptr<Solar> g_solar;
int g_iread = 0;
std::shared_ptr<GasCon> g_gascon;

// ch4cont.f:44-57, nh3cont.f:46-70: the coefficient at lambda = 1e4/nu
// micron (10 micron for nu <= 1000), 0 off the table, linear between the
// points frecipes/locate.f brackets (the last interval at the last point)
double tabcont(const std::vector<double> &wl, const std::vector<double> &kt, double nu) {
  const int n = (int)wl.size();
  const double lambda = nu > 1000.0 ? 1e4 / nu : 10.0;
  if (n < 2 || lambda < wl[0] || lambda > wl[n - 1]) return 0.0;
  // locate's JL: the number of points below lambda
  int k = (int)(std::lower_bound(wl.begin(), wl.end(), lambda) - wl.begin());
  k = std::max(1, std::min(k, n - 1));
  const double fr = (lambda - wl[k - 1]) / (wl[k] - wl[k - 1]);
  return kt[k - 1] + fr * (kt[k] - kt[k - 1]);
}

// hmin_bf.f:23-41 with a zero starting sum: the H- bound-free cross-section
double hmin_bf(double v0) {
  static const double coeff[6] = {152.519, 49.534, -118.858, 92.536, -34.194, 4.982};
  const double lambda0 = 1.6419, wv = 1e4 / v0;
  if (!(wv < lambda0)) return 0.0;
  const double x = 1.0 / wv - 1.0 / lambda0;
  double s = 0.0;
  for (int i = 0; i < 6; i++) s += coeff[i] * std::pow(x, 0.5 * i);
  return 1e-18 * std::pow(wv, 3.0) * std::pow(x, 1.5) * s;
}

// CKD 2.1 self-continuum modification factors at 700, 710, ... 1200 cm-1
// (ngascon.f:103-113)
const double CKD_XFAC[51] = {
    1.00000, 1.01792, 1.03767, 1.05749, 1.07730, 1.09708, 1.10489, 1.11268, 1.12047,
    1.12822, 1.13597, 1.14367, 1.15135, 1.15904, 1.16669, 1.17431, 1.18786, 1.20134,
    1.21479, 1.22821, 1.24158, 1.26580, 1.28991, 1.28295, 1.27600, 1.26896, 1.25550,
    1.24213, 1.22879, 1.21560, 1.20230, 1.18162, 1.16112, 1.14063, 1.12016, 1.10195,
    1.09207, 1.08622, 1.08105, 1.07765, 1.07398, 1.06620, 1.05791, 1.04905, 1.03976,
    1.02981, 1.00985, 1.00000, 1.00000, 1.00000, 1.00000};

// The wavenumber-only factors of the CKD H2O continuum, ngascon.f:147-201:
// out[0] = 1 in the table's range (else the continuum is 0), then the self
// continuum at 296 and 260 K, the foreign one times its CKD 2.1 scaling,
// and the self continuum's scaling (SFAC S1 S2)
void h2o_factors(const GasCon &g, double ff, double *out) {
  const int n = (int)g.s296.size();
  const double vt1 = g.h2o_v0 + g.h2o_dv * (n - 1);
  for (int i = 0; i < 5; i++) out[i] = 0.0;
  if (!(ff >= g.h2o_v0 && ff < vt1)) return;
  int mk = 1 + (int)((ff - g.h2o_v0) / g.h2o_dv);
  if (mk == n) mk = n - 1;
  const double gr = (ff - (g.h2o_v0 + g.h2o_dv * (mk - 1))) / g.h2o_dv;
  auto lin = [&](const std::vector<double> &y) { return y[mk - 1] + gr * (y[mk] - y[mk - 1]); };
  double sfac = 1.0;
  if (ff >= 700.0 && ff < 1200.0) {
    const int ic = (int)(0.1 * (ff - 700.0));
    const double gc = (ff - (700.0 + ic * 10.0)) / 10.0;
    sfac = CKD_XFAC[ic] + gc * (CKD_XFAC[ic + 1] - CKD_XFAC[ic]);
  }
  const double alpha2 = 200.0 * 200.0, alphs2 = 120.0 * 120.0, betas = 5.0e-6, v0s = 1310.0,
               factrs = 0.15;
  const double hwsqf = 330. * 330., betaf = 8.e-11, v0f = 1130., factrf = 0.97, v0f2 = 1900.,
               hwsqf2 = 150. * 150., beta2 = 3.e-06;
  const double vs2 = (ff - v0s) * (ff - v0s);
  const double s1 = 1.0 - 0.2333 * (alpha2 / ((ff - 1050.0) * (ff - 1050.0) + alpha2));
  const double s2 = 1.0 - factrs * (alphs2 / (vs2 + betas * (vs2 * vs2) + alphs2));
  double vf2 = (ff - v0f) * (ff - v0f);
  double fscal = 1. - factrf * (hwsqf / (vf2 + betaf * (vf2 * vf2 * vf2) + hwsqf));
  vf2 = (ff - v0f2) * (ff - v0f2);
  fscal = fscal * (1. - 0.6 * (hwsqf2 / (vf2 + beta2 * (vf2 * vf2) + hwsqf2)));
  out[0] = 1.0;
  out[1] = lin(g.s296);
  out[2] = lin(g.s260);
  out[3] = lin(g.f296) * fscal;
  out[4] = sfac * s1 * s2;
}

// get_solar_wave.f:19-40 without the distance: interp.f on the spectrum,
// zero outside its range (ascending or descending)
double solar_at(const Solar &s, double x) {
  const int n = (int)s.wave.size();
  const bool asc = s.wave[0] < s.wave[n - 1];
  // interp.f: the first grid point past x (above it on an ascending grid,
  // below it on a descending one; 1-based), N if none, at least 2
  int i = 1;
  while (i <= n && !(asc ? s.wave[i - 1] > x : s.wave[i - 1] < x)) i++;
  i = std::max(2, std::min(i, n));
  double y = s.rad[i - 2] + (s.rad[i - 1] - s.rad[i - 2]) * (x - s.wave[i - 2]) / (s.wave[i - 1] - s.wave[i - 2]);
  if (asc) {
    if (x < s.wave[0] || x > s.wave[n - 1]) y = 0.0;
  } else if (x > s.wave[0] || x < s.wave[n - 1]) {
    y = 0.0;
  }
  return y;
}

// verint.f:24-48: linear interpolation in the interval above the first grid
// point past xin (ascending or descending), the end intervals extrapolated;
// a zero-width interval or a NaN result gives y(i)
double verint(const double *x, const double *y, int n, double xin) {
  int i = 1;
  if (x[0] < x[n - 1]) {
    while (i <= n && !(x[i - 1] > xin)) i++;
  } else {
    while (i <= n && !(x[i - 1] < xin)) i++;
  }
  if (i > n) i = n;
  if (i == 1) i = 2;
  double yout = y[i - 2] + (y[i - 1] - y[i - 2]) * (xin - x[i - 2]) / (x[i - 1] - x[i - 2]);
  if (x[i - 2] == x[i - 1]) yout = y[i - 1];
  if (std::isnan(yout)) yout = y[i - 1];
  return yout;
}

// n2n2cia.f / n2h2cia.f: zero off the knots' range, verint inside
double n2nir_coef(const std::vector<double> &x, const std::vector<double> &y, double v) {
  const int n = (int)x.size();
  if (n < 2 || v < x[0] || v > x[n - 1]) return 0.0;
  return verint(x.data(), y.data(), n, v);
}

// /DUST/ for the fixed-argument entry points (cirsradg_wave.f:225,312)
struct Dust {
  int nsec = 0, ncont = 0;
  std::vector<double> vsec, xsec;  // [nsec], [ncont][nsec]
};
std::shared_ptr<Dust> g_dust;

// /SOLARDAT/ iform for the fixed-argument entry points
int g_iform = 0;

// /NLTE_FLAGS/ for the fixed-argument entry points
int g_inlte = 0;
double g_nlte_k1 = 0.0, g_nlte_a = 0.0;

// The tables a batch made now runs on (call with g_mu held): the
// line-by-line tables under ILBL = 2, else the k-tables.
std::shared_ptr<TableSet> active_tables() { return g_ilbl == 2 ? g_lbl : g_tables; }
int no_tables() {
  return g_ilbl == 2 ? fail(RADG_ENOTABLES, "ILBL=2 and no line-by-line tables registered (radg_register_lbltables)")
                     : fail(RADG_ENOTABLES, "no k-tables registered (radg_register_ktables)");
}

bool valid_mode(int mode) {
  return mode == RADG_MODE_FP64 || mode == RADG_MODE_MIXED || mode == RADG_MODE_BF16 ||
         mode == RADG_MODE_F16;
}

// Make the table form a mode reads (FP32 copy or BF16 log10 codes) from the
// FP64 master on the device, once per registered table set.
int ensure_form(TableSet &ts, int mode) {
  if (mode == RADG_MODE_FP64) return RADG_OK;
  std::lock_guard<std::mutex> lk(ts.form_mu);
  if (mode == RADG_MODE_MIXED && !ts.d_k32) {
    float *d = nullptr;
    HIPCHK(hipMalloc(&d, ts.total * sizeof(float)));
    hipError_t e = radg_convert_f32(ts.d_k64, d, ts.total, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(RADG_EHIP, "FP32 table conversion: %s", hipGetErrorString(e));
    }
    ts.d_k32 = d;
  }
  const int ntab = ts.np * ts.nt * ts.ng;
  if ((mode == RADG_MODE_BF16 || mode == RADG_MODE_F16) && !ts.d_koff) {
    float *d = nullptr;
    HIPCHK(hipMalloc(&d, (ts.total / ntab) * ts.ng * sizeof(float)));
    hipError_t e = radg_slice_offsets(ts.d_k64, d, ts.total / ntab, ntab, ts.ng, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(RADG_EHIP, "table offsets: %s", hipGetErrorString(e));
    }
    ts.d_koff = d;
  }
  unsigned short **form = mode == RADG_MODE_BF16 ? &ts.d_k16 : mode == RADG_MODE_F16 ? &ts.d_kh : nullptr;
  if (form && !*form) {
    unsigned short *d = nullptr;
    HIPCHK(hipMalloc(&d, ts.total * sizeof(unsigned short)));
    hipError_t e = radg_convert_code16(mode == RADG_MODE_BF16 ? 1 : 2, ts.d_k64, ts.d_koff, d, ts.total,
                                       ntab, ts.ng, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(RADG_EHIP, "16-bit table conversion: %s", hipGetErrorString(e));
    }
    *form = d;
  }
  return RADG_OK;
}

// frecipes/locate.f: 1-based JL in 0..n
int locate(const double *xx, int n, double x) {
  int jl = 0, ju = n + 1;
  const bool asc = xx[n - 1] > xx[0];
  while (ju - jl > 1) {
    const int jm = (ju + jl) / 2;
    if (asc == (x > xx[jm - 1])) jl = jm;
    else ju = jm;
  }
  return jl;
}

// ------------------------------------------------------------- packing
struct Packed {
  DevReq d{};
  std::vector<LayerConst> lay;
  std::vector<double> amo, press, scale, emtemp, vwave, esurf, pweight;
  std::vector<double> dext, cont;  // dust: [nwave][ncont] extinction, [L][ncont] columns
  std::vector<double> n2nir;       // [nwave][2]: N2-N2, N2-H2 near-infrared CIA (cm-1 amagat-2)
  std::vector<double> solar;       // [nwave]: the registered stellar spectrum at vwave
  std::vector<double> ppg;         // [L][ngas] partial pressures (gas continua)
  std::vector<double> gcw;         // [nwave][RADG_GCW] gas-continuum factors (or empty)
  std::vector<PathInfo> paths;
  std::vector<int> layinc;
  std::vector<WaveGas> wg;
  double alg_bytes[3] = {0, 0, 0};  // table nodes, outputs, inputs
};

bool finite_pos(double x) { return std::isfinite(x) && x > 0; }

int pack_request(const radg_request &q, const TableSet &ts, int mode, const AuxTables &aux,
                 Packed &P) {
  const ParaCIA *para = aux.para.get();
  const int L = q.nlayer, ngas = q.ngas, npath = q.npath, nwave = q.nwave;
  if (L < 1 || L > RADG_MAXLAY) return fail(RADG_ELIMITS, "nlayer=%d outside 1..%d", L, RADG_MAXLAY);
  if (npath < 1 || npath > RADG_MAXPAT) return fail(RADG_ELIMITS, "npath=%d outside 1..%d", npath, RADG_MAXPAT);
  if (ngas < 1 || ngas > RADG_MAXGAS) return fail(RADG_ELIMITS, "ngas=%d outside 1..%d", ngas, RADG_MAXGAS);
  if (nwave < 1) return fail(RADG_ELIMITS, "nwave=%d", nwave);
  if (q.ispace != 0 && q.ispace != 1) return fail(RADG_EINPUT, "ispace=%d", q.ispace);
  // dust, check_limits.f:24-33 and cirsradg_wave.f:439-453
  const int ncont = q.ncont;
  if (ncont < 0 || ncont > RADG_MAXCON) return fail(RADG_ELIMITS, "ncont=%d outside 0..%d", ncont, RADG_MAXCON);
  if (ncont > 0) {
    if (q.nsec < 1 || q.nsec > RADG_MAXSEC) return fail(RADG_ELIMITS, "nsec=%d outside 1..%d", q.nsec, RADG_MAXSEC);
    if (!q.cont || !q.vsec || !q.xsec) return fail(RADG_EINPUT, "ncont=%d without cont/vsec/xsec", ncont);
    for (int i = 0; i < q.nsec; i++)
      if (!std::isfinite(q.vsec[i]) || (i > 0 && !(q.vsec[i] > q.vsec[i - 1])))
        return fail(RADG_EINPUT, "dust wavenumbers vsec must be finite and increasing");
    for (long long i = 0; i < (long long)ncont * q.nsec; i++)
      if (!std::isfinite(q.xsec[i])) return fail(RADG_EINPUT, "dust cross-section not finite");
    for (long long i = 0; i < (long long)L * ncont; i++)
      if (!std::isfinite(q.cont[i])) return fail(RADG_EINPUT, "dust column cont not finite");
  }
  // calc_nlte_t.f:63-67 stops for k1 outside [1e-20, 1e-3]
  if (q.inlte == 1 && !(q.nlte_k1 >= 1.0e-20 && q.nlte_k1 <= 1.0e-3))
    return fail(RADG_EINPUT, "inlte=1 needs 1e-20 <= k1 <= 1e-3 (k1=%g)", q.nlte_k1);
  if (q.inlte == 1 && !std::isfinite(q.nlte_a)) return fail(RADG_EINPUT, "nlte A=%g", q.nlte_a);
  // para-H2 (cirsradg_wave.f:376,633-637): the fraction per layer is hfp
  const int h2p = q.flagh2p == 1 ? 1 : 0;
  if (h2p) {
    if (!para) return fail(RADG_ENOTABLES, "FLAGH2P=1 needs a para-H2 table (radg_register_parah2)");
    if (!q.hfp) return fail(RADG_EINPUT, "FLAGH2P=1 without hfp");
    for (int j = 0; j < L; j++)
      if (!std::isfinite(q.hfp[j])) return fail(RADG_EINPUT, "hfp(%d)=%g", j + 1, q.hfp[j]);
  }
  const int nparam = ngas + 1 + ncont + h2p;  // cirsradg_wave.f:375-376
  if (q.nsw < 0 || q.nsw > nparam) return fail(RADG_ELIMITS, "nsw=%d > nparam=%d", q.nsw, nparam);
  for (int k = 0; k < q.nsw; k++)
    if (q.isw[k] < 1 || q.isw[k] > nparam) return fail(RADG_EINPUT, "isw(%d)=%d outside 1..%d", k + 1, q.isw[k], nparam);
  bool has_n2 = false;
  for (int k = 0; k < ngas; k++) has_n2 = has_n2 || q.idgas[k] == 22;
  for (int j = 0; j < L; j++) {
    if (!finite_pos(q.press[j]) || !finite_pos(q.temp[j]) || !std::isfinite(q.delh[j]))
      return fail(RADG_EINPUT, "layer %d: press=%g temp=%g delh=%g", j + 1, q.press[j], q.temp[j], q.delh[j]);
    for (int k = 0; k < ngas; k++)
      if (!std::isfinite(q.amount[j * ngas + k]) || !std::isfinite(q.pp[j * ngas + k]) ||
          q.amount[j * ngas + k] < 0 || q.pp[j * ngas + k] < 0)
        return fail(RADG_EINPUT, "layer %d gas %d: amount/pp invalid", j + 1, k + 1);
  }
  int maxlays = 0;
  for (int p = 0; p < npath; p++) {
    const int nl = q.nlayin[p];
    if (nl < 1 || nl > L) return fail(RADG_ELIMITS, "nlayin(%d)=%d outside 1..nlayer", p + 1, nl);
    const int im = q.imod[p];
    if (im != 0 && im != 1 && im != 3 && im != 8 && im != 13 && im != 14 && im != 18)
      return fail(RADG_EUNSUPPORTED, "imod(%d)=%d: only 0, 1, 3, 8, 13, 14 and 18 are provided", p + 1, im);
    // imod 8 (cirsradg_wave.f:1037-1082): path layinc(1) must be a cell (imod
    // 13 or 14; the reference stops otherwise), both it and the atmosphere
    // layinc(2) must come before this path (the reference reads their values
    // for this g-ordinate), the atmosphere not itself an imod 8 product, and
    // it must cover the nlayin(1) rows that are summed and mapped
    if (im == 8) {
      if (nl < 2) return fail(RADG_ELIMITS, "imod(%d)=8 needs nlayin >= 2 (its two paths)", p + 1);
      const int i1 = q.layinc[(long long)p * L] - 1, i2 = q.layinc[(long long)p * L + 1] - 1;
      if (i1 < 0 || i2 < 0 || i1 >= p || i2 >= p)
        return fail(RADG_EUNSUPPORTED, "imod(%d)=8: its paths %d and %d must precede it", p + 1, i1 + 1, i2 + 1);
      if (q.imod[i1] != 13 && q.imod[i1] != 14)
        return fail(RADG_EINPUT, "imod(%d)=8: path %d must be a cell (imod 13 or 14)", p + 1, i1 + 1);
      if (q.imod[i2] == 8 || q.nlayin[0] > q.nlayin[i2])
        return fail(RADG_EUNSUPPORTED, "imod(%d)=8: atmosphere path %d must be no product and cover nlayin(1)",
                    p + 1, i2 + 1);
    }
    // imod 18 interpolates E3 (e3interp.f) and always adds the surface at
    // tsurf (cirsradg_wave.f:1159,1213)
    if (q.imod[p] == 18 && !aux.e3)
      return fail(RADG_ENOTABLES, "imod(%d)=18 needs the E3 table (radg_register_e3)", p + 1);
    if (q.imod[p] == 18 && !finite_pos(q.tsurf))
      return fail(RADG_EUNSUPPORTED, "imod(%d)=18 needs tsurf > 0 (tsurf=%g)", p + 1, q.tsurf);
    for (int m = 0; m < nl; m++) {
      const int li = q.layinc[(long long)p * L + m];
      if (li < 1 || li > L) return fail(RADG_ELIMITS, "layinc(%d,%d)=%d exceeds nlayer", m + 1, p + 1, li);
      if (!std::isfinite(q.scale[(long long)p * L + m]))
        return fail(RADG_EINPUT, "scale(%d,%d) not finite", m + 1, p + 1);
      if ((q.imod[p] == 3 || q.imod[p] == 18) && !finite_pos(q.emtemp[(long long)p * L + m]))
        return fail(RADG_EINPUT, "emtemp(%d,%d)=%g", m + 1, p + 1, q.emtemp[(long long)p * L + m]);
    }
    maxlays = std::max(maxlays, nl);
  }

  DevReq &d = P.d;
  d.L = L;
  d.ngas = ngas;
  d.npath = npath;
  d.nparam = nparam;
  d.ncont = ncont;
  d.flagh2p = h2p;
  d.inlte = q.inlte;
  d.nlte_k1 = q.nlte_k1;
  d.nlte_a = q.nlte_a;
  // cirsradg_wave.f:912-925 with no solar file; pi is constdef.f's 3.1415927
  d.xfac = 1.0;
  if (q.iform == 1 || q.iform == 3) {
    const double pi = 3.1415927, rr = (q.radius1 + q.radextra) * 1e5;
    d.xfac = d.xfac * pi * 4. * pi * (rr * rr);
  }
  if (q.iform == 3) d.xfac = d.xfac * 1e-18;
  if (!std::isfinite(d.xfac)) return fail(RADG_EINPUT, "radius1=%g radextra=%g", q.radius1, q.radextra);
  // imod 18 (cirsradg_wave.f:1134-1149): 4 pi r^2 without imod 3's further
  // pi, and the 2 pi that every imod 18 term carries
  {
    const double pi = 3.1415927;
    double x = 1.0;
    if (q.iform == 1 || q.iform == 3) {
      const double rr = (q.radius1 + q.radextra) * 1e5;
      x = x * 4. * pi * (rr * rr);
    }
    if (q.iform == 3) x = x * 1e-18;
    d.xfac18 = x * 2.0 * pi;
  }
  // with a stellar spectrum (iread = 999): iform 1 divides imod 3 and 18
  // paths by it (cirsradg_wave.f:921-925,1144-1149), iform 4 multiplies imod
  // 0 transmissions by its flux at dist AU (:866-869, get_solar_wave.f:41-46)
  d.solar = aux.solar && (q.iform == 1 || q.iform == 4) ? q.iform : 0;
  d.solar_area = 1.0;
  if (d.solar == 4 && q.dist > 0.0) {
    const double pi = 3.1415927, au = 1.49598e13;
    d.solar_area = 4 * pi * (q.dist * au) * (q.dist * au);
  }
  d.nsw = q.nsw;
  d.nwave = nwave;
  d.ispace = q.ispace;
  d.inormal = q.inormal;
  // cirsradg_wave.f:663: the Rayleigh branch runs for iray > 0 and itype != 0;
  // iray 1, 2, 3 pick rayleighj/v/a, any larger value rayleighls.
  d.iray = (q.iray > 0 && q.itype != 0) ? std::min(q.iray, 4) : 0;
  d.tsurf = q.tsurf;
  d.tsurf_le0 = q.tsurf <= 0.0;
  for (int k = 0; k < RADG_MAXPARAM; k++) d.isw[k] = k < q.nsw ? q.isw[k] - 1 : -1;
  for (int l = 0; l < 5; l++) d.iabs[l] = -1;

  // gas continua (cirsradg_wave.f:590-611, ngascon.f:144-318): each gas's
  // kind under the registered /GASCONDEF/ flags.  H- free-free reads the
  // electron state (electron.f) that earlier gases of the same layer set:
  // e- (80) partial pressure, H- (91) amount and partial pressure.
  const GasCon gnone;
  const GasCon &gc = aux.gascon ? *aux.gascon : gnone;
  bool any_gc = false;
  {
    int k80 = -1, k91 = -1, first48 = -1, last8091 = -1;
    for (int k = 0; k < ngas; k++) {
      if (q.idgas[k] == 80) k80 = last8091 = k;
      if (q.idgas[k] == 91) k91 = last8091 = k;
      if (q.idgas[k] == 48 && q.isogas[k] == 0 && first48 < 0) first48 = k;
    }
    static const int uvid[10] = {11, 28, 26, 27, 30, 6, 5, 33, 1, 34};  // ngascon.f:274-318
    for (int k = 0; k < ngas; k++) {
      const int id = q.idgas[k], iso = q.isogas[k];
      const bool iso01 = iso == 0 || iso == 1;
      signed char kind = RADG_GC_NONE;
      for (int b = 0; b < 10 && gc.iuvscat > 0; b++)
        if (id == uvid[b] && (gc.iuvscat & (1 << b)))
          return fail(RADG_EUNSUPPORTED, "gas %d (id %d): UV cross-sections (iuvscat=%d) are not provided",
                      k + 1, id, gc.iuvscat);
      if (id == 1 && gc.ih2o == 1) {
        if (gc.s296.size() < 2) return fail(RADG_ENOTABLES, "IH2O=1 without the CKD H2O tables");
        kind = RADG_GC_H2O;
      } else if (id == 1 && iso01 && gc.ih2o == 2) {
        return fail(RADG_EUNSUPPORTED, "IH2O=2: ngascon.f leaves the H2O absorption unassigned");
      } else if (id == 6 && iso01 && gc.ich4 != 0) {
        if (gc.ch4_wl.size() < 2) return fail(RADG_ENOTABLES, "ICH4 set without Karkoschka's CH4 table");
        kind = RADG_GC_CH4;
      } else if (id == 11 && iso01 && gc.inh3 != 0) {
        if (gc.nh3_wl.size() < 2) return fail(RADG_ENOTABLES, "INH3 set without Lutz & Owen's NH3 table");
        kind = RADG_GC_NH3;
      } else if (id == 3 && iso01 && gc.io3 != 0) {
        return fail(RADG_EUNSUPPORTED, "IO3 set: the ozone UV cross-sections are not provided");
      } else if (id == 48 && iso == 0 && k80 >= 0) {
        if (k91 < 0 || last8091 > first48)
          return fail(RADG_EUNSUPPORTED,
                      "H- free-free (gas 48) needs e- (80) and H- (91) before it in the gas list");
        for (int j = 0; j < L; j++)
          if (!(q.pp[j * ngas + k91] > 0))
            return fail(RADG_EINPUT, "layer %d: H- free-free divides by the H- partial pressure %g", j + 1,
                        q.pp[j * ngas + k91]);
        kind = RADG_GC_HFF;
      } else if (id == 91 && iso == 0) {
        kind = RADG_GC_HBF;
      }
      d.gcon[k] = kind;
      any_gc = any_gc || kind != RADG_GC_NONE;
    }
    for (int k = ngas; k < RADG_MAXGAS; k++) d.gcon[k] = RADG_GC_NONE;
    d.ke80 = (short)std::max(k80, 0);
    d.ke91 = (short)std::max(k91, 0);
  }
  P.ppg.assign(q.pp, q.pp + (size_t)L * ngas);
  P.gcw.assign(any_gc ? (size_t)nwave * RADG_GCW : 0, 0.0);
  for (int iw = 0; any_gc && iw < nwave; iw++) {
    const double vn = q.ispace == 1 ? 1e4 / q.vwave[iw] : q.vwave[iw];
    double *gw = P.gcw.data() + (size_t)iw * RADG_GCW;
    if (gc.ich4 != 0) gw[RADG_GCW_CH4] = tabcont(gc.ch4_wl, gc.ch4_k, vn) / 26850.0;
    if (gc.inh3 != 0) gw[RADG_GCW_NH3] = 1000 * tabcont(gc.nh3_wl, gc.nh3_k, vn) / 26850.0;
    gw[RADG_GCW_HBF] = hmin_bf(vn);
    if (gc.ih2o == 1) h2o_factors(gc, vn, gw + RADG_GCW_H2OIN);
  }

  // cirsradg_wave.f:393-414: the gases whose fractions feed rayleighls (the
  // last H2, He, CH4, NH3 of isotope 0 or 1); an absent one contributes 0.
  int jh2 = -1, jhe = -1, jch4 = -1, jnh3 = -1;
  for (int k = 0; k < ngas; k++) {
    if (q.isogas[k] != 0 && q.isogas[k] != 1) continue;
    if (q.idgas[k] == 39) jh2 = k;
    if (q.idgas[k] == 40) jhe = k;
    if (q.idgas[k] == 6) jch4 = k;
    if (q.idgas[k] == 11) jnh3 = k;
  }

  // ---- per layer
  const int np = ts.np, nt = ts.nt, NG = ts.ng;
  P.lay.resize(L);
  P.amo.resize((size_t)L * ngas);
  P.press.assign(q.press, q.press + L);
  const double MODBOLTZA = 10. * KBOLTZMANN / 1.013;
  std::set<long long> nodes;  // distinct (cp,ct) corners touched
  for (int j = 0; j < L; j++) {
    LayerConst &c = P.lay[j];
    // get_kg.f:113-144
    double p1 = std::log(q.press[j]);
    p1 = std::min(std::max(p1, ts.lnp[0]), ts.lnp[np - 1]);
    int cp = locate(ts.lnp.data(), np, p1);
    cp = std::min(std::max(cp, 1), np - 1);
    c.cp = cp;
    c.v = (p1 - ts.lnp[cp - 1]) / (ts.lnp[cp] - ts.lnp[cp - 1]);
    // the T bracket on a temperature grid (get_kg.f:126-139); with one grid
    // per pressure (NT < 0) on the grids of both bracketing pressures
    // (:145-173)
    auto tbracket = [&](const double *tn, int &ct, double &u, double &dudt) {
      const double t1 = std::min(std::max(q.temp[j], tn[0]), tn[nt - 1]);
      ct = std::min(std::max(locate(tn, nt, t1), 1), nt - 1);
      u = (t1 - tn[ct - 1]) / (tn[ct] - tn[ct - 1]);
      dudt = 1.0 / (tn[ct] - tn[ct - 1]);
    };
    if (!ts.t2) {
      tbracket(ts.tk.data(), c.ct, c.u, c.dudt);
      c.ct2 = c.ct;
      c.u2 = c.u;
      c.dudt2 = c.dudt;
    } else {
      tbracket(ts.tk.data() + (size_t)(cp - 1) * nt, c.ct, c.u, c.dudt);
      tbracket(ts.tk.data() + (size_t)cp * nt, c.ct2, c.u2, c.dudt2);
    }
    c.pad2 = 0;
    for (int b = 0; b < 2; b++) {
      nodes.insert((long long)(cp - 1) * nt + (c.ct - 1 + b));
      nodes.insert((long long)cp * nt + (c.ct2 - 1 + b));
    }
    double ut = 0;
    for (int k = 0; k < ngas; k++) {
      ut += q.amount[j * ngas + k];
      P.amo[(size_t)j * ngas + k] = q.amount[j * ngas + k] * 1.0e-20;
    }
    c.utotl = ut * 1.e-20;
    c.temp = q.temp[j];
    c.press = q.press[j];
    c.fpara = h2p ? q.hfp[j] : 0.0;
    {
      const double *ppj = q.pp + (size_t)j * ngas;
      const double fh2 = jh2 >= 0 ? ppj[jh2] / q.press[j] : 0.0;
      const double fhe = jhe >= 0 ? ppj[jhe] / q.press[j] : 0.0;
      const double fch4 = jch4 >= 0 ? ppj[jch4] / q.press[j] : 0.0;
      c.fnh3 = jnh3 >= 0 ? ppj[jnh3] / q.press[j] : 0.0;
      c.fheh2 = fh2 > 0.0 ? fhe / fh2 : 0.0;
      c.fch4h2 = fh2 > 0.0 ? fch4 / fh2 : 0.0;
    }
    // nciacon.f:65-176
    const double p = q.press[j], t = q.temp[j];
    const double *pp = q.pp + (size_t)j * ngas;
    const double *am = q.amount + (size_t)j * ngas;
    c.qh2 = c.qhe = c.qn2 = c.qch4 = c.qco2 = 0;
    for (int i = 0; i < ngas; i++) {
      const int id = q.idgas[i], iso = q.isogas[i];
      if (id == 39 && (iso == 0 || iso == 1)) { c.qh2 = pp[i] / p; d.iabs[0] = i; }
      if (id == 40) { c.qhe = pp[i] / p; d.iabs[1] = i; }
      if (id == 22) { c.qn2 = pp[i] / p; d.iabs[2] = i; }
      if (id == 6 && (iso == 0 || iso == 1)) { c.qch4 = pp[i] / p; d.iabs[3] = i; }
      if (id == 2) { c.qco2 = pp[i] / p; d.iabs[4] = i; }
    }
    const double xlen = q.delh[j];
    if (xlen == 0.0) {
      double xp1 = 0.0, height = 0.0;
      for (int i = 0; i < ngas; i++)
        if (pp[i] > xp1) {
          height = am[i] * MODBOLTZA * t / pp[i];
          xp1 = pp[i];
        }
      c.tau = ((p / 1.0) * (p / 1.0)) * ((273.15 / t) * (273.15 / t)) * height;
    } else {
      const double height = xlen * 1e5;
      double xp1 = -1.0, totam = 0.0;
      for (int i = 0; i < ngas; i++)
        if (pp[i] > xp1 && pp[i] > 0.0) {
          totam = am[i] * p / pp[i];
          xp1 = pp[i];
        }
      const double amag1 = (totam / height) / AMAGAT;
      c.tau = height * amag1 * amag1;
    }
    if (c.utotl <= 0.0 && (c.qh2 > 0 || c.qhe > 0))
      return fail(RADG_EINPUT, "layer %d: zero total amount with CIA gases present", j + 1);
  }

  // ---- per path
  P.paths.resize(npath);
  P.layinc.assign((size_t)npath * L, 0);
  P.scale.assign((size_t)npath * L, 0.0);
  P.emtemp.assign((size_t)npath * L, 0.0);
  P.pweight.assign(npath, 0.0);
  for (int p = 0; p < npath; p++) {
    const int nl = q.nlayin[p];
    PathInfo &pi = P.paths[p];
    pi.nlays = nl;
    pi.imod = q.imod[p];
    pi.cell = pi.atm = -1;
    int j1 = (int)(0.5 * nl);
    if (j1 < 1) j1 = 1;
    const double pa = q.press[q.layinc[(long long)p * L + j1 - 1] - 1];
    const double pb = q.press[q.layinc[(long long)p * L + nl - 1] - 1];
    pi.surf = pb > pa || pi.imod == 18;  // cirsradg_wave.f:992-999; imod 18 always
    // an imod 8 path's layer list is two path numbers; its rows are
    // nlayin(1) rows on path 1's layers, as map2pro.f:52-57 maps them
    const int src = pi.imod == 8 ? 0 : p;
    if (pi.imod == 8) {
      pi.cell = q.layinc[(long long)p * L] - 1;
      pi.atm = q.layinc[(long long)p * L + 1] - 1;
      pi.nlays = q.nlayin[0];
      pi.surf = 0;
    }
    for (int m = 0; m < pi.nlays; m++) {
      P.layinc[(size_t)p * L + m] = q.layinc[(long long)src * L + m] - 1;
      P.scale[(size_t)p * L + m] = q.scale[(long long)src * L + m];
      P.emtemp[(size_t)p * L + m] = q.emtemp[(long long)src * L + m];
    }
    P.pweight[p] = q.path_weight ? q.path_weight[p] : 0.0;
  }

  // ---- per wavenumber
  // iform 1 divides imod 3 and 18 paths by the star (cirsradg_wave.f:921-925,
  // 1144-1149); other paths never read the spectrum
  bool divides_by_star = false;
  for (int p = 0; p < npath; p++) divides_by_star = divides_by_star || q.imod[p] == 3 || q.imod[p] == 18;
  P.vwave.assign(q.vwave, q.vwave + nwave);
  P.esurf.resize(nwave);
  P.dext.assign((size_t)nwave * ncont, 0.0);
  P.n2nir.assign((size_t)nwave * 2, 0.0);
  P.solar.assign((size_t)nwave, 0.0);
  P.cont.assign(q.cont, q.cont + (size_t)L * ncont);
  P.wg.resize((size_t)nwave * ngas);
  const long long ntab = (long long)np * nt * NG;
  const size_t kbytes = (mode == RADG_MODE_FP64) ? 8 : (mode == RADG_MODE_MIXED) ? 4 : 2;
  const size_t obytes_el = (mode == RADG_MODE_FP64) ? 8 : 4;
  double tbytes = 0;
  for (int iw = 0; iw < nwave; iw++) {
    const double v = q.vwave[iw];
    if (!std::isfinite(v) || v <= 0) return fail(RADG_EINPUT, "vwave(%d)=%g", iw + 1, v);
    // N2's near-infrared CIA (n2n2cia.f, n2h2cia.f; nciacon.f:287-296,
    // nparacon_all.f:312-320), at the wavenumber the CIA sees: tabulated for
    // 3995-4995 cm-1 and zero elsewhere, so with N2 present there the tables
    // must be registered
    {
      const double vn = q.ispace == 1 ? 1e4 / v : v;
      if (aux.n2nir) {
        P.n2nir[(size_t)iw * 2] = n2nir_coef(aux.n2nir->wn1, aux.n2nir->k1, vn);
        P.n2nir[(size_t)iw * 2 + 1] = n2nir_coef(aux.n2nir->wn2, aux.n2nir->k2, vn);
      } else if (has_n2 && vn >= 3995.0 && vn <= 4995.0) {
        return fail(RADG_ENOTABLES, "N2 at %g cm-1 needs its near-infrared CIA (radg_register_n2nir)", vn);
      }
    }
    // the stellar spectrum at this wavenumber; iform 1 divides by it, which
    // the reference would do at 0 too (an infinite result): refused
    if (aux.solar && (q.iform == 1 || q.iform == 4)) {
      P.solar[iw] = solar_at(*aux.solar, v);
      if (q.iform == 1 && divides_by_star && !(P.solar[iw] != 0.0 && std::isfinite(P.solar[iw])))
        return fail(RADG_EINPUT, "iform=1: wavenumber %g outside the stellar spectrum", v);
    }
    // interpem.f:17-30 (called unconditionally by cirsradg_wave.f:381)
    const int nem = q.nem;
    if (nem < 1 || v < q.vem[0] || v > q.vem[nem - 1])
      return fail(RADG_EINPUT, "wavenumber %g outside the emissivity grid", v);
    double es = 0.0;
    for (int i = 0; i < nem - 1; i++)
      if (v >= q.vem[i] && v < q.vem[i + 1]) {
        const double f = (v - q.vem[i]) / (q.vem[i + 1] - q.vem[i]);
        es = (1 - f) * q.emissivity[i] + f * q.emissivity[i + 1];
      }
    if (v == q.vem[nem - 1]) es = q.emissivity[nem - 1];
    P.esurf[iw] = es;
    // dust extinction at this wavenumber (cirsradg_wave.f:559-569): the one
    // value for nsec = 1, else interp.f on the cross-section grid, which
    // must cover the wavenumber (:446-452)
    if (ncont > 0) {
      const int ns = q.nsec;
      if (ns > 1 && (v < q.vsec[0] || v > q.vsec[ns - 1]))
        return fail(RADG_EINPUT, "wavenumber %g outside the dust cross-section grid [%g, %g]", v,
                    q.vsec[0], q.vsec[ns - 1]);
      int i = 1;  // interp.f: first grid point above v (1-based), at least 2
      if (ns > 1) {
        while (i <= ns && !(q.vsec[i - 1] > v)) i++;
        i = std::max(2, std::min(i, ns));
      }
      for (int k = 0; k < ncont; k++) {
        const double *y = q.xsec + (size_t)k * ns;
        P.dext[(size_t)iw * ncont + k] =
            ns == 1 ? y[0]
                    : y[i - 2] + (y[i - 1] - y[i - 2]) * (v - q.vsec[i - 2]) / (q.vsec[i - 1] - q.vsec[i - 2]);
      }
    }
    for (int k = 0; k < ngas; k++) {
      WaveGas &w = P.wg[(size_t)iw * ngas + k];
      w.base = -1;
      w.base2 = -1;
      w.w2 = 0.0;
      w.coinc = 1;
      w.ob = w.ob2 = 0;
      w.pad = 0;
      if (ts.lbl) {
        // get_klblg.f:167-222: the last table of the gas (read_klbllist.f:
        // 190-198), slice 1 + INT((v + 0.02 delv - vmin) / delv), which must
        // lie on the grid (:186-189 read past the table otherwise); never
        // merged; an all-zero slice gives k = 0
        const Table *tl = nullptr;
        for (const Table &t : ts.tabs)
          if (t.idgas == q.idgas[k] && t.isogas == q.isogas[k]) tl = &t;
        if (!tl) continue;
        const double x = (v + 0.02 * tl->delv - tl->vmin) / tl->delv;
        if (!(x > -1.0 && x < (double)tl->npoint))
          return fail(RADG_EINPUT, "wavenumber %g outside the line-by-line tables' grid [%g, %g]", v, tl->vmin,
                      tl->vmin + (tl->npoint - 1) * tl->delv);
        const long long n1 = (long long)x;  // INT truncates toward zero
        if (tl->zero[n1]) continue;
        w.base = tl->off + n1 * ntab;
        w.ob = w.base / ntab * NG;
        tbytes += (double)nodes.size() * kbytes;
        continue;
      }
      const Table *tb = nullptr;
      for (const Table &t : ts.tabs) {  // read_klist.f:260-290: first covering match
        if (t.idgas != q.idgas[k] || t.isogas != q.isogas[k]) continue;
        if (t.delv <= 0) {  // an irregular grid covers every wavenumber (:292-330)
          tb = &t;
          continue;
        }
        if (tb) continue;
        // the grid ends are widened by get_kg's coincidence tolerance so that
        // a REAL(4) wavenumber rounded past the last point still matches
        const double xmax = t.vmin + (t.npoint - 1) * t.delv;
        const double tol = (t.delv >= 0.0001) ? 0.02 * t.delv : 0.25 * t.delv;
        if (v >= t.vmin - tol && v <= xmax + tol) tb = &t;
      }
      if (!tb) continue;
      if (tb->delv <= 0) {
        // read_klist.f:292-330: DELV < 0 takes the point nearest in per cent
        // (get_kg.f:267-273: coincident); DELV = 0 the nearest at or below,
        // merged with the next one (:274-279, :386-389)
        long long ipo = -1;
        double best = 1e10, frack = 1.0;
        for (int i = 0; i < tb->npoint; i++) {
          const double dx = 100 * std::fabs((v - tb->vcen[i]) / tb->vcen[i]);
          if (tb->delv == 0) {
            if (dx < best && tb->vcen[i] <= v) {
              ipo = i;
              frack = i < tb->npoint - 1 ? 1.0 - (v - tb->vcen[i]) / (tb->vcen[i + 1] - tb->vcen[i]) : 1.;
              best = dx;
            }
          } else if (dx < best) {
            ipo = i;
            best = dx;
          }
        }
        if (ipo < 0)
          return fail(RADG_EINPUT, "wavenumber %g lies below table %d's irregular grid (DELV=0)", v,
                      (int)(tb - ts.tabs.data()) + 1);
        if (tb->zero[ipo]) continue;
        w.base = tb->off + ipo * ntab;
        w.ob = w.base / ntab * NG;
        if (tb->delv == 0) {
          w.coinc = 0;
          w.base2 = (ipo + 1 < tb->npoint) ? w.base + ntab : w.base;
          w.ob2 = w.base2 / ntab * NG;
          w.w2 = 1.0 - frack;
          tbytes += (double)nodes.size() * NG * kbytes;
        }
        tbytes += (double)nodes.size() * NG * kbytes;
        continue;
      }
      const double eps = (tb->delv >= 0.0001) ? 0.02 * tb->delv : 0.25 * tb->delv;
      long long n1 = (long long)((v + eps - tb->vmin) / tb->delv);
      n1 = std::min<long long>(std::max<long long>(n1, 0), tb->npoint - 1);
      const double tmp = tb->vmin + tb->delv * n1;
      // get_kg.f:224-250 tests the slice at or below the wavenumber only
      if (tb->zero[n1]) continue;  // all-zero slice: k = 0
      w.base = tb->off + n1 * ntab;
      w.ob = w.base / ntab * NG;
      if (!(std::fabs(v - tmp) < eps)) {
        // between table points: interpolate the two slices and merge them
        // with rankk (get_kg.f:379-440); the last point has no slice above
        // and reuses its own
        w.coinc = 0;
        w.base2 = (n1 + 1 < tb->npoint) ? w.base + ntab : w.base;
        w.ob2 = w.base2 / ntab * NG;
        w.w2 = std::min(1.0, std::max(0.0, (v - tmp) / tb->delv));
        tbytes += (double)nodes.size() * NG * kbytes;
      }
      tbytes += (double)nodes.size() * NG * kbytes;
    }
  }
  // gases with a k-table anywhere on the spectrum get a dk row
  P.d.nks = 0;
  for (int k = 0; k < ngas; k++) {
    bool any = false;
    for (int iw = 0; iw < nwave && !any; iw++) any = P.wg[(size_t)iw * ngas + k].base >= 0;
    P.d.kslot[k] = any ? (signed char)P.d.nks++ : (signed char)-1;
  }
  for (int k1 = 0; k1 < RADG_MAXPARAM; k1++) {
    const int k = k1 < P.d.nsw ? P.d.isw[k1] : -1;
    P.d.kdrow[k1] = (k >= 0 && k < ngas) ? P.d.kslot[k] : (k == ngas ? (signed char)P.d.nks : (signed char)-1);
  }
  // algorithmic bytes: table nodes + layer inputs + CIA nodes + outputs
  const double obytes = (double)nwave * ((double)npath * (2.0 + (double)L * nparam) * obytes_el +
                                         (2.0 + (double)L * nparam) * 8.0);
  const double ibytes = (double)L * (4 + 2 * ngas + ncont) * 8.0 + (double)nwave * (4 * 9 + ncont) * 8.0;
  P.alg_bytes[0] = tbytes;
  P.alg_bytes[1] = obytes;
  P.alg_bytes[2] = ibytes;
  (void)maxlays;
  return RADG_OK;
}

}  // namespace

// Failures of the other sources (radg_oe.hip) land in the same per-thread
// message radg_last_error() returns.
int radg_fail_msg(int code, const char *msg) { return fail(code, "%s", msg); }

// ============================================================ batches
struct radg_batch {
  int mode = 0, nreq = 0;
  std::shared_ptr<TableSet> ts;
  std::shared_ptr<CIA> cia;
  AuxTables aux;
  std::vector<DevReq> reqs;
  char *dmem = nullptr;
  size_t bytes = 0;
  KernelArgs args{};
  int E = 1, RL = 1, maxngas = 1, maxnsw = 1;
  size_t ldsA = 0, ldsB = 0;
  // launches per class of requests with one emission layers-per-lane count
  // (chosen per request from its own layer count, so a request's arithmetic
  // never depends on the other requests of its batch).  A class whose
  // layer results fit the LDS of one workgroup runs as the fused kernel;
  // the others as the continuum, layer and emission kernels.  Both forms do
  // the same arithmetic.
  struct EmClass {
    int RL, nreq;
    bool disc;  // a member has an imod 8, 13, 14 or 18 path or uses the stellar spectrum
    long long nA, nB, nW;                      // layer, path, wavenumber items
    const int *reqmap;
    const long long *startA, *start, *startW;  // class-local item prefix sums
    size_t lds, ldsA;                          // emission, layer kernels
    int lwaves, lwstride;                      // layer kernel: items per workgroup, LDS per item
    bool fused;
    size_t ldsF;                               // fused kernel
    int wsbytes;
  };
  std::vector<EmClass> em;
  std::vector<long long> sig;  // layout of the plan the device block was laid out by
  double alg_bytes[3] = {0, 0, 0};
  hipStream_t stream = nullptr;
  bool stream_used = false;
  // event ring: 4 events per launch, the last RING launches are kept
  static constexpr int RING = 128;
  std::vector<hipEvent_t> evs;
  long long launches = 0;
  hipEvent_t *ev = nullptr;  // events of the last launch
  // output regions (byte offsets in dmem) and their pinned host copies
  size_t o_out = 0, o_gts = 0, o_jac = 0, o_wout = 0, o_wgts = 0, o_wjac = 0;
  size_t span_p0 = 0, span_p1 = 0, span_w0 = 0, span_w1 = 0;
  int fetch = 0;           // bit 0: weighted outputs, bit 1: per-path outputs
  char *hst = nullptr;     // pinned: [weighted span][per-path span]
  // host copy of each request's path layers (0-based layinc, zero past nlays)
  struct HostReq {
    int L = 0, npath = 0, ngas = 0, nwave = 0, ncont = 0, h2p = 0;
    std::vector<int> layinc, nlays;
  };
  std::vector<HostReq> hreq;
  // pinned staging of the inputs (and of the state maps), uploaded by the
  // next launch on its stream
  char *hup = nullptr, *sup = nullptr;
  size_t hup_bytes = 0, sup_bytes = 0;
  bool up_pending = false, sup_pending = false;
  // layer -> profile -> state maps (map2pro.f:49-84, map2xvec.f:41-54), set
  // by radg_batch_set_statemaps and applied after the quadrature of every
  // launch: one GEMM per (request, path class, parameter) forms the combined
  // matrix C[(j, k)][iv] = sum_level D_k[layinc(j)][level] xmap[iv][k][level],
  // then kk = wjac C when all paths of a request share their layers (else
  // the per-path products summed with the path weights)
  struct SMap {
    int nv = 0, npro = 0, ncls = 0, per_path = 0;
    std::vector<int> pcls;           // path -> class
    size_t o_cm = 0, o_kk = 0, o_dx = 0;
  };
  std::vector<SMap> sm;
  char *smem = nullptr;
  size_t smem_bytes = 0;
  GemmProblem *d_pc = nullptr, *d_pk64 = nullptr, *d_pkT = nullptr;
  void *d_ps = nullptr;
  int n_pc = 0, n_pk64 = 0, n_pkT = 0, n_ps = 0;
  int mx_pc[2] = {0, 0}, mx_pk64[2] = {0, 0}, mx_pkT[2] = {0, 0};
  long long mx_ps = 0;
  ~radg_batch();
};

namespace {

// Caching allocators.  Batches come and go at serving rates; a fresh
// hipMalloc / hipHostMalloc of hundreds of MB per batch would dominate, so
// freed blocks are kept (up to a cap) and handed out again for requests of
// between half and all of their size.  A block returned to a full cache
// evicts cached blocks, largest first, to make room, so that the blocks of
// an earlier, larger workload (which the current one cannot use) do not
// hold the cache and send every batch back to hipHostMalloc.  The caches
// live until the process ends.
struct BlockCache {
  std::mutex mu;
  std::multimap<size_t, void *> freed;
  std::map<void *, size_t> live;
  size_t cached = 0, cap = 0;
  bool pinned = false;
  BlockCache(size_t c, bool p) : cap(c), pinned(p) {}
  // Sizes are rounded up to classes (powers of two to 256 MiB, then steps of
  // 128 MiB) so that batches of different sizes reuse each other's blocks.
  static size_t size_class(size_t n) {
    const size_t big = (size_t)256 << 20;
    if (n <= big) {
      size_t c = 4096;
      while (c < n) c <<= 1;
      return c;
    }
    const size_t step = (size_t)128 << 20;
    return (n + step - 1) / step * step;
  }
  void *get(size_t n0) {
    const size_t n = size_class(n0);
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = freed.lower_bound(n);
      if (it != freed.end() && it->first <= 2 * n + ((size_t)64 << 20)) {
        void *p = it->second;
        const size_t sz = it->first;
        freed.erase(it);
        cached -= sz;
        live[p] = sz;
        return p;
      }
    }
    void *p = nullptr;
    hipError_t e = pinned ? hipHostMalloc(&p, n, hipHostMallocDefault) : hipMalloc(&p, n);
    if (e != hipSuccess) {
      // drop the cache and retry once
      release_all();
      e = pinned ? hipHostMalloc(&p, n, hipHostMallocDefault) : hipMalloc(&p, n);
      if (e != hipSuccess) return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    live[p] = n;
    return p;
  }
  void put(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu);
    auto it = live.find(p);
    if (it == live.end()) return;
    const size_t sz = it->second;
    live.erase(it);
    while (cached + sz > cap && !freed.empty()) {
      auto big = std::prev(freed.end());
      if (pinned) (void)hipHostFree(big->second);
      else (void)hipFree(big->second);
      cached -= big->first;
      freed.erase(big);
    }
    if (cached + sz <= cap) {
      freed.insert({sz, p});
      cached += sz;
    } else {
      if (pinned) (void)hipHostFree(p);
      else (void)hipFree(p);
    }
  }
  void release_all() {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &kv : freed) {
      if (pinned) (void)hipHostFree(kv.second);
      else (void)hipFree(kv.second);
    }
    freed.clear();
    cached = 0;
  }
};

BlockCache &dev_cache() {
  static BlockCache *c = new BlockCache((size_t)48 << 30, false);
  return *c;
}
BlockCache &host_cache() {
  static BlockCache *c = new BlockCache((size_t)8 << 30, true);
  return *c;
}

}  // namespace

radg_batch::~radg_batch() {
  if (launches > 0 && stream_used) (void)hipStreamSynchronize(stream);
  host_cache().put(hup);
  host_cache().put(sup);
  dev_cache().put(smem);
  dev_cache().put(dmem);
  host_cache().put(hst);
  for (auto &e : evs)
    if (e) (void)hipEventDestroy(e);
}

namespace {

struct Layout {
  size_t off = 0;
  template <typename X>
  size_t take(size_t n) {
    off = (off + 255) & ~size_t(255);
    const size_t o = off;
    off += n * sizeof(X);
    return o;
  }
};

struct Plan {
  std::vector<Packed> packs;
  size_t o_reqs, o_sA, o_sB, o_sC, o_lay, o_amo, o_press, o_paths, o_layinc, o_scale, o_emt,
      o_vw, o_es, o_n2, o_sol, o_ppg, o_gcw, o_wg, o_pw, o_dx, o_ct, o_inter, o_out, o_gts, o_jac, o_wout,
      o_wgts, o_wjac, o_err;
  size_t total = 0;
  long long nA = 0, nB = 0, nC = 0;
  long long nlay = 0, nlg = 0, npath = 0, npl = 0, nwv = 0, nwg = 0, ninter = 0, nout = 0, njac = 0,
            nwout = 0, nwjac = 0, ndx = 0, nct = 0, ngcw = 0;
  std::vector<long long> sA, sB, sC;
  int maxngas = 1, maxnsw = 1, maxlays = 1;
  std::vector<int> cls_rl, cls_ngas;
  std::vector<std::vector<int>> cls_members;
  size_t o_clsmap = 0, o_clsstart = 0, o_clsstartA = 0, o_clsstartW = 0;
};

int plan_batch(int nreq, const radg_request *reqs, int mode, const TableSet &ts,
               const AuxTables &aux, Plan &pl) {
  if (nreq < 1) return fail(RADG_EINPUT, "nreq=%d", nreq);
  pl.packs.resize(nreq);
  pl.sA.assign(nreq + 1, 0);
  pl.sB.assign(nreq + 1, 0);
  pl.sC.assign(nreq + 1, 0);
  for (int r = 0; r < nreq; r++) {
    int rc = pack_request(reqs[r], ts, mode, aux, pl.packs[r]);
    if (rc) return rc;
    DevReq &d = pl.packs[r].d;
    const radg_request &q = reqs[r];
    d.lay0 = pl.nlay;
    d.lg0 = pl.nlg;
    d.path0 = pl.npath;
    d.pl0 = pl.npl;
    d.wave0 = pl.nwv;
    d.wg0 = pl.nwg;
    d.itemA0 = pl.nA;
    d.itemB0 = pl.nB;
    d.itemC0 = pl.nC;
    d.inter0 = pl.ninter;
    d.out0 = pl.nout;
    d.jac0 = pl.njac;
    d.wout0 = pl.nwout;
    d.wjac0 = pl.nwjac;
    d.dx0 = pl.ndx;
    d.ct0 = pl.nct;

    d.gcw0 = pl.packs[r].gcw.empty() ? -1 : pl.ngcw;
    pl.ngcw += (long long)pl.packs[r].gcw.size();
    pl.ndx += (long long)q.nwave * q.ncont;
    pl.nct += (long long)q.nlayer * q.ncont;
    pl.nlay += q.nlayer;
    pl.nlg += (long long)q.nlayer * q.ngas;
    pl.npath += q.npath;
    pl.npl += (long long)q.npath * q.nlayer;
    pl.nwv += q.nwave;
    pl.nwg += (long long)q.nwave * q.ngas;
    pl.nA += (long long)q.nwave * q.nlayer;
    pl.nB += (long long)q.nwave * q.npath;
    pl.nC += q.nwave;
    pl.ninter += (long long)q.nwave * inter_block(ts.ng, q.nlayer, d.nks, d.nparam);
    pl.nout += (long long)q.nwave * q.npath;
    pl.njac += (long long)q.nwave * q.npath * q.nlayer * d.nparam;
    pl.nwout += q.nwave;
    pl.nwjac += (long long)q.nwave * q.nlayer * d.nparam;
    pl.sA[r + 1] = pl.nA;
    pl.sB[r + 1] = pl.nB;
    pl.sC[r + 1] = pl.nC;
    pl.maxngas = std::max(pl.maxngas, q.ngas);
    pl.maxnsw = std::max(pl.maxnsw, std::max(q.nsw, 1));
    pl.maxlays = std::max(pl.maxlays, q.nlayer);
  }
  // launch classes: requests grouped by their own layers-per-lane and gas
  // count, so that each launch sizes its LDS (and so its occupancy) by its
  // members alone
  pl.cls_rl.clear();
  pl.cls_ngas.clear();
  pl.cls_members.clear();
  for (int r = 0; r < nreq; r++) {
    const int rl = radg_emission_RL(reqs[r].nlayer), ng = reqs[r].ngas;
    size_t c = 0;
    while (c < pl.cls_rl.size() && (pl.cls_rl[c] != rl || pl.cls_ngas[c] != ng)) c++;
    if (c == pl.cls_rl.size()) {
      pl.cls_rl.push_back(rl);
      pl.cls_ngas.push_back(ng);
      pl.cls_members.emplace_back();
    }
    pl.cls_members[c].push_back(r);
  }
  const size_t ts_ = (mode == RADG_MODE_FP64) ? 8 : 4;
  Layout lo;
  pl.o_reqs = lo.take<DevReq>(nreq);
  pl.o_sA = lo.take<long long>(nreq + 1);
  pl.o_sB = lo.take<long long>(nreq + 1);
  pl.o_sC = lo.take<long long>(nreq + 1);
  pl.o_lay = lo.take<LayerConst>(pl.nlay);
  pl.o_amo = lo.take<double>(pl.nlg);
  pl.o_press = lo.take<double>(pl.nlay);
  pl.o_paths = lo.take<PathInfo>(pl.npath);
  pl.o_layinc = lo.take<int>(pl.npl);
  pl.o_scale = lo.take<double>(pl.npl);
  pl.o_emt = lo.take<double>(pl.npl);
  pl.o_vw = lo.take<double>(pl.nwv);
  pl.o_es = lo.take<double>(pl.nwv);
  pl.o_n2 = lo.take<double>(2 * pl.nwv);
  pl.o_sol = lo.take<double>(pl.nwv);
  pl.o_ppg = lo.take<double>(pl.nlg);
  pl.o_gcw = lo.take<double>(pl.ngcw);
  pl.o_wg = lo.take<WaveGas>(pl.nwg);
  pl.o_pw = lo.take<double>(pl.npath);
  pl.o_dx = lo.take<double>(pl.ndx);
  pl.o_ct = lo.take<double>(pl.nct);
  pl.o_clsmap = lo.take<int>(nreq);
  pl.o_clsstart = lo.take<long long>(nreq + pl.cls_rl.size());
  pl.o_clsstartA = lo.take<long long>(nreq + pl.cls_rl.size());
  pl.o_clsstartW = lo.take<long long>(nreq + pl.cls_rl.size());
  pl.o_inter = lo.take<char>((size_t)pl.ninter * ts_);
  pl.o_out = lo.take<char>((size_t)pl.nout * ts_);
  pl.o_gts = lo.take<char>((size_t)pl.nout * ts_);
  pl.o_jac = lo.take<char>((size_t)pl.njac * ts_);
  pl.o_wout = lo.take<double>(pl.nwout);
  pl.o_wgts = lo.take<double>(pl.nwout);
  pl.o_wjac = lo.take<double>(pl.nwjac);
  pl.o_err = lo.take<int>(4);
  pl.total = lo.off;
  return RADG_OK;
}

template <typename X>
void put(char *h, size_t off, const X *src, size_t n) {
  if (n) std::memcpy(h + off, src, n * sizeof(X));
}

// The host image of a batch's device block up to its intermediates: the
// request records, item starts, emission-class maps and every packed input.
void stage_inputs(const Plan &pl, int nreq, const radg_request *reqs, char *h) {
  std::memset(h, 0, pl.o_inter);
  std::vector<DevReq> dr(nreq);
  for (int r = 0; r < nreq; r++) dr[r] = pl.packs[r].d;
  put(h, pl.o_reqs, dr.data(), nreq);
  put(h, pl.o_sA, pl.sA.data(), nreq + 1);
  put(h, pl.o_sB, pl.sB.data(), nreq + 1);
  put(h, pl.o_sC, pl.sC.data(), nreq + 1);
  {
    std::vector<int> cmap;
    std::vector<long long> cstart, cstartA, cstartW;
    for (const auto &mem : pl.cls_members) {
      long long acc = 0, accA = 0, accW = 0;
      cstart.push_back(0);
      cstartA.push_back(0);
      cstartW.push_back(0);
      for (int r : mem) {
        cmap.push_back(r);
        acc += (long long)reqs[r].nwave * reqs[r].npath;
        accA += (long long)reqs[r].nwave * reqs[r].nlayer;
        accW += reqs[r].nwave;
        cstart.push_back(acc);
        cstartA.push_back(accA);
        cstartW.push_back(accW);
      }
    }
    put(h, pl.o_clsmap, cmap.data(), cmap.size());
    put(h, pl.o_clsstart, cstart.data(), cstart.size());
    put(h, pl.o_clsstartA, cstartA.data(), cstartA.size());
    put(h, pl.o_clsstartW, cstartW.data(), cstartW.size());
  }
  for (int r = 0; r < nreq; r++) {
    const Packed &P = pl.packs[r];
    const DevReq &d = P.d;
    put(h, pl.o_lay + d.lay0 * sizeof(LayerConst), P.lay.data(), P.lay.size());
    put(h, pl.o_amo + d.lg0 * sizeof(double), P.amo.data(), P.amo.size());
    put(h, pl.o_press + d.lay0 * sizeof(double), P.press.data(), P.press.size());
    put(h, pl.o_paths + d.path0 * sizeof(PathInfo), P.paths.data(), P.paths.size());
    put(h, pl.o_layinc + d.pl0 * sizeof(int), P.layinc.data(), P.layinc.size());
    put(h, pl.o_scale + d.pl0 * sizeof(double), P.scale.data(), P.scale.size());
    put(h, pl.o_emt + d.pl0 * sizeof(double), P.emtemp.data(), P.emtemp.size());
    put(h, pl.o_vw + d.wave0 * sizeof(double), P.vwave.data(), P.vwave.size());
    put(h, pl.o_es + d.wave0 * sizeof(double), P.esurf.data(), P.esurf.size());
    put(h, pl.o_n2 + 2 * d.wave0 * sizeof(double), P.n2nir.data(), P.n2nir.size());
    put(h, pl.o_sol + d.wave0 * sizeof(double), P.solar.data(), P.solar.size());
    put(h, pl.o_ppg + d.lg0 * sizeof(double), P.ppg.data(), P.ppg.size());
    if (d.gcw0 >= 0) put(h, pl.o_gcw + d.gcw0 * sizeof(double), P.gcw.data(), P.gcw.size());
    put(h, pl.o_wg + d.wg0 * sizeof(WaveGas), P.wg.data(), P.wg.size());
    put(h, pl.o_pw + d.path0 * sizeof(double), P.pweight.data(), P.pweight.size());
    put(h, pl.o_dx + d.dx0 * sizeof(double), P.dext.data(), P.dext.size());
    put(h, pl.o_ct + d.ct0 * sizeof(double), P.cont.data(), P.cont.size());
  }
}

// Kernel arguments for a batch whose device block starts at m (laid out by
// plan_batch).
void set_kernel_args(KernelArgs &a, char *m, const Plan &pl, int nreq, const TableSet &ts_,
                     const CIA &cia_, const AuxTables &aux, int mode) {
  const TableSet *ts = &ts_;
  const CIA *cia = &cia_;
  a.reqs = reinterpret_cast<const DevReq *>(m + pl.o_reqs);
  a.nreq = nreq;
  a.startA = reinterpret_cast<const long long *>(m + pl.o_sA);
  a.startB = reinterpret_cast<const long long *>(m + pl.o_sB);
  a.startC = reinterpret_cast<const long long *>(m + pl.o_sC);
  a.nA = pl.nA;
  a.nB = pl.nB;
  a.nC = pl.nC;
  a.ng = ts->ng;
  {
    // RADG_REBIN=records keeps every overlap round on the record rebin (the
    // general form the row-count rebin falls back to)
    const char *rb = std::getenv("RADG_REBIN");
    a.rebin = (rb && std::strcmp(rb, "records") == 0) ? 0 : 1;
  }
  a.lbl = ts->lbl ? 1 : 0;
  a.koff = ts->d_koff;
  a.ktab = (mode == RADG_MODE_FP64)   ? (const void *)ts->d_k64
           : (mode == RADG_MODE_BF16) ? (const void *)ts->d_k16
           : (mode == RADG_MODE_F16)  ? (const void *)ts->d_kh
                                      : (const void *)ts->d_k32;
  a.np = ts->np;
  a.nt = ts->nt;
  a.t2 = ts->t2 ? 1 : 0;
  a.G = ts->d_G;
  a.delg64 = ts->d_delg64;
  a.delg32 = ts->d_delg32;
  a.hdr_img = ts->d_hdr[mode == RADG_MODE_FP64 ? 1 : 0];
  a.kcia = cia->d_kcia;
  a.cia_temps = cia->d_temps;
  a.cia_dnu = cia->dnu;
  a.lay = reinterpret_cast<const LayerConst *>(m + pl.o_lay);
  a.amo = reinterpret_cast<const double *>(m + pl.o_amo);
  a.press = reinterpret_cast<const double *>(m + pl.o_press);
  a.paths = reinterpret_cast<const PathInfo *>(m + pl.o_paths);
  a.layinc = reinterpret_cast<const int *>(m + pl.o_layinc);
  a.scale = reinterpret_cast<const double *>(m + pl.o_scale);
  a.emtemp = reinterpret_cast<const double *>(m + pl.o_emt);
  a.vwave = reinterpret_cast<const double *>(m + pl.o_vw);
  a.esurf = reinterpret_cast<const double *>(m + pl.o_es);
  a.n2nir = reinterpret_cast<const double *>(m + pl.o_n2);
  a.solar = reinterpret_cast<const double *>(m + pl.o_sol);
  a.ppg = reinterpret_cast<const double *>(m + pl.o_ppg);
  a.gcw = reinterpret_cast<const double *>(m + pl.o_gcw);
  a.h2o_c2 = aux.gascon ? aux.gascon->h2o_c2 : 0.0;
  a.h2o_ts = aux.gascon ? aux.gascon->h2o_ts : 0.0;
  a.h2o_ps = aux.gascon ? aux.gascon->h2o_ps : 0.0;
  a.wg = reinterpret_cast<const WaveGas *>(m + pl.o_wg);
  a.pweight = reinterpret_cast<const double *>(m + pl.o_pw);
  a.dext = reinterpret_cast<const double *>(m + pl.o_dx);
  a.cont = reinterpret_cast<const double *>(m + pl.o_ct);
  a.e3tab = aux.e3 ? aux.e3->d_tab : nullptr;
  const ParaCIA *para = aux.para.get();
  if (para) {
    a.pkh2h2 = para->d_kh2h2;
    a.pkh2he = para->d_kh2he;
    a.pkcia5 = para->d_kcia5;
    a.ptemps = para->d_temps;
    a.pfp = para->d_fp;
    a.pdnu = para->dnu;
    a.npara = para->npara;
  } else {
    a.pkh2h2 = a.pkh2he = a.pkcia5 = a.ptemps = a.pfp = nullptr;
    a.pdnu = 0.0;
    a.npara = 0;
  }
  a.inter = m + pl.o_inter;
  a.out = m + pl.o_out;
  a.gts = m + pl.o_gts;
  a.jac = m + pl.o_jac;
  a.wout = reinterpret_cast<double *>(m + pl.o_wout);
  a.wgts = reinterpret_cast<double *>(m + pl.o_wgts);
  a.wjac = reinterpret_cast<double *>(m + pl.o_wjac);
  a.err = reinterpret_cast<int *>(m + pl.o_err);
}

// The launches of each layers-per-lane class (plan_batch's classes), with
// the class maps in the device block at m.  Classes run as the continuum,
// layer and emission kernels.  RADG_FUSED=1 in the environment runs each
// class whose workgroup (16 wavefronts) fits the CU's 160 KiB of LDS as the
// fused kernel instead: the same arithmetic with the layer results kept in
// LDS, slower at the 120-layer workload because a CU then holds one
// workgroup and waits at its phase barriers.
std::vector<radg_batch::EmClass> emission_classes(const Plan &pl, const radg_request *reqs, char *m,
                                                  int mode, int ng, bool lbl) {
  const int fp64 = mode == RADG_MODE_FP64;
  const int E = radg_layer_E(ng);
  const char *fz = std::getenv("RADG_FUSED");
  const bool allow_fused = fz && fz[0] == '1';
  const size_t fused_cap = 160 * 1024;
  std::vector<radg_batch::EmClass> em;
  size_t off = 0, soff = 0;
  for (size_t c = 0; c < pl.cls_rl.size(); c++) {
    radg_batch::EmClass ec;
    ec.RL = pl.cls_rl[c];
    ec.nreq = (int)pl.cls_members[c].size();
    ec.nA = ec.nB = ec.nW = 0;
    int maxnsw = 1, maxgas = 1, maxkrow = 1;
    long long maxblock = 0;
    ec.disc = false;
    for (int r : pl.cls_members[c]) {
      ec.nA += (long long)reqs[r].nwave * reqs[r].nlayer;
      ec.nB += (long long)reqs[r].nwave * reqs[r].npath;
      ec.nW += reqs[r].nwave;
      maxnsw = std::max(maxnsw, reqs[r].nsw);
      maxgas = std::max(maxgas, reqs[r].ngas);
      maxkrow = std::max(maxkrow, pl.packs[r].d.nks + 1);
      maxblock = std::max(maxblock, inter_block(ng, reqs[r].nlayer, pl.packs[r].d.nks, pl.packs[r].d.nparam));
      for (int p = 0; p < reqs[r].npath; p++) {
        const int im = reqs[r].imod[p];
        ec.disc = ec.disc || im == 8 || im == 13 || im == 14 || im == 18;
      }
      ec.disc = ec.disc || pl.packs[r].d.solar != 0;
    }
    ec.reqmap = reinterpret_cast<const int *>(m + pl.o_clsmap) + off;
    ec.start = reinterpret_cast<const long long *>(m + pl.o_clsstart) + soff;
    ec.startA = reinterpret_cast<const long long *>(m + pl.o_clsstartA) + soff;
    ec.startW = reinterpret_cast<const long long *>(m + pl.o_clsstartW) + soff;
    ec.lds = radg_lds_emission_bytes(ec.RL, maxkrow, fp64);
    ec.ldsA = radg_lds_layer_bytes(ng, maxgas, E, fp64);
    ec.lwaves = radg_layer_waves(ng, maxgas, E, fp64);
    ec.lwstride = radg_layer_stride(ng, maxgas, E, fp64);
    ec.wsbytes = 0;
    ec.ldsF = radg_lds_fused_bytes(mode, ng, E, ec.RL, maxgas, maxnsw, maxblock, &ec.wsbytes);
    ec.fused = allow_fused && !lbl && ec.ldsF > 0 && ec.ldsF <= fused_cap;
    off += ec.nreq;
    soff += ec.nreq + 1;
    em.push_back(ec);
  }
  return em;
}

// The kernels of a batch on stream st: the layer pass of every class (the
// fused kernel, or the continuum and layer kernels), event ev1, the
// emission kernel of the classes not fused, event ev2, the path quadrature.
int launch_classes(const KernelArgs &base, const std::vector<radg_batch::EmClass> &em, int mode,
                   int E, hipStream_t st, hipEvent_t ev1, hipEvent_t ev2) {
  const int fp64 = mode == RADG_MODE_FP64;
  for (const auto &ec : em) {
    KernelArgs a = base;
    a.reqmapA = a.reqmapB = ec.reqmap;
    a.nreqA = a.nreqB = ec.nreq;
    if (ec.fused) {
      a.startB = ec.startW;
      a.nB = ec.nW;
      a.wsbytes = ec.wsbytes;
      HIPCHK(radg_launch_fused(a, mode, E, ec.RL, ec.ldsF, st));
    } else {
      if (ec.ldsA > 160 * 1024 || ec.lds > 160 * 1024)
        return fail(RADG_ELIMITS, "LDS need %zu/%zu bytes exceeds 160 KiB", ec.ldsA, ec.lds);
      a.startA = ec.startA;
      a.nA = ec.nA;
      a.lwaves = ec.lwaves;
      a.lwstride = ec.lwstride;
      HIPCHK(radg_launch_layer_lds(a, mode, E, ec.ldsA, st));
    }
  }
  if (ev1) HIPCHK(hipEventRecord(ev1, st));
  for (const auto &ec : em) {
    if (ec.fused) continue;
    KernelArgs a = base;
    a.startB = ec.start;
    a.nB = ec.nB;
    a.reqmapB = ec.reqmap;
    a.nreqB = ec.nreq;
    HIPCHK(radg_launch_emission(a, fp64, ec.RL, ec.disc, ec.lds, st));
  }
  if (ev2) HIPCHK(hipEventRecord(ev2, st));
  HIPCHK(radg_launch_quadrature(base, fp64, st));
  return RADG_OK;
}

}  // namespace

extern "C" {

int radg_set_mode(int mode) {
  if (!valid_mode(mode)) return fail(RADG_EINPUT, "mode=%d", mode);
  g_mode = mode;
  // the table form the mode reads, for the tables registered now
  std::shared_ptr<TableSet> ts, lb;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ts = g_tables;
    lb = g_lbl;
  }
  int rc = ts ? ensure_form(*ts, mode) : RADG_OK;
  if (rc == RADG_OK && lb) rc = ensure_form(*lb, mode);
  return rc;
}
int radg_get_mode(void) { return g_mode; }
const char *radg_last_error(void) { return g_err.c_str(); }

// The layer kernel's LDS header (header_layout): G[0..ng] then +inf to
// ng + 5 (rankg.f:59-65 edges; the +inf pad lets searches read past ng), the
// g weights in FP64 and in T, the FP64 weight delg_i delg_j of the element in
// each sort slot (0 for padding) and the unforced cumulative weights summed
// in order (rankg.f:59-62).  Every workgroup copies it into LDS.
static std::vector<char> layer_header_image(int ng, int E, int ts, const double *G,
                                            const double *delg) {
  const HeaderLayout H = header_layout(ng, E, ts);
  std::vector<char> img((size_t)H.total, 0);
  double *g = reinterpret_cast<double *>(img.data() + H.G);
  double *dl64 = reinterpret_cast<double *>(img.data() + H.dl64);
  double *w64 = reinterpret_cast<double *>(img.data() + H.w64);
  double *gw = reinterpret_cast<double *>(img.data() + H.Gw);
  for (int i = 0; i < ng + 6; i++) g[i] = (i <= ng) ? G[i] : HUGE_VAL;
  for (int i = 0; i < ng; i++) {
    dl64[i] = delg[i];
    if (ts == 8) reinterpret_cast<double *>(img.data() + H.dlg)[i] = delg[i];
    else reinterpret_cast<float *>(img.data() + H.dlg)[i] = (float)delg[i];
  }
  double s = 0.0;
  gw[0] = 0.0;
  for (int i = 0; i < ng; i++) {
    s += delg[i];
    gw[i + 1] = s;
  }
  const int lpr = layer_lpr(ng, E);
  for (int sl = 0; sl < RADG_WAVE * E; sl++) {
    const int l = sl / E, i = l / lpr, jx = (l - i * lpr) * E + (sl & (E - 1));
    w64[sl] = (i < ng && jx < ng) ? delg[i] * delg[jx] : 0.0;
  }
  return img;
}

// The k-table registry (lbl false) or the line-by-line one (lbl true: one
// g-ordinate of weight 1, every table on one regular wavenumber grid).
static int register_tables(bool lbl, int ntab, const int *idgas, const int *isogas, const int *npoint,
                           const double *vmin, const double *delv, int np, int nt_signed, int ng,
                           const double *press_atm, const double *temp, const double *g_ord,
                           const double *delg, const double *const *k,
                           const double *const *vcen) {
  if (ntab < 1) return fail(RADG_EINPUT, "ntab=%d", ntab);
  const bool t2 = nt_signed < 0;
  const int nt = t2 ? -nt_signed : nt_signed;
  if (np < 2 || np > RADG_MAXK || nt < 2 || nt > RADG_MAXK)
    return fail(RADG_ELIMITS, "np=%d nt=%d outside 2..%d", np, nt_signed, RADG_MAXK);
  if (ng < 1 || ng > RADG_GPU_MAXG)
    return fail(RADG_ELIMITS, "ng=%d outside 1..%d (the overlap sort keeps NG*NG keys in one wavefront)",
                ng, RADG_GPU_MAXG);
  auto ts = std::make_shared<TableSet>();
  ts->np = np;
  ts->nt = nt;
  ts->t2 = t2;
  ts->ng = ng;
  ts->lbl = lbl;
  for (int i = 0; i < np; i++) {
    if (!finite_pos(press_atm[i])) return fail(RADG_EINPUT, "table pressure %d = %g", i + 1, press_atm[i]);
    ts->lnp.push_back(std::log(press_atm[i]));  // read_khead.f:80-85
  }
  ts->tk.assign(temp, temp + (t2 ? (size_t)np * nt : (size_t)nt));
  for (int r = 0; r < (t2 ? np : 1); r++)
    for (int j = 1; j < nt; j++)
      if (!(ts->tk[(size_t)r * nt + j] > ts->tk[(size_t)r * nt + j - 1]))
        return fail(RADG_EINPUT, "table temperatures must increase (row %d)", r + 1);
  ts->g_ord.assign(g_ord, g_ord + ng);
  ts->delg.assign(delg, delg + ng);
  double wmin = 1e300, wmax = 0;
  for (int i = 0; i < ng; i++) {
    if (!(delg[i] > 0)) return fail(RADG_EINPUT, "delg(%d)=%g must be positive", i + 1, delg[i]);
    wmin = std::min(wmin, delg[i]);
    wmax = std::max(wmax, delg[i]);
  }
  // Every mixed weight delg_i*delg_j must be narrower than every g-bin, so
  // that one element never spans two bin edges; rankg.f advances one bin
  // per element and the parallel rebin relies on that never mattering.
  if (ng > 1 && wmax * wmax >= wmin)
    return fail(RADG_EUNSUPPORTED, "g-weights with max(delg)^2 >= min(delg) are not supported");
  ts->G.assign(ng + 1, 0.0);
  for (int i = 1; i <= ng; i++) ts->G[i] = ts->G[i - 1] + delg[i - 1];
  if (ts->G[ng] < 1.0) ts->G[ng] = 1.0;  // rankg.f:65
  const long long per = (long long)np * nt * ng;
  long long total = 0;
  // read_klist.f:117-160 stops when the tables differ in DELV, so regular and
  // irregular grids never share a list
  for (int t = 1; t < ntab; t++)
    if ((delv[t] > 0) != (delv[0] > 0))
      return fail(RADG_EINPUT, "tables 1 and %d mix regular (DELV > 0) and irregular grids", t + 1);
  for (int t = 0; t < ntab; t++) {
    if (npoint[t] < 1 || !std::isfinite(delv[t])) return fail(RADG_EINPUT, "table %d: npoint=%d delv=%g", t + 1, npoint[t], delv[t]);
    Table tb;
    tb.idgas = idgas[t];
    tb.isogas = isogas[t];
    tb.npoint = npoint[t];
    tb.vmin = vmin[t];
    tb.delv = delv[t];
    if (!(delv[t] > 0)) {  // read_khead.f:93-100: the wavenumbers follow the header
      if (!vcen || !vcen[t]) return fail(RADG_EINPUT, "table %d: DELV=%g needs its wavenumbers", t + 1, delv[t]);
      tb.vcen.assign(vcen[t], vcen[t] + npoint[t]);
      for (int j = 0; j < npoint[t]; j++)
        if (!finite_pos(tb.vcen[j]) || (j > 0 && !(tb.vcen[j] > tb.vcen[j - 1])))
          return fail(RADG_EINPUT, "table %d: wavenumbers must be positive and increasing", t + 1);
    }
    tb.off = total;
    tb.zero.assign(npoint[t], 1);
    for (long long s = 0; s < npoint[t]; s++) {
      const double *sl = k[t] + s * per;
      for (long long i = 0; i < per; i++) {
        if (!std::isfinite(sl[i]) || std::fabs(sl[i]) > 1e30)
          return fail(RADG_EINPUT, "table %d: value %g at wavenumber %lld is not finite or > 1e30",
                      t + 1, sl[i], s + 1);
        if (sl[i] > 0.0) tb.zero[s] = 0;
      }
    }
    total += (long long)npoint[t] * per;
    ts->tabs.push_back(std::move(tb));
  }
  ts->total = total;
  HIPCHK(hipMalloc(&ts->d_k64, total * sizeof(double)));
  for (int t = 0; t < ntab; t++) {
    const long long n = (long long)npoint[t] * per;
    HIPCHK(hipMemcpy(ts->d_k64 + ts->tabs[t].off, k[t], n * sizeof(double), hipMemcpyHostToDevice));
  }
  std::vector<float> dg32(delg, delg + ng);
  HIPCHK(hipMalloc(&ts->d_G, (ng + 1) * sizeof(double)));
  HIPCHK(hipMalloc(&ts->d_delg64, ng * sizeof(double)));
  HIPCHK(hipMalloc(&ts->d_delg32, ng * sizeof(float)));
  HIPCHK(hipMemcpy(ts->d_G, ts->G.data(), (ng + 1) * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ts->d_delg64, delg, ng * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ts->d_delg32, dg32.data(), ng * sizeof(float), hipMemcpyHostToDevice));
  ts->hdr_E = radg_layer_E(ng);
  for (int f = 0; f < 2; f++) {
    const std::vector<char> img = layer_header_image(ng, ts->hdr_E, f ? 8 : 4, ts->G.data(), delg);
    HIPCHK(hipMalloc(&ts->d_hdr[f], img.size()));
    HIPCHK(hipMemcpy(ts->d_hdr[f], img.data(), img.size(), hipMemcpyHostToDevice));
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (lbl) g_lbl = ts;
  else g_tables = ts;
  return RADG_OK;
}

int radg_register_ktables(int ntab, const int *idgas, const int *isogas, const int *npoint,
                          const double *vmin, const double *delv, int np, int nt_signed, int ng,
                          const double *press_atm, const double *temp, const double *g_ord,
                          const double *delg, const double *const *k,
                          const double *const *vcen) {
  return register_tables(false, ntab, idgas, isogas, npoint, vmin, delv, np, nt_signed, ng, press_atm,
                         temp, g_ord, delg, k, vcen);
}

// /INTERPKLBL/ (read_klbllist.f:30-198): every table on the wavenumbers
// vmin + i delv, i < npoint, and one P/T grid; one g-ordinate of weight 1
// (cirsradg_wave.f:336-339).
int radg_register_lbltables(int ntab, const int *idgas, const int *isogas, int npoint, double vmin,
                            double delv, int np, int nt, const double *press_atm, const double *temp,
                            const double *const *k) {
  if (ntab < 1) return fail(RADG_EINPUT, "ntab=%d", ntab);
  if (npoint < 1 || !std::isfinite(vmin) || !std::isfinite(delv) || !(delv > 0))
    return fail(RADG_EINPUT, "LBL grid npoint=%d vmin=%g delv=%g (get_klblg.f:182-185 needs delv > 0)", npoint,
                vmin, delv);
  const std::vector<int> npts(ntab, npoint);
  const std::vector<double> vmins(ntab, vmin), delvs(ntab, delv);
  const double one = 1.0;
  return register_tables(true, ntab, idgas, isogas, npts.data(), vmins.data(), delvs.data(), np, nt, 1,
                         press_atm, temp, &one, &one, k, nullptr);
}

// /LBLTABLE/ ILBL: 0 the k-tables, 2 the line-by-line tables; ILBL = 1
// reaches noverlapg1 with the k-tables' NG, where noverlapg1.f:64-67 stops.
int radg_set_ilbl(int ilbl) {
  if (ilbl == 1)
    return fail(RADG_EUNSUPPORTED, "ILBL=1: noverlapg1.f:64-67 stops unless NG = 1; ILBL=2 is the LBL form");
  if (ilbl != 0 && ilbl != 2) return fail(RADG_EINPUT, "ilbl=%d", ilbl);
  std::shared_ptr<TableSet> ts;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_ilbl = ilbl;
    ts = g_lbl;
  }
  return (ilbl == 2 && ts) ? ensure_form(*ts, g_mode) : RADG_OK;
}
int radg_get_ilbl(void) { return g_ilbl; }

int radg_ktable_codes(int mode, int itab, unsigned short *codes, float *offsets) {
  if (mode != RADG_MODE_BF16 && mode != RADG_MODE_F16) return fail(RADG_EINPUT, "mode=%d is not a 16-bit form", mode);
  std::shared_ptr<TableSet> ts;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ts = g_tables;
  }
  if (!ts) return fail(RADG_ENOTABLES, "no k-tables registered (radg_register_ktables)");
  if (itab < 0 || itab >= (int)ts->tabs.size()) return fail(RADG_EINPUT, "itab=%d", itab);
  int rc = ensure_form(*ts, mode);
  if (rc) return rc;
  const Table &tb = ts->tabs[itab];
  const long long ntab = (long long)ts->np * ts->nt * ts->ng;
  const long long n = (long long)tb.npoint * ntab;
  const unsigned short *src = mode == RADG_MODE_BF16 ? ts->d_k16 : ts->d_kh;
  HIPCHK(hipMemcpy(codes, src + tb.off, n * sizeof(unsigned short), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(offsets, ts->d_koff + tb.off / ntab * ts->ng, (size_t)tb.npoint * ts->ng * sizeof(float),
                   hipMemcpyDeviceToHost));
  return RADG_OK;
}

int radg_register_cia(double dnu, const double *temps, const double *kcia) {
  if (!(dnu > 0)) return fail(RADG_EINPUT, "dnu=%g", dnu);
  for (int i = 1; i < 25; i++)
    if (!(temps[i] > temps[i - 1])) return fail(RADG_EINPUT, "CIA temperatures must increase");
  auto c = std::make_shared<CIA>();
  c->dnu = dnu;
  const size_t n = (size_t)1501 * 25 * 9;
  HIPCHK(hipMalloc(&c->d_kcia, n * sizeof(double)));
  HIPCHK(hipMalloc(&c->d_temps, 25 * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_kcia, kcia, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_temps, temps, 25 * sizeof(double), hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> lk(g_mu);
  g_cia = c;
  return RADG_OK;
}

int radg_read_cia(const char *path, double dnu) {
  // ciaread.f:77-89: two unformatted sequential records, the 25 REAL*8
  // temperatures, then kcia(9,25,1501) in REAL (4 bytes) or REAL*8
  if (!path) return fail(RADG_EINPUT, "null path");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(RADG_EIO, "cannot open %s", path);
  std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<std::pair<size_t, size_t>> rec;  // (offset, length)
  for (size_t o = 0; o < raw.size();) {
    int32_t n = 0, m = -1;
    if (o + 4 > raw.size()) return fail(RADG_EIO, "%s: truncated record", path);
    std::memcpy(&n, raw.data() + o, 4);
    if (n < 0 || o + 8 + (size_t)n > raw.size()) return fail(RADG_EIO, "%s: bad record length", path);
    std::memcpy(&m, raw.data() + o + 4 + n, 4);
    if (m != n) return fail(RADG_EIO, "%s: record markers differ", path);
    rec.emplace_back(o + 4, (size_t)n);
    o += 8 + (size_t)n;
  }
  const size_t nk = (size_t)1501 * 25 * 9;
  if (rec.size() < 2 || rec[0].second != 25 * 8 || (rec[1].second != 4 * nk && rec[1].second != 8 * nk))
    return fail(RADG_EIO, "%s: not a CIA table (25 REAL*8 temperatures, then kcia(9,25,1501))", path);
  double temps[25];
  std::memcpy(temps, raw.data() + rec[0].first, sizeof temps);
  std::vector<double> k(nk);
  const char *src = raw.data() + rec[1].first;
  for (size_t i = 0; i < nk; i++) {
    if (rec[1].second == 4 * nk) {
      float v;
      std::memcpy(&v, src + 4 * i, 4);
      k[i] = v;
    } else {
      std::memcpy(&k[i], src + 8 * i, 8);
    }
  }
  return radg_register_cia(dnu, temps, k.data());
}

int radg_register_parah2(double dnu, const double *temps, int npara, const double *fp,
                         const double *kh2h2, const double *kh2he, const double *kcia5) {
  if (npara == 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_para.reset();
    return RADG_OK;
  }
  if (npara < 3 || npara > 64) return fail(RADG_ELIMITS, "npara=%d outside 3..64", npara);
  if (!(dnu > 0) || !temps || !fp || !kh2h2 || !kh2he || !kcia5)
    return fail(RADG_EINPUT, "para-H2 table: dnu=%g or a null array", dnu);
  for (int i = 1; i < 25; i++)
    if (!(temps[i] > temps[i - 1])) return fail(RADG_EINPUT, "para-H2 temperatures must increase");
  if (!(fp[1] > fp[0])) return fail(RADG_EINPUT, "para-H2 fractions must increase");
  auto c = std::make_shared<ParaCIA>();
  c->dnu = dnu;
  c->npara = npara;
  const size_t nk = (size_t)1501 * 25 * npara, n5 = (size_t)1501 * 25 * 5;
  HIPCHK(hipMalloc(&c->d_kh2h2, nk * sizeof(double)));
  HIPCHK(hipMalloc(&c->d_kh2he, nk * sizeof(double)));
  HIPCHK(hipMalloc(&c->d_kcia5, n5 * sizeof(double)));
  HIPCHK(hipMalloc(&c->d_temps, 25 * sizeof(double)));
  HIPCHK(hipMalloc(&c->d_fp, npara * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_kh2h2, kh2h2, nk * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_kh2he, kh2he, nk * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_kcia5, kcia5, n5 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_temps, temps, 25 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_fp, fp, npara * sizeof(double), hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> lk(g_mu);
  g_para = c;
  return RADG_OK;
}

int radg_register_n2nir(int n1, const double *wn1, const double *k1, int n2, const double *wn2,
                        const double *k2) {
  if (n1 == 0 && n2 == 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_n2nir.reset();
    return RADG_OK;
  }
  if (n1 < 2 || n2 < 2 || n1 > 100000 || n2 > 100000)
    return fail(RADG_ELIMITS, "N2 near-infrared CIA: n1=%d n2=%d, each needs 2..100000 knots", n1, n2);
  if (!wn1 || !k1 || !wn2 || !k2) return fail(RADG_EINPUT, "N2 near-infrared CIA: null array");
  for (const auto &t : {std::make_pair(wn1, n1), std::make_pair(wn2, n2)})
    for (int i = 1; i < t.second; i++)
      if (!(t.first[i] > t.first[i - 1]))
        return fail(RADG_EINPUT, "N2 near-infrared CIA: knots must increase");
  for (int i = 0; i < n1; i++)
    if (!std::isfinite(k1[i])) return fail(RADG_EINPUT, "N2-N2 value %d not finite", i + 1);
  for (int i = 0; i < n2; i++)
    if (!std::isfinite(k2[i])) return fail(RADG_EINPUT, "N2-H2 value %d not finite", i + 1);
  auto t = std::make_shared<N2Nir>();
  t->wn1.assign(wn1, wn1 + n1);
  t->k1.assign(k1, k1 + n1);
  t->wn2.assign(wn2, wn2 + n2);
  t->k2.assign(k2, k2 + n2);
  std::lock_guard<std::mutex> lk(g_mu);
  g_n2nir = t;
  return RADG_OK;
}

int radg_register_e3(int n, const double *trans, const double *e3) {
  if (n == 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_e3.reset();
    return RADG_OK;
  }
  if (n != RADG_E3N) return fail(RADG_ELIMITS, "E3 table: n=%d, e3interp.f holds %d points", n, RADG_E3N);
  if (!trans || !e3) return fail(RADG_EINPUT, "E3 table: null array");
  for (int i = 0; i < n; i++)
    if (!std::isfinite(trans[i]) || !std::isfinite(e3[i]))
      return fail(RADG_EINPUT, "E3 table: entry %d not finite", i + 1);
  // e3interp.f:25-51: the table's spacing is 1/(mtab-1) and the gradient its
  // one-sided differences at the ends, central differences inside
  std::vector<double> h((size_t)3 * n);
  const double dtrans = 1.0 / (double)(n - 1);
  for (int i = 0; i < n; i++) {
    h[i] = trans[i];
    h[n + i] = e3[i];
  }
  double *g = h.data() + 2 * n;
  g[0] = (e3[1] - e3[0]) / dtrans;
  g[n - 1] = (e3[n - 1] - e3[n - 2]) / dtrans;
  for (int i = 1; i < n - 1; i++) g[i] = 0.5 * (e3[i + 1] - e3[i - 1]) / dtrans;
  auto c = std::make_shared<E3Tab>();
  HIPCHK(hipMalloc(&c->d_tab, h.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->d_tab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> lk(g_mu);
  g_e3 = c;
  return RADG_OK;
}

int radg_register_solar(int npt, const double *wave, const double *rad) {
  if (npt == 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    // /SOLARDAT/ must stay consistent: iread = 999 names this spectrum
    // (get_solar_wave.f:14-17 stops without one)
    if (g_iread == 999)
      return fail(RADG_EINPUT, "iread=999 is set: clear it (radg_set_iform) before the spectrum");
    g_solar.reset();
    return RADG_OK;
  }
  if (npt < 2 || npt > 1000000) return fail(RADG_ELIMITS, "stellar spectrum: npt=%d outside 2..1000000", npt);
  if (!wave || !rad) return fail(RADG_EINPUT, "stellar spectrum: null array");
  for (int i = 0; i < npt; i++)
    if (!std::isfinite(wave[i]) || !std::isfinite(rad[i]))
      return fail(RADG_EINPUT, "stellar spectrum: entry %d not finite", i + 1);
  for (int i = 1; i < npt; i++)
    if ((wave[i] - wave[i - 1]) * (wave[npt - 1] - wave[0]) <= 0)
      return fail(RADG_EINPUT, "stellar spectrum: wavelengths must be monotonic");
  auto s = std::make_shared<Solar>();
  s->wave.assign(wave, wave + npt);
  s->rad.assign(rad, rad + npt);
  std::lock_guard<std::mutex> lk(g_mu);
  g_solar = s;
  return RADG_OK;
}

int radg_register_gascon(const radg_gascon *g) {
  if (!g) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_gascon.reset();
    return RADG_OK;
  }
  auto c = std::make_shared<GasCon>();
  c->ih2o = g->ih2o;
  c->ich4 = g->ich4;
  c->io3 = g->io3;
  c->inh3 = g->inh3;
  c->iuvscat = g->iuvscat;
  auto table = [](int n, const double *x, const double *y, const char *what, std::vector<double> &ox,
                  std::vector<double> &oy) {
    if (n == 0) return RADG_OK;
    if (n < 2 || n > 1000000 || !x || !y) return fail(RADG_EINPUT, "%s table: n=%d", what, n);
    // lutzowen.dat repeats a wavelength: non-decreasing is accepted, and
    // the interval locate.f brackets is never a repeated pair
    for (int i = 0; i < n; i++)
      if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || (i > 0 && x[i] < x[i - 1]))
        return fail(RADG_EINPUT, "%s table: entry %d not finite or decreasing", what, i + 1);
    if (!(x[n - 1] > x[0])) return fail(RADG_EINPUT, "%s table: no wavelength range", what);
    ox.assign(x, x + n);
    oy.assign(y, y + n);
    return RADG_OK;
  };
  int rc = table(g->nch4, g->ch4_wl, g->ch4_k, "CH4", c->ch4_wl, c->ch4_k);
  if (!rc) rc = table(g->nnh3, g->nh3_wl, g->nh3_k, "NH3", c->nh3_wl, c->nh3_k);
  if (rc) return rc;
  if (g->nh2o != 0) {
    const int n = g->nh2o;
    if (n < 2 || n > 1000000 || !g->h2o_s296 || !g->h2o_s260 || !g->h2o_f296)
      return fail(RADG_EINPUT, "CKD H2O tables: n=%d", n);
    if (!(g->h2o_dv > 0) || !std::isfinite(g->h2o_v0) || !std::isfinite(g->h2o_c2) ||
        !(g->h2o_ts > 0) || !(g->h2o_ps > 0))
      return fail(RADG_EINPUT, "CKD H2O tables: grid or constants invalid");
    for (int i = 0; i < n; i++)
      if (!std::isfinite(g->h2o_s296[i]) || !std::isfinite(g->h2o_s260[i]) || !std::isfinite(g->h2o_f296[i]))
        return fail(RADG_EINPUT, "CKD H2O tables: entry %d not finite", i + 1);
    c->h2o_v0 = g->h2o_v0;
    c->h2o_dv = g->h2o_dv;
    c->h2o_c2 = g->h2o_c2;
    c->h2o_ts = g->h2o_ts;
    c->h2o_ps = g->h2o_ps;
    c->s296.assign(g->h2o_s296, g->h2o_s296 + n);
    c->s260.assign(g->h2o_s260, g->h2o_s260 + n);
    c->f296.assign(g->h2o_f296, g->h2o_f296 + n);
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_gascon = c;
  return RADG_OK;
}

int radg_set_iform(int iform, int iread) {
  std::lock_guard<std::mutex> lk(g_mu);
  // iread = 999 reads the registered stellar spectrum (radg_register_solar);
  // another iread leaves a registered spectrum unused by these entry points
  if (iread == 999 && !g_solar)
    return fail(RADG_EINPUT, "iread=999 but no stellar spectrum is registered");
  g_iform = iform;
  g_iread = iread;
  return RADG_OK;
}

int radg_set_nlte(int flag, double k1, double a) {
  if (flag == 1 && !(k1 >= 1.0e-20 && k1 <= 1.0e-3))
    return fail(RADG_EINPUT, "inlte=1 needs 1e-20 <= k1 <= 1e-3 (k1=%g)", k1);
  std::lock_guard<std::mutex> lk(g_mu);
  g_inlte = flag;
  g_nlte_k1 = k1;
  g_nlte_a = a;
  return RADG_OK;
}

int radg_register_dust(int nsec, int ncont, const double *vsec, const double *xsec) {
  if (ncont == 0) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_dust.reset();
    return RADG_OK;
  }
  if (ncont < 0 || ncont > RADG_MAXCON || nsec < 1 || nsec > RADG_MAXSEC)
    return fail(RADG_ELIMITS, "dust: nsec=%d ncont=%d outside 1..%d, 0..%d", nsec, ncont, RADG_MAXSEC,
                RADG_MAXCON);
  if (!vsec || !xsec) return fail(RADG_EINPUT, "dust: null vsec or xsec");
  auto d = std::make_shared<Dust>();
  d->nsec = nsec;
  d->ncont = ncont;
  d->vsec.assign(vsec, vsec + nsec);
  d->xsec.assign(xsec, xsec + (size_t)ncont * nsec);
  std::lock_guard<std::mutex> lk(g_mu);
  g_dust = d;
  return RADG_OK;
}

int radg_batch_bytes(int nreq, const radg_request *reqs, int mode, size_t *bytes) {
  std::shared_ptr<TableSet> ts;
  AuxTables aux;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ts = active_tables();
    aux = {g_para, g_e3, g_n2nir, g_solar, g_gascon};
  }
  if (!ts) return no_tables();
  Plan pl;
  int rc = plan_batch(nreq, reqs, mode, *ts, aux, pl);
  if (rc) return rc;
  *bytes = pl.total;
  return RADG_OK;
}

namespace {
// Batches see the registered stellar spectrum as iread = 999; the
// fixed-argument entry points see it only when /SOLARDAT/'s iread is 999
// (radg_set_iform), as cirsradg_wave.f:921,866 test iread.
int batch_create(int nreq, const radg_request *reqs, int mode, size_t budget, radg_batch **out,
                 bool fixed_args);
}  // namespace

int radg_batch_create(int nreq, const radg_request *reqs, int mode, size_t budget, radg_batch **out) {
  return batch_create(nreq, reqs, mode, budget, out, false);
}

namespace {
// The layout of a plan: its offsets and item counts.  Two plans with the
// same layout fit the same device block.
extern "C++" std::vector<long long> plan_sig(const Plan &pl) {
  std::vector<long long> v = {(long long)pl.total, pl.nA, pl.nB, pl.nC, (long long)pl.o_reqs,
                              (long long)pl.o_inter, (long long)pl.o_out, (long long)pl.o_jac,
                              (long long)pl.o_wout, (long long)pl.o_wjac, (long long)pl.o_err,
                              (long long)pl.o_clsstart, (long long)pl.cls_rl.size()};
  for (int rl : pl.cls_rl) v.push_back(rl);
  for (int ng : pl.cls_ngas) v.push_back(ng);
  for (long long x : pl.sA) v.push_back(x);
  for (long long x : pl.sB) v.push_back(x);
  return v;
}

// Everything of a batch but its device block and pinned staging: the staged
// inputs (uploaded by the next launch), the host copies, the kernel
// arguments and the launch classes.  Events already made are kept.
int batch_fill(radg_batch *b, const Plan &pl, int nreq, const radg_request *reqs, int mode,
               const std::shared_ptr<TableSet> &ts, const std::shared_ptr<CIA> &cia, const AuxTables &aux) {
  b->mode = mode;
  b->nreq = nreq;
  b->ts = ts;
  b->cia = cia;
  b->aux = aux;
  b->bytes = pl.total;
  b->o_out = pl.o_out;
  b->o_gts = pl.o_gts;
  b->o_jac = pl.o_jac;
  b->o_wout = pl.o_wout;
  b->o_wgts = pl.o_wgts;
  b->o_wjac = pl.o_wjac;
  const size_t tsz = (mode == RADG_MODE_FP64) ? 8 : 4;
  b->span_p0 = pl.o_out;
  b->span_p1 = pl.o_jac + (size_t)pl.njac * tsz;
  b->span_w0 = pl.o_wout;
  b->span_w1 = pl.o_wjac + (size_t)pl.nwjac * 8;
  b->hup_bytes = pl.o_inter;
  b->up_pending = true;
  stage_inputs(pl, nreq, reqs, b->hup);
  for (int i = 0; i < 3; i++) b->alg_bytes[i] = 0.0;
  for (int r = 0; r < nreq; r++)
    for (int i = 0; i < 3; i++) b->alg_bytes[i] += pl.packs[r].alg_bytes[i];
  b->hreq.resize(nreq);
  for (int r = 0; r < nreq; r++) {
    const Packed &P = pl.packs[r];
    radg_batch::HostReq &hq = b->hreq[r];
    hq.L = reqs[r].nlayer;
    hq.npath = reqs[r].npath;
    hq.ngas = reqs[r].ngas;
    hq.nwave = reqs[r].nwave;
    hq.ncont = reqs[r].ncont;
    hq.h2p = P.d.flagh2p;
    hq.layinc = P.layinc;
    hq.nlays.resize(hq.npath);
    for (int p = 0; p < hq.npath; p++) hq.nlays[p] = P.paths[p].nlays;
  }
  b->reqs.resize(nreq);
  for (int r = 0; r < nreq; r++) b->reqs[r] = pl.packs[r].d;
  set_kernel_args(b->args, b->dmem, pl, nreq, *ts, *cia, aux, mode);
  b->E = radg_layer_E(ts->ng);
  b->maxngas = pl.maxngas;
  b->maxnsw = pl.maxnsw;
  b->RL = radg_emission_RL(pl.maxlays);
  b->em = emission_classes(pl, reqs, b->dmem, mode, ts->ng, ts->lbl);
  for (const auto &ec : b->em)
    if (!ec.fused && (ec.ldsA > 160 * 1024 || ec.lds > 160 * 1024))
      return fail(RADG_ELIMITS, "LDS need %zu/%zu bytes exceeds 160 KiB", ec.ldsA, ec.lds);
  // events are created on first use of each ring slot
  if (b->evs.empty()) {
    b->evs.assign(4 * radg_batch::RING, nullptr);
    b->ev = &b->evs[0];
  }
  b->sig = plan_sig(pl);
  return RADG_OK;
}

int batch_create(int nreq, const radg_request *reqs, int mode, size_t budget, radg_batch **out,
                 bool fixed_args) {
  *out = nullptr;
  if (!valid_mode(mode)) return fail(RADG_EINPUT, "mode=%d", mode);
  std::shared_ptr<TableSet> ts;
  std::shared_ptr<CIA> cia;
  AuxTables aux;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ts = active_tables();
    cia = g_cia;
    aux = {g_para, g_e3, g_n2nir, g_solar, g_gascon};
    if (fixed_args && g_iread != 999) aux.solar.reset();
  }
  if (!ts) return no_tables();
  if (!cia) return fail(RADG_ENOTABLES, "no CIA table registered (radg_register_cia)");
  Plan pl;
  int rc = plan_batch(nreq, reqs, mode, *ts, aux, pl);
  if (rc) return rc;
  rc = ensure_form(*ts, mode);
  if (rc) return rc;
  if (budget && pl.total > budget)
    return fail(RADG_EBUDGET, "batch needs %zu bytes, budget %zu", pl.total, budget);
  std::unique_ptr<radg_batch> b(new radg_batch);
  b->dmem = static_cast<char *>(dev_cache().get(pl.total));
  if (!b->dmem) return fail(RADG_EHIP, "device allocation of %zu bytes failed", pl.total);
  // Host staging of everything but the intermediates and outputs, in pinned
  // memory.  Nothing here touches the device: the staged inputs go up with
  // the first launch, on the launch's stream, so that packing a batch never
  // waits for the kernels of batches already running.
  b->hup = static_cast<char *>(host_cache().get(pl.o_inter));
  if (!b->hup) return fail(RADG_EHIP, "pinned allocation of %zu bytes failed", pl.o_inter);
  rc = batch_fill(b.get(), pl, nreq, reqs, mode, ts, cia, aux);
  if (rc) return rc;
  *out = b.release();
  return RADG_OK;
}

// Re-stage a batch for new requests of the same layout, in the tables,
// CIA and mode it was made for, after its last launch has finished: 0 done,
// 1 the requests need another layout (the batch is left as it was), else an
// error status.
int batch_restage(radg_batch *b, int nreq, const radg_request *reqs, int mode, bool fixed_args) {
  std::shared_ptr<TableSet> ts;
  std::shared_ptr<CIA> cia;
  AuxTables aux;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ts = active_tables();
    cia = g_cia;
    aux = {g_para, g_e3, g_n2nir, g_solar, g_gascon};
    if (fixed_args && g_iread != 999) aux.solar.reset();
  }
  if (!ts || !cia || ts != b->ts || cia != b->cia || mode != b->mode || b->smem) return 1;
  if (aux.para != b->aux.para || aux.e3 != b->aux.e3 || aux.n2nir != b->aux.n2nir ||
      aux.solar != b->aux.solar || aux.gascon != b->aux.gascon)
    return 1;
  Plan pl;
  int rc = plan_batch(nreq, reqs, mode, *ts, aux, pl);
  if (rc) return rc;
  if (plan_sig(pl) != b->sig) return 1;
  if (b->launches > 0) HIPCHK(hipEventSynchronize(b->ev[3]));
  return batch_fill(b, pl, nreq, reqs, mode, ts, cia, aux);
}

}  // namespace

// ------------------------------------------ the entry point that allocates
// nothing: every buffer beyond the caller's inputs and outputs lies in this
// file-scope array, one copy per GPU, held until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress returns on
// each device (noalloc_scratch); the array's name is a host-side stand-in.
static __device__ __align__(256) char scratch[RADG_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_dev[kMaxDevices] = {};

// The current device's copy of scratch, looked up on the first call there.
int noalloc_scratch(char **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(RADG_ELIMITS, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(RADG_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<char *>(p);
  }
  *out = g_scratch_dev[dev];
  return RADG_OK;
}
}  // namespace

int radg_spectrum_noalloc(const radg_request *req, int mode, double *wout, double *wgts, double *wjac,
                          void *hip_stream) {
  if (!req || !wout || !wgts || !wjac) return fail(RADG_EINPUT, "null request or output");
  if (!valid_mode(mode)) return fail(RADG_EINPUT, "mode=%d", mode);
  std::shared_ptr<TableSet> ts;
  std::shared_ptr<CIA> cia;
  AuxTables aux;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ts = active_tables();
    cia = g_cia;
    aux = {g_para, g_e3, g_n2nir, g_solar, g_gascon};
  }
  if (!ts) return no_tables();
  if (!cia) return fail(RADG_ENOTABLES, "no CIA table registered (radg_register_cia)");
  // making a table form allocates: it must exist already
  if ((mode == RADG_MODE_MIXED && !ts->d_k32) || (mode == RADG_MODE_BF16 && !ts->d_k16) ||
      (mode == RADG_MODE_F16 && !ts->d_kh))
    return fail(RADG_EINPUT, "mode %d: its table form is made by radg_set_mode(%d) after registration", mode, mode);
  Plan pl;
  int rc = plan_batch(1, req, mode, *ts, aux, pl);
  if (rc) return rc;
  // inputs, intermediates and per-path outputs; the weighted outputs are the caller's
  const size_t need = pl.o_wout;
  if (need > RADG_NOALLOC_BYTES)
    return fail(RADG_ELIMITS, "the request needs %zu bytes of scratch; the library holds %zu "
                "(split its wavenumbers over calls)", need, (size_t)RADG_NOALLOC_BYTES);
  const int E = radg_layer_E(ts->ng);
  char *m = nullptr;
  rc = noalloc_scratch(&m);
  if (rc) return rc;
  std::vector<char> h(pl.o_inter);
  stage_inputs(pl, 1, req, h.data());
  hipStream_t st = reinterpret_cast<hipStream_t>(hip_stream);
  // from pageable memory: the copy is staged before the call returns
  HIPCHK(hipMemcpyAsync(m, h.data(), pl.o_inter, hipMemcpyHostToDevice, st));
  KernelArgs a{};
  set_kernel_args(a, m, pl, 1, *ts, *cia, aux, mode);
  a.wout = wout;
  a.wgts = wgts;
  a.wjac = wjac;
  a.err = nullptr;
  return launch_classes(a, emission_classes(pl, req, m, mode, ts->ng, ts->lbl), mode, E, st, nullptr, nullptr);
}

int radg_batch_launch(radg_batch *b, void *stream) {
  if (!b) return fail(RADG_EINPUT, "null batch");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  b->stream = st;
  const int fp64 = b->mode == RADG_MODE_FP64;
  b->ev = &b->evs[4 * (b->launches % radg_batch::RING)];
  for (int i = 0; i < 4; i++)
    if (!b->ev[i]) HIPCHK(hipEventCreate(&b->ev[i]));
  b->launches++;
  b->stream_used = true;
  if (b->up_pending) {
    HIPCHK(hipMemcpyAsync(b->dmem, b->hup, b->hup_bytes, hipMemcpyHostToDevice, st));
    b->up_pending = false;
  }
  if (b->sup_pending) {
    HIPCHK(hipMemcpyAsync(b->smem, b->sup, b->sup_bytes, hipMemcpyHostToDevice, st));
    b->sup_pending = false;
  }
  HIPCHK(hipEventRecord(b->ev[0], st));
  {
    const int rc = launch_classes(b->args, b->em, b->mode, b->E, st, b->ev[1], b->ev[2]);
    if (rc) return rc;
  }
  if (b->smem) {
    HIPCHK(radg_launch_gemm(b->d_pc, b->n_pc, b->mx_pc[0], b->mx_pc[1], 0, st));
    HIPCHK(radg_launch_gemm(b->d_pk64, b->n_pk64, b->mx_pk64[0], b->mx_pk64[1], 0, st));
    HIPCHK(radg_launch_gemm(b->d_pkT, b->n_pkT, b->mx_pkT[0], b->mx_pkT[1], fp64 ? 0 : 1, st));
    HIPCHK(radg_launch_pathsum(b->d_ps, b->n_ps, b->mx_ps, st));
  }
  HIPCHK(hipEventRecord(b->ev[3], st));
  // queue the copies of the requested outputs to pinned host memory behind
  // the kernels, so they overlap the next batch's work
  const size_t wlen = b->span_w1 - b->span_w0;
  if (b->fetch & 1)
    HIPCHK(hipMemcpyAsync(b->hst, b->dmem + b->span_w0, wlen, hipMemcpyDeviceToHost, st));
  if (b->fetch & 2)
    HIPCHK(hipMemcpyAsync(b->hst + wlen, b->dmem + b->span_p0, b->span_p1 - b->span_p0,
                          hipMemcpyDeviceToHost, st));
  return RADG_OK;
}

int radg_batch_set_fetch(radg_batch *b, int what) {
  if (!b || what < 0 || what > 3) return fail(RADG_EINPUT, "bad batch or fetch flags");
  host_cache().put(b->hst);
  b->hst = nullptr;
  b->fetch = 0;
  if (what == 0) return RADG_OK;
  const size_t need = (b->span_w1 - b->span_w0) + (b->span_p1 - b->span_p0);
  b->hst = static_cast<char *>(host_cache().get(need));
  if (!b->hst) return fail(RADG_EHIP, "pinned allocation of %zu bytes failed", need);
  b->fetch = what;
  return RADG_OK;
}

int radg_batch_sync(radg_batch *b) {
  if (!b) return fail(RADG_EINPUT, "null batch");
  if (b->launches == 0) return RADG_OK;
  HIPCHK(hipEventSynchronize(b->ev[3]));
  return RADG_OK;
}

int radg_batch_timings(radg_batch *b, float *ms3) {
  if (!b) return fail(RADG_EINPUT, "null batch");
  if (b->launches == 0) return fail(RADG_EINPUT, "batch was never launched");
  HIPCHK(hipEventSynchronize(b->ev[3]));
  for (int i = 0; i < 3; i++) HIPCHK(hipEventElapsedTime(&ms3[i], b->ev[i], b->ev[i + 1]));
  return RADG_OK;
}

int radg_batch_timing_stats(radg_batch *b, int *nlaunch, float *sum_ms3) {
  if (!b) return fail(RADG_EINPUT, "null batch");
  const long long n = std::min<long long>(b->launches, radg_batch::RING);
  for (int i = 0; i < 3; i++) sum_ms3[i] = 0.f;
  for (long long l = b->launches - n; l < b->launches; l++) {
    hipEvent_t *e = &b->evs[4 * (l % radg_batch::RING)];
    HIPCHK(hipEventSynchronize(e[3]));
    for (int i = 0; i < 3; i++) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e[i], e[i + 1]));
      sum_ms3[i] += ms;
    }
  }
  *nlaunch = (int)n;
  return RADG_OK;
}

int radg_batch_info(radg_batch *b, double *alg_bytes3, int *nspectra, int *nitems_a) {
  if (!b) return fail(RADG_EINPUT, "null batch");
  if (alg_bytes3)
    for (int i = 0; i < 3; i++) alg_bytes3[i] = b->alg_bytes[i];
  if (nspectra) *nspectra = b->nreq;
  if (nitems_a) *nitems_a = (int)b->args.nA;
  return RADG_OK;
}

int radg_batch_read(radg_batch *b, int r, double *output, double *gtsurf, double *doutputdq,
                    double *wout, double *wgts, double *wjac) {
  if (!b || r < 0 || r >= b->nreq) return fail(RADG_EINPUT, "bad batch/request index");
  if (b->launches == 0) return fail(RADG_EINPUT, "batch was never launched");
  HIPCHK(hipEventSynchronize(b->ev[3]));
  const DevReq &d = b->reqs[r];
  const KernelArgs &a = b->args;
  const bool fp64 = b->mode == RADG_MODE_FP64;
  auto fetch = [&](const void *base, long long off, long long n, double *dst) -> int {
    if (!dst) return RADG_OK;
    if (fp64) {
      HIPCHK(hipMemcpy(dst, reinterpret_cast<const double *>(base) + off, n * sizeof(double),
                       hipMemcpyDeviceToHost));
    } else {
      std::vector<float> t(n);
      HIPCHK(hipMemcpy(t.data(), reinterpret_cast<const float *>(base) + off, n * sizeof(float),
                       hipMemcpyDeviceToHost));
      for (long long i = 0; i < n; i++) dst[i] = t[i];
    }
    return RADG_OK;
  };
  const long long nout = (long long)d.nwave * d.npath;
  const size_t nwj = (size_t)d.nwave * d.L * d.nparam;
  if (b->fetch) {
    // results were copied to pinned memory behind the kernels
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t wlen = b->span_w1 - b->span_w0;
    const size_t ts_ = fp64 ? 8 : 4;
    auto from_pinned = [&](size_t region, long long off, long long n, double *dst) {
      if (!dst) return;
      const char *src = b->hst + wlen + (region - b->span_p0) + (size_t)off * ts_;
      if (fp64) std::memcpy(dst, src, n * sizeof(double));
      else {
        const float *f = reinterpret_cast<const float *>(src);
        for (long long i = 0; i < n; i++) dst[i] = f[i];
      }
    };
    auto wfrom = [&](size_t region, long long off, size_t n, double *dst) {
      if (dst) std::memcpy(dst, b->hst + (region - b->span_w0) + off * 8, n * 8);
    };
    const bool need_p = output || gtsurf || doutputdq;
    const bool need_w = wout || wgts || wjac;
    if ((!need_p || (b->fetch & 2)) && (!need_w || (b->fetch & 1))) {
      if (b->fetch & 2) {
        from_pinned(b->o_out, d.out0, nout, output);
        from_pinned(b->o_gts, d.out0, nout, gtsurf);
        if (doutputdq)
          radg_jac_to_host(b->hst + wlen + (b->o_jac - b->span_p0) + (size_t)d.jac0 * ts_, fp64,
                           nout * d.L * d.nparam, d.nparam, d.ngas, doutputdq);
      }
      if (b->fetch & 1) {
        wfrom(b->o_wout, d.wout0, d.nwave, wout);
        wfrom(b->o_wgts, d.wout0, d.nwave, wgts);
        wfrom(b->o_wjac, d.wjac0, nwj, wjac);
      }
      return RADG_OK;
    }
  }
  int rc;
  if ((rc = fetch(a.out, d.out0, nout, output))) return rc;
  if ((rc = fetch(a.gts, d.out0, nout, gtsurf))) return rc;
  if (doutputdq) {
    const size_t nj = (size_t)nout * d.L * d.nparam, ts_ = fp64 ? 8 : 4;
    std::vector<char> t(nj * ts_);
    HIPCHK(hipMemcpy(t.data(), static_cast<const char *>(a.jac) + (size_t)d.jac0 * ts_, nj * ts_,
                     hipMemcpyDeviceToHost));
    radg_jac_to_host(t.data(), fp64, nj, d.nparam, d.ngas, doutputdq);
  }
  if (wout) HIPCHK(hipMemcpy(wout, a.wout + d.wout0, d.nwave * sizeof(double), hipMemcpyDeviceToHost));
  if (wgts) HIPCHK(hipMemcpy(wgts, a.wgts + d.wout0, d.nwave * sizeof(double), hipMemcpyDeviceToHost));
  if (wjac) HIPCHK(hipMemcpy(wjac, a.wjac + d.wjac0, nwj * sizeof(double), hipMemcpyDeviceToHost));
  return RADG_OK;
}

}  // extern "C"

int radg_batch_pinned_view(radg_batch *b, int r, PinnedView *v) {
  if (!b || !v || r < 0 || r >= b->nreq) return fail(RADG_EINPUT, "bad batch/request index");
  if (!b->hst || b->launches == 0) return fail(RADG_EINPUT, "batch has no fetched results");
  const DevReq &d = b->reqs[r];
  const bool fp64 = b->mode == RADG_MODE_FP64;
  const size_t ts = fp64 ? 8 : 4;
  const size_t wlen = b->span_w1 - b->span_w0;
  *v = PinnedView{};
  v->fp64 = fp64 ? 1 : 0;
  v->nwave = d.nwave;
  v->npath = d.npath;
  v->L = d.L;
  v->nparam = d.nparam;
  v->ngas = d.ngas;
  if (b->fetch & 1) {
    v->wout = reinterpret_cast<const double *>(b->hst + (b->o_wout - b->span_w0)) + d.wout0;
    v->wgts = reinterpret_cast<const double *>(b->hst + (b->o_wgts - b->span_w0)) + d.wout0;
    v->wjac = reinterpret_cast<const double *>(b->hst + (b->o_wjac - b->span_w0)) + d.wjac0;
  }
  if (b->fetch & 2) {
    const char *pb = b->hst + wlen;
    v->out = pb + (b->o_out - b->span_p0) + (size_t)d.out0 * ts;
    v->gts = pb + (b->o_gts - b->span_p0) + (size_t)d.out0 * ts;
    v->jac = pb + (b->o_jac - b->span_p0) + (size_t)d.jac0 * ts;
  }
  return RADG_OK;
}

std::shared_ptr<char> radg_batch_take_pinned(radg_batch *b) {
  char *p = b ? b->hst : nullptr;
  if (b) {
    b->hst = nullptr;
    b->fetch = 0;
  }
  return std::shared_ptr<char>(p, [](char *q) { host_cache().put(q); });
}

extern "C" {

int radg_batch_set_statemaps(radg_batch *b, const radg_statemap *maps, int per_path) {
  if (!b || !maps) return fail(RADG_EINPUT, "null batch or maps");
  if (b->launches > 0) HIPCHK(hipEventSynchronize(b->ev[3]));
  dev_cache().put(b->smem);
  b->smem = nullptr;
  b->smem_bytes = 0;
  b->sm.clear();
  b->n_pc = b->n_pk64 = b->n_pkT = b->n_ps = 0;
  const int nreq = b->nreq;
  const int fp64 = b->mode == RADG_MODE_FP64;
  const size_t tsz = fp64 ? 8 : 4;
  std::vector<radg_batch::SMap> sm(nreq);
  // ---- sizes and path classes
  Layout lo;
  std::vector<size_t> o_dam(nreq), o_dte(nreq), o_dco(nreq), o_dfp(nreq), o_xm(nreq), o_rm(nreq);
  int n_pc = 0, n_pk64 = 0, n_pkT = 0, n_ps = 0;
  for (int r = 0; r < nreq; r++) {
    const radg_statemap &mp = maps[r];
    if (mp.nv == 0) continue;
    const radg_batch::HostReq &hq = b->hreq[r];
    const int L = hq.L, npar = hq.ngas + 1 + hq.ncont + hq.h2p;
    if (mp.nv < 0 || mp.npro < 1 || !mp.dam || !mp.dte || !mp.xmap)
      return fail(RADG_EINPUT, "request %d: statemap needs nv >= 0, npro >= 1 and dam, dte, xmap", r);
    if (hq.ncont > 0 && !mp.dco)
      return fail(RADG_EINPUT, "request %d: statemap of a request with dust needs dco", r);
    if (hq.h2p && !mp.dfp)
      return fail(RADG_EINPUT, "request %d: statemap of a FLAGH2P=1 request needs dfp", r);
    radg_batch::SMap &S = sm[r];
    S.nv = mp.nv;
    S.npro = mp.npro;
    S.pcls.assign(hq.npath, -1);
    std::vector<int> rep;  // representative path of each class
    for (int p = 0; p < hq.npath; p++) {
      for (size_t c = 0; c < rep.size(); c++) {
        const int q = rep[c];
        if (hq.nlays[q] == hq.nlays[p] &&
            std::equal(hq.layinc.begin() + (size_t)p * L, hq.layinc.begin() + (size_t)p * L + hq.nlays[p],
                       hq.layinc.begin() + (size_t)q * L)) {
          S.pcls[p] = (int)c;
          break;
        }
      }
      if (S.pcls[p] < 0) {
        S.pcls[p] = (int)rep.size();
        rep.push_back(p);
      }
    }
    S.ncls = (int)rep.size();
    S.per_path = (per_path || S.ncls > 1) ? 1 : 0;
    o_dam[r] = lo.take<double>((size_t)L * mp.npro);
    o_dte[r] = lo.take<double>((size_t)L * mp.npro);
    o_dco[r] = lo.take<double>(hq.ncont > 0 ? (size_t)L * mp.npro : 0);
    o_dfp[r] = lo.take<double>(hq.h2p ? (size_t)L * mp.npro : 0);
    o_xm[r] = lo.take<double>((size_t)mp.nv * npar * mp.npro);
    o_rm[r] = lo.take<int>((size_t)S.ncls * L);
    n_pc += S.ncls * npar;
    if (S.ncls == 1) n_pk64 += 1;
    else n_ps += 1;
    if (S.per_path) n_pkT += hq.npath;
  }
  const size_t o_pc = lo.take<GemmProblem>(n_pc), o_pk64 = lo.take<GemmProblem>(n_pk64),
               o_pkT = lo.take<GemmProblem>(n_pkT);
  const size_t o_ps = lo.take<char>((size_t)n_ps * radg_pathsum_bytes());
  // the inputs and descriptors above are uploaded; the products below are
  // written on the device
  const size_t staged = lo.off;
  for (int r = 0; r < nreq; r++) {
    radg_batch::SMap &S = sm[r];
    if (S.nv == 0) continue;
    const radg_batch::HostReq &hq = b->hreq[r];
    const int L = hq.L, npar = hq.ngas + 1 + hq.ncont + hq.h2p;
    S.o_cm = lo.take<double>((size_t)S.ncls * L * npar * S.nv);
    S.o_kk = lo.take<double>((size_t)hq.nwave * S.nv);
    if (S.per_path) S.o_dx = lo.take<double>((size_t)hq.nwave * hq.npath * S.nv);
  }
  const size_t total = lo.off;
  if (n_pc == 0) {
    b->sm = sm;
    return RADG_OK;
  }
  char *dm = static_cast<char *>(dev_cache().get(total));
  if (!dm) return fail(RADG_EHIP, "device allocation of %zu bytes failed", total);
  // ---- inputs and problem descriptors, staged in pinned memory and
  // uploaded by the next launch
  host_cache().put(b->sup);
  b->sup = static_cast<char *>(host_cache().get(staged));
  if (!b->sup) {
    dev_cache().put(dm);
    return fail(RADG_EHIP, "pinned allocation of %zu bytes failed", staged);
  }
  b->sup_bytes = staged;
  char *hs = b->sup;
  std::memset(hs, 0, staged);
  std::vector<GemmProblem> pc, pk64, pkT;
  std::vector<char> ps((size_t)n_ps * radg_pathsum_bytes());
  int mpc[2] = {0, 0}, mk64[2] = {0, 0}, mkT[2] = {0, 0};
  long long mps = 0;
  int ips = 0;
  const KernelArgs &a = b->args;
  auto gp = [](const void *A, const int *rm, const double *B, double *C, long long sAm,
               long long sAk, long long sBk, long long sBn, long long sCm, long long sCn, int M,
               int N, int K) {
    GemmProblem P{};
    P.A = A;
    P.rowmapA = rm;
    P.B = B;
    P.scaleK = nullptr;
    P.C = C;
    P.sAm = sAm;
    P.sAk = sAk;
    P.sBk = sBk;
    P.sBn = sBn;
    P.sCm = sCm;
    P.sCn = sCn;
    P.M = M;
    P.N = N;
    P.K = K;
    P.accumulate = 0;
    P.lower_only = 0;
    P.alpha = 1.0;
    return P;
  };
  for (int r = 0; r < nreq; r++) {
    const radg_batch::SMap &S = sm[r];
    if (S.nv == 0) continue;
    const radg_statemap &mp = maps[r];
    const radg_batch::HostReq &hq = b->hreq[r];
    const DevReq &d = b->reqs[r];
    const int L = hq.L, npar = hq.ngas + 1 + hq.ncont + hq.h2p, nv = S.nv, npro = S.npro;
    std::memcpy(hs + o_dam[r], mp.dam, (size_t)L * npro * 8);
    std::memcpy(hs + o_dte[r], mp.dte, (size_t)L * npro * 8);
    if (hq.ncont > 0) std::memcpy(hs + o_dco[r], mp.dco, (size_t)L * npro * 8);
    if (hq.h2p) std::memcpy(hs + o_dfp[r], mp.dfp, (size_t)L * npro * 8);
    std::memcpy(hs + o_xm[r], mp.xmap, (size_t)nv * npar * npro * 8);
    int *rm = reinterpret_cast<int *>(hs + o_rm[r]);
    std::vector<int> rep(S.ncls, -1);
    for (int p = 0; p < hq.npath; p++)
      if (rep[S.pcls[p]] < 0) rep[S.pcls[p]] = p;
    for (int c = 0; c < S.ncls; c++)
      for (int j = 0; j < L; j++)  // rows past the path's layers read level row 0; their Jacobian rows are 0
        rm[(size_t)c * L + j] = j < hq.nlays[rep[c]] ? hq.layinc[(size_t)rep[c] * L + j] : 0;
    const double *ddam = reinterpret_cast<const double *>(dm + o_dam[r]);
    const double *ddte = reinterpret_cast<const double *>(dm + o_dte[r]);
    const double *ddco = reinterpret_cast<const double *>(dm + o_dco[r]);
    const double *ddfp = reinterpret_cast<const double *>(dm + o_dfp[r]);
    const double *dxm = reinterpret_cast<const double *>(dm + o_xm[r]);
    const int *drm = reinterpret_cast<const int *>(dm + o_rm[r]);
    double *dcm = reinterpret_cast<double *>(dm + S.o_cm);
    double *dkk = reinterpret_cast<double *>(dm + S.o_kk);
    const long long csz = (long long)L * npar * nv;
    for (int c = 0; c < S.ncls; c++)
      for (int k = 0; k < npar; k++) {
        // map2pro.f:62-80: DAM for the gases, DTE for T, DCO for every dust
        // type, DFP for the para-H2 fraction
        const double *D = k < hq.ngas ? ddam : k == hq.ngas ? ddte : k - hq.ngas - 1 < hq.ncont ? ddco : ddfp;
        pc.push_back(gp(D, drm + (size_t)c * L, dxm + (size_t)k * npro,
                        dcm + c * csz + (long long)k * nv, npro, 1, 1, (long long)npar * npro,
                        (long long)npar * nv, 1, L, nv, npro));
        mpc[0] = std::max(mpc[0], L);
        mpc[1] = std::max(mpc[1], nv);
      }
    if (S.ncls == 1) {
      pk64.push_back(gp(a.wjac + d.wjac0, nullptr, dcm, dkk, (long long)L * npar, 1, nv, 1, nv, 1,
                        hq.nwave, nv, L * npar));
      mk64[0] = std::max(mk64[0], hq.nwave);
      mk64[1] = std::max(mk64[1], nv);
    }
    if (S.per_path) {
      double *ddx = reinterpret_cast<double *>(dm + S.o_dx);
      for (int p = 0; p < hq.npath; p++) {
        const char *ja = reinterpret_cast<const char *>(a.jac) +
                         (size_t)(d.jac0 + (long long)p * L * npar) * tsz;
        GemmProblem q = gp(ja, nullptr, dcm + S.pcls[p] * csz, ddx + (long long)p * nv,
                           (long long)hq.npath * L * npar, 1, nv, 1, (long long)hq.npath * nv, 1,
                           hq.nwave, nv, L * npar);
        q.kper = npar;  // the stored Jacobian's gas columns are per 1e20 molecules
        q.kgas = hq.ngas;
        pkT.push_back(q);
      }
      mkT[0] = std::max(mkT[0], hq.nwave);
      mkT[1] = std::max(mkT[1], nv);
      if (S.ncls > 1) {
        radg_pathsum_fill(ps.data() + (size_t)ips * radg_pathsum_bytes(), ddx, a.pweight + d.path0,
                          dkk, hq.nwave, hq.npath, nv);
        ips++;
        mps = std::max(mps, (long long)hq.nwave * nv);
      }
    }
  }
  std::memcpy(hs + o_pc, pc.data(), pc.size() * sizeof(GemmProblem));
  if (!pk64.empty()) std::memcpy(hs + o_pk64, pk64.data(), pk64.size() * sizeof(GemmProblem));
  if (!pkT.empty()) std::memcpy(hs + o_pkT, pkT.data(), pkT.size() * sizeof(GemmProblem));
  if (!ps.empty()) std::memcpy(hs + o_ps, ps.data(), ps.size());
  b->sup_pending = true;
  b->smem = dm;
  b->smem_bytes = total;
  b->sm = sm;
  b->d_pc = reinterpret_cast<GemmProblem *>(dm + o_pc);
  b->d_pk64 = reinterpret_cast<GemmProblem *>(dm + o_pk64);
  b->d_pkT = reinterpret_cast<GemmProblem *>(dm + o_pkT);
  b->d_ps = dm + o_ps;
  b->n_pc = n_pc;
  b->n_pk64 = n_pk64;
  b->n_pkT = n_pkT;
  b->n_ps = n_ps;
  std::copy(mpc, mpc + 2, b->mx_pc);
  std::copy(mk64, mk64 + 2, b->mx_pk64);
  std::copy(mkT, mkT + 2, b->mx_pkT);
  b->mx_ps = mps;
  return RADG_OK;
}

int radg_batch_read_mapped(radg_batch *b, int r, double *doutdx, double *kk) {
  if (!b || r < 0 || r >= b->nreq) return fail(RADG_EINPUT, "bad batch or request index");
  if ((int)b->sm.size() != b->nreq || b->sm[r].nv == 0)
    return fail(RADG_EINPUT, "request %d has no state map", r);
  if (b->launches == 0) return fail(RADG_EINPUT, "batch was never launched");
  const radg_batch::SMap &S = b->sm[r];
  const radg_batch::HostReq &hq = b->hreq[r];
  if (doutdx && !S.per_path)
    return fail(RADG_EINPUT, "request %d: per-path mapped output was not requested", r);
  HIPCHK(hipEventSynchronize(b->ev[3]));
  if (kk) HIPCHK(hipMemcpy(kk, b->smem + S.o_kk, (size_t)hq.nwave * S.nv * 8, hipMemcpyDeviceToHost));
  if (doutdx)
    HIPCHK(hipMemcpy(doutdx, b->smem + S.o_dx, (size_t)hq.nwave * hq.npath * S.nv * 8,
                     hipMemcpyDeviceToHost));
  return RADG_OK;
}

int radg_batch_device_outputs(radg_batch *b, int r, const double **wout, const double **kk) {
  if (!b || r < 0 || r >= b->nreq) return fail(RADG_EINPUT, "bad batch or request index");
  if (wout) *wout = b->args.wout + b->reqs[r].wout0;
  if (kk) {
    const bool has = (int)b->sm.size() == b->nreq && b->sm[r].nv > 0;
    *kk = has ? reinterpret_cast<const double *>(b->smem + b->sm[r].o_kk) : nullptr;
  }
  return RADG_OK;
}

int radg_batch_destroy(radg_batch *b) {
  delete b;
  return RADG_OK;
}

}  // extern "C"

// ====================================================== Fortran surface
namespace {

template <typename R>
int run_fixed(const R *delh, int L, int npath, int ngas, const R *press, const R *temp,
              const R *pp, const R *amount, int ispace, int inormal, int iray, int itype,
              int flagh2p, const int *nlayin, const int *layinc, const R *scale, const int *imod,
              const int *idgas, const int *isogas, const R *emtemp, int nem, const R *vem,
              const R *emis, R tsurf, int nsw, const int *isw, int nwave, const R *vwave,
              const R *cont, R radius1, R radextra, const R *hfp, R dist, std::vector<double> &out,
              std::vector<double> &gts, std::vector<double> &jac) {
  if (L < 1 || L > RADG_MAXLAY || npath < 1 || npath > RADG_MAXPAT || ngas < 1 || ngas > RADG_MAXGAS)
    return fail(RADG_ELIMITS, "nlayer=%d npath=%d ngas=%d exceed the array limits", L, npath, ngas);
  if (nem < 1 || nem > RADG_MAXSEC) return fail(RADG_ELIMITS, "nem=%d", nem);
  // fixed leading dimensions (arrdef.f) -> compact
  std::vector<double> pr(press, press + L), tp(temp, temp + L), dh(delh, delh + L);
  std::vector<double> ppc((size_t)L * ngas), amc((size_t)L * ngas);
  for (int j = 0; j < L; j++)
    for (int k = 0; k < ngas; k++) {
      ppc[(size_t)j * ngas + k] = pp[(size_t)k * RADG_MAXLAY + j];
      amc[(size_t)j * ngas + k] = amount[(size_t)k * RADG_MAXLAY + j];
    }
  std::vector<int> li((size_t)npath * L, 1);
  std::vector<double> sc((size_t)npath * L, 0.0), et((size_t)npath * L, 0.0);
  for (int p = 0; p < npath; p++)
    for (int m = 0; m < L; m++) {
      li[(size_t)p * L + m] = layinc[(size_t)p * RADG_MAXLAY + m];
      sc[(size_t)p * L + m] = scale[(size_t)p * RADG_MAXLAY + m];
      et[(size_t)p * L + m] = emtemp[(size_t)p * RADG_MAXLAY + m];
    }
  std::vector<double> vw(vwave, vwave + nwave), ve(vem, vem + nem), em(emis, emis + nem);
  // /DUST/ as registered, cont(maxcon,maxlay) -> [L][ncont]
  std::shared_ptr<Dust> dust;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    dust = g_dust;
  }
  const int ncont = dust ? dust->ncont : 0;
  int inlte, iform;
  double nk1, na;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    iform = g_iform;
    inlte = g_inlte;
    nk1 = g_nlte_k1;
    na = g_nlte_a;
  }
  std::vector<double> ct((size_t)L * ncont);
  for (int j = 0; j < L; j++)
    for (int k = 0; k < ncont; k++) ct[(size_t)j * ncont + k] = cont[(size_t)j * RADG_MAXCON + k];
  radg_request q{};
  q.nlayer = L;
  q.npath = npath;
  q.ngas = ngas;
  q.nsw = nsw;
  q.nwave = nwave;
  q.nem = nem;
  q.inormal = inormal;
  q.iray = iray;
  q.ispace = ispace;
  q.itype = itype;
  q.flagh2p = flagh2p;
  q.tsurf = tsurf;
  q.press = pr.data();
  q.temp = tp.data();
  q.delh = dh.data();
  q.pp = ppc.data();
  q.amount = amc.data();
  q.idgas = idgas;
  q.isogas = isogas;
  q.nlayin = nlayin;
  q.imod = imod;
  q.layinc = li.data();
  q.scale = sc.data();
  q.emtemp = et.data();
  q.isw = isw;
  q.vwave = vw.data();
  q.vem = ve.data();
  q.emissivity = em.data();
  q.path_weight = nullptr;
  q.inlte = inlte;
  q.nlte_k1 = nk1;
  q.nlte_a = na;
  q.iform = iform;
  q.radius1 = radius1;
  q.radextra = radextra;
  q.dist = dist;
  std::vector<double> hf;
  if (flagh2p == 1) {
    hf.assign(hfp, hfp + L);
    q.hfp = hf.data();
  }
  q.ncont = ncont;
  if (ncont > 0) {
    q.nsec = dust->nsec;
    q.cont = ct.data();
    q.vsec = dust->vsec.data();
    q.xsec = dust->xsec.data();
  }
  // The calling thread keeps its last one-wavenumber batch: a call whose
  // request has the same layout (a caller's wavenumber loop) re-stages its
  // inputs into it instead of making a new one.  RADG_FIXED_CACHE=0 in the
  // environment makes a batch per call.  The batch is never freed: it is
  // held until the process ends.
  static thread_local radg_batch *cached = nullptr;
  const char *env_cache = std::getenv("RADG_FIXED_CACHE");
  const bool use_cache = !(env_cache && env_cache[0] == '0');
  radg_batch *b = nullptr;
  int rc;
  if (use_cache && cached) {
    rc = batch_restage(cached, 1, &q, g_mode, true);
    if (rc == RADG_OK) b = cached;
    else if (rc != 1) return rc;
  }
  std::unique_ptr<radg_batch> guard;
  if (!b) {
    rc = batch_create(1, &q, g_mode, 0, &b, true);
    if (rc) return rc;
    if (use_cache) {
      delete cached;
      cached = b;
    } else {
      guard.reset(b);
    }
  }
  if ((rc = radg_batch_launch(b, nullptr))) return rc;
  if ((rc = radg_batch_sync(b))) return rc;
  out.assign((size_t)nwave * npath, 0.0);
  gts.assign((size_t)nwave * npath, 0.0);
  // every column radg_batch_read writes: gases, T, dust, the para-H2 fraction
  jac.assign((size_t)nwave * npath * L * b->reqs[0].nparam, 0.0);
  return radg_batch_read(b, 0, out.data(), gts.data(), jac.data(), nullptr, nullptr, nullptr);
}

template <typename R>
int spectrum_fixed(const R *dist, const int *inormal, const int *iray, const R *delh,
                   const int *nlayer, const int *npath, const int *ngas, const R *press,
                   const R *temp, const R *pp, const R *amount, const int *ispace,
                   const int *nlayin, const int *layinc, const R *cont, const R *scale,
                   const int *imod, const int *idgas, const int *isogas, const R *emtemp,
                   const int *itype, const int *nem, const R *vem, const R *emissivity,
                   const R *tsurf, const R *radius1, const R *radextra, const int *flagh2p,
                   const R *hfp, const int *nsw, const int *isw, int nwave, const R *vwave,
                   R *output, R *gtsurf, R *doutputdq) {
  std::vector<double> o, g, j;
  int rc = run_fixed<R>(delh, *nlayer, *npath, *ngas, press, temp, pp, amount, *ispace, *inormal,
                        *iray, *itype, *flagh2p, nlayin, layinc, scale, imod, idgas, isogas,
                        emtemp, *nem, vem, emissivity, *tsurf, *nsw, isw, nwave, vwave, cont,
                        *radius1, *radextra, hfp, *dist, o, g, j);
  if (rc) return rc;
  for (size_t i = 0; i < o.size(); i++) {
    output[i] = (R)o[i];
    gtsurf[i] = (R)g[i];
  }
  for (size_t i = 0; i < j.size(); i++) doutputdq[i] = (R)j[i];
  return RADG_OK;
}

}  // namespace

extern "C" {

void cirsradg_wave_(const float *dist, const int *inormal, const int *iray, const float *delh,
                    const int *nlayer, const int *npath, const int *ngas, const float *press,
                    const float *temp, const float *pp, const float *amount, const int *iwave,
                    const int *ispace, const int *amform, const float *vwave, const int *nlayin,
                    const int *layinc, const float *cont, const float *scale, const int *imod,
                    const int *idgas, const int *isogas, const float *emtemp, const int *itype,
                    const int *nem, const float *vem, const float *emissivity, const float *tsurf,
                    float *gtsurf, const float *radius1, const int *flagh2p, const float *hfp,
                    const float *radextra, const int *nsw, const int *isw, float *output,
                    float *doutputdq) {
  (void)iwave; (void)amform;
  std::vector<double> o, g, j;
  const int L = *nlayer, np_ = *npath, ng = *ngas;
  int rc = run_fixed<float>(delh, L, np_, ng, press, temp, pp, amount, *ispace, *inormal, *iray,
                            *itype, *flagh2p, nlayin, layinc, scale, imod, idgas, isogas, emtemp,
                            *nem, vem, emissivity, *tsurf, *nsw, isw, 1, vwave, cont, *radius1,
                            *radextra, hfp, *dist, o, g, j);
  g_cirs_status = rc;
  if (rc) return;
  const int nparam = (int)(j.size() / ((size_t)np_ * L));  // ngas + 1 + ncont
  for (int p = 0; p < np_; p++) {
    output[p] = (float)o[p];
    gtsurf[p] = (float)g[p];
    // doutputdq(maxpat,maxlay,maxgas+2+maxcon): zero every row, then fill
    for (int m = 0; m < RADG_MAXLAY; m++)
      for (int k = 0; k < nparam; k++)
        doutputdq[p + (size_t)RADG_MAXPAT * (m + (size_t)RADG_MAXLAY * k)] =
            m < L ? (float)j[((size_t)p * L + m) * nparam + k] : 0.0f;
  }
}

int radg_cirsradg_status(void) { return g_cirs_status; }

int radg_spectrum_r4(const float *dist, const int *inormal, const int *iray, const float *delh,
                     const int *nlayer, const int *npath, const int *ngas, const float *press,
                     const float *temp, const float *pp, const float *amount, const int *ispace,
                     const int *nlayin, const int *layinc, const float *cont, const float *scale,
                     const int *imod, const int *idgas, const int *isogas, const float *emtemp,
                     const int *itype, const int *nem, const float *vem, const float *emissivity,
                     const float *tsurf, const float *radius1, const float *radextra,
                     const int *flagh2p, const float *hfp, const int *nsw, const int *isw,
                     int nwave, const float *vwave, float *output, float *gtsurf,
                     float *doutputdq) {
  return spectrum_fixed<float>(dist, inormal, iray, delh, nlayer, npath, ngas, press, temp, pp,
                               amount, ispace, nlayin, layinc, cont, scale, imod, idgas, isogas,
                               emtemp, itype, nem, vem, emissivity, tsurf, radius1, radextra,
                               flagh2p, hfp, nsw, isw, nwave, vwave, output, gtsurf, doutputdq);
}

int radg_spectrum_r8(const double *dist, const int *inormal, const int *iray, const double *delh,
                     const int *nlayer, const int *npath, const int *ngas, const double *press,
                     const double *temp, const double *pp, const double *amount, const int *ispace,
                     const int *nlayin, const int *layinc, const double *cont,
                     const double *scale, const int *imod, const int *idgas, const int *isogas,
                     const double *emtemp, const int *itype, const int *nem, const double *vem,
                     const double *emissivity, const double *tsurf, const double *radius1,
                     const double *radextra, const int *flagh2p, const double *hfp,
                     const int *nsw, const int *isw, int nwave, const double *vwave,
                     double *output, double *gtsurf, double *doutputdq) {
  return spectrum_fixed<double>(dist, inormal, iray, delh, nlayer, npath, ngas, press, temp, pp,
                                amount, ispace, nlayin, layinc, cont, scale, imod, idgas, isogas,
                                emtemp, itype, nem, vem, emissivity, tsurf, radius1, radextra,
                                flagh2p, hfp, nsw, isw, nwave, vwave, output, gtsurf, doutputdq);
}

// .kta reader (read_khead.f:18-120; regular grids, NT > 0)
// /INTERPKLBL/ from the .lta files an .lls list names (read_klbllist.f:
// 30-198, read_klblhead.f:19-75).  Every table must share NP, NT, DELV,
// NPOINT, VMIN and IREC0, and its pressure and temperature grid ends must
// lie within 1e-4 of the first table's (:131-168); the grid used is the last
// file's, as each header read overwrites the one grid.
int radg_read_klbllist(const char *list_path, int recl) {
  if (recl != 4 && recl != 8) return fail(RADG_EINPUT, "recl=%d (4 or 8)", recl);
  std::ifstream kl(list_path);
  if (!kl) return fail(RADG_EIO, "cannot open %s", list_path);
  std::vector<std::string> names;
  std::string line;
  while (std::getline(kl, line)) {
    std::string nm;
    for (char c : line)
      if (c != ' ' && c != '\t' && c != '\r') nm += c;  // remsp.f
    if (nm.empty()) break;                               // read_klbllist.f:83: a blank line ends the list
    // file.f: a '.' among the last four characters starts the extension,
    // which becomes lta; else .lta is appended
    size_t dot = std::string::npos;
    for (size_t i = nm.size(); i-- > (nm.size() > 4 ? nm.size() - 4 : 0);) {
      if (nm[i] == '/' || nm[i] == ']') break;
      if (nm[i] == '.') {
        dot = i;
        break;
      }
    }
    names.push_back(dot == std::string::npos ? nm + ".lta" : nm.substr(0, dot + 1) + "lta");
  }
  if (names.empty()) return fail(RADG_EIO, "%s lists no tables", list_path);
  std::vector<int> ids, isos;
  std::vector<double> press, temps;
  std::vector<std::vector<double>> data;
  int irec00 = 0, npoint0 = 0, np0 = 0, nt0 = 0;
  double vmin0 = 0, delv0 = 0, lp1 = 0, lpn = 0, ta1 = 0, ta2 = 0;
  for (const auto &nm : names) {
    std::ifstream f(nm, std::ios::binary);
    if (!f) return fail(RADG_EIO, "cannot open %s", nm.c_str());
    std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const size_t nrec = raw.size() / recl;
    auto gi = [&](size_t r) {
      int32_t v;
      std::memcpy(&v, raw.data() + (r - 1) * recl, 4);
      return (int)v;
    };
    auto gr = [&](size_t r) {
      if (recl == 4) {
        float v;
        std::memcpy(&v, raw.data() + (r - 1) * recl, 4);
        return (double)v;
      }
      double v;
      std::memcpy(&v, raw.data() + (r - 1) * recl, 8);
      return v;
    };
    if (nrec < 8) return fail(RADG_EIO, "%s: truncated header", nm.c_str());
    const int irec0 = gi(1), npoint = gi(2), np = gi(5), nts = gi(6);
    const int nt = std::abs(nts);
    const double vmin = gr(3), delv = gr(4);
    if (np < 2 || np > RADG_MAXK || nt < 2 || nt > RADG_MAXK || npoint < 1 || irec0 < 1)
      return fail(RADG_ELIMITS, "%s: np=%d nt=%d npoint=%d irec0=%d", nm.c_str(), np, nts, npoint, irec0);
    const size_t ntemp = nts < 0 ? (size_t)np * nt : (size_t)nt;
    const size_t per = (size_t)np * nt;
    if (8 + np + ntemp > nrec || (size_t)irec0 - 1 + (size_t)npoint * per > nrec)
      return fail(RADG_EIO, "%s: truncated", nm.c_str());
    press.clear();
    temps.clear();
    for (int j = 0; j < np; j++) press.push_back(gr(9 + j));
    for (size_t j = 0; j < ntemp; j++) temps.push_back(gr(9 + np + j));
    for (double p : press)
      if (!finite_pos(p)) return fail(RADG_EINPUT, "%s: pressure %g", nm.c_str(), p);
    const double l1 = std::log(press[0]), ln = std::log(press[np - 1]);
    const double t1 = temps[0], tn = temps[ntemp - 1];
    if (ids.empty()) {
      irec00 = irec0, npoint0 = npoint, np0 = np, nt0 = nts, vmin0 = vmin, delv0 = delv;
      lp1 = l1, lpn = ln, ta1 = t1, ta2 = tn;
    } else {
      if (np != np0 || nts != nt0 || delv != delv0 || npoint != npoint0 || vmin != vmin0 || irec0 != irec00)
        return fail(RADG_EINPUT, "%s: array sizes differ from the first table (read_klbllist.f:131-145)",
                    nm.c_str());
      if (std::fabs(l1 - lp1) > 1e-4 || std::fabs(ln - lpn) > 1e-4 || std::fabs(t1 - ta1) > 1e-4 ||
          std::fabs(tn - ta2) > 1e-4)
        return fail(RADG_EINPUT, "%s: P/T grid ends differ from the first table (read_klbllist.f:155-168)",
                    nm.c_str());
    }
    ids.push_back(gi(7));
    isos.push_back(gi(8));
    std::vector<double> d(per * npoint);
    for (size_t i = 0; i < d.size(); i++) d[i] = gr(irec0 + i);
    data.push_back(std::move(d));
  }
  std::vector<const double *> ptrs;
  for (auto &d : data) ptrs.push_back(d.data());
  return radg_register_lbltables((int)ids.size(), ids.data(), isos.data(), npoint0, vmin0, delv0, np0, nt0,
                                 press.data(), temps.data(), ptrs.data());
}

int radg_read_klist(const char *klist_path, int recl) {
  if (recl != 4 && recl != 8) return fail(RADG_EINPUT, "recl=%d (4 or 8)", recl);
  std::ifstream kl(klist_path);
  if (!kl) return fail(RADG_EIO, "cannot open %s", klist_path);
  std::vector<std::string> names;
  std::string line;
  while (std::getline(kl, line)) {
    const auto a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos) break;  // read_klist.f:80: a blank line ends the list
    const auto z = line.find_last_not_of(" \t\r");
    std::string nm = line.substr(a, z - a + 1);
    if (nm.size() < 4 || nm.substr(nm.size() - 4) != ".kta") nm += ".kta";
    names.push_back(nm);
  }
  if (names.empty()) return fail(RADG_EIO, "%s lists no tables", klist_path);
  std::vector<int> ids, isos, npts;
  std::vector<double> vmins, delvs, press, temps, gord, delg;
  std::vector<std::vector<double>> data, vcens;
  int np0 = 0, nt0 = 0, ng0 = 0;
  for (const auto &nm : names) {
    std::ifstream f(nm, std::ios::binary);
    if (!f) return fail(RADG_EIO, "cannot open %s", nm.c_str());
    std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const size_t nrec = raw.size() / recl;
    auto gi = [&](size_t r) {
      int32_t v;
      std::memcpy(&v, raw.data() + (r - 1) * recl, 4);
      return (int)v;
    };
    auto gr = [&](size_t r) {
      if (recl == 4) {
        float v;
        std::memcpy(&v, raw.data() + (r - 1) * recl, 4);
        return (double)v;
      }
      double v;
      std::memcpy(&v, raw.data() + (r - 1) * recl, 8);
      return v;
    };
    if (nrec < 10) return fail(RADG_EIO, "%s: truncated header", nm.c_str());
    const int irec0 = gi(1), npoint = gi(2), np = gi(6), nts = gi(7), ng = gi(8);
    const int nt = std::abs(nts);
    const double vmin = gr(3), delv = gr(4);
    if (np < 2 || np > RADG_MAXK || nt < 2 || nt > RADG_MAXK || ng < 1 || ng > RADG_MAXG - 1 || npoint < 1)
      return fail(RADG_ELIMITS, "%s: np=%d nt=%d ng=%d npoint=%d", nm.c_str(), np, nts, ng, npoint);
    const size_t per = (size_t)np * nt * ng;
    const size_t ntemp = nts < 0 ? (size_t)np * nt : (size_t)nt;
    const size_t r0 = 11 + 2 * ng + 2;  // read_khead.f:71-100
    if (r0 + np + ntemp + (delv <= 0 ? npoint : 0) > nrec ||
        (size_t)irec0 - 1 + (size_t)npoint * per > nrec)
      return fail(RADG_EIO, "%s: truncated", nm.c_str());
    if (ids.empty()) {
      np0 = np; nt0 = nts; ng0 = ng;
      for (int j = 0; j < ng; j++) gord.push_back(gr(11 + j));
      for (int j = 0; j < ng; j++) delg.push_back(gr(11 + ng + j));
      for (int j = 0; j < np; j++) press.push_back(gr(r0 + j));
      for (size_t j = 0; j < ntemp; j++) temps.push_back(gr(r0 + np + j));
    } else if (np != np0 || nts != nt0 || ng != ng0) {
      return fail(RADG_ELIMITS, "%s: grid differs from the first table (read_klist.f:117-130)", nm.c_str());
    }
    std::vector<double> vc;
    if (delv <= 0)
      for (int j = 0; j < npoint; j++) vc.push_back(gr(r0 + np + ntemp + j));
    vcens.push_back(std::move(vc));
    ids.push_back(gi(9));
    isos.push_back(gi(10));
    npts.push_back(npoint);
    vmins.push_back(vmin);
    delvs.push_back(delv);
    std::vector<double> d(per * npoint);
    for (size_t i = 0; i < d.size(); i++) d[i] = gr(irec0 + i);
    data.push_back(std::move(d));
  }
  std::vector<const double *> ptrs, vptrs;
  for (auto &d : data) ptrs.push_back(d.data());
  for (auto &v : vcens) vptrs.push_back(v.empty() ? nullptr : v.data());
  return radg_register_ktables((int)ids.size(), ids.data(), isos.data(), npts.data(), vmins.data(),
                               delvs.data(), np0, nt0, ng0, press.data(), temps.data(), gord.data(),
                               delg.data(), ptrs.data(), vptrs.data());
}

}  // extern "C"
