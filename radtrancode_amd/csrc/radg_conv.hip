// cirsconv.f (radtran/cirsrad/cirsconv.f:12-246) for many spectra at once.
//
// The reference convolves each spectrum -- every path's radiance and surface
// gradient and every state-vector column of the Jacobian
// (radtran/cirsradg/cirsrtfg_wave.f:468-516) -- separately.  Each of its
// three forms is linear in the spectrum, so it is a matrix C[nconv][nwave]
// that depends only on the grids: built here once on the host in FP64 and
// applied on the device to all ncol spectra in one launch.
//   fwhm > 0: the mean of the natural cubic spline through the spectrum
//     (cspline.f, csplint.f) over [vconv - w/2, vconv + w/2] by a 20-point
//     trapezoid, divided by fwhm (w = fwhm, or the FWHM table's value at
//     vconv), with linearly extrapolated end points where the boxcar passes
//     the grid.  A row of C is the adjoint of the spline recursion applied to
//     the sample weights; the spline's influence decays by ~0.27 per knot, so
//     coefficients below 1e-17 of a row's largest are dropped and C is
//     stored as one column band per row.
//   fwhm = 0: yout(i) = y(i).
//   fwhm < 0: the filter functions of the .fil file, interpolated on the
//     wavenumber grid (interp.f) and normalised over the grid's intervals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/radg.h"

int radg_fail_msg(int code, const char *msg);  // radg_api.hip

namespace {

int conv_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int conv_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return radg_fail_msg(code, buf);
}

// interp.f / verint.f interval choice: the first grid point past x (1-based),
// N if none, at least 2
int past(const double *x, int n, double xin) {
  int i = 1;
  if (x[0] < x[n - 1]) {
    while (i <= n && !(x[i - 1] > xin)) i++;
  } else {
    while (i <= n && !(x[i - 1] < xin)) i++;
  }
  if (i > n) i = n;
  if (i == 1) i = 2;
  return i;
}

double verint(const double *x, const double *y, int n, double xin) {
  const int i = past(x, n, xin);
  double yout = y[i - 2] + (y[i - 1] - y[i - 2]) * (xin - x[i - 2]) / (x[i - 1] - x[i - 2]);
  if (x[i - 2] == x[i - 1]) yout = y[i - 1];
  if (std::isnan(yout)) yout = y[i - 1];
  return yout;
}

double interp(const double *x, const double *y, int n, double xin) {
  const int i = past(x, n, xin);
  return y[i - 2] + (y[i - 1] - y[i - 2]) * (xin - x[i - 2]) / (x[i - 1] - x[i - 2]);
}

// Banded rows: row i holds coef[start[i] .. start[i+1]) for columns lo[i]...
struct Band {
  std::vector<int> lo;
  std::vector<long long> start;
  std::vector<double> coef;
  void push(const std::vector<double> &row) {
    double mx = 0.0;
    for (double v : row) mx = std::max(mx, std::fabs(v));
    int a = 0, b = (int)row.size() - 1;
    const double tol = 1e-17 * mx;
    while (a < b && std::fabs(row[a]) <= tol) a++;
    while (b > a && std::fabs(row[b]) <= tol) b--;
    if (start.empty()) start.push_back(0);
    lo.push_back(a);
    coef.insert(coef.end(), row.begin() + a, row.begin() + b + 1);
    start.push_back((long long)coef.size());
  }
};

// fwhm > 0, cirsconv.f:79-140
int spline_rows(int nwave, const double *vw, int nconv, const double *vc, double fwhm, int nfwhm,
                const double *vfwhm, const double *xfwhm, Band &band) {
  // the extended grid (:81-99) and its map E from the spectrum: point k of
  // the grid is e0[k] y[i0[k]] + e1[k] y[i0[k] + 1]
  std::vector<double> xc, e0, e1;
  std::vector<int> i0;
  const bool lo_ext = vc[0] < vw[0] + fwhm / 2.;
  const bool hi_ext = vc[nconv - 1] > vw[nwave - 1] - fwhm / 2.;
  if (lo_ext) {
    const double x = vc[0] - fwhm, s = (x - vw[0]) / (vw[1] - vw[0]);
    xc.push_back(x);
    i0.push_back(0);
    e0.push_back(1.0 - s);
    e1.push_back(s);
  }
  for (int i = 0; i < nwave; i++) {
    xc.push_back(vw[i]);
    i0.push_back(std::min(i, nwave - 2));
    e0.push_back(i < nwave - 1 ? 1.0 : 0.0);
    e1.push_back(i < nwave - 1 ? 0.0 : 1.0);
  }
  if (hi_ext) {
    const double x = vc[nconv - 1] + fwhm;
    const double s = (x - vw[nwave - 1]) / (vw[nwave - 1] - vw[nwave - 2]);
    xc.push_back(x);
    i0.push_back(nwave - 2);
    e0.push_back(-s);
    e1.push_back(1.0 + s);
  }
  const int n = (int)xc.size();
  // cspline.f's forward-sweep factors, which depend on the grid only
  std::vector<double> sig(n, 0.0), p(n, 1.0), q(n, 0.0);
  std::vector<char> dup(n, 0);
  for (int i = 1; i < n - 1; i++) {
    sig[i] = (xc[i] - xc[i - 1]) / (xc[i + 1] - xc[i - 1]);
    p[i] = sig[i] * q[i - 1] + 2.;
    q[i] = (sig[i] - 1.) / p[i];
    dup[i] = xc[i - 1] == xc[i] || xc[i] == xc[i + 1];
  }
  const int nstep = 20;
  std::vector<double> gy(n), g2(n), ub(n), row(nwave);
  for (int ic = 0; ic < nconv; ic++) {
    std::fill(gy.begin(), gy.end(), 0.0);
    std::fill(g2.begin(), g2.end(), 0.0);
    const double yf = nfwhm > 0 ? verint(vfwhm, xfwhm, nfwhm, vc[ic]) : fwhm;
    const double x1 = vc[ic] - yf / 2.;
    const double delx = ((vc[ic] + yf / 2.) - x1) / (double)(nstep - 1);
    for (int j = 1; j <= nstep; j++) {
      const double xi = x1 + (j - 1) * delx;
      const double w = ((j == 1 || j == nstep) ? delx / 2. : delx) / fwhm;
      int klo = 1, khi = n;  // csplint.f:15-24
      while (khi - klo > 1) {
        const int k = (khi + klo) / 2;
        if (xc[k - 1] > xi) khi = k;
        else klo = k;
      }
      const double h = xc[khi - 1] - xc[klo - 1];
      const double a = (xc[khi - 1] - xi) / h, b = (xi - xc[klo - 1]) / h;
      gy[klo - 1] += w * a;
      gy[khi - 1] += w * b;
      g2[klo - 1] += w * (a * a * a - a) * (h * h) / 6.;
      g2[khi - 1] += w * (b * b * b - b) * (h * h) / 6.;
    }
    // adjoint of the back substitution y2(k) = q(k) y2(k+1) + u(k),
    // k = n-1..1 (y2(n) = 0 for natural ends)
    std::fill(ub.begin(), ub.end(), 0.0);
    for (int k = 0; k < n - 1; k++) {
      ub[k] += g2[k];
      g2[k + 1] += q[k] * g2[k];
    }
    // adjoint of u(i) = (6 d(i) - sig(i) u(i-1)) / p(i), i = 2..n-1 (u(1) = 0),
    // d(i) the divided second difference of the spectrum
    for (int i = n - 2; i >= 1; i--) {
      if (dup[i]) {
        ub[i - 1] += ub[i];
        continue;
      }
      const double dbar = 6. * ub[i] / p[i];
      ub[i - 1] += -sig[i] * ub[i] / p[i];
      const double span = xc[i + 1] - xc[i - 1];
      const double hr = xc[i + 1] - xc[i], hl = xc[i] - xc[i - 1];
      gy[i + 1] += dbar / (span * hr);
      gy[i] += -dbar / (span * hr) - dbar / (span * hl);
      gy[i - 1] += dbar / (span * hl);
    }
    std::fill(row.begin(), row.end(), 0.0);
    for (int k = 0; k < n; k++) {
      row[i0[k]] += e0[k] * gy[k];
      row[i0[k] + 1] += e1[k] * gy[k];
    }
    band.push(row);
  }
  return RADG_OK;
}

// fwhm < 0, cirsconv.f:148-243
int filter_rows(int nwave, const double *vw, int nconv, const double *vc, int nfil, const int *fn,
                const double *fvc, const double *fv, const double *ff, Band &band) {
  std::vector<std::vector<double>> rows(nconv);
  long long off = 0;
  for (int k = 0; k < nfil; k++) {
    const int ns = fn[k];
    if (ns < 2) return conv_fail(RADG_EINPUT, "filter %d has %d points", k + 1, ns);
    for (int i = 0; i < nconv; i++) {
      if (!(std::fabs(fvc[k] - vc[i]) < 0.00001)) continue;
      std::vector<double> r(nwave, 0.0);
      double sumf = 0.0;
      for (int j = 0; j < nwave - 1; j++) {
        double yy = 0.0;
        if (vw[j] >= fv[off] && vw[j] <= fv[off + ns - 1]) yy = interp(fv + off, ff + off, ns, vw[j]);
        if (yy != 0) {
          const double delv = vw[j + 1] - vw[j];
          r[j] = yy * delv;
          sumf += yy * delv;
        }
      }
      for (double &x : r) x /= sumf;
      rows[i] = std::move(r);
    }
    off += ns;
  }
  for (int i = 0; i < nconv; i++) {
    if (rows[i].empty()) return conv_fail(RADG_EINPUT, "no filter is centred on vconv(%d)=%g", i + 1, vc[i]);
    band.push(rows[i]);
  }
  return RADG_OK;
}

// yout[i][c] = sum_j C[i][lo_i + j] y[lo_i + j][c]: one block row per conv
// point, threads over the spectra (coalesced along c)
__global__ __launch_bounds__(256) void band_apply_kernel(const int *lo, const long long *start,
                                                         const double *coef, const double *y,
                                                         double *yout, int ncol) {
  const int i = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const long long s0 = start[i], s1 = start[i + 1];
  const double *yc = y + (long long)lo[i] * ncol + c;
  double s = 0.0;
  for (long long k = s0; k < s1; k++) s += coef[k] * yc[(k - s0) * ncol];
  yout[(long long)i * ncol + c] = s;
}

}  // namespace

extern "C" int radg_cirsconv(int nwave, const double *vwave, int nconv, const double *vconv,
                             const radg_conv *cv, int ncol, const double *y, double *yout,
                             void *hip_stream) {
  if (!vwave || !vconv || !cv || !y || !yout) return conv_fail(RADG_EINPUT, "null argument");
  if (nwave < 1 || nconv < 1 || ncol < 1 || ncol > 256 * 65535)
    return conv_fail(RADG_EINPUT, "nwave=%d nconv=%d ncol=%d", nwave, nconv, ncol);
  const double fwhm = cv->fwhm;
  if (!std::isfinite(fwhm)) return conv_fail(RADG_EINPUT, "fwhm=%g", fwhm);
  if (fwhm != 0.0 && nwave < 2)  // cirsconv.f:70-77 returns 1 here
    return conv_fail(RADG_EINPUT, "fwhm=%g needs at least 2 wavenumbers", fwhm);
  if (fwhm == 0.0 && nconv > nwave) return conv_fail(RADG_EINPUT, "fwhm=0 copies: nconv=%d > nwave=%d", nconv, nwave);
  for (int i = 1; i < nwave; i++)
    if (!(vwave[i] > vwave[i - 1])) return conv_fail(RADG_EINPUT, "vwave must increase (entry %d)", i + 1);
  Band band;
  int rc = RADG_OK;
  if (fwhm > 0.0) {
    if (cv->nfwhm < 0 || (cv->nfwhm > 0 && (!cv->vfwhm || !cv->xfwhm)))
      return conv_fail(RADG_EINPUT, "FWHM table: nfwhm=%d", cv->nfwhm);
    rc = spline_rows(nwave, vwave, nconv, vconv, fwhm, cv->nfwhm, cv->vfwhm, cv->xfwhm, band);
  } else if (fwhm == 0.0) {
    for (int i = 0; i < nconv; i++) {
      std::vector<double> r(i + 1, 0.0);
      r[i] = 1.0;
      band.push(r);
    }
  } else {
    if (cv->nfil < 1 || !cv->fil_n || !cv->fil_vcen || !cv->fil_v || !cv->fil_f)
      return conv_fail(RADG_EINPUT, "fwhm < 0 needs the filter functions");
    rc = filter_rows(nwave, vwave, nconv, vconv, cv->nfil, cv->fil_n, cv->fil_vcen, cv->fil_v,
                     cv->fil_f, band);
  }
  if (rc) return rc;
  // device copy of the band: lo | start | coef, one allocation
  const size_t b_lo = band.lo.size() * sizeof(int), b_st = band.start.size() * sizeof(long long);
  const size_t o_st = (b_lo + 255) & ~size_t(255), o_cf = (o_st + b_st + 255) & ~size_t(255);
  const size_t bytes = o_cf + band.coef.size() * sizeof(double);
  std::vector<char> h(bytes);
  std::memcpy(h.data(), band.lo.data(), b_lo);
  std::memcpy(h.data() + o_st, band.start.data(), b_st);
  std::memcpy(h.data() + o_cf, band.coef.data(), band.coef.size() * sizeof(double));
  char *d = nullptr;
  hipStream_t st = reinterpret_cast<hipStream_t>(hip_stream);
  hipError_t e = hipMalloc(&d, bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(d, h.data(), bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(band_apply_kernel, dim3(nconv, (ncol + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<const int *>(d), reinterpret_cast<const long long *>(d + o_st),
                       reinterpret_cast<const double *>(d + o_cf), y, yout, ncol);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d) (void)hipFree(d);
  if (e != hipSuccess) return conv_fail(RADG_EHIP, "cirsconv on the device: %s", hipGetErrorString(e));
  return RADG_OK;
}
