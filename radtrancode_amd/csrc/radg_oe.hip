// gfx950 kernels for the consumers of the CIRSRADG_WAVE Jacobian.
//
//   gemm_kernel      batched strided FP64 GEMM (radg_dev.h GemmProblem):
//                    the layer -> profile -> state maps (map2pro.f:49-84,
//                    map2xvec.f:41-54) and the normal matrix of the
//                    optimal-estimation step
//   oe_*             the optimal-estimation proposal of nemesis/coreret.f
//                    for many independent retrievals, all in FP64:
//                      r  = (y - yn) + K (xn - xa)
//                      H  = Sa^-1 + K^T Se^-1 K
//                      z  = H^-1 K^T Se^-1 r            (Cholesky, one solve)
//                      x_out = xa + z,  xn1 = xn + (x_out - xn) / (1 + lambda)
//                    x_out equals calcnextxn.f's xa + dd (y - yn) - aa (xa - xn)
//                    with the gain matrix of calc_gain_matrix.f's nx < ny form
//                    (dd = H^-1 K^T Se^-1, aa = dd K), here obtained by one
//                    factorisation and solve instead of the explicit inverse of
//                    dinvertm.f; the cost function is calc_phiret.f's.
#include <hip/hip_runtime.h>
#include <math.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/radg.h"
#include "radg_dev.h"

namespace {

constexpr int TM = 64, TN = 64, TK = 16;

// 256 threads, a 64 x 64 output tile, 4 x 4 outputs per thread at rows
// ty + 16 r and columns tx + 16 c, K staged through LDS 16 at a time.
template <typename TA>
__global__ __launch_bounds__(256) void gemm_kernel(const GemmProblem *probs, int tiles_n) {
  const GemmProblem P = probs[blockIdx.y];
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x - tm * tiles_n;
  const int m0 = tm * TM, n0 = tn * TN;
  if (m0 >= P.M || n0 >= P.N) return;
  if (P.lower_only && n0 > m0 + TM - 1) return;
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const TA *A = reinterpret_cast<const TA *>(P.A);
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = 0.0;
  for (int k0 = 0; k0 < P.K; k0 += TK) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int m, kk;
      if (P.sAm == 1) {  // coalesce along m
        m = t & 63;
        kk = (t >> 6) + 4 * i;
      } else {
        m = (t >> 4) + 16 * i;
        kk = t & 15;
      }
      const int gm = m0 + m, gk = k0 + kk;
      double v = 0.0;
      if (gm < P.M && gk < P.K) {
        const long long row = P.rowmapA ? P.rowmapA[gm] : gm;
        v = (double)A[row * P.sAm + (long long)gk * P.sAk];
        if (P.kper > 0 && gk % P.kper < P.kgas) v *= 1e-20;
      }
      As[kk][m] = v;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int kk, n;
      if (P.sBn == 1) {  // coalesce along n
        n = t & 63;
        kk = (t >> 6) + 4 * i;
      } else {
        n = (t >> 4) + 16 * i;
        kk = t & 15;
      }
      const int gn = n0 + n, gk = k0 + kk;
      double v = 0.0;
      if (gn < P.N && gk < P.K) {
        v = P.B[(long long)gk * P.sBk + (long long)gn * P.sBn];
        if (P.scaleK) v *= P.scaleK[gk];
      }
      Bs[kk][n] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; r++) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; c++) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int gm = m0 + ty + 16 * r;
    if (gm >= P.M) continue;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int gn = n0 + tx + 16 * c;
      if (gn >= P.N) continue;
      double *cp = P.C + (long long)gm * P.sCm + (long long)gn * P.sCn;
      const double v = P.alpha * acc[r][c];
      *cp = P.accumulate ? *cp + v : v;
    }
  }
}

// out[s][w][n] = sum_p pw[s][p] * in[s][w][p][n]  (angle-weighted sum)
struct PathSum {
  const double *in;
  const double *pw;
  double *out;
  int nwave, npath, n, pad;
};
__global__ __launch_bounds__(256) void pathsum_kernel(const PathSum *ps) {
  const PathSum P = ps[blockIdx.y];
  const long long tot = (long long)P.nwave * P.n;
  for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < tot;
       x += (long long)gridDim.x * blockDim.x) {
    const long long w = x / P.n, nn = x - w * P.n;
    double s = 0.0;
    for (int p = 0; p < P.npath; p++) s += P.pw[p] * P.in[(w * P.npath + p) * P.n + nn];
    P.out[x] = s;
  }
}

// ------------------------------------------------------------- OE kernels
__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; i++) s += red[i];
  return s;
}

// wres[r][i] = ((y - yn)_i + sum_j K_ij (xn - xa)_j) / se_i; one wave per row
__global__ __launch_bounds__(1024) void oe_resid_kernel(const double *const *kk,
                                                        const double *const *yn, const double *y,
                                                        const double *se, const double *xa,
                                                        const double *xn, double *wres, int nx,
                                                        int ny) {
  const int r = blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = blockIdx.x * 16 + wv;
  if (i >= ny) return;
  const double *K = kk[r] + (long long)i * nx;
  const double *xar = xa + (long long)r * nx, *xnr = xn + (long long)r * nx;
  double s = 0.0;
  for (int j = lane; j < nx; j += 64) s += K[j] * (xnr[j] - xar[j]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) {
    const long long q = (long long)r * ny + i;
    wres[q] = ((y[q] - yn[r][i]) + s) / se[q];
  }
}

// g[r][j] = sum_i K_ij wres_i; one thread per column
__global__ __launch_bounds__(256) void oe_grad_kernel(const double *const *kk, const double *wres,
                                                      double *g, int nx, int ny) {
  const int r = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nx) return;
  const double *K = kk[r];
  const double *wr = wres + (long long)r * ny;
  double s = 0.0;
  for (int i = 0; i < ny; i++) s += K[(long long)i * nx + j] * wr[i];
  g[(long long)r * nx + j] = s;
}

// calc_phiret.f: chisq = (yn - y)^T Se^-1 (yn - y), phi = chisq + (xn - xa)^T Sa^-1 (xn - xa)
__global__ __launch_bounds__(256) void oe_cost_kernel(const double *const *yn, const double *y,
                                                      const double *se, const double *xa,
                                                      const double *xn, const double *sai,
                                                      long long sai_stride, double *phi,
                                                      double *chisq, int nx, int ny) {
  __shared__ double red[16];
  const int r = blockIdx.x;
  double c = 0.0;
  for (int i = threadIdx.x; i < ny; i += blockDim.x) {
    const long long q = (long long)r * ny + i;
    const double d = yn[r][i] - y[q];
    c += d * d / se[q];
  }
  const double chi = block_sum(c, red);
  const double *S = sai + (long long)r * sai_stride;
  const double *xar = xa + (long long)r * nx, *xnr = xn + (long long)r * nx;
  double p2 = 0.0;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < nx; j++) s += S[(long long)i * nx + j] * (xnr[j] - xar[j]);
    p2 += (xnr[i] - xar[i]) * s;
  }
  const double ph2 = block_sum(p2, red);
  if (threadIdx.x == 0) {
    chisq[r] = chi;
    phi[r] = chi + ph2;
  }
}

// In-place Cholesky of the lower triangle of A (nx x nx, row-major), right
// looking, one column per step; a non-positive pivot becomes 1 as in
// dcholdc.f:33-48.  L is left in the lower triangle, its diagonal included.
__device__ void chol_lower(double *A, int nx, double *col, double *piv) {
  const int tid = threadIdx.x, bd = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = bd >> 6;
  for (int k = 0; k < nx; k++) {
    if (tid == 0) {
      const double d = A[(long long)k * nx + k];
      const double p = d > 0.0 ? sqrt(d) : 1.0;
      A[(long long)k * nx + k] = p;
      *piv = p;
    }
    __syncthreads();
    const double p = *piv;
    for (int i = k + 1 + tid; i < nx; i += bd) {
      const double v = A[(long long)i * nx + k] / p;
      A[(long long)i * nx + k] = v;
      col[i] = v;
    }
    __syncthreads();
    for (int i = k + 1 + wv; i < nx; i += nw) {
      const double ci = col[i];
      double *row = A + (long long)i * nx;
      for (int jj = k + 1 + lane; jj <= i; jj += 64) row[jj] -= ci * col[jj];
    }
    __syncthreads();
  }
}

// L L^T z = b in LDS vector z (dcholsl.f's forward and back substitution)
__device__ void chol_solve(const double *A, int nx, double *z, double *piv) {
  const int tid = threadIdx.x, bd = blockDim.x;
  for (int k = 0; k < nx; k++) {
    if (tid == 0) z[k] /= A[(long long)k * nx + k];
    __syncthreads();
    const double zk = z[k];
    for (int i = k + 1 + tid; i < nx; i += bd) z[i] -= A[(long long)i * nx + k] * zk;
    __syncthreads();
  }
  for (int k = nx - 1; k >= 0; k--) {
    if (tid == 0) z[k] /= A[(long long)k * nx + k];
    __syncthreads();
    const double zk = z[k];
    for (int i = tid; i < k; i += bd) z[i] -= A[(long long)k * nx + i] * zk;
    __syncthreads();
  }
  (void)piv;
}

// Factor H, solve H z = g, x_out = xa + z, xn1 = xn + (x_out - xn)/(1 + lambda)
__global__ __launch_bounds__(1024) void oe_solve_kernel(double *H, const double *g,
                                                        const double *xa, const double *xn,
                                                        const double *lambda, double *xn1,
                                                        int nx) {
  extern __shared__ double sh[];
  double *col = sh, *z = sh + nx;
  __shared__ double piv;
  const int r = blockIdx.x;
  double *A = H + (long long)r * nx * nx;
  chol_lower(A, nx, col, &piv);
  for (int i = threadIdx.x; i < nx; i += blockDim.x) z[i] = g[(long long)r * nx + i];
  __syncthreads();
  chol_solve(A, nx, z, &piv);
  const double lam = lambda[r];
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    const long long q = (long long)r * nx + i;
    const double xo = xa[q] + z[i];
    xn1[q] = xn[q] + (xo - xn[q]) / (1.0 + lam);
  }
}

// Sa^-1 from Sa (coreret.f:225-250 inverts Sa by dinvertm, JMOD = 2): factor
// a copy in place, then one thread per column of the identity runs the two
// substitutions; the inverse is written row-major (it is symmetric).
__global__ __launch_bounds__(1024) void oe_inverse_kernel(double *F, double *Ainv, int nx) {
  extern __shared__ double sh[];
  __shared__ double piv;
  const int r = blockIdx.x;
  double *A = F + (long long)r * nx * nx;
  double *X = Ainv + (long long)r * nx * nx;
  chol_lower(A, nx, sh, &piv);
  for (int c = threadIdx.x; c < nx; c += blockDim.x) {
    // column c of the inverse, kept in X[.][c]
    for (int i = 0; i < nx; i++) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) s -= A[(long long)i * nx + k] * X[(long long)k * nx + c];
      X[(long long)i * nx + c] = s / A[(long long)i * nx + i];
    }
    for (int i = nx - 1; i >= 0; i--) {
      double s = X[(long long)i * nx + c];
      for (int k = i + 1; k < nx; k++) s -= A[(long long)k * nx + i] * X[(long long)k * nx + c];
      X[(long long)i * nx + c] = s / A[(long long)i * nx + i];
    }
  }
}

// H = Sa^-1 (shared or per retrieval) before the K^T Se^-1 K accumulation
__global__ __launch_bounds__(256) void oe_init_h_kernel(const double *sai, long long sai_stride,
                                                        double *H, long long n2) {
  const int r = blockIdx.y;
  for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < n2;
       x += (long long)gridDim.x * blockDim.x)
    H[(long long)r * n2 + x] = sai[(long long)r * sai_stride + x];
}

}  // namespace

int radg_fail_msg(int code, const char *msg);  // radg_api.hip

namespace {
int oe_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return radg_fail_msg(code, buf);
}
}  // namespace

// Shared with radg_api.hip (state maps of a batch).
hipError_t radg_launch_gemm(const GemmProblem *probs_dev, int nprob, int maxM, int maxN,
                            int a_fp32, hipStream_t st) {
  if (nprob <= 0 || maxM <= 0 || maxN <= 0) return hipSuccess;
  const int tiles_n = (maxN + TN - 1) / TN, tiles_m = (maxM + TM - 1) / TM;
  for (int p0 = 0; p0 < nprob; p0 += 65535) {
    const int np = nprob - p0 < 65535 ? nprob - p0 : 65535;
    const dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)np);
    if (a_fp32)
      hipLaunchKernelGGL(gemm_kernel<float>, grid, dim3(256), 0, st, probs_dev + p0, tiles_n);
    else
      hipLaunchKernelGGL(gemm_kernel<double>, grid, dim3(256), 0, st, probs_dev + p0, tiles_n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t radg_launch_pathsum(const void *ps_dev, int nps, long long maxn, hipStream_t st) {
  if (nps <= 0 || maxn <= 0) return hipSuccess;
  long long nb = (maxn + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(pathsum_kernel, dim3((unsigned)nb, (unsigned)nps), dim3(256), 0, st,
                     reinterpret_cast<const PathSum *>(ps_dev));
  return hipGetLastError();
}
size_t radg_pathsum_bytes() { return sizeof(PathSum); }
void radg_pathsum_fill(void *dst, const double *in, const double *pw, double *out, int nwave,
                       int npath, int n) {
  PathSum p{in, pw, out, nwave, npath, n, 0};
  std::memcpy(dst, &p, sizeof(p));
}

// ================================================================== C ABI
struct radg_oe {
  int nret = 0, nx = 0, ny = 0, sa_shared = 0;
  double *d_y = nullptr, *d_se = nullptr, *d_xa = nullptr, *d_sai = nullptr;
  double *d_xn = nullptr, *d_xn1 = nullptr, *d_lam = nullptr, *d_wres = nullptr;
  double *d_g = nullptr, *d_H = nullptr, *d_phi = nullptr, *d_chi = nullptr, *d_sei = nullptr;
  const double **d_kk = nullptr, **d_yn = nullptr;
  GemmProblem *d_prob = nullptr;
  hipStream_t st = nullptr;
  ~radg_oe();
};

radg_oe::~radg_oe() {
  for (void *p : {(void *)d_y, (void *)d_se, (void *)d_xa, (void *)d_sai, (void *)d_xn,
                  (void *)d_xn1, (void *)d_lam, (void *)d_wres, (void *)d_g, (void *)d_H,
                  (void *)d_phi, (void *)d_chi, (void *)d_sei, (void *)d_kk, (void *)d_yn,
                  (void *)d_prob})
    if (p) (void)hipFree(p);
}

#define OECHK(x)                                                                  \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess)                                                         \
      return oe_fail(RADG_EHIP, "%s: %s", #x, hipGetErrorString(e_));             \
  } while (0)

extern "C" {

int radg_oe_create(int nret, int nx, int ny, const double *y, const double *se,
                   const double *xa, const double *sa, int sa_shared, radg_oe **out) {
  if (!out) return oe_fail(RADG_EINPUT, "null output handle");
  *out = nullptr;
  if (nret < 1 || nx < 1 || ny < 1) return oe_fail(RADG_EINPUT, "nret, nx, ny must be >= 1");
  if (!y || !se || !xa || !sa) return oe_fail(RADG_EINPUT, "null input array");
  for (long long i = 0; i < (long long)nret * ny; i++)
    if (!(se[i] > 0.0)) return oe_fail(RADG_EINPUT, "se[%lld] = %g is not positive", i, se[i]);
  // LDS of the solve kernels: two vectors of nx doubles
  if ((size_t)2 * nx * sizeof(double) > 64 * 1024)
    return oe_fail(RADG_ELIMITS, "nx = %d exceeds the solver's limit of 4096", nx);
  radg_oe *h = new radg_oe();
  h->nret = nret;
  h->nx = nx;
  h->ny = ny;
  h->sa_shared = sa_shared ? 1 : 0;
  const size_t nxy = (size_t)nret * ny, nxx = (size_t)nret * nx;
  const size_t nsa = (sa_shared ? 1 : (size_t)nret) * nx * nx;
  auto fin = [&](int code) {
    delete h;
    return code;
  };
#define OEALLOC(p, n)                                                              \
  if (hipMalloc(&(p), (n)) != hipSuccess)                                          \
    return fin(oe_fail(RADG_EHIP, "device allocation of %zu bytes failed", (size_t)(n)));
  OEALLOC(h->d_y, nxy * 8);
  OEALLOC(h->d_se, nxy * 8);
  OEALLOC(h->d_wres, nxy * 8);
  OEALLOC(h->d_sei, nxy * 8);
  OEALLOC(h->d_xa, nxx * 8);
  OEALLOC(h->d_xn, nxx * 8);
  OEALLOC(h->d_xn1, nxx * 8);
  OEALLOC(h->d_g, nxx * 8);
  OEALLOC(h->d_lam, (size_t)nret * 8);
  OEALLOC(h->d_phi, (size_t)nret * 8);
  OEALLOC(h->d_chi, (size_t)nret * 8);
  OEALLOC(h->d_sai, nsa * 8);
  OEALLOC(h->d_H, (size_t)nret * nx * nx * 8);
  OEALLOC(h->d_kk, (size_t)nret * sizeof(double *));
  OEALLOC(h->d_yn, (size_t)nret * sizeof(double *));
  OEALLOC(h->d_prob, (size_t)nret * sizeof(GemmProblem));
#undef OEALLOC
  if (hipStreamCreate(&h->st) != hipSuccess) return fin(oe_fail(RADG_EHIP, "hipStreamCreate failed"));
  if (hipMemcpy(h->d_y, y, nxy * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_se, se, nxy * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_xa, xa, nxx * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_H, sa, nsa * 8, hipMemcpyHostToDevice) != hipSuccess)
    return fin(oe_fail(RADG_EHIP, "upload of the retrieval constants failed"));
  {
    std::vector<double> sei(nxy);
    for (size_t i = 0; i < nxy; i++) sei[i] = 1.0 / se[i];
    if (hipMemcpy(h->d_sei, sei.data(), nxy * 8, hipMemcpyHostToDevice) != hipSuccess)
      return fin(oe_fail(RADG_EHIP, "upload of Se^-1 failed"));
  }
  const int nb = sa_shared ? 1 : nret;
  hipLaunchKernelGGL(oe_inverse_kernel, dim3(nb), dim3(512), (size_t)nx * 8, h->st, h->d_H,
                     h->d_sai, nx);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->st) != hipSuccess)
    return fin(oe_fail(RADG_EHIP, "Sa inversion failed"));
  *out = h;
  return RADG_OK;
}

int radg_oe_cost(radg_oe *h, const double *xn, const double *const *yn_dev, double *phi,
                 double *chisq) {
  if (!h || !xn || !yn_dev || !phi) return oe_fail(RADG_EINPUT, "null argument");
  const size_t nxx = (size_t)h->nret * h->nx;
  OECHK(hipMemcpyAsync(h->d_xn, xn, nxx * 8, hipMemcpyHostToDevice, h->st));
  OECHK(hipMemcpyAsync(h->d_yn, yn_dev, h->nret * sizeof(double *), hipMemcpyHostToDevice, h->st));
  const long long sstride = h->sa_shared ? 0 : (long long)h->nx * h->nx;
  hipLaunchKernelGGL(oe_cost_kernel, dim3(h->nret), dim3(256), 0, h->st, h->d_yn, h->d_y,
                     h->d_se, h->d_xa, h->d_xn, h->d_sai, sstride, h->d_phi, h->d_chi, h->nx,
                     h->ny);
  OECHK(hipGetLastError());
  OECHK(hipMemcpyAsync(phi, h->d_phi, h->nret * 8, hipMemcpyDeviceToHost, h->st));
  if (chisq) OECHK(hipMemcpyAsync(chisq, h->d_chi, h->nret * 8, hipMemcpyDeviceToHost, h->st));
  OECHK(hipStreamSynchronize(h->st));
  return RADG_OK;
}

int radg_oe_propose(radg_oe *h, const double *xn, const double *const *yn_dev,
                    const double *const *kk_dev, const double *lambda, double *xn1) {
  if (!h || !xn || !yn_dev || !kk_dev || !lambda || !xn1)
    return oe_fail(RADG_EINPUT, "null argument");
  const int nret = h->nret, nx = h->nx, ny = h->ny;
  for (int r = 0; r < nret; r++)
    if (!(lambda[r] >= 0.0)) return oe_fail(RADG_EINPUT, "lambda[%d] = %g", r, lambda[r]);
  const size_t nxx = (size_t)nret * nx;
  OECHK(hipMemcpyAsync(h->d_xn, xn, nxx * 8, hipMemcpyHostToDevice, h->st));
  OECHK(hipMemcpyAsync(h->d_lam, lambda, nret * 8, hipMemcpyHostToDevice, h->st));
  OECHK(hipMemcpyAsync(h->d_yn, yn_dev, nret * sizeof(double *), hipMemcpyHostToDevice, h->st));
  OECHK(hipMemcpyAsync(h->d_kk, kk_dev, nret * sizeof(double *), hipMemcpyHostToDevice, h->st));
  hipLaunchKernelGGL(oe_resid_kernel, dim3((ny + 15) / 16, nret), dim3(1024), 0, h->st,
                     h->d_kk, h->d_yn, h->d_y, h->d_se, h->d_xa, h->d_xn, h->d_wres, nx, ny);
  OECHK(hipGetLastError());
  hipLaunchKernelGGL(oe_grad_kernel, dim3((nx + 255) / 256, nret), dim3(256), 0, h->st, h->d_kk,
                     h->d_wres, h->d_g, nx, ny);
  OECHK(hipGetLastError());
  const long long n2 = (long long)nx * nx;
  hipLaunchKernelGGL(oe_init_h_kernel, dim3(256, nret), dim3(256), 0, h->st, h->d_sai,
                     h->sa_shared ? 0 : n2, h->d_H, n2);
  OECHK(hipGetLastError());
  // H += K^T Se^-1 K, lower tiles only (the factorisation reads the lower
  // triangle): A(m=i, k=row) = K[row][i], B(k=row, n=j) = K[row][j] / se_row
  std::vector<GemmProblem> pr(nret);
  for (int r = 0; r < nret; r++) {
    GemmProblem &P = pr[r];
    P.A = kk_dev[r];
    P.rowmapA = nullptr;
    P.B = kk_dev[r];
    P.scaleK = nullptr;
    P.C = h->d_H + (long long)r * n2;
    P.sAm = 1;
    P.sAk = nx;
    P.sBk = nx;
    P.sBn = 1;
    P.sCm = nx;
    P.sCn = 1;
    P.M = nx;
    P.N = nx;
    P.K = ny;
    P.accumulate = 1;
    P.lower_only = 1;
    P.pad = 0;
    P.alpha = 1.0;
  }
  // the Se^-1 weights scale the K axis
  for (int r = 0; r < nret; r++) pr[r].scaleK = h->d_sei + (long long)r * ny;
  OECHK(hipMemcpyAsync(h->d_prob, pr.data(), nret * sizeof(GemmProblem), hipMemcpyHostToDevice,
                       h->st));
  OECHK(radg_launch_gemm(h->d_prob, nret, nx, nx, 0, h->st));
  hipLaunchKernelGGL(oe_solve_kernel, dim3(nret), dim3(1024), (size_t)2 * nx * 8, h->st, h->d_H,
                     h->d_g, h->d_xa, h->d_xn, h->d_lam, h->d_xn1, nx);
  OECHK(hipGetLastError());
  OECHK(hipMemcpyAsync(xn1, h->d_xn1, nxx * 8, hipMemcpyDeviceToHost, h->st));
  OECHK(hipStreamSynchronize(h->st));
  return RADG_OK;
}

int radg_oe_destroy(radg_oe *h) {
  if (h && h->st) (void)hipStreamSynchronize(h->st);
  if (h && h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return RADG_OK;
}

}  // extern "C"
