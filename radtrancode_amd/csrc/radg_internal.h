// Library-internal interfaces between the batch code (radg_api.hip) and the
// serving loop (radg_server.hip).  Not part of the C ABI.
#pragma once
#include <memory>

struct radg_batch;

// Where request r's results lie in a batch's pinned host copy (the copies
// radg_batch_set_fetch queues behind the kernels).  The angle-weighted
// arrays are FP64; the per-path arrays are in the batch's compute type.
struct PinnedView {
  const double *wout = nullptr, *wgts = nullptr, *wjac = nullptr;
  const void *out = nullptr, *gts = nullptr, *jac = nullptr;
  int fp64 = 0;
  long long nwave = 0, npath = 0, L = 0, nparam = 0, ngas = 0;
};

// The stored per-path Jacobian [.. ][nparam] keeps its first ngas columns
// (the gases) per 1e20 molecules; this copies n values of it to FP64 per
// molecule.
inline void radg_jac_to_host(const void *src, int fp64, size_t n, int nparam, int ngas,
                             double *dst) {
  const double *d = static_cast<const double *>(src);
  const float *f = static_cast<const float *>(src);
  for (size_t i = 0, k = 0; i < n; i++) {
    const double x = fp64 ? d[i] : (double)f[i];
    dst[i] = (int)k < ngas ? x * 1e-20 : x;
    if (++k == (size_t)nparam) k = 0;
  }
}

// Records a failure: the message radg_last_error() returns on this thread.
int radg_fail_msg(int code, const char *msg);

// Fills *v for request r of a launched batch whose copies have completed.
int radg_batch_pinned_view(radg_batch *b, int r, PinnedView *v);

// Detaches the pinned copy from the batch; the block goes back to the
// pinned-memory cache when its last holder releases it.
std::shared_ptr<char> radg_batch_take_pinned(radg_batch *b);
