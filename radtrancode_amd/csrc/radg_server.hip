// Persistent serving loop for forward + Jacobian requests.
//
// A worker thread owns up to `max_inflight` device batches, each on its own
// non-blocking HIP stream.  Submitted requests are deep-copied, validated
// and sized at submit time, then queued.  Every pass of the worker retires
// the batches whose completion event has fired (results are copied into
// each request's ticket), then admits queued requests in arrival order into
// a new batch while the sum of in-flight batch bytes stays within the HBM
// budget.  Requests that arrive while earlier batches run are therefore
// admitted as soon as a stream slot and memory free up, and their kernels
// overlap the running ones on the device.  One launch per batch per step.
//
// This replaces the reference's process-pool pattern
// (nemesisPY/nemesis.py:4586-4664: independent forward-model subprocesses that
// communicate through files) for many independent solves.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/radg.h"
#include "radg_internal.h"

namespace {

using clk = std::chrono::steady_clock;

// The counts and pointers OwnedRequest's copy reads: each count within the
// arrdef.f limits (check_limits.f) and every array it copies present.
int request_shape(const radg_request &s, std::string &err) {
  char buf[160];
  auto bad = [&](int code, const char *fmt, long long v) {
    std::snprintf(buf, sizeof buf, fmt, v);
    err = buf;
    return code;
  };
  if (s.nlayer < 1 || s.nlayer > RADG_MAXLAY) return bad(RADG_ELIMITS, "nlayer=%lld outside 1..400", s.nlayer);
  if (s.npath < 1 || s.npath > RADG_MAXPAT) return bad(RADG_ELIMITS, "npath=%lld outside 1..200", s.npath);
  if (s.ngas < 1 || s.ngas > RADG_MAXGAS) return bad(RADG_ELIMITS, "ngas=%lld outside 1..30", s.ngas);
  if (s.nsw < 0 || s.nsw > RADG_MAXGAS + 2 + RADG_MAXCON) return bad(RADG_ELIMITS, "nsw=%lld", s.nsw);
  if (s.nwave < 1 || s.nwave > (1 << 24)) return bad(RADG_ELIMITS, "nwave=%lld", s.nwave);
  if (s.nem < 1 || s.nem > RADG_MAXSEC) return bad(RADG_ELIMITS, "nem=%lld outside 1..16000", s.nem);
  const void *need[] = {s.press, s.temp, s.delh, s.pp, s.amount, s.idgas, s.isogas, s.nlayin, s.imod,
                        s.layinc, s.scale, s.emtemp, s.vwave, s.vem, s.emissivity};
  for (size_t i = 0; i < sizeof need / sizeof need[0]; i++)
    if (!need[i]) return bad(RADG_EINPUT, "request array %lld is null", (long long)i + 1);
  if (s.nsw > 0 && !s.isw) return bad(RADG_EINPUT, "isw is null with nsw=%lld", s.nsw);
  return RADG_OK;
}

// Deep copy of one radg_request, made once request_shape accepts it.
struct OwnedRequest {
  radg_request q{};
  std::vector<double> press, temp, delh, pp, amount, scale, emtemp, vwave, vem, emis, pw;
  std::vector<double> cont, vsec, xsec, hfp;
  std::vector<int> idgas, isogas, nlayin, imod, layinc, isw;

  explicit OwnedRequest(const radg_request &s) {
    q = s;
    const size_t L = s.nlayer, G = s.ngas, P = s.npath;
    press.assign(s.press, s.press + L);
    temp.assign(s.temp, s.temp + L);
    delh.assign(s.delh, s.delh + L);
    pp.assign(s.pp, s.pp + L * G);
    amount.assign(s.amount, s.amount + L * G);
    idgas.assign(s.idgas, s.idgas + G);
    isogas.assign(s.isogas, s.isogas + G);
    nlayin.assign(s.nlayin, s.nlayin + P);
    imod.assign(s.imod, s.imod + P);
    layinc.assign(s.layinc, s.layinc + P * L);
    scale.assign(s.scale, s.scale + P * L);
    emtemp.assign(s.emtemp, s.emtemp + P * L);
    isw.assign(s.isw, s.isw + s.nsw);
    vwave.assign(s.vwave, s.vwave + s.nwave);
    vem.assign(s.vem, s.vem + s.nem);
    emis.assign(s.emissivity, s.emissivity + s.nem);
    if (s.path_weight) pw.assign(s.path_weight, s.path_weight + P);
    else pw.assign(P, 0.0);
    q.press = press.data();
    q.temp = temp.data();
    q.delh = delh.data();
    q.pp = pp.data();
    q.amount = amount.data();
    q.idgas = idgas.data();
    q.isogas = isogas.data();
    q.nlayin = nlayin.data();
    q.imod = imod.data();
    q.layinc = layinc.data();
    q.scale = scale.data();
    q.emtemp = emtemp.data();
    q.isw = isw.data();
    q.vwave = vwave.data();
    q.vem = vem.data();
    q.emissivity = emis.data();
    q.path_weight = pw.data();
    // dust arrays are copied only when their sizes are in range; otherwise the
    // pointers stay as given and validation at submit rejects the request
    if (s.ncont > 0 && s.ncont <= RADG_MAXCON && s.nsec > 0 && s.nsec <= RADG_MAXSEC && s.cont &&
        s.vsec && s.xsec) {
      cont.assign(s.cont, s.cont + L * s.ncont);
      vsec.assign(s.vsec, s.vsec + s.nsec);
      xsec.assign(s.xsec, s.xsec + (size_t)s.ncont * s.nsec);
      q.cont = cont.data();
      q.vsec = vsec.data();
      q.xsec = xsec.data();
    } else {
      q.cont = q.vsec = q.xsec = nullptr;
    }
    if (s.flagh2p == 1 && s.hfp && s.nlayer > 0 && s.nlayer <= RADG_MAXLAY) {
      hfp.assign(s.hfp, s.hfp + L);
      q.hfp = hfp.data();
    } else {
      q.hfp = nullptr;
    }
  }
};

struct Ticket {
  long long id = 0;
  std::unique_ptr<OwnedRequest> req;
  size_t bytes = 0;
  int status = -1;  // -1 pending, else a RADG status
  bool cancelled = false;
  std::string err;
  // results: a view into the pinned copy of the finished batch, which the
  // tickets of that batch share until the last of them is collected
  std::shared_ptr<char> block;
  PinnedView view;
  clk::time_point t_submit, t_admit, t_done;
};

struct Inflight {
  radg_batch *b = nullptr;
  hipStream_t stream = nullptr;
  std::vector<long long> ids;
  size_t bytes = 0;
};

}  // namespace

struct radg_server {
  int mode = RADG_MODE_MIXED;
  size_t budget = 0;
  int max_batch = 64;
  int max_inflight = 2;
  int fetch = 1;  // bit 1: also keep the per-path outputs
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<long long> queue;
  std::map<long long, std::shared_ptr<Ticket>> tickets;
  std::vector<Inflight> slots;
  std::atomic<bool> running{true};
  std::thread worker;
  long long next_id = 1;
  size_t inflight_bytes = 0;
  long long steps = 0, admitted = 0, completed = 0, max_batch_seen = 0, peak_bytes = 0;
  // worker time (microseconds): packing + upload + launch, retiring, kernels
  long long us_create = 0, us_retire = 0, us_gpu = 0;
  int device = 0;
};

namespace {

void retire(radg_server *s, Inflight &f) {
  const auto t0 = clk::now();
  float ms3[3] = {0, 0, 0};
  if (radg_batch_timings(f.b, ms3) == RADG_OK) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->us_gpu += (long long)(1e3 * (ms3[0] + ms3[1] + ms3[2]));
  }
  // results for every request of a finished batch: views into the batch's
  // pinned copy, which outlives the batch's device memory
  std::vector<PinnedView> views(f.ids.size());
  std::vector<int> rcs(f.ids.size());
  std::vector<std::string> errs(f.ids.size());
  for (size_t r = 0; r < f.ids.size(); r++) {
    rcs[r] = radg_batch_pinned_view(f.b, (int)r, &views[r]);
    if (rcs[r]) errs[r] = radg_last_error();
  }
  std::shared_ptr<char> block = radg_batch_take_pinned(f.b);
  for (size_t r = 0; r < f.ids.size(); r++) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(f.ids[r]);
    if (it == s->tickets.end()) continue;
    Ticket &t = *it->second;
    t.status = rcs[r];
    if (rcs[r]) t.err = errs[r];
    else {
      t.block = block;
      t.view = views[r];
    }
    t.t_done = clk::now();
    t.req.reset();  // inputs no longer needed
    s->completed++;
  }
  radg_batch_destroy(f.b);
  f.b = nullptr;
  f.ids.clear();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->inflight_bytes -= f.bytes;
    s->us_retire += std::chrono::duration_cast<std::chrono::microseconds>(clk::now() - t0).count();
  }
  f.bytes = 0;
  s->cv_done.notify_all();
}

void worker_loop(radg_server *s) {
  (void)hipSetDevice(s->device);
  Inflight ready;  // a batch packed ahead of a free slot (no stream yet)
  while (s->running.load()) {
    bool progressed = false;
    // retire finished batches
    for (auto &f : s->slots) {
      if (f.b && hipStreamQuery(f.stream) == hipSuccess) {
        retire(s, f);
        progressed = true;
      }
    }
    // pack the next batch from the queue (also while every slot is busy, so
    // that a freed slot is refilled at once)
    if (!ready.b) {
      std::vector<std::shared_ptr<Ticket>> take;
      size_t bytes = 0;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        while (!s->queue.empty() && (int)take.size() < s->max_batch) {
          auto it = s->tickets.find(s->queue.front());
          if (it == s->tickets.end() || it->second->cancelled) {
            s->queue.pop_front();
            continue;
          }
          const size_t need = it->second->bytes;
          if (s->budget && s->inflight_bytes + bytes + need > s->budget) break;
          take.push_back(it->second);
          bytes += need;
          s->queue.pop_front();
        }
        s->inflight_bytes += bytes;
        s->peak_bytes = std::max<long long>(s->peak_bytes, (long long)s->inflight_bytes);
      }
      if (!take.empty()) {
        std::vector<radg_request> qs;
        for (auto &t : take) qs.push_back(t->req->q);
        radg_batch *nb = nullptr;
        const auto tc0 = clk::now();
        int rc = radg_batch_create((int)qs.size(), qs.data(), s->mode, 0, &nb);
        if (rc == RADG_OK) rc = radg_batch_set_fetch(nb, 1 | (s->fetch & 2));
        const auto now = clk::now();
        if (rc != RADG_OK) {
          const std::string msg = radg_last_error();
          if (nb) radg_batch_destroy(nb);
          std::lock_guard<std::mutex> lk(s->mu);
          for (auto &t : take) {
            t->status = rc;
            t->err = msg;
            t->t_done = now;
          }
          s->inflight_bytes -= bytes;
          s->cv_done.notify_all();
        } else {
          ready.b = nb;
          ready.bytes = bytes;
          for (auto &t : take) ready.ids.push_back(t->id);
          std::lock_guard<std::mutex> lk(s->mu);
          s->us_create += std::chrono::duration_cast<std::chrono::microseconds>(now - tc0).count();
        }
        progressed = true;
      }
    }
    // launch the packed batch into a free slot
    Inflight *slot = nullptr;
    for (auto &f : s->slots)
      if (!f.b) {
        slot = &f;
        break;
      }
    if (slot && ready.b) {
      const int rc = radg_batch_launch(ready.b, slot->stream);
      const auto now = clk::now();
      std::lock_guard<std::mutex> lk(s->mu);
      if (rc != RADG_OK) {
        const std::string msg = radg_last_error();
        for (long long id : ready.ids) {
          auto it = s->tickets.find(id);
          if (it == s->tickets.end()) continue;
          it->second->status = rc;
          it->second->err = msg;
          it->second->t_done = now;
        }
        radg_batch_destroy(ready.b);
        s->inflight_bytes -= ready.bytes;
        s->cv_done.notify_all();
      } else {
        slot->b = ready.b;
        slot->bytes = ready.bytes;
        slot->ids = ready.ids;
        for (long long id : ready.ids) {
          auto it = s->tickets.find(id);
          if (it != s->tickets.end()) it->second->t_admit = now;
        }
        s->steps++;
        s->admitted += (long long)ready.ids.size();
        s->max_batch_seen = std::max<long long>(s->max_batch_seen, (long long)ready.ids.size());
      }
      ready.b = nullptr;
      ready.ids.clear();
      ready.bytes = 0;
      progressed = true;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(s->mu);
      bool busy = false;
      for (auto &f : s->slots) busy |= (f.b != nullptr);
      if (busy || ready.b || !s->queue.empty())
        s->cv_work.wait_for(lk, std::chrono::microseconds(50));
      else
        s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
  for (auto &f : s->slots)
    if (f.b) {
      (void)hipStreamSynchronize(f.stream);
      retire(s, f);
    }
  if (ready.b) {  // packed but never launched: fail its requests
    std::lock_guard<std::mutex> lk(s->mu);
    for (long long id : ready.ids) {
      auto it = s->tickets.find(id);
      if (it == s->tickets.end()) continue;
      it->second->status = RADG_ECANCELLED;
      it->second->err = "server destroyed before launch";
      it->second->t_done = clk::now();
    }
    radg_batch_destroy(ready.b);
    s->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

int radg_server_create(int mode, size_t budget_bytes, int max_batch, int max_inflight,
                       int per_path_outputs, radg_server **out) {
  *out = nullptr;
  if (mode != RADG_MODE_FP64 && mode != RADG_MODE_MIXED && mode != RADG_MODE_BF16 && mode != RADG_MODE_F16)
    return RADG_EINPUT;
  if (max_batch < 1 || max_inflight < 1 || max_inflight > 16) return RADG_EINPUT;
  auto *s = new radg_server;
  s->mode = mode;
  s->fetch = per_path_outputs ? 3 : 1;
  s->budget = budget_bytes;
  s->max_batch = max_batch;
  s->max_inflight = max_inflight;
  (void)hipGetDevice(&s->device);
  s->slots.resize(max_inflight);
  for (auto &f : s->slots) {
    if (hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking) != hipSuccess) {
      delete s;
      return RADG_EHIP;
    }
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return RADG_OK;
}

int radg_server_submit(radg_server *s, const radg_request *req, long long *ticket) {
  if (!s || !req || !ticket) return RADG_EINPUT;
  auto t = std::make_shared<Ticket>();
  t->t_submit = clk::now();
  size_t bytes = 0;
  // the sizes and pointers the deep copy reads are checked first, so a
  // malformed request fails alone instead of reading past the caller's arrays
  std::string bad;
  int rc = request_shape(*req, bad);
  if (rc == RADG_OK) {
    t->req.reset(new OwnedRequest(*req));
    rc = radg_batch_bytes(1, &t->req->q, s->mode, &bytes);
  } else {
    radg_fail_msg(rc, bad.c_str());
  }
  std::lock_guard<std::mutex> lk(s->mu);
  t->id = s->next_id++;
  *ticket = t->id;
  s->tickets[t->id] = t;
  if (rc != RADG_OK) {
    t->status = rc;
    t->err = radg_last_error();
    t->t_done = t->t_submit;
    return rc;
  }
  if (s->budget && bytes > s->budget) {
    t->status = RADG_EBUDGET;
    t->err = "request needs more device memory than the whole budget";
    t->t_done = t->t_submit;
    return RADG_EBUDGET;
  }
  t->bytes = bytes;
  s->queue.push_back(t->id);
  s->cv_work.notify_one();
  return RADG_OK;
}

int radg_server_cancel(radg_server *s, long long ticket) {
  if (!s) return RADG_EINPUT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return RADG_EINPUT;
  bool queued = false;
  for (long long id : s->queue) queued |= (id == ticket);
  if (queued && it->second->status == -1) {
    it->second->cancelled = true;
    it->second->status = RADG_ECANCELLED;
    it->second->err = "cancelled before admission";
    it->second->t_done = clk::now();
    s->cv_done.notify_all();
  }
  return RADG_OK;
}

int radg_server_wait(radg_server *s, long long ticket, double timeout_s, double *output,
                     double *gtsurf, double *doutputdq, double *wout, double *wgts, double *wjac,
                     int *status, double *latency_s) {
  if (!s) return RADG_EINPUT;
  std::shared_ptr<Ticket> t;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return RADG_EINPUT;
    t = it->second;
    const auto deadline = clk::now() + std::chrono::duration_cast<clk::duration>(
                                           std::chrono::duration<double>(timeout_s));
    while (t->status == -1) {
      if (timeout_s >= 0 && clk::now() >= deadline) return RADG_ETIMEOUT;
      s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    }
    s->tickets.erase(it);
  }
  if (status) *status = t->status;
  if (latency_s) *latency_s = std::chrono::duration<double>(t->t_done - t->t_submit).count();
  if (t->status != RADG_OK) return RADG_OK;
  const PinnedView &v = t->view;
  const size_t nw = (size_t)v.nwave, nout = nw * v.npath;
  const size_t nwj = nw * v.L * v.nparam, nj = nout * v.L * v.nparam;
  auto cpw = [](double *dst, const double *src, size_t n) {
    if (dst && src) std::memcpy(dst, src, n * sizeof(double));
  };
  auto cpp = [&](double *dst, const void *src, size_t n) {
    if (!dst || !src) return;
    if (v.fp64) std::memcpy(dst, src, n * sizeof(double));
    else {
      const float *f = static_cast<const float *>(src);
      for (size_t i = 0; i < n; i++) dst[i] = f[i];
    }
  };
  cpw(wout, v.wout, nw);
  cpw(wgts, v.wgts, nw);
  cpw(wjac, v.wjac, nwj);
  cpp(output, v.out, nout);
  cpp(gtsurf, v.gts, nout);
  if (doutputdq && v.jac)
    radg_jac_to_host(v.jac, v.fp64, nj, (int)v.nparam, (int)v.ngas, doutputdq);
  return RADG_OK;
}

int radg_server_stats(radg_server *s, long long *stats9) {
  if (!s || !stats9) return RADG_EINPUT;
  std::lock_guard<std::mutex> lk(s->mu);
  stats9[0] = s->steps;
  stats9[1] = s->admitted;
  stats9[2] = s->completed;
  stats9[3] = s->max_batch_seen;
  stats9[4] = s->peak_bytes;
  stats9[5] = (long long)s->queue.size();
  stats9[6] = s->us_create;
  stats9[7] = s->us_retire;
  stats9[8] = s->us_gpu;
  return RADG_OK;
}

int radg_server_destroy(radg_server *s) {
  if (!s) return RADG_OK;
  s->running = false;
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto &f : s->slots)
    if (f.stream) (void)hipStreamDestroy(f.stream);
  delete s;
  return RADG_OK;
}

}  // extern "C"
