"""Serving loop front end: many forward + Jacobian requests, continuously batched.

:class:`Server` wraps the library's persistent loop (``radg_server_*``,
include/radg.h): requests are admitted by device-memory budget into batches
while earlier batches are still running, one launch per batch, with the
k-tables and CIA table resident and shared.  :func:`poisson_workload` builds
the uneven request mix (layers, wavenumber windows, gas counts) and
:func:`run_poisson` drives it with Poisson arrivals and reports throughput
and latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from . import synth
from ._native import RadgError, check, lib
from .cirsradg import MODE_MIXED, _requests


class Server:
    def __init__(self, mode=MODE_MIXED, budget_bytes=0, max_batch=64, max_inflight=2,
                 per_path_outputs=True):
        h = C.c_void_p()
        check(lib().radg_server_create(int(mode), C.c_size_t(budget_bytes), int(max_batch),
                                       int(max_inflight), int(bool(per_path_outputs)),
                                       C.byref(h)))
        self.h = h
        self.per_path = bool(per_path_outputs)
        self._meta = {}
        self._lock = threading.Lock()

    def submit(self, req) -> int:
        """Queue one request; returns its ticket (validation errors raise)."""
        keep, arr = _requests([req])
        t = C.c_longlong()
        rc = lib().radg_server_submit(self.h, C.cast(arr, C.c_void_p), C.byref(t))
        with self._lock:
            self._meta[t.value] = (req.nwave, req.npath, req.nlayer, req.nparam)
        if rc != 0:
            raise RadgError(rc, lib().radg_last_error().decode())
        return t.value

    def cancel(self, ticket: int) -> None:
        check(lib().radg_server_cancel(self.h, C.c_longlong(ticket)))

    def wait(self, ticket: int, timeout=-1.0, outputs=True):
        """Results of a ticket: dict with status, latency_s and the arrays."""
        with self._lock:
            nw, npath, L, npar = self._meta.pop(ticket)
        res = {}
        ptr = lambda a: C.c_void_p(a.ctypes.data)
        if outputs:
            res = dict(wout=np.zeros(nw), wgts=np.zeros(nw), wjac=np.zeros((nw, L, npar)))
            args = [None, None, None]
            if self.per_path:
                res.update(output=np.zeros((nw, npath)), gtsurf=np.zeros((nw, npath)),
                           doutputdq=np.zeros((nw, npath, L, npar)))
                args = [ptr(res[k]) for k in ("output", "gtsurf", "doutputdq")]
            args += [ptr(res[k]) for k in ("wout", "wgts", "wjac")]
        else:
            args = [None] * 6
        st = C.c_int()
        lat = C.c_double()
        rc = lib().radg_server_wait(self.h, C.c_longlong(ticket), C.c_double(timeout), *args,
                                    C.byref(st), C.byref(lat))
        if rc != 0:
            with self._lock:
                self._meta[ticket] = (nw, npath, L, npar)
            raise RadgError(rc, "wait failed")
        res.update(status=st.value, latency_s=lat.value)
        return res

    def stats(self):
        s = np.zeros(9, dtype=np.int64)
        check(lib().radg_server_stats(self.h, C.c_void_p(s.ctypes.data)))
        return dict(steps=int(s[0]), admitted=int(s[1]), completed=int(s[2]),
                    max_batch=int(s[3]), peak_bytes=int(s[4]), queued=int(s[5]),
                    create_s=s[6] * 1e-6, retire_s=s[7] * 1e-6, gpu_s=s[8] * 1e-6)

    def close(self):
        if getattr(self, "h", None):
            lib().radg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serving_tables(npoint=5000, nkgas=6, ng=20, v0=600.0, dv=0.16, seed=2):
    """Shared resident k-tables for the serving mix: ``nkgas`` gases on one
    ``npoint``-point wavenumber grid (requests take contiguous windows)."""
    return synth.make_ktables(synth.KGAS_IDS[:nkgas], v0, dv, npoint, ng=ng, seed=seed)


def poisson_workload(n=4096, rate=200.0, seed=5, npoint=5000, l_range=(40, 200),
                     w_range=(200, 5000), g_range=(1, 6), v0=600.0, dv=0.16):
    """``n`` requests with Poisson arrival times (``rate`` per second).

    Per request: layers ~ U[l_range], wavenumbers ~ U[w_range] taken as a
    contiguous window of the shared grid, k-table gases ~ U[g_range] plus
    H2/He, a seeded atmosphere.  Returns (arrival_times, requests).
    """
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for i in range(n):
        L = int(rng.integers(l_range[0], l_range[1] + 1))
        nw = int(rng.integers(w_range[0], min(w_range[1], npoint) + 1))
        nk = int(rng.integers(g_range[0], g_range[1] + 1))
        off = int(rng.integers(0, npoint - nw + 1))
        reqs.append(synth.make_request(nlayer=L, nkgas=nk, nwave=nw, seed=1000 + i,
                                       wave_offset=off, table_vmin=v0, table_delv=dv))
    return arrivals, reqs


def run_poisson(server: Server, arrivals, reqs, collect=None):
    """Submit each request at its arrival time; returns per-request latency
    (s), wall time and the requests' total work in units of the 120-layer x
    1000-wavenumber, 3 k-gas spectrum: nwave/1000 x nlayer/120 x (0.33 +
    0.335 r), where a request with k k-gases has r = max(k - 1, 0) overlap
    rounds.  The weights are the measured costs per (wavenumber, layer) at
    the 120-layer workload: the layer kernel's interpolation and stores
    (654 VALU instructions, 2.46 ms of a 16-spectrum step) and each overlap
    round (about 1390 instructions, 5.24 ms), with the emission kernel
    (2.63 ms) counted with the fixed part, over the 15.57 ms step."""
    n = len(reqs)
    tickets = [None] * n
    t0 = time.perf_counter()

    def submitter():
        for i in range(n):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets[i] = server.submit(reqs[i])

    th = threading.Thread(target=submitter)
    th.start()
    lat = np.zeros(n)
    status = np.zeros(n, dtype=np.int64)
    for i in range(n):
        while tickets[i] is None:
            time.sleep(1e-4)
        r = server.wait(tickets[i], outputs=collect is not None and i in collect)
        lat[i] = r["latency_s"]
        status[i] = r["status"]
        if collect is not None and i in collect:
            collect[i] = r
    th.join()
    wall = time.perf_counter() - t0
    return dict(latency=lat, status=status, wall=wall, work=work_units(reqs))


def work_units(reqs):
    """Total work of requests in units of the 120-layer x 1000-wavenumber,
    3 k-gas spectrum (the cost model run_poisson describes)."""
    def units(q):
        nk = int(sum(1 for g in q.idgas if g not in (synth.H2_ID, synth.HE_ID)))
        return q.nwave / 1000.0 * q.nlayer / 120.0 * (0.33 + 0.335 * max(nk - 1, 0))

    return float(sum(units(q) for q in reqs))
