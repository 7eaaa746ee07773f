"""Serving loop: continuous batching, admission by device-memory budget,
overlap of in-flight batches, error isolation and cancellation.

The central check: every request's result is bit-for-bit the same computed
alone and computed inside a mix of uneven requests that arrive while other
batches are running (the kernels never let one request's data touch
another's, and every reduction has a fixed order).
"""
import threading
import time

import numpy as np
import pytest

from radtrancode_amd import serving, synth
from tests.golden.golden_io import load_cia
from tests.tolerance import TOL_JAC_MIXED, TOL_RAD_MIXED, jac_err, rad_err

pytestmark = pytest.mark.gpu

NPOINT = 400


@pytest.fixture(scope="module")
def tables(gpu_lib):
    t = serving.serving_tables(npoint=NPOINT, nkgas=4, ng=12)
    gpu_lib.register_ktables(t)
    return t


def _mix(n, seed):
    arr, reqs = serving.poisson_workload(n=n, rate=1e9, seed=seed, npoint=NPOINT,
                                         l_range=(40, 130), w_range=(20, NPOINT),
                                         g_range=(1, 4))
    return reqs


def _alone(gpu_lib, req, mode):
    b = gpu_lib.Batch([req], mode=mode)
    b.launch()
    r = b.read(0)
    b.close()
    return r


@pytest.mark.parametrize("mode", [0, 1], ids=["fp64", "mixed"])
def test_alone_equals_mixed_bitwise(gpu_lib, tables, mode):
    reqs = _mix(14, seed=11)
    print(f"serving mix of {len(reqs)} requests, mode {mode}", flush=True)
    srv = serving.Server(mode=mode, budget_bytes=0, max_batch=5, max_inflight=2)
    tickets = []

    def submit_staggered():
        for q in reqs:
            tickets.append(srv.submit(q))
            time.sleep(0.002)

    th = threading.Thread(target=submit_staggered)
    th.start()
    th.join()
    mixed = [srv.wait(t, timeout=60.0) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["steps"] >= 3
    for q, m in zip(reqs, mixed):
        assert m["status"] == 0
        a = _alone(gpu_lib, q, mode)
        for k in ("output", "gtsurf", "doutputdq", "wout", "wjac"):
            assert np.array_equal(a[k], m[k]), k


def test_serving_results_match_oracle(gpu_lib, tables):
    from oracle import oracle
    reqs = _mix(3, seed=12)
    srv = serving.Server(mode=1, max_batch=3)
    ts = [srv.submit(q) for q in reqs]
    res = [srv.wait(t, timeout=60.0) for t in ts]
    srv.close()
    for q, r in zip(reqs, res):
        sel = np.linspace(0, q.nwave - 1, 6).astype(int)
        o, g, d = oracle.spectrum(tables, load_cia(), q, vwave=q.vwave[sel], nthreads=8)
        assert rad_err(r["output"][sel], o) < TOL_RAD_MIXED
        assert jac_err(r["doutputdq"][sel], d) < TOL_JAC_MIXED


def test_budget_limits_inflight_memory(gpu_lib, tables):
    reqs = _mix(8, seed=13)
    sizes = [gpu_lib.batch_bytes([q], mode=1) for q in reqs]
    budget = int(max(sizes) * 1.5)
    srv = serving.Server(mode=1, budget_bytes=budget, max_batch=8, max_inflight=3)
    ts = [srv.submit(q) for q in reqs]
    res = [srv.wait(t, timeout=60.0) for t in ts]
    st = srv.stats()
    srv.close()
    assert all(r["status"] == 0 for r in res)
    assert st["peak_bytes"] <= budget
    assert st["steps"] >= 4


def test_bad_requests_fail_alone(gpu_lib, tables):
    reqs = _mix(4, seed=14)
    bad_t = synth.make_request(nlayer=50, nkgas=2, nwave=30, seed=3, table_vmin=600.0,
                               table_delv=0.16)
    bad_t.temp[7] = np.nan
    bad_imod = synth.make_request(nlayer=50, nkgas=2, nwave=30, seed=4, table_vmin=600.0,
                                  table_delv=0.16)
    bad_imod.imod[2] = 15
    srv = serving.Server(mode=1, budget_bytes=gpu_lib.batch_bytes([reqs[0]], mode=1) * 100)
    good = [srv.submit(q) for q in reqs[:2]]
    for bad, code in ((bad_t, 4), (bad_imod, 2)):
        with pytest.raises(gpu_lib.RadgError) as ei:
            srv.submit(bad)
        assert ei.value.code == code
    good += [srv.submit(q) for q in reqs[2:]]
    res = [srv.wait(t, timeout=60.0) for t in good]
    srv.close()
    assert [r["status"] for r in res] == [0, 0, 0, 0]


def test_malformed_requests_fail_before_any_copy(gpu_lib, tables):
    """Counts outside the arrdef.f limits and null arrays are refused at
    submit, before the server copies the request, and the server goes on."""
    import ctypes as C
    from radtrancode_amd._native import lib
    good = _mix(1, seed=15)[0]
    srv = serving.Server(mode=1)
    for field, value, code in (("nlayer", -1, 1), ("nlayer", 401, 1), ("ngas", 0, 1),
                               ("nwave", -5, 1), ("press", None, 4), ("layinc", None, 4)):
        keep, arr = gpu_lib._requests([good])
        setattr(arr[0], field, value)
        t = C.c_longlong()
        rc = lib().radg_server_submit(srv.h, C.cast(arr, C.c_void_p), C.byref(t))
        print(f"malformed {field}={value}: status {rc}", flush=True)
        assert rc == code, lib().radg_last_error()
    t = srv.submit(good)
    res = srv.wait(t, timeout=60.0)
    srv.close()
    assert res["status"] == 0


def test_cancel_before_admission(gpu_lib, tables):
    big = synth.make_request(nlayer=200, nkgas=4, nwave=NPOINT, seed=5, table_vmin=600.0,
                             table_delv=0.16)
    small = synth.make_request(nlayer=40, nkgas=1, nwave=50, seed=6, table_vmin=600.0,
                               table_delv=0.16)
    srv = serving.Server(mode=0, max_batch=1, max_inflight=1)
    ta = [srv.submit(big) for _ in range(3)]
    tb = srv.submit(small)
    srv.cancel(tb)
    rb = srv.wait(tb, timeout=60.0, outputs=False)
    ra = [srv.wait(t, timeout=60.0, outputs=False) for t in ta]
    srv.close()
    assert rb["status"] == 8  # RADG_ECANCELLED
    assert all(r["status"] == 0 for r in ra)
