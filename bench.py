#!/usr/bin/env python
"""Jacobian-spectra/s of the CIRSRADG_WAVE path on one MI355X.

Workload (default): the mixed-precision form at 120 layers x 1000
wavenumbers (600-1400 cm-1) x 20 g-ordinates, 3 k-table gases + H2/He CIA,
5 emission-angle paths.  One step = one device-resident batch of ``--batch``
independent spectra (seeded atmospheres, shared resident k-tables), i.e. the
layer/overlap kernel, the emission kernel and the angle quadrature over every
wavenumber of every spectrum.  The plain FP64 form is timed on the same batch.

Prints one JSON line (see the keys below).  ``--gpus N`` with N > 1 exits:
this build runs on one GPU.
"""
import argparse
import copy
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def make_batch_inputs(nspec, nwave=1000, nlayer=120, nkgas=3, ng=20):
    from radtrancode_amd import synth

    tables, cia, req0 = synth.config1(nwave=nwave, nlayer=nlayer, nkgas=nkgas, ng=ng)
    reqs = [req0] + [synth.make_request(nlayer=nlayer, nkgas=nkgas, nwave=nwave, seed=100 + i)
                     for i in range(1, nspec)]
    return tables, cia, reqs


def time_batch(rad, reqs, mode, steps, warmup, torch):
    b = rad.Batch(reqs, mode=mode)
    st = torch.cuda.current_stream()
    for _ in range(warmup):
        b.launch(st)
    torch.cuda.synchronize()
    n0, s0 = b.timing_stats()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(st)
        b.launch(st)
        ev[i][1].record(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n1, s1 = b.timing_stats()
    lat = np.array([a.elapsed_time(z) for a, z in ev])
    kern = (s1 - s0) / max(1, n1 - n0) if n1 - n0 == steps else s1 / max(1, n1)
    return b, t1 - t0, lat, kern


def cpu_baseline(tables, cia, req, seconds):
    """The reference's own cirsradg_wave (REAL(8) build) if it was built,
    else the oracle restatement, on one host core over the first wavenumbers
    of one spectrum until about ``seconds`` of CPU time."""
    from oracle import ref

    n = 8
    kind = "reference" if ref.available(64) else "port"
    total_t, total_n = 0.0, 0
    while total_t < seconds and total_n < req.nwave:
        n = min(n, req.nwave - total_n)
        ts = [copy.copy(t) for t in tables]
        for t, t0 in zip(ts, tables):
            t.k = t0.k[total_n:total_n + n].copy()
            t.vmin = t0.vmin + total_n * t0.delv
        r2 = copy.copy(req)
        r2.vwave = req.vwave[total_n:total_n + n].copy()
        if kind == "reference":
            rr = ref.RefRunner(ts, cia, r2, prec=64)
            t0 = time.perf_counter()
            rr.spectrum()
            dt = time.perf_counter() - t0
            rr.close()
        else:
            from oracle import oracle
            t0 = time.perf_counter()
            oracle.spectrum(ts, cia, r2, nthreads=1)
            dt = time.perf_counter() - t0
        total_t += dt
        total_n += n
        n = min(4 * n, 256)
    value = (total_n / req.nwave) / total_t
    what = ("reference cirsradg_wave, REAL(8) build, called once per wavenumber with .kta "
            "files on the host page cache" if kind == "reference" else "FP64 C restatement")
    return dict(value=value, unit="Jacobian-spectra/s", cores=1, kind=kind,
                sample=f"{total_n} of {req.nwave} wavenumbers of one 120-layer x 20 g spectrum "
                       f"({what}) in {total_t:.1f} s")


_CPU_JOB = {}


def _cpu_worker(job):
    """One host process of the parallel CPU baseline: the reference's
    cirsradg_wave over consecutive wavenumbers of the spectrum from its own
    starting point (wrapping), chunks of up to 64, until ``seconds`` of
    compute; returns (wavenumbers, seconds)."""
    start, seconds = job
    tables, cia, req = _CPU_JOB["inputs"]
    from oracle import ref

    nw = req.nwave
    pos, n, done, used = start, 8, 0, 0.0
    while used < seconds:
        n = min(n, nw - pos)
        ts = [copy.copy(t) for t in tables]
        for t, t0 in zip(ts, tables):
            t.k = t0.k[pos:pos + n].copy()
            t.vmin = t0.vmin + pos * t0.delv
        r2 = copy.copy(req)
        r2.vwave = req.vwave[pos:pos + n].copy()
        rr = ref.RefRunner(ts, cia, r2, prec=64)
        t0 = time.perf_counter()
        rr.spectrum()
        used += time.perf_counter() - t0
        rr.close()
        done += n
        pos = (pos + n) % nw
        n = min(2 * n, 64)
    return done, used


def cpu_baseline_parallel(tables, cia, req, seconds, nproc):
    """The reference's cirsradg_wave (REAL(8) build) in ``nproc``
    independent host processes at once, as NEMESIS runs independent
    retrievals side by side; forked before the GPU is touched.  The rate is
    the sum of the processes' own rates."""
    import multiprocessing as mp

    from oracle import ref

    if not ref.available(64):
        return None
    _CPU_JOB["inputs"] = (tables, cia, req)
    jobs = [(int(w * req.nwave // nproc), seconds) for w in range(nproc)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(nproc) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    _CPU_JOB.clear()
    value = sum((d / req.nwave) / u for d, u in res)
    return dict(value=value, unit="Jacobian-spectra/s", cores=nproc, kind="reference",
                sample=f"{nproc} processes, {sum(d for d, _ in res)} wavenumbers of one "
                       f"120-layer x 20 g spectrum in total (reference cirsradg_wave, REAL(8) "
                       f"build, once per wavenumber), about {seconds:.0f} s of compute each, "
                       f"{wall:.1f} s wall")


def kernel_source_hash():
    """SHA-256 (first 16 hex digits) of the layer kernel's sources: PMC
    counters are merged into the bench line only when they were taken on
    the same kernel code."""
    import hashlib

    h = hashlib.sha256()
    for name in ("radg_kernels.hip", "radg_dev.h"):
        with open(os.path.join(ROOT, "radtrancode_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def network_op_count(ng, nkgas):
    """The overlap's sort arithmetic per (wavenumber, layer) item, from the
    algorithm alone: each of the nkgas - 1 random-overlap rounds sorts 64 E
    slots (E per lane, 64 E >= ng^2) with a bitonic network whose in-lane
    merge levels a lane's already sorted row skips; every stage is
    64 E / 2 compare-exchanges, each a min and a max."""
    import math

    E = 1
    while 64 * E < ng * ng or ng * (-(-ng // E)) > 64:
        E *= 2
    levels, skip = int(math.log2(64 * E)), int(math.log2(E))
    stages = levels * (levels + 1) // 2 - skip * (skip + 1) // 2
    ce = stages * 64 * E // 2 * max(0, nkgas - 1)
    return {"slots": 64 * E, "network_stages_per_round": stages,
            "compare_exchanges_per_item": ce,
            "minmax_wave_instructions_per_item": 2 * ce / 64}


def pmc_figures(batch, t_layer_s):
    """HBM bytes per launch and VALU issue of the layer kernel from the
    committed rocprofv3 PMC passes (profiles/layer_kernel_traffic.json), when
    they were taken on this batch size and on these kernel sources; the
    VALU rate uses this run's kernel time."""
    path = os.path.join(ROOT, "profiles", "layer_kernel_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        p = json.load(f)
    if p.get("spectra_per_launch") != batch:
        return {}
    if p.get("kernel_source_sha256") != kernel_source_hash():
        return {"traffic": None,
                "traffic_source": "not merged: the committed PMC passes were taken on other "
                                  "kernel sources (" + str(p.get("kernel_source_sha256")) + ")"}
    traffic = (p["fetch_kib"] * p["fetch_scale"] + p["write_kib"] * p["write_scale"]) * 1024
    rate = p["valu_wave_instructions"] / t_layer_s
    # the operation count beside the issue rate: the sort's compare-exchange
    # arithmetic (20 g-ordinates, 3 k-gases) against the issued instructions
    ops = network_op_count(20, 3)
    if p.get("valu_per_item"):
        ops["frac_of_issued_valu"] = ops["minmax_wave_instructions_per_item"] / p["valu_per_item"]
    return {"traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (" + p["source"] + ")",
            "kernel_source_sha256": p["kernel_source_sha256"],
            "valu_issue": {"wave_instructions_per_launch": p["valu_wave_instructions"],
                           "instructions_per_item": p.get("valu_per_item"),
                           "per_item_breakdown": p.get("valu_per_item_breakdown"),
                           "achieved_per_s": rate, "peak_per_s": p["valu_issue_peak_per_s"],
                           "frac": rate / p["valu_issue_peak_per_s"],
                           "peak_note": "nominal: one wave64 VALU instruction per 4 cycles per "
                                        "SIMD at 2.4 GHz; not a ceiling for every mix (the FP64 "
                                        "form's layer kernel issues 656 G/s at 2.33 GHz)"},
            "op_count": ops}


def layer_sweep():
    """One batch per k-gas count (1, 2, 3: no overlap round, one, two) of 4
    spectra x 200 wavenumbers x 120 layers x 20 g, launched once in the
    mixed and the FP64 form.  Under rocprofv3 --pmc the layer-kernel
    dispatches give the instructions of the interpolation and of each
    overlap round per (wavenumber, layer) item."""
    import torch

    from radtrancode_amd import cirsradg as rad
    from radtrancode_amd import synth

    torch.cuda.set_device(0)
    for nk in (1, 2, 3):
        tables, cia, req0 = synth.config1(nwave=200, nlayer=120, nkgas=nk, ng=20)
        rad.register_ktables(tables)
        rad.register_cia(cia)
        reqs = [req0] + [synth.make_request(nlayer=120, nkgas=nk, nwave=200, seed=100 + i)
                         for i in range(1, 4)]
        for mode in (rad.MODE_MIXED, rad.MODE_FP64):
            b = rad.Batch(reqs, mode=mode)
            b.launch()
            b.sync()
            b.close()
        print(f"layer sweep: {nk} k-gases done", flush=True)


def weighted(out, dq, pw):
    """Angle-weighted radiance and Jacobian from per-path results."""
    return np.einsum("wp,p->w", out, pw), np.einsum("wplq,p->wlq", dq, pw)


def config5_run(rad, n, cia, nthreads):
    """The large-spectrum stress: ``n`` requests of 200 layers x 10000
    wavenumbers (0.1 cm-1) with 8 k-gases + H2/He, k-tables on a 30 P x 15 T
    x 20 g grid stored as BF16 log10(k), all submitted at once to the serving
    loop (device-memory budget admission).  The BF16-storage error of request
    0 is measured against the FP64 oracle on a sample of wavenumbers."""
    from oracle import oracle
    from radtrancode_amd import serving, synth

    t0 = time.perf_counter()
    tabs = synth.config5_tables()
    rad.register_ktables(tabs)
    t_reg = time.perf_counter() - t0
    reqs = [synth.config5_request(seed=5000 + i) for i in range(n)]
    budget = 96 << 30
    srv = serving.Server(mode=rad.MODE_BF16, budget_bytes=budget, max_batch=16, max_inflight=2,
                         per_path_outputs=False)
    t0 = time.perf_counter()
    tickets = [srv.submit(r) for r in reqs]
    lat, res0 = [], None
    for i, t in enumerate(tickets):
        r = srv.wait(t, outputs=(i == 0))
        lat.append(r["latency_s"])
        if i == 0:
            res0 = r
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    sample = np.arange(0, reqs[0].nwave, 500)
    o_ref, _, d_ref = oracle.spectrum(tabs, cia, reqs[0], vwave=reqs[0].vwave[sample],
                                      nthreads=nthreads)
    wo_ref, wj_ref = weighted(o_ref, d_ref, reqs[0].path_weight)
    wo, wj = res0["wout"][sample], res0["wjac"][sample]
    rad_e = float(np.max(np.abs(wo - wo_ref) / np.abs(wo_ref)))
    sc = np.max(np.abs(wj_ref), axis=1, keepdims=True)
    jac_e = float(np.max(np.abs(wj - wj_ref) / np.where(sc > 0, sc, 1.0)))
    # the same request with the IEEE-half codes
    bh = rad.Batch([reqs[0]], mode=rad.MODE_F16)
    bh.launch()
    rh = bh.read(0)
    bh.close()
    wo_h, wj_h = rh["wout"][sample], rh["wjac"][sample]
    rad_h = float(np.max(np.abs(wo_h - wo_ref) / np.abs(wo_ref)))
    jac_h = float(np.max(np.abs(wj_h - wj_ref) / np.where(sc > 0, sc, 1.0)))
    lat_ms = np.array(lat) * 1e3
    del tabs
    return {
        "requests": n, "wall_s": wall, "requests_per_s": n / wall,
        "latency_ms_p50": float(np.percentile(lat_ms, 50)),
        "latency_ms_p99": float(np.percentile(lat_ms, 99)),
        "batches": st["steps"], "largest_batch": st["max_batch"],
        "peak_inflight_bytes": st["peak_bytes"], "budget_bytes": budget,
        "worker_s": {"pack_upload_launch": st["create_s"], "retire": st["retire_s"],
                     "kernels": st["gpu_s"]},
        "table_setup_s": t_reg,
        "bf16_storage_error_vs_fp64_oracle": {
            "radiance_rel": rad_e, "jacobian_col_scaled": jac_e,
            "sample": f"{len(sample)} wavenumbers of request 0, angle-weighted"},
        "f16_storage_error_vs_fp64_oracle": {
            "radiance_rel": rad_h, "jacobian_col_scaled": jac_h,
            "sample": f"{len(sample)} wavenumbers of request 0 with IEEE-half codes, angle-weighted"},
        "shape": "200 layers x 10000 wavenumbers (0.1 cm-1), 8 k-gases + H2/He, 30 P x 15 T "
                 "x 20 g, BF16 log10(k) tables (1.44 GB), 5 angles, weighted outputs",
    }


def percall_run(rad, req, ncalls, cache):
    """The unchanged Fortran call path: ``ncalls`` calls of cirsradg_wave_,
    one wavenumber each, as cirsrtfg_wave.f:361-448 makes them, with the
    REAL(4) arrays of the caller (arrdef.f leading dimensions) built once.
    ``cache``: whether a call may re-stage the previous call's batch
    (RADG_FIXED_CACHE)."""
    import ctypes as C

    os.environ["RADG_FIXED_CACHE"] = "1" if cache else "0"
    a = rad.fortran_arrays(req)
    out = np.zeros(rad.MAXPAT, np.float32)
    gts = np.zeros(rad.MAXPAT, np.float32)
    dq = np.zeros((rad.MAXPAT, rad.MAXLAY, rad.NPARMAX), np.float32, order="F")
    f = lambda x: C.byref(C.c_float(float(x)))
    i = lambda x: C.byref(C.c_int(int(x)))
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    vv = C.c_float(0.0)
    args = [f(req.dist), i(req.inormal), i(req.iray), p(a["delh"]), i(req.nlayer),
            i(req.npath), i(req.ngas), p(a["press"]), p(a["temp"]), p(a["pp"]), p(a["amount"]),
            i(1), i(req.ispace), i(req.amform), C.byref(vv), p(a["nlayin"]), p(a["layinc"]),
            p(a["cont"]), p(a["scale"]), p(a["imod"]), p(a["idgas"]), p(a["isogas"]),
            p(a["emtemp"]), i(req.itype), i(len(req.vem)), p(a["vem"]), p(a["emissivity"]),
            f(req.tsurf), p(gts), f(req.radius1), i(req.flagh2p), p(a["hfp"]), f(req.radextra),
            i(len(req.isw)), p(a["isw"]), p(out), p(dq)]
    fn = rad.lib().cirsradg_wave_
    lat = []
    for k in range(ncalls + 10):
        vv.value = float(req.vwave[k % req.nwave])
        t0 = time.perf_counter()
        fn(*args)
        t1 = time.perf_counter()
        rad.check(rad.lib().radg_cirsradg_status())
        if k >= 10:
            lat.append(t1 - t0)
    os.environ.pop("RADG_FIXED_CACHE", None)
    lat = np.array(lat) * 1e3
    return {"calls": ncalls, "ms_mean": float(lat.mean()), "ms_p50": float(np.percentile(lat, 50)),
            "ms_p99": float(np.percentile(lat, 99)),
            "spectra_per_s": 1e3 / (float(lat.mean()) * req.nwave)}


def serving_run(rad, n, rate, budget_gib=16, max_batch=32):
    """The serving mix: ``n`` requests (layers 40-200, wavenumber windows of
    200-5000 points on a shared 5000-point grid, 1-6 k-gases + H2/He)
    arriving Poisson at ``rate``/s, mixed form, admitted by device-memory
    budget into continuously formed batches on 2 streams."""
    from radtrancode_amd import serving

    tabs = serving.serving_tables()
    rad.register_ktables(tabs)
    arrivals, reqs = serving.poisson_workload(n=n, rate=rate)
    budget = budget_gib << 30
    srv = serving.Server(mode=rad.MODE_MIXED, budget_bytes=budget, max_batch=max_batch, max_inflight=2,
                         per_path_outputs=False)
    r = serving.run_poisson(srv, arrivals, reqs)
    st = srv.stats()
    srv.close()
    lat = r["latency"] * 1e3
    ok = int(np.sum(r["status"] == 0))
    return {
        "requests": n, "completed_ok": ok, "arrival_rate_per_s": rate,
        "wall_s": r["wall"], "requests_per_s": n / r["wall"],
        "work_units_per_s": r["work"] / r["wall"],
        "work_unit": "one 120-layer x 1000-wavenumber x 3 k-gas spectrum, scaled by "
                     "wavenumbers x layers x (0.33 + 0.335 x overlap rounds), the measured "
                     "costs of the interpolation + emission and of an overlap round",
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "latency_ms_max": float(lat.max()),
        "batches": st["steps"], "largest_batch": st["max_batch"],
        "peak_inflight_bytes": st["peak_bytes"], "budget_bytes": budget,
        "worker_s": {"pack_upload_launch": st["create_s"], "retire": st["retire_s"],
                     "kernels": st["gpu_s"]},
        "mix": "layers U[40,200], wavenumbers U[200,5000] (window of a 5000-point grid), "
               "k-gases U[1,6] + H2/He, 20 g, 5 angles; mixed form",
    }


def mix_batch_rate(rad, n, bs=32):
    """Batch mode on the serving mix: the first ``n`` requests of the
    serving workload in batches of ``bs``, each launched once to warm and
    once timed (launch to completion); work units per second of GPU time."""
    from radtrancode_amd import serving

    _, reqs = serving.poisson_workload(n=n, rate=1e6)
    work, gpu = 0.0, 0.0
    for i in range(0, n, bs):
        b = rad.Batch(reqs[i:i + bs], mode=rad.MODE_MIXED)
        b.launch()
        b.sync()
        t0 = time.perf_counter()
        b.launch()
        b.sync()
        gpu += time.perf_counter() - t0
        b.close()
    work = serving.work_units(reqs)
    return work / gpu


def retrieval_run(rad, nret, iters):
    """The batched-retrieval workload: ``nret`` independent retrievals of T
    and two gas VMRs on 121 levels (nx = 363) from 1000 wavenumbers (ny),
    measurements made by the mixed-form forward model at a seeded truth plus
    1% noise, a priori correlation length 1.5 scale heights; ``iters``
    Levenberg-Marquardt iterations (coreret.f's brake and convergence test)
    with the forward model + Jacobian of every running retrieval in one
    device batch per iteration and the FP64 normal-equation solve on the
    device."""
    from radtrancode_amd import retrieval, synth

    tables, _, _ = synth.config1(nwave=1000, nlayer=120, nkgas=3, ng=20)
    rad.register_ktables(tables)
    prob = retrieval.retrieval_problem(nret=nret, nwave=1000)
    t0 = time.perf_counter()
    y, se = retrieval.measure(prob)
    t_meas = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = retrieval.retrieve(prob, y, se, kiter=iters)
    wall = time.perf_counter() - t0
    nlev = prob.nlev
    e0 = float(np.abs(prob.xa[:, :nlev] - prob.x_true[:, :nlev]).mean())
    e1 = float(np.abs(out["x"][:, :nlev] - prob.x_true[:, :nlev]).mean())
    return {
        "retrievals": nret, "iterations": iters, "wall_s": wall,
        "retrievals_per_s": nret / wall,
        "forward_evaluations": int(nret + out["iters"].sum()),
        "converged": int(out["converged"].sum()),
        "mean_iterations": float(out["iters"].mean()),
        "median_chisq_over_ny": float(np.median(out["phi"]) / y.shape[1]),
        "mean_abs_T_error_K": {"a_priori": e0, "retrieved": e1},
        "timings_s": {k: float(v) for k, v in out["timings"].items()},
        "measurement_s": t_meas,
        "shape": "nx = 363 (T and 2 ln VMR on 121 levels), ny = 1000 wavenumbers, 120 layers, "
                 "3 k-gases + H2/He, 20 g, 5 angles; mixed-form forward model, FP64 OE solve",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="host processes of the parallel CPU baseline (at most the host's cores)")
    ap.add_argument("--skip-fp64", action="store_true", help="time only the mixed form")
    ap.add_argument("--serving-requests", type=int, default=4096,
                    help="requests in the Poisson serving run (0: skip it)")
    ap.add_argument("--no-serving-capacity", action="store_true",
                    help="skip the saturated serving run (all requests queued at once)")
    ap.add_argument("--serving-rate", type=float, default=100.0,
                    help="Poisson arrival rate of the serving run, requests/s")
    ap.add_argument("--bf16-steps", type=int, default=3)
    ap.add_argument("--config5-requests", type=int, default=1024,
                    help="requests in the large-spectrum BF16 stress (0: skip it)")
    ap.add_argument("--percall", type=int, default=1000,
                    help="calls of the unchanged cirsradg_wave_ entry, one wavenumber each (0: skip)")
    ap.add_argument("--retrievals", type=int, default=512,
                    help="retrievals in the batched optimal-estimation run (0: skip it)")
    ap.add_argument("--retrieval-iters", type=int, default=10)
    ap.add_argument("--layer-sweep", action="store_true",
                    help="only launch the layer pass once per k-gas count 1, 2, 3 (0, 1, 2 overlap "
                         "rounds), mixed and FP64 forms, 4 spectra of 200 wavenumbers each: the "
                         "dispatches a rocprofv3 --pmc pass splits the work per round by")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a kernel of known "
                         "traffic, to scale rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by wavenumber or by request with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.layer_sweep:
        layer_sweep()
        return
    tables, cia, reqs = make_batch_inputs(args.batch)
    # the host baselines first, before anything touches the GPU (the
    # parallel one forks its processes)
    cpu = None
    if not args.no_cpu:
        single = cpu_baseline(tables, cia, reqs[0], args.cpu_seconds)
        nproc = max(1, min(args.cpu_procs, os.cpu_count() or 1))
        par = cpu_baseline_parallel(tables, cia, reqs[0], args.cpu_seconds, nproc)
        cpu = single
        if par is not None:
            cpu = dict(par, single_core=single)

    import torch

    from oracle import oracle
    from radtrancode_amd import cirsradg as rad

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rad.register_ktables(tables)
    rad.register_cia(cia)

    # ---- mixed-precision form: the workload
    b, wall, lat, kern = time_batch(rad, reqs, rad.MODE_MIXED, args.steps, args.warmup, torch)
    info = b.info()
    nspec = args.batch * args.steps
    value = nspec / wall
    # accuracy of this run against the FP64 oracle on a sample of wavenumbers
    sample = np.arange(0, 1000, 50)
    res = b.read(0, weighted=False)
    o_ref, g_ref, d_ref = oracle.spectrum(tables, cia, reqs[0], vwave=reqs[0].vwave[sample],
                                          nthreads=os.cpu_count() or 1)
    o_m = res["output"][sample]
    d_m = res["doutputdq"][sample]
    rad_err = float(np.max(np.abs(o_m - o_ref) / np.abs(o_ref)))
    sc = np.max(np.abs(d_ref), axis=2, keepdims=True)
    jac_err = float(np.max(np.abs(d_m - d_ref) / np.where(sc > 0, sc, 1.0)))
    b.close()

    # ---- plain FP64 form, same batch
    if args.skip_fp64:
        wall64, kern64, value64 = float("nan"), np.full(3, np.nan), float("nan")
    else:
        b64, wall64, lat64, kern64 = time_batch(rad, reqs, rad.MODE_FP64, args.fp64_steps, 1,
                                                torch)
        b64.close()
        value64 = args.batch * args.fp64_steps / wall64

    # ---- the 16-bit log10(k) table forms (BF16 and IEEE-half codes), same batch
    forms16 = {}
    for name, mode16 in (("bf16", rad.MODE_BF16), ("f16", rad.MODE_F16)):
        steps16 = max(1, args.bf16_steps)
        b16, wall16, lat16, kern16 = time_batch(rad, reqs, mode16, steps16, 1, torch)
        info16 = b16.info()
        res16 = b16.read(0, weighted=False)
        b16.close()
        o16, d16 = res16["output"][sample], res16["doutputdq"][sample]
        alg16 = info16["table_bytes"] + info16["input_bytes"]
        forms16[name] = {
            "value": args.batch * steps16 / wall16, "ms_per_step": wall16 / steps16 * 1e3,
            "kernel_ms": {"layer": float(kern16[0]), "emission": float(kern16[1]),
                          "quadrature": float(kern16[2])},
            "layer_kernel_alg_GBps": alg16 / (kern16[0] / 1e3) / 1e9,
            "error_vs_fp64_oracle": {
                "radiance_rel": float(np.max(np.abs(o16 - o_ref) / np.abs(o_ref))),
                "jacobian_col_scaled": float(np.max(np.abs(d16 - d_ref) / np.where(sc > 0, sc, 1.0)))},
            "codes": "log10(k) minus an FP32 offset per (wavenumber, gas, g-ordinate), rounded to "
                     + ("BF16" if name == "bf16" else "IEEE half"),
        }

    # ---- roofline of the dominant kernel (layer/overlap kernel): it reads
    # the table nodes and the layer inputs of every spectrum in the batch
    t_layer_s = float(kern[0]) / 1e3
    alg_read = info["table_bytes"] + info["input_bytes"]
    achieved = alg_read / t_layer_s / 1e9
    roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=achieved / HBM_PEAK_GBS, traffic=None,
                kernel="layer_kernel<float,8>", kernel_ms=float(kern[0]),
                alg_bytes_per_launch=alg_read)
    roof.update(pmc_figures(args.batch, t_layer_s))

    percall = None
    if args.percall > 0:
        rad.set_mode(rad.MODE_MIXED)
        percall = {
            "reused_batch": percall_run(rad, reqs[0], args.percall, True),
            "batch_per_call": percall_run(rad, reqs[0], min(args.percall, 200), False),
            "call": "cirsradg_wave_ (REAL(4) arrays, one wavenumber per call), mixed form, "
                    "120 layers x 5 paths x 5 gases, wall clock per call",
        }

    c5 = None
    if args.config5_requests > 0:
        c5 = config5_run(rad, args.config5_requests, cia, os.cpu_count() or 1)

    serve = None
    if args.serving_requests > 0:
        serve = serving_run(rad, args.serving_requests, args.serving_rate)
        if not args.no_serving_capacity:
            # the same requests all queued at once: the loop's capacity,
            # against batch mode on the same mix (batches of 32, timed on
            # the GPU alone) and against the config-1 batch above
            cap = serving_run(rad, args.serving_requests, 1e6)
            mix = mix_batch_rate(rad, min(1024, args.serving_requests))
            serve["saturated"] = {k: cap[k] for k in (
                "wall_s", "requests_per_s", "work_units_per_s", "latency_ms_p50",
                "latency_ms_p99", "batches", "largest_batch", "peak_inflight_bytes", "worker_s")}
            serve["saturated"]["batch_mode_on_mix_work_units_per_s"] = mix
            serve["saturated"]["efficiency_vs_batch_on_mix"] = cap["work_units_per_s"] / mix
            serve["saturated"]["efficiency_vs_config1_batch"] = cap["work_units_per_s"] / value

    oe_run = None
    if args.retrievals > 0:
        oe_run = retrieval_run(rad, args.retrievals, args.retrieval_iters)

    out = {
        "metric": "Jacobian spectra/s (120 lyr,1k wn,20 g) + p50/p99 latency vs plain FP64",
        "value": value,
        "unit": "Jacobian-spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mixed-precision CIRSRADG_WAVE forward + analytic Jacobian, "
                               "120 layers x 1000 wavenumbers x 20 g, 3 k-gases + H2/He CIA, "
                               "5 emission angles",
                   "spectra_per_step": args.batch, "parallelism": "single GPU"},
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "latency_unit": "one step (batch of spectra), HIP events",
        "kernel_ms": {"layer": float(kern[0]), "emission": float(kern[1]),
                      "quadrature": float(kern[2])},
        "fp64_form": {"value": value64, "ms_per_step": wall64 / args.fp64_steps * 1e3,
                      "kernel_ms": {"layer": float(kern64[0]), "emission": float(kern64[1]),
                                    "quadrature": float(kern64[2])}},
        "bf16_form": forms16["bf16"],
        "f16_form": forms16["f16"],
        "speedup_vs_fp64": value / value64,
        "error_vs_fp64_oracle": {"radiance_rel": rad_err, "jacobian_col_scaled": jac_err,
                                 "sample": f"{len(sample)} wavenumbers of spectrum 0"},
        "roofline": roof,
        "per_call_cirsradg_wave": percall,
        "cpu_baseline": cpu,
        "serving": serve,
        "config5_bf16_stress": c5,
        "batched_retrieval": oe_run,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
