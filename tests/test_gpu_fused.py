"""The fused kernel (continuum, layer pass and emission of one wavenumber in
one workgroup, layer results in LDS) against the three-kernel form: both
run the same device functions in the same order, so every output must be
bit-for-bit the same.  RADG_FUSED=1 selects the fused form for each class of
requests whose workgroup fits 160 KiB of LDS (here: the mixed and BF16 forms
at 20 g-ordinates and up to 128 layers).
"""
import os

import numpy as np
import pytest

from radtrancode_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload(gpu_lib):
    tables, _, _ = synth.config1(nwave=24, nlayer=60, nkgas=3, ng=20)
    gpu_lib.register_ktables(tables)
    reqs = [synth.make_request(nlayer=L, nkgas=3, nwave=24, seed=40 + L) for L in (40, 60, 100)]
    # one request with a disc-integrated path and a transmission path
    q = synth.make_request(nlayer=50, nkgas=3, nwave=24, seed=77)
    q.imod[1] = 18
    q.imod[2] = 1
    q.tsurf = 150.0  # imod 18 takes the surface at tsurf > 0
    reqs.append(q)
    return reqs


def _run(gpu_lib, reqs, mode, fused):
    old = os.environ.get("RADG_FUSED")
    os.environ["RADG_FUSED"] = "1" if fused else "0"
    try:
        b = gpu_lib.Batch(reqs, mode=mode)
        b.launch()
        res = [b.read(i) for i in range(len(reqs))]
        b.close()
    finally:
        if old is None:
            del os.environ["RADG_FUSED"]
        else:
            os.environ["RADG_FUSED"] = old
    return res


@pytest.mark.parametrize("mode", [1, 2], ids=["mixed", "bf16"])
def test_fused_equals_three_kernels(gpu_lib, workload, mode):
    from tests.golden.golden_io import load_cia
    if mode == 2:
        gpu_lib.set_mode(2)
    cia = load_cia()
    cia.e3 = synth.make_e3_table()
    gpu_lib.register_cia(cia)
    try:
        print(f"fused vs three-kernel, mode {mode}, {len(workload)} requests", flush=True)
        a = _run(gpu_lib, workload, mode, fused=True)
        b = _run(gpu_lib, workload, mode, fused=False)
    finally:
        gpu_lib.register_cia(load_cia())
    for i, (x, y) in enumerate(zip(a, b)):
        for k in ("output", "gtsurf", "doutputdq", "wout", "wjac"):
            assert np.all(np.isfinite(x[k])), (i, k)
            assert np.array_equal(x[k], y[k]), (i, k)
