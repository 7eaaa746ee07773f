"""Batched optimal-estimation retrievals on the GPU forward model.

The reference drives one retrieval at a time from ``nemesis/coreret.f``: a
forward model plus Jacobian (``forwardavfovX`` -> ``CIRSrtfg_wave`` ->
``cirsradg_wave``), the layer -> profile -> state maps (``map2pro.f``,
``map2xvec.f``), the gain matrix (``calc_gain_matrix.f``), the next state
(``calcnextxn.f``), the cost function (``calc_phiret.f``) and the
Levenberg-Marquardt brake (``coreret.f:591-1068``).  Here many retrievals
advance together: every iteration is one device batch of forward models with
their state maps, followed by one batched FP64 proposal on the device
(``radg_oe_*``, include/radg.h).

* :class:`OE` -- the ``radg_oe`` handle.
* :func:`retrieval_problem` -- the batched-retrieval workload: T and the VMRs of
  two gases on 121 levels (nx = 363), 1000 wavenumbers, noise 1% of the
  radiance, a-priori correlation length 1.5 scale heights.
* :func:`retrieve` -- the iteration loop for all retrievals at once.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import cirsradg as rad
from . import synth
from ._native import check, lib


class OE:
    """``radg_oe_create`` / ``_cost`` / ``_propose`` / ``_destroy``."""

    def __init__(self, y, se, xa, sa, sa_shared=False):
        y, se, xa = (np.ascontiguousarray(v, dtype=np.float64) for v in (y, se, xa))
        sa = np.ascontiguousarray(sa, dtype=np.float64)
        self.nret, self.ny = y.shape
        self.nx = xa.shape[1]
        h = C.c_void_p()
        check(lib().radg_oe_create(self.nret, self.nx, self.ny, y.ctypes.data, se.ctypes.data,
                                   xa.ctypes.data, sa.ctypes.data, int(bool(sa_shared)),
                                   C.byref(h)))
        self.h = h

    def _ptrs(self, addrs):
        if len(addrs) != self.nret:
            raise ValueError("one device address per retrieval")
        return (C.c_void_p * self.nret)(*addrs)

    def cost(self, xn, yn_dev):
        """calc_phiret.f per retrieval: (phi, chisq)."""
        xn = np.ascontiguousarray(xn, dtype=np.float64)
        phi, chi = np.zeros(self.nret), np.zeros(self.nret)
        check(lib().radg_oe_cost(self.h, xn.ctypes.data, self._ptrs(yn_dev), phi.ctypes.data,
                                 chi.ctypes.data))
        return phi, chi

    def propose(self, xn, yn_dev, kk_dev, lam):
        """coreret.f:609-632 per retrieval: xn1 = xn + (x_out - xn)/(1 + lambda)."""
        xn = np.ascontiguousarray(xn, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        xn1 = np.zeros_like(xn)
        check(lib().radg_oe_propose(self.h, xn.ctypes.data, self._ptrs(yn_dev),
                                    self._ptrs(kk_dev), lam.ctypes.data, xn1.ctypes.data))
        return xn1

    def close(self):
        if getattr(self, "h", None):
            lib().radg_oe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- the workload
@dataclass
class Problem:
    """Levels, the fixed parts of every forward model and the OE constants."""
    plev_bar: np.ndarray       # [nlev] level pressures, bottom first
    ncol: np.ndarray           # [L] molecules cm-2 per layer
    base: list                 # per retrieval: synth.Request with the fixed inputs
    vmr_fixed: np.ndarray      # [nret, L, ngas] layer VMRs of the gases not retrieved
    xa: np.ndarray             # [nret, nx]
    sa: np.ndarray             # [nx, nx] (shared)
    x_true: np.ndarray         # [nret, nx]
    rgas: tuple                # 0-based gas slots retrieved (log VMR)

    @property
    def nlev(self):
        return self.plev_bar.shape[0]

    @property
    def nx(self):
        return self.xa.shape[1]


def _layer_arrays(prob: Problem, r: int, x):
    """Levels -> layers for one state: layer j spans levels j and j+1; its
    T and VMRs are the level means, amount = VMR x column.  Returns
    (temp, vmr [L, ngas], dte, dam) with dte, dam [L, nlev] the layer
    derivatives with respect to the level quantities (the DTE and DAM of
    radtran/includes/laygrad.f, which radtran/path/layerg.f fills for the
    reference's own layering)."""
    nlev = prob.nlev
    L = nlev - 1
    tl = x[:nlev]
    temp = 0.5 * (tl[:-1] + tl[1:])
    vmr = prob.vmr_fixed[r].copy()
    for s, g in enumerate(prob.rgas):
        v = np.exp(x[nlev * (1 + s):nlev * (2 + s)])
        vmr[:, g] = 0.5 * (v[:-1] + v[1:])
    w = np.zeros((L, nlev))
    j = np.arange(L)
    w[j, j] = 0.5
    w[j, j + 1] = 0.5
    return temp, vmr, w, w * prob.ncol[:, None]


def state_request(prob: Problem, r: int, x):
    """The forward-model request and state map of retrieval r at state x."""
    q = prob.base[r]
    temp, vmr, dte, dam = _layer_arrays(prob, r, x)
    req = synth.Request(**{**q.__dict__})
    req.temp = temp
    req.amount = vmr * prob.ncol[:, None]
    req.pp = vmr * q.press[:, None]
    req.emtemp = temp[q.layinc - 1]
    nlev, ngas = prob.nlev, q.ngas
    xmap = np.zeros((prob.nx, ngas + 1, nlev))
    lev = np.arange(nlev)
    xmap[lev, ngas, lev] = 1.0                             # temperature
    for s, g in enumerate(prob.rgas):                      # d vmr / d ln vmr
        xmap[nlev * (1 + s) + lev, g, lev] = np.exp(x[nlev * (1 + s):nlev * (2 + s)])
    # parameters with a non-zero map only (cirsrtfg_wave.f:301-313)
    req.isw = np.array([g + 1 for g in prob.rgas] + [ngas + 1], dtype=np.int32)
    return req, dict(dam=dam, dte=dte, xmap=xmap)


def retrieval_problem(nret=512, nwave=1000, seed=40, t_sigma=5.0, lnvmr_sigma=0.5,
                      corr_scale_heights=1.5):
    """The batched-retrieval workload on the 120-layer, 3 k-gas + H2/He,
    600-1400 cm-1 atmosphere of :func:`synth.config1` (register the tables of
    ``synth.config1(nwave=nwave)`` first):

    * state: T and ln VMR of k-gases 1 and 2 on the 121 levels (nx = 363);
    * truth: a seeded perturbation of the a priori drawn from Sa;
    * a priori: T = 110 + 60 tanh(log10(p / 0.1 bar)) plus a seeded offset
      per retrieval, ln VMR uniform per retrieval;
    * Sa: sigma_T = 5 K, sigma_lnVMR = 0.5, correlation exp(-|ln p_i - ln p_j|
      / 1.5) within each block (1.5 scale heights), no cross-block terms.
    Measurements are made by the caller from x_true with the GPU forward
    model, plus 1% noise (:func:`measure`)."""
    rng = np.random.default_rng(seed)
    L = 120
    nlev = L + 1
    plev = np.geomspace(10.0, 1e-4, nlev)
    q0 = synth.make_request(nlayer=L, nkgas=3, nwave=nwave)
    dp_pa = (plev[:-1] - plev[1:]) * 1e5
    ncol = dp_pa / (2.3 * synth.AMU_KG * 24.79) * 1e-4
    lnp = np.log(plev)
    corr = np.exp(-np.abs(lnp[:, None] - lnp[None, :]) / corr_scale_heights)
    nx = 3 * nlev
    sa = np.zeros((nx, nx))
    for b, sg in enumerate((t_sigma, lnvmr_sigma, lnvmr_sigma)):
        sa[b * nlev:(b + 1) * nlev, b * nlev:(b + 1) * nlev] = sg * sg * corr
    chol = np.linalg.cholesky(sa + 1e-12 * np.eye(nx))
    base, vfix, xa, xt = [], [], [], []
    tprof = 110.0 + 60.0 * np.tanh(np.log10(plev / 0.1))
    for r in range(nret):
        q = synth.Request(**{**q0.__dict__})
        base.append(q)
        v = q.amount / ncol[:, None]
        vfix.append(v)
        a = np.concatenate([tprof + rng.normal(0.0, 2.0),
                            np.full(nlev, np.log(rng.uniform(1e-6, 1e-4))),
                            np.full(nlev, np.log(rng.uniform(1e-6, 1e-4)))])
        xa.append(a)
        xt.append(a + chol @ rng.normal(size=nx))
    return Problem(plev_bar=plev, ncol=ncol, base=base, vmr_fixed=np.array(vfix),
                   xa=np.array(xa), sa=sa, x_true=np.array(xt), rgas=(0, 1))


def _chunk_batch(prob, xs, idx, mode):
    reqs, maps = [], []
    for r in idx:
        q, m = state_request(prob, r, xs[r])
        reqs.append(q)
        maps.append(m)
    b = rad.Batch(reqs, mode=mode)
    b.set_statemaps(maps)
    b.launch()
    return b


_POOL = None


def forward(prob: Problem, xs, idx, mode=rad.MODE_MIXED, chunk=16, workers=8):
    """The forward models and state maps of retrievals ``idx`` at states
    ``xs[idx]``: device batches of ``chunk`` requests, packed by a pool of
    host threads and each launched as soon as it is packed, so packing
    overlaps the kernels of the batches before it.  Returns the list of
    (batch, index in batch) for each entry of ``idx``."""
    global _POOL
    from concurrent.futures import ThreadPoolExecutor
    if _POOL is None:
        _POOL = ThreadPoolExecutor(max_workers=workers)
    idx = list(idx)
    parts = [idx[i:i + chunk] for i in range(0, len(idx), chunk)]
    futs = [_POOL.submit(_chunk_batch, prob, xs, p, mode) for p in parts]
    out = []
    for p, f in zip(parts, futs):
        b = f.result()
        out += [(b, i) for i in range(len(p))]
    return out


def measure(prob: Problem, mode=rad.MODE_MIXED, noise=0.01, seed=41):
    """y = F(x_true) (1 + noise N(0, 1)), se = (noise F(x_true))^2."""
    rng = np.random.default_rng(seed)
    nret = prob.xa.shape[0]
    held = forward(prob, prob.x_true, range(nret), mode)
    y = np.array([b.read(i)["wout"] for b, i in held])
    del held
    se = (noise * y) ** 2
    return y * (1.0 + noise * rng.normal(size=y.shape)), se


def retrieve(prob: Problem, y, se, kiter=10, phlimit=0.1, mode=rad.MODE_MIXED, log=None):
    """coreret.f's iteration for every retrieval at once.  Each retrieval
    keeps its own state, cost, brake and convergence flag; an iteration
    evaluates the forward model of every retrieval still running at its
    proposed state in one device batch.  Returns a dict with the final
    states, costs, iteration counts and timings."""
    nret = prob.xa.shape[0]
    oe = OE(y, se, prob.xa, prob.sa, sa_shared=True)
    xn = prob.xa.copy()
    t = dict(forward=0.0, propose=0.0, cost=0.0)
    def settle(held):
        for bb in {id(x[0]): x[0] for x in held}.values():
            bb.sync()

    t0 = time.perf_counter()
    cur = forward(prob, xn, range(nret), mode)  # batch holding retrieval r's yn, kk
    settle(cur)
    t["forward"] += time.perf_counter() - t0

    def ptrs(state):
        w, k = zip(*[bb.device_outputs(i) for bb, i in state])
        return list(w), list(k)

    t0 = time.perf_counter()
    wp, kp = ptrs(cur)
    phi, chisq = oe.cost(xn, wp)
    t["cost"] += time.perf_counter() - t0
    ophi = phi.copy()
    lam = np.ones(nret)
    done = np.zeros(nret, dtype=bool)
    iters = np.zeros(nret, dtype=int)
    history = [float(np.median(phi))]
    for it in range(1, kiter + 1):
        act = np.flatnonzero(~done)
        if act.size == 0:
            break
        t0 = time.perf_counter()
        wp, kp = ptrs(cur)
        xn1 = oe.propose(xn, wp, kp, lam)
        t["propose"] += time.perf_counter() - t0
        bad = ~np.all(np.isfinite(xn1), axis=1)
        run = [r for r in act if not bad[r]]
        for r in act:
            if bad[r]:                        # coreret.f:640-650: brake and retry
                lam[r] = min(lam[r] * 10.0, 1e10)
        if not run:
            continue
        t0 = time.perf_counter()
        held = forward(prob, xn1, run, mode)
        settle(held)
        trial = list(cur)
        for r, h in zip(run, held):
            trial[r] = h
        t["forward"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        wp, _ = ptrs(trial)
        xc = xn.copy()
        xc[run] = xn1[run]
        phi1, chi1 = oe.cost(xc, wp)
        t["cost"] += time.perf_counter() - t0
        for r in run:
            tphi = 100.0 * (ophi[r] - phi1[r]) / ophi[r]
            iters[r] = it
            if phi1[r] <= ophi[r]:
                xn[r] = xn1[r]
                cur[r] = trial[r]
                if 0.0 <= tphi <= phlimit and lam[r] < 1.0:
                    done[r] = True
                else:
                    lam[r] *= 0.3
                ophi[r] = phi1[r]
            elif it >= 5 and abs(tphi) <= phlimit:
                if lam[r] < 0.1:
                    lam[r] *= 10.0
                else:
                    xn[r] = xn1[r]
                    cur[r] = trial[r]
                    ophi[r] = phi1[r]
                    done[r] = True
            else:
                lam[r] = min(lam[r] * 10.0, 1e10)
        history.append(float(np.median(ophi)))
        if log:
            log(f"iteration {it}: running {len(run)}, median phi {history[-1]:.4g}, "
                f"accepted {int(sum(cur[r] is trial[r] for r in run))}")
    out = dict(x=xn, phi=ophi, iters=iters, converged=done, history=history, timings=t)
    oe.close()
    return out
