"""Batched chains and device parallel tempering.

``GPUBatch`` runs many chains — the rungs of a temperature ladder on one model,
or independent chains of different models and sizes — in one launch per colour
(sce_batch_* in include/sce_hip.h).  ``gpu_run_pt`` is the device counterpart of
the reference's ``gpu_run_pt`` (src/gpu/gpu_pt.jl:261-319): thermalisation with
step adaptation, a renormalised boundary, a measurement phase, adjacent-rung
exchange every ``exchange_interval`` sweeps — with the exchange decided on the
device instead of on the host.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L
from .gpu import DEFAULT_LANES, GPUTiledHamiltonian
from .hamiltonian import ChainState, resolve_kt
from .models import random_config
from .observables import ObservableStat, standard_stats


class GPUBatch:
    def __init__(self, gHs: Sequence[GPUTiledHamiltonian], states: Sequence[ChainState],
                 seeds: Sequence[int], kts: Sequence[float], exchange_seed: int = 0):
        R = len(gHs)
        if not (len(states) == len(seeds) == len(kts) == R) or R < 1:
            raise ValueError("gHs, states, seeds and kts must have the same nonzero length")
        self.gHs = list(gHs)
        self.R = R
        self.kts = np.asarray(kts, np.float64)
        self._keep = [np.ascontiguousarray(st.config, np.float64) for st in states]
        self._keepz = [np.ascontiguousarray(st.zrows, np.float64) for st in states]
        tabs = (C.c_void_p * R)(*[g.handle.value for g in gHs])
        cfgs = (C.c_void_p * R)(*[a.ctypes.data for a in self._keep])
        zrs = (C.c_void_p * R)(*[a.ctypes.data for a in self._keepz])
        en = np.asarray([st.energy for st in states], np.float64)
        sd = np.asarray([int(x) & (2**64 - 1) for x in seeds], np.uint64)
        stp = np.asarray([st.step for st in states], np.float64)
        h = C.c_void_p()
        L.check(L.lib().sce_batch_create(R, tabs, cfgs, zrs, L.ptr(en), L.ptr(sd), L.ptr(stp),
                                         L.ptr(self.kts), C.c_uint64(int(exchange_seed) & (2**64 - 1)),
                                         C.byref(h)))
        self._h = h
        self.phase_done = 0

    def run(self, nsweeps: int, lanes: int = DEFAULT_LANES, renorm_interval: int = 1000,
            measure_interval: int = 0, exchange_interval: int = 0, phase_done: int = 0,
            phase_total: Optional[int] = None) -> np.ndarray:
        """Returns the measurements, shape (R, n_meas, 4): (E, m_x, m_y, m_z)."""
        if phase_total is None:
            phase_total = phase_done + nsweeps
        nmeas = ((phase_done + nsweeps) // measure_interval - phase_done // measure_interval
                 if measure_interval > 0 else 0)
        obs = np.zeros((self.R, max(nmeas, 1), 4))
        n = C.c_int64()
        L.check(L.lib().sce_batch_run(self._h, int(nsweeps), int(lanes), int(renorm_interval),
                                      int(measure_interval), int(exchange_interval), int(phase_done),
                                      int(phase_total), L.ptr(obs), int(max(nmeas, 1)), C.byref(n)))
        return obs[:, : n.value]

    def download(self, k: int):
        H = self.gHs[k].host
        cfg = np.empty((H.n_sites, 3))
        z = np.empty((H.n_sites, H.nlm))
        e = C.c_double()
        L.check(L.lib().sce_batch_download(self._h, int(k), L.ptr(cfg), L.ptr(z), C.byref(e)))
        return cfg, z, e.value

    def info(self, k: int) -> L.ChainInfo:
        inf = L.ChainInfo()
        L.check(L.lib().sce_batch_info(self._h, int(k), C.byref(inf)))
        return inf

    def set_step(self, k: int, step: float) -> None:
        L.check(L.lib().sce_batch_set_step(self._h, int(k), float(step)))

    def reset_counters(self) -> None:
        L.check(L.lib().sce_batch_reset_counters(self._h))

    def renormalize(self) -> None:
        L.check(L.lib().sce_batch_renormalize(self._h))

    def swap_stats(self):
        att = np.zeros(self.R, np.int64)
        acc = np.zeros(self.R, np.int64)
        L.check(L.lib().sce_batch_swap_stats(self._h, L.ptr(att), L.ptr(acc)))
        return att[: self.R - 1], acc[: self.R - 1]

    def set_graph(self, on: bool) -> None:
        L.check(L.lib().sce_batch_set_graph(self._h, 1 if on else 0))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.sce_batch_destroy(h)
            self._h = None


def derive_seeds(seed: int, R: int):
    """Seed derivation of a ladder run: one numpy PCG64 master stream gives, in
    order, the R initial-configuration seeds, the exchange seed and the R device
    (Philox key) seeds — the order is part of the reproducibility contract, like
    the reference's master Xoshiro derivation (gpu_pt.jl:17-29)."""
    master = np.random.default_rng(seed)
    init = [int(x) for x in master.integers(0, 2**63, R)]
    exch = int(master.integers(0, 2**63))
    dev = [int(x) for x in master.integers(0, 2**63, R)]
    return init, exch, dev


@dataclass
class TempResult:
    kT: float
    stats: Dict[str, ObservableStat]
    acceptance_metropolis: float
    final_step: float
    max_drift: float


@dataclass
class PTResult:
    points: List[TempResult]
    swap_acceptance: np.ndarray
    final_configs: List[np.ndarray]
    seed: int
    wall_s: float = 0.0


def _adapt(step, acc, att, target):
    if att > 0:
        step = min(max(step * math.exp(0.5 * (acc / att - target)), 1e-3), math.pi)
    return step


def gpu_run_pt(gH: GPUTiledHamiltonian, temperature=None, kT=None, exchange_interval: int = 10,
               sweeps_therm: int = 2000, sweeps_measure: int = 10000, measure_interval: int = 1,
               step: float = 0.6, adapt_target: float = 0.5, adapt_interval: int = 50,
               renorm_interval: int = 1000, nbins: int = 32, init=None, seed: int = 0,
               lanes: int = DEFAULT_LANES) -> PTResult:
    """Replica exchange on the device — gpu_run_pt (src/gpu/gpu_pt.jl:261-319)."""
    kts = resolve_kt(temperature, kT)
    R = len(kts)
    if R < 2:
        raise ValueError(f"parallel tempering needs a ladder of >= 2 temperatures; got {R}")
    d = np.diff(kts)
    if not (np.all(d > 0) or np.all(d < 0)):
        raise ValueError("the temperature ladder must be strictly monotone")
    if exchange_interval < 1:
        raise ValueError("exchange_interval must be >= 1")
    if sweeps_measure < 1 or not (1 <= measure_interval <= sweeps_measure):
        raise ValueError("measure_interval must be in 1:sweeps_measure")
    H = gH.host
    init_seeds, exch_seed, dev_seeds = derive_seeds(seed, R)
    states = []
    for r in range(R):
        cfg = init if init is not None else random_config(H.n_sites, np.random.default_rng(init_seeds[r]))
        states.append(ChainState(H, cfg, step))
    b = GPUBatch([gH] * R, states, dev_seeds, kts, exch_seed)
    t0 = time.perf_counter()
    steps = [step] * R
    done = 0
    while done < sweeps_therm:
        n = min(adapt_interval - done % adapt_interval, sweeps_therm - done)
        b.run(n, lanes, renorm_interval, 0, exchange_interval, done, sweeps_therm)
        done += n
        if done % adapt_interval == 0:
            for r in range(R):
                inf = b.info(r)
                steps[r] = _adapt(steps[r], inf.acc_metro, inf.att_metro, adapt_target)
                b.set_step(r, steps[r])
            b.reset_counters()
    b.renormalize()
    b.reset_counters()
    obs = b.run(sweeps_measure, lanes, renorm_interval, measure_interval, exchange_interval, 0,
                sweeps_measure)
    wall = time.perf_counter() - t0
    points = []
    for r in range(R):
        inf = b.info(r)
        points.append(TempResult(float(kts[r]), standard_stats(obs[r], float(kts[r]), H.n_active, nbins),
                                 inf.acc_metro / inf.att_metro if inf.att_metro else math.nan,
                                 inf.step, inf.max_drift))
    att, acc = b.swap_stats()
    with np.errstate(invalid="ignore", divide="ignore"):
        swaps = np.where(att > 0, acc / np.maximum(att, 1), np.nan)
    finals = [b.download(r)[0] for r in range(R)]
    return PTResult(points, swaps, finals, seed, wall)


def autotune_lanes(make_batch, candidates=(16, 32, 64), sweeps: int = 30, cache: Optional[dict] = None,
                   key=None) -> int:
    """The project autotuner: times `sweeps` graph-replayed sweeps of a scratch
    batch for each lanes-per-site candidate and returns the fastest.  `lanes` is
    part of the determinism scope, so tuning runs on a throwaway batch built by
    `make_batch()` before the production run.  Results are memoised in `cache`
    under `key` (e.g. a size-class signature)."""
    if cache is not None and key is not None and key in cache:
        return cache[key]
    best, best_t = None, math.inf
    for lanes in candidates:
        b = make_batch()
        try:
            b.run(3, lanes, 0)          # capture + warm
            t0 = time.perf_counter()
            b.run(sweeps, lanes, 0)
            dt = time.perf_counter() - t0
        except L.SceError:
            continue
        finally:
            del b
        if dt < best_t:
            best, best_t = lanes, dt
    if best is None:
        raise RuntimeError("no lanes candidate could run this batch")
    if cache is not None and key is not None:
        cache[key] = best
    return best
