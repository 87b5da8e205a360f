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
import threading
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
        self.exchange_seed = int(exchange_seed) & (2**64 - 1)
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
        self._graph, self._persistent = True, False
        # every chain on one tables handle: replica exchange and shared walks
        self.homogeneous = all(g is self.gHs[0] for g in self.gHs)

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
        # exactly n_chains - 1 entries, as the C ABI documents (pair k, k + 1)
        att = np.zeros(max(self.R - 1, 0), np.int64)
        acc = np.zeros(max(self.R - 1, 0), np.int64)
        L.check(L.lib().sce_batch_swap_stats(self._h, L.ptr(att), L.ptr(acc)))
        return att, acc

    def set_graph(self, on: bool) -> None:
        L.check(L.lib().sce_batch_set_graph(self._h, 1 if on else 0))
        self._graph = bool(on)

    def set_persistent(self, on: bool) -> None:
        """One persistent launch per renormalisation segment for every chain,
        exchange included; bit-identical to the per-colour launches at the same
        lanes per site."""
        L.check(L.lib().sce_batch_set_persistent(self._h, 1 if on else 0))
        self._persistent = bool(on)

    def set_site_pairs(self, on: bool) -> None:
        """At 64 lanes per site, walk two same-sublattice sites of one chain's
        colour class per wavefront (any batch; used where chains do not share
        walks); the results are those of 64 lanes per site."""
        L.check(L.lib().sce_batch_set_site_pairs(self._h, 1 if on else 0))

    def set_shared_walk(self, chains: int) -> None:
        """At 64 lanes per site, walk each site's sublattice walk once for
        `chains` (2 or 4) chains of a batch on one model; the results are
        those of 64 lanes per site.  0 (the default) shares between two chains
        where the neighbour stencil fits, 1 turns it off."""
        L.check(L.lib().sce_batch_set_shared_walk(self._h, int(chains)))

    def set_counters(self, k: int, sweep_index: int, acc_metro: int, att_metro: int,
                     acc_or: int = 0, att_or: int = 0, max_drift: float = 0.0) -> None:
        """Restore chain k's keyed-RNG sweep index, counters and drift record."""
        L.check(L.lib().sce_batch_set_counters(self._h, int(k), int(sweep_index), int(acc_metro),
                                               int(att_metro), int(acc_or), int(att_or), float(max_drift)))

    def exchange_state(self):
        """(exchanges attempted so far, parity of the next attempt)."""
        n = C.c_int64()
        par = C.c_int32()
        L.check(L.lib().sce_batch_get_exchange(self._h, C.byref(n), C.byref(par)))
        return n.value, par.value

    def set_exchange(self, count: int, parity: int, att, acc) -> None:
        att = np.ascontiguousarray(att, np.int64)
        acc = np.ascontiguousarray(acc, np.int64)
        if att.shape != (self.R - 1,) or acc.shape != (self.R - 1,):
            raise ValueError(f"exchange statistics need {self.R - 1} entries each")
        L.check(L.lib().sce_batch_set_exchange(self._h, int(count), int(parity), L.ptr(att), L.ptr(acc)))

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
    lanes: int = 0     # lanes per site the run swept with (part of the determinism scope)
    persistent: bool = False  # one persistent launch per segment (results identical either way)


def _adapt(step, acc, att, target):
    if att > 0:
        step = min(max(step * math.exp(0.5 * (acc / att - target)), 1e-3), math.pi)
    return step


@dataclass
class PTPlan:
    """What a ladder run is, as the checkpoint stores it (the reference's
    UpdatePlan fields a gpu_pt file holds, src/checkpoint.jl:257-288)."""
    kts: np.ndarray
    exchange_interval: int
    sweeps_therm: int
    sweeps_measure: int
    measure_interval: int
    step: float
    adapt_target: float
    adapt_interval: int
    renorm_interval: int
    nbins: int
    seed: int
    lanes: int


class _Checkpointer:
    """Writes gpu_pt checkpoints every `interval` sweeps (0: only at the
    thermalisation -> measurement boundary) — _Checkpointer/_ck_gpu_pt!
    (src/checkpoint.jl:16-42, 428-437)."""

    def __init__(self, path: str, interval: int):
        if interval < 0:
            raise ValueError(f"checkpoint_interval must be >= 0; got {interval}")
        self.path = str(path)
        self.interval = int(interval)
        self.since = 0

    def room(self, n: int) -> int:
        """Sweeps the next run may take so that a due write lands at its end."""
        return min(n, self.interval - self.since) if self.interval > 0 else n

    def tick(self, n: int, write) -> None:
        if self.interval <= 0:
            return
        self.since += n
        if self.since >= self.interval:
            self.write(write)

    def write(self, write) -> None:
        self.since = 0
        write(self.path)


def _validate_pt(kts, exchange_interval, sweeps_measure, measure_interval):
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


def gpu_run_pt(gH: GPUTiledHamiltonian, temperature=None, kT=None, exchange_interval: int = 10,
               sweeps_therm: int = 2000, sweeps_measure: int = 10000, measure_interval: int = 1,
               step: float = 0.6, adapt_target: float = 0.5, adapt_interval: int = 50,
               renorm_interval: int = 1000, nbins: int = 32, init=None, seed: int = 0,
               lanes: int = DEFAULT_LANES, checkpoint: Optional[str] = None,
               checkpoint_interval: int = 0, persistent: Optional[bool] = None) -> PTResult:
    """Replica exchange on the device — gpu_run_pt (src/gpu/gpu_pt.jl:261-319).

    `checkpoint` / `checkpoint_interval` write restartable checkpoints (kind
    "gpu_pt", interval in sweeps, 0: only at the thermalisation -> measurement
    boundary); `resume_pt(path, gH)` continues one bit-identically.  As in the
    reference there is no end-of-run write.  `lanes="auto"` tunes lanes per
    site (pinned in the plan) and, unless `persistent` is given, whether the
    ladder runs as one persistent launch per segment or as per-colour
    launches; the launch form never changes the results."""
    kts = resolve_kt(temperature, kT)
    _validate_pt(kts, exchange_interval, sweeps_measure, measure_interval)
    plan = PTPlan(kts, int(exchange_interval), int(sweeps_therm), int(sweeps_measure),
                  int(measure_interval), float(step), float(adapt_target), int(adapt_interval),
                  int(renorm_interval), int(nbins), int(seed), 0 if lanes == "auto" else int(lanes))
    if lanes == "auto":
        # tuned on a throwaway copy of the ladder, then pinned in the plan (and
        # so in every checkpoint and the result) like the reference's workgroupsize
        cands = (_BATCH_CANDIDATES if persistent is None else
                 tuple(c for c in _BATCH_CANDIDATES if c[1] == bool(persistent)))
        plan.lanes, persistent, walk = autotune_batch(lambda: _pt_batch(gH, plan, init), cands, sweeps=20)
    ck = _Checkpointer(checkpoint, checkpoint_interval) if checkpoint is not None else None
    b = _pt_batch(gH, plan, init)
    b.set_persistent(bool(persistent))
    if lanes == "auto":
        set_walk_form(b, walk)
    res = _pt_run(b, gH, plan, "therm", 0, None, ck)
    res.persistent = bool(persistent)
    return res


def _pt_batch(gH: GPUTiledHamiltonian, plan: PTPlan, init=None) -> GPUBatch:
    """The ladder's batch as a fresh run starts it (seed derivation of derive_seeds)."""
    H = gH.host
    R = len(plan.kts)
    init_seeds, exch_seed, dev_seeds = derive_seeds(plan.seed, R)
    states = []
    for r in range(R):
        cfg = init if init is not None else random_config(H.n_sites, np.random.default_rng(init_seeds[r]))
        states.append(ChainState(H, cfg, plan.step))
    return GPUBatch([gH] * R, states, dev_seeds, plan.kts, exch_seed)


def _pt_therm(b: GPUBatch, plan: PTPlan, done: int, ck: Optional[_Checkpointer], write) -> None:
    """The thermalisation phase from sweep `done`: exchanges, renormalisation,
    and step adaptation on the counters every adapt_interval sweeps."""
    total = plan.sweeps_therm
    while done < total:
        n = min(plan.adapt_interval - done % plan.adapt_interval, total - done)
        if ck is not None:
            n = ck.room(n)
        b.run(n, plan.lanes, plan.renorm_interval, 0, plan.exchange_interval, done, total)
        done += n
        if done % plan.adapt_interval == 0:
            for r in range(b.R):
                inf = b.info(r)
                b.set_step(r, _adapt(inf.step, inf.acc_metro, inf.att_metro, plan.adapt_target))
            b.reset_counters()
        if ck is not None:
            ck.tick(n, lambda path, d=done: write(path, "therm", d, None))


def _pt_run(b: GPUBatch, gH: GPUTiledHamiltonian, plan: PTPlan, phase: str, done: int,
            rows: Optional[np.ndarray], ck: Optional[_Checkpointer], wall0: float = 0.0) -> PTResult:
    """The phase driver shared by gpu_run_pt and resume_pt — _gpu_pt_run!
    (src/gpu/gpu_pt.jl:141-224): `phase`/`done` enter mid-flight from a
    checkpoint (a fresh run: "therm", 0); `rows` are the measurements a
    measurement-phase checkpoint held."""
    from .checkpoint import save_pt
    H = gH.host
    R = b.R
    t0 = time.perf_counter()

    def write(path, ph, d, obs):
        save_pt(path, b, gH, plan, ph, d, obs, ck.interval if ck is not None else 0)

    if phase == "therm":
        _pt_therm(b, plan, done, ck, write)
        # thermalisation -> measurement boundary: renormalise, fresh counters
        b.renormalize()
        b.reset_counters()
        done = 0
        rows = np.zeros((R, 0, 4))
        if ck is not None:
            ck.write(lambda path: write(path, "measure", 0, rows))
    parts = [rows]
    while done < plan.sweeps_measure:
        n = plan.sweeps_measure - done
        if ck is not None:
            n = ck.room(n)
        parts.append(b.run(n, plan.lanes, plan.renorm_interval, plan.measure_interval,
                           plan.exchange_interval, done, plan.sweeps_measure))
        done += n
        if ck is not None:
            ck.tick(n, lambda path, d=done: write(path, "measure", d, np.concatenate(parts, axis=1)))
    obs = np.concatenate(parts, axis=1)
    wall = wall0 + time.perf_counter() - t0
    points = []
    for r in range(R):
        inf = b.info(r)
        points.append(TempResult(float(plan.kts[r]),
                                 standard_stats(obs[r], float(plan.kts[r]), H.n_active, plan.nbins),
                                 inf.acc_metro / inf.att_metro if inf.att_metro else math.nan,
                                 inf.step, inf.max_drift))
    att, acc = b.swap_stats()
    with np.errstate(invalid="ignore", divide="ignore"):
        swaps = np.where(att > 0, acc / np.maximum(att, 1), np.nan)
    finals = [b.download(r)[0] for r in range(R)]
    return PTResult(points, swaps, finals, plan.seed, wall, plan.lanes)


def resume_pt(path: str, gH: GPUTiledHamiltonian, checkpoint: Optional[str] = "same",
              checkpoint_interval: Optional[int] = None) -> PTResult:
    """Continue a checkpointed gpu_run_pt and return the full run's result,
    bit-identical to the uninterrupted run — resume(path, gH)
    (src/checkpoint.jl:555-635).  By default the resumed run keeps writing to
    `path` with the stored cadence (`checkpoint=None` disables,
    `checkpoint_interval` overrides)."""
    from .checkpoint import load_pt
    st = load_pt(path, gH)
    if checkpoint == "same":
        checkpoint = path
    interval = st["checkpoint_interval"] if checkpoint_interval is None else int(checkpoint_interval)
    ck = _Checkpointer(checkpoint, interval) if checkpoint is not None else None
    return _pt_run(st["batch"], gH, st["plan"], st["phase"], st["done"], st["rows"], ck)


# (lanes per site, persistent launch, walk form) forms the batch autotuner
# tries.  Walk forms at 64 lanes: "" one site per walk, "s" two replicas of a
# batch on one model per walk (set_shared_walk(2)), "q" two same-sublattice
# sites of one chain per walk (set_site_pairs)
_BATCH_CANDIDATES = ((16, False, ""), (32, False, ""), (64, False, ""), (64, False, "s"), (64, False, "q"),
                     (16, True, ""), (32, True, ""), (64, True, ""))


def set_walk_form(b: GPUBatch, walk: str) -> None:
    """Apply an autotuned walk form to a batch (none of them changes results)."""
    b.set_shared_walk(2 if walk == "s" else 1)
    b.set_site_pairs(walk == "q")


def autotune_batch(make_batch, candidates=_BATCH_CANDIDATES, sweeps: int = 30,
                   timings: Optional[dict] = None):
    """The project autotuner for batches: times `sweeps` sweeps of a scratch
    batch built by `make_batch()` in each (lanes per site, persistent, walk
    form) form — per-colour launches replayed as hipGraphs, at 64 lanes also
    with two sites per walk (two replicas of a batch on one model, "s", or two
    same-sublattice sites of one chain, "q"), or one persistent launch per
    segment, each at 16/32/64 lanes (sites per wavefront = 64 / lanes) — and
    returns the fastest as (lanes, persistent, walk); set_walk_form applies
    the walk form.  `lanes` is part of the determinism scope, so tuning runs on
    a throwaway batch before the production run and the caller records the
    choice with the run; the launch and walk forms are not.  `timings`
    receives the seconds per sweep of every candidate that ran, keyed "16",
    "32", "64", "64s", "64q" and "64p"."""
    best, best_t = None, math.inf
    for lanes, pers, walk in candidates:
        b = make_batch()
        if walk == "s" and not b.homogeneous:
            del b
            continue
        try:
            b.set_persistent(pers)
            set_walk_form(b, walk)
            b.run(3, lanes, 0)          # capture + warm
            t0 = time.perf_counter()
            b.run(sweeps, lanes, 0)
            dt = time.perf_counter() - t0
        except L.SceError:
            continue
        finally:
            del b
        if timings is not None:
            timings[f"{lanes}{'p' if pers else ''}{walk}"] = dt / sweeps
        if dt < best_t:
            best, best_t = (int(lanes), bool(pers), walk), dt
    if best is None:
        raise RuntimeError("no candidate form could run this batch")
    return best


def run_concurrent(runs):
    """Run several batches or chains at once: `runs` holds (target, args,
    kwargs), where target is a GPUBatch (its run method is called) or a
    function such as gpu_run_sweeps whose first argument is the chain it
    sweeps; each batch or chain appears at most once.  Every run sweeps on its
    batch's or chain's own stream from its own host thread (the library calls
    release the GIL), so they share the GPU; each one's results are those of
    running it alone.  For the duration the runs launch their kernels directly
    instead of replaying captured graphs (a stream capture in one thread is
    invalidated by a blocking HIP call in another; both forms give the same
    bits), and persistent launches are refused: a cooperative grid needs all
    its blocks resident, which work sharing the GPU can prevent.  Returns the
    results in order; re-raises the first error."""
    owners = [t if hasattr(t, "run") else (a[0] if a else t) for t, a, _ in runs]
    if len({id(o) for o in owners}) != len(owners):
        raise ValueError("a batch or chain can appear only once in a concurrent run")
    if any(getattr(o, "_persistent", False) for o in owners):
        raise ValueError("persistent launches cannot share the GPU; use per-colour launches")
    graphs = [getattr(o, "_graph", None) for o in owners]
    out = [None] * len(runs)
    err = [None] * len(runs)

    def work(i, t, args, kwargs):
        try:
            out[i] = t.run(*args, **kwargs) if hasattr(t, "run") else t(*args, **kwargs)
        except BaseException as e:  # re-raised on the calling thread
            err[i] = e
    try:
        for o, g in zip(owners, graphs):
            if g:
                o.set_graph(False)
        threads = [threading.Thread(target=work, args=(i, t, tuple(a), dict(k)))
                   for i, (t, a, k) in enumerate(runs)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    finally:
        for o, g in zip(owners, graphs):
            if g:
                o.set_graph(True)
    for e in err:
        if e is not None:
            raise e
    return out


def autotune_lanes(make_batch, candidates=(16, 32, 64), sweeps: int = 30, cache: Optional[dict] = None,
                   key=None, timings: Optional[dict] = None) -> int:
    """Lanes per site only (per-colour launches): autotune_batch over the
    candidate lane counts.  Results are memoised in `cache` under `key` (e.g. a
    size-class signature); `timings` receives the seconds per sweep of every
    candidate that ran, keyed by lanes."""
    if cache is not None and key is not None and key in cache:
        return cache[key]
    t = {} if timings is not None else None
    lanes, _, _ = autotune_batch(make_batch, tuple((n, False, "") for n in candidates), sweeps, t)
    if timings is not None:
        timings.update({int(k): v for k, v in t.items()})
    if cache is not None and key is not None:
        cache[key] = lanes
    return lanes
