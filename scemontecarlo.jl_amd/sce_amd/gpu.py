"""Device tables and chains — the drop-in names of the reference's GPU path.

=========================  ==============================================
reference (Julia)          here
=========================  ==============================================
GPUTiledHamiltonian(b, H)  GPUTiledHamiltonian(H)       gpu_hamiltonian.jl:61-88
GPUChainState(gH, st)      GPUChainState(gH, st, seed)  gpu_state.jl:49-61
gpu_metropolis_sweep!      gpu_metropolis_sweep         gpu_sweep.jl:158-190
gpu_run_sweeps!            gpu_run_sweeps               gpu_sweep.jl:204-218
to_host!                   to_host                      gpu_state.jl:71-76
_from_host!                from_host                    gpu_state.jl:81-86
_swap_payload!             swap_payload                 gpu_pt.jl:46-51
=========================  ==============================================

``workgroupsize`` of the reference is ``lanes`` here: the number of lanes that
walk one site (16, 32 or 64; default 64 = one wavefront; 128 and 256, two
and four wavefronts, for plain Metropolis sweeps), part of the
determinism scope exactly like the reference's pinned workgroup size.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .hamiltonian import ChainState, TiledHamiltonian

DEFAULT_LANES = 64


class GPUTiledHamiltonian:
    """Device copy of the tables of ``H`` plus its per-sublattice walk layout."""

    def __init__(self, H: TiledHamiltonian):
        self.host = H
        h = C.c_void_p()
        L.check(L.lib().sce_tables_create(C.byref(H.desc), C.byref(h)))
        self._h = h
        info = np.zeros(5, np.int64)
        L.check(L.lib().sce_tables_layout_info(h, L.ptr(info)))
        (self.n_sublattices, self.nb_max, self.walk_max, self.walk_total,
         self.fmax) = (int(x) for x in info)

    @property
    def handle(self):
        return self._h

    @property
    def n_sites(self) -> int:
        return self.host.n_sites

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.sce_tables_destroy(h)
            self._h = None


class GPUChainState:
    """Device-resident chain: config, rows, staged ΔE, and the chain scalars
    (incremental energy, keyed-RNG seed and sweep index, step, counters) —
    all kept on the device between calls."""

    def __init__(self, gH: GPUTiledHamiltonian, st: ChainState, seed: Optional[int] = None):
        if seed is None:
            seed = int(np.random.default_rng().integers(0, 2**63))
        self.gH = gH
        self.seed = int(seed) & (2**64 - 1)
        cfg = np.ascontiguousarray(st.config, np.float64)
        z = np.ascontiguousarray(st.zrows, np.float64)
        h = C.c_void_p()
        L.check(L.lib().sce_chain_create(gH.handle, L.ptr(cfg), L.ptr(z), float(st.energy),
                                         C.c_uint64(self.seed), float(st.step), C.byref(h)))
        self._h = h
        self._graph, self._persistent = True, False

    @property
    def handle(self):
        return self._h

    def info(self) -> L.ChainInfo:
        inf = L.ChainInfo()
        L.check(L.lib().sce_chain_info_get(self._h, C.byref(inf)))
        return inf

    @property
    def energy(self) -> float:
        return self.info().energy

    @property
    def sweep_index(self) -> int:
        return self.info().sweep_index

    @property
    def step(self) -> float:
        return self.info().step

    @step.setter
    def step(self, v: float) -> None:
        L.check(L.lib().sce_chain_set_step(self._h, float(v)))

    @property
    def acc_metro(self) -> int:
        return self.info().acc_metro

    @property
    def att_metro(self) -> int:
        return self.info().att_metro

    def time_walk(self, lanes: int = DEFAULT_LANES, reps: int = 20):
        """(average colour-walk kernel duration in ms, launches timed); the chain
        state is left unchanged."""
        ms, n = C.c_double(), C.c_int64()
        L.check(L.lib().sce_chain_time_walk(self._h, int(lanes), int(reps), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_counters(self, sweep_index: int, acc_metro: int = 0, att_metro: int = 0,
                     acc_or: int = 0, att_or: int = 0) -> None:
        """Restore the lane-owned counters (resume; the next sweep is sweep_index + 1)."""
        L.check(L.lib().sce_chain_set_counters(self._h, int(sweep_index), int(acc_metro),
                                               int(att_metro), int(acc_or), int(att_or)))

    def set_drift(self, max_drift: float) -> None:
        """Restore the drift record (resume)."""
        L.check(L.lib().sce_chain_set_drift(self._h, float(max_drift)))

    def reset_counters(self) -> None:
        L.check(L.lib().sce_chain_reset_counters(self._h))

    def download(self):
        H = self.gH.host
        cfg = np.empty((H.n_sites, 3))
        z = np.empty((H.n_sites, H.nlm))
        e = C.c_double()
        L.check(L.lib().sce_download(self._h, L.ptr(cfg), L.ptr(z), C.byref(e)))
        return cfg, z, e.value

    def set_persistent(self, on: bool) -> None:
        """Persistent cooperative launches (opt-in) or per-colour launches (default)."""
        L.check(L.lib().sce_chain_set_persistent(self._h, 1 if on else 0))
        self._persistent = bool(on)

    def binner_reset(self, bin_size: int = 0, nbins: int = 0) -> None:
        """Zero the device binner and feed it every later measurement row
        (LogBinner of the five standard series; a BinStore of `nbins` bins of
        `bin_size` rows when both are given)."""
        L.check(L.lib().sce_chain_binner_reset(self._h, int(bin_size), int(nbins)))

    def binner_off(self) -> None:
        L.check(L.lib().sce_chain_binner_off(self._h))

    def binner_state(self):
        """(LogBinner per series, BinStore means per series) as the device holds
        them, in the host classes of ``observables``."""
        from .observables import SERIES, LogBinner, MAX_LEVELS
        n = C.c_int64()
        nfull = C.c_int32()
        count = np.zeros(MAX_LEVELS, np.int64)
        sums = np.zeros((MAX_LEVELS, len(SERIES)))
        sums2 = np.zeros((MAX_LEVELS, len(SERIES)))
        means = np.zeros((1024, len(SERIES)))
        L.check(L.lib().sce_chain_binner_get(self._h, C.byref(n), L.ptr(count), L.ptr(sums), L.ptr(sums2),
                                             C.byref(nfull), L.ptr(means), means.shape[0]))
        binners, bin_means = {}, {}
        for j, name in enumerate(SERIES):
            b = LogBinner(1)
            b.n = n.value
            b.count[:] = count
            b.sums[:, 0] = sums[:, j]
            b.sums2[:, 0] = sums2[:, j]
            binners[name] = b
            bin_means[name] = means[: nfull.value, j].copy()
        return binners, bin_means

    def device_stats(self, kT: float):
        """The standard observables from the device binner (no measurement rows
        are downloaded)."""
        from .observables import stats_from_binners
        binners, means = self.binner_state()
        return stats_from_binners(binners, means, kT, self.gH.host.n_active)

    def set_graph(self, on: bool) -> None:
        """Replay sweeps as hipGraphs (default) or launch their kernels directly."""
        L.check(L.lib().sce_chain_set_graph(self._h, 1 if on else 0))
        self._graph = bool(on)

    def set_timing(self, on: bool) -> None:
        L.check(L.lib().sce_chain_set_timing(self._h, 1 if on else 0))

    def timing(self):
        tot, walk, n = C.c_double(), C.c_double(), C.c_int64()
        L.check(L.lib().sce_chain_timing(self._h, C.byref(tot), C.byref(walk), C.byref(n)))
        return tot.value, walk.value, n.value

    def total_energy(self) -> float:
        e = C.c_double()
        L.check(L.lib().sce_chain_total_energy(self._h, C.byref(e)))
        return e.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.sce_chain_destroy(h)
            self._h = None


def gpu_metropolis_sweep(gst: GPUChainState, gH: GPUTiledHamiltonian, beta: float,
                         lanes: int = DEFAULT_LANES) -> int:
    """One device Metropolis sweep; returns the number of accepted moves."""
    if gst.gH is not gH:
        raise ValueError("chain was created on other tables")
    n = C.c_int64()
    L.check(L.lib().sce_metropolis_sweep(gst.handle, float(beta), int(lanes), C.byref(n)))
    return n.value


def gpu_overrelaxation_sweep(gst: GPUChainState, gH: GPUTiledHamiltonian, beta: float,
                             or_pass: int = 0, lanes: int = DEFAULT_LANES) -> int:
    n = C.c_int64()
    L.check(L.lib().sce_overrelaxation_sweep(gst.handle, float(beta), int(or_pass), int(lanes),
                                             C.byref(n)))
    return n.value


def gpu_run_sweeps(gst: GPUChainState, gH: GPUTiledHamiltonian, st: Optional[ChainState],
                   beta: float, nsweeps: int, renorm_interval: int = 1000,
                   lanes: int = DEFAULT_LANES, or_per_metropolis: int = 0,
                   measure_interval: int = 0):
    """``gpu_run_sweeps!``: ``nsweeps`` sweeps with device renormalisation every
    ``renorm_interval`` sweeps (<= 0 disables); downloads the final state into
    ``st`` when given.  Returns the (E, m_x, m_y, m_z) rows of every
    ``measure_interval``-th sweep (empty when measure_interval <= 0)."""
    nobs = nsweeps // measure_interval if measure_interval > 0 else 0
    obs = np.zeros((max(nobs, 1), 4))
    n = C.c_int64()
    L.check(L.lib().sce_run_sweeps(gst.handle, float(beta), int(nsweeps), int(or_per_metropolis),
                                   int(renorm_interval), int(measure_interval), int(lanes),
                                   L.ptr(obs), int(max(nobs, 1)), C.byref(n)))
    if st is not None:
        to_host(st, gst)
    return obs[: n.value]


def autotune_chain(gH: GPUTiledHamiltonian, st: ChainState, beta: float, sweeps: int = 200,
                   candidates=((16, False), (32, False), (64, False), (128, False), (256, False), (64, True))):
    """The project autotuner for single chains: times `sweeps` sweeps (measured
    every sweep, as production runs are) of a throwaway copy of the chain for
    every (lanes per site, persistent) candidate — per-colour graph launches at
    16/32/64/128/256 lanes (128, 256: two, four wavefronts per site), or the dependency-driven
    persistent sweep at 64 — and returns
    ``(lanes, persistent, timings)`` of the fastest.  Lanes are part of the
    determinism scope: a caller pins the choice for its run and records it."""
    import time
    timings = {}
    best, best_t = None, float("inf")
    for lanes, pers in candidates:
        g = GPUChainState(gH, st, seed=1)
        g.set_persistent(pers)
        try:
            gpu_run_sweeps(g, gH, None, beta, 10, renorm_interval=0, lanes=lanes, measure_interval=1)
            t0 = time.perf_counter()
            gpu_run_sweeps(g, gH, None, beta, sweeps, renorm_interval=0, lanes=lanes, measure_interval=1)
            dt = (time.perf_counter() - t0) / sweeps
        except L.SceError:
            continue
        finally:
            del g
        timings[f"{lanes}{'p' if pers else ''}"] = dt
        if dt < best_t:
            best, best_t = (lanes, pers), dt
    if best is None:
        raise RuntimeError("no candidate could run this chain")
    return best[0], best[1], timings


GRADIENT_LANES = 128  # the reference's default workgroup size for the gradient


def gpu_energy_gradient(gH: GPUTiledHamiltonian, config, lanes: int = GRADIENT_LANES) -> np.ndarray:
    """gpu_energy_gradient!(dG, gH, dconfig, gsc; workgroupsize = lanes)
    (src/gpu/gpu_gradient.jl:187-212): the tangent-projected dE/de_s of every
    site (n_sites x 3) for unit spins `config`, rows rebuilt on the device."""
    cfg = np.ascontiguousarray(config, np.float64)
    H = gH.host
    if cfg.shape != (H.n_sites, 3):
        raise ValueError(f"config has shape {cfg.shape}; expected ({H.n_sites}, 3)")
    G = np.zeros((H.n_sites, 3))
    L.check(L.lib().sce_energy_gradient(gH.handle, L.ptr(cfg), L.ptr(G), int(lanes)))
    return G


def gpu_chain_energy_gradient(gst: "GPUChainState", lanes: int = GRADIENT_LANES) -> np.ndarray:
    """gpu_energy_gradient!(dG, gst, gH, gsc) (:222-226): the same on the
    chain's device spins and rows."""
    G = np.zeros((gst.gH.host.n_sites, 3))
    L.check(L.lib().sce_chain_energy_gradient(gst.handle, L.ptr(G), int(lanes)))
    return G


def gpu_time_gradient(gst: "GPUChainState", lanes: int = GRADIENT_LANES, reps: int = 20) -> float:
    """Average gradient-kernel duration (ms) over `reps` back-to-back launches."""
    ms = C.c_double()
    L.check(L.lib().sce_chain_time_gradient(gst.handle, int(lanes), int(reps), C.byref(ms)))
    return ms.value


def gpu_renormalize(gst: GPUChainState) -> float:
    d = C.c_double()
    L.check(L.lib().sce_renormalize(gst.handle, C.byref(d)))
    return d.value


def to_host(st: ChainState, gst: GPUChainState) -> ChainState:
    cfg, z, e = gst.download()
    st.config[...] = cfg
    st.zrows[...] = z
    st.energy = e
    return st


def from_host(gst: GPUChainState, st: ChainState) -> GPUChainState:
    cfg = np.ascontiguousarray(st.config, np.float64)
    z = np.ascontiguousarray(st.zrows, np.float64)
    L.check(L.lib().sce_upload(gst.handle, L.ptr(cfg), L.ptr(z), float(st.energy)))
    return gst


def swap_payload(a: GPUChainState, b: GPUChainState) -> None:
    L.check(L.lib().sce_swap_payload(a.handle, b.handle))


# ---- device arithmetic exposed for parity tests -------------------------------

def dev_philox(seeds, ctrs):
    seeds = np.ascontiguousarray(seeds, np.uint64)
    ctrs = np.ascontiguousarray(ctrs, np.uint32).reshape(-1, 4)
    out = np.zeros_like(ctrs)
    L.check(L.lib().sce_test_philox(len(seeds), L.ptr(seeds), L.ptr(ctrs), L.ptr(out)))
    return out


def chain_plan(color_ptr, next_prop: bool, pending: int, carry_last: bool, walk_props: bool = True):
    """Where a chained sweep makes its proposals (host only): (records, kf) with
    records rows (launch, lo, n, sweep offset), launch -1 before the walks,
    0..K-1 the non-empty classes' walks, K the reduction, K + 1 after it."""
    cp = np.ascontiguousarray(color_ptr, np.int32)
    out = np.zeros((4 * len(cp) + 8, 4), np.int32)
    n = C.c_int32()
    kf = C.c_int32()
    L.check(L.lib().sce_test_chain_plan(len(cp) - 1, L.ptr(cp), int(next_prop), int(pending), int(carry_last),
                                        int(walk_props), L.ptr(out), len(out), C.byref(n), C.byref(kf)))
    return out[: n.value], kf.value


def dev_detmath(op: int, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    L.check(L.lib().sce_test_detmath(len(x), op, L.ptr(x), L.ptr(out)))
    return out


def dev_zlm(dirs, lmax):
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    out = np.zeros((len(d), (lmax + 1) ** 2))
    L.check(L.lib().sce_test_zlm(len(d), lmax, L.ptr(d), L.ptr(out)))
    return out


def dev_proposal(seed, sites, sweeps, e, step):
    sites = np.ascontiguousarray(sites, np.int32)
    sweeps = np.ascontiguousarray(sweeps, np.int32)
    e = np.ascontiguousarray(e, np.float64).reshape(-1, 3)
    e2 = np.zeros_like(e)
    u = np.zeros(len(sites))
    L.check(L.lib().sce_test_proposal(len(sites), C.c_uint64(seed), L.ptr(sites), L.ptr(sweeps),
                                      L.ptr(e), float(step), L.ptr(e2), L.ptr(u)))
    return e2, u


def dev_walk_delta(gH: GPUTiledHamiltonian, zrows, sites, znew, lanes):
    z = np.ascontiguousarray(zrows, np.float64)
    sites = np.ascontiguousarray(sites, np.int32)
    zn = np.ascontiguousarray(znew, np.float64)
    out = np.zeros(len(sites))
    L.check(L.lib().sce_test_walk_delta(gH.handle, L.ptr(z), len(sites), L.ptr(sites), L.ptr(zn),
                                        int(lanes), L.ptr(out)))
    return out


def dev_walk_coeffs(gH: GPUTiledHamiltonian, zrows, sites, lanes):
    z = np.ascontiguousarray(zrows, np.float64)
    sites = np.ascontiguousarray(sites, np.int32)
    out = np.zeros((len(sites), gH.host.nlm))
    L.check(L.lib().sce_test_walk_coeffs(gH.handle, L.ptr(z), len(sites), L.ptr(sites), int(lanes),
                                         L.ptr(out)))
    return out


def dev_reduce(x):
    x = np.ascontiguousarray(x, np.float64)
    o = C.c_double()
    L.check(L.lib().sce_test_reduce(len(x), L.ptr(x), C.byref(o)))
    return o.value


def dev_mfma(a4, b4, c):
    """One matrix-core instruction per row: c + sum_k a[k] b[k] (diagnostic)."""
    a = np.ascontiguousarray(a4, np.float32).reshape(-1, 4)
    b = np.ascontiguousarray(b4, np.float32).reshape(-1, 4)
    cc = np.ascontiguousarray(c, np.float32)
    out = np.zeros(max(len(cc), 1), np.float32)
    L.check(L.lib().sce_test_mfma(len(cc), L.ptr(a), L.ptr(b), L.ptr(cc), L.ptr(out)))
    return out[: len(cc)]
