"""Streaming error analysis and the standard observables.

Restates src/binning.jl (LogBinner :25-128, BinStore :138-188, jackknife
:198-218) and the standard evaluables of src/observables.jl:160-166 (specific
heat, |m|-connected susceptibility, Binder ratio) for the device measurement
stream (E, m) that ``sce_run_sweeps`` produces.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, List, Sequence

import numpy as np

MIN_BINS = 32
MAX_LEVELS = 62


class LogBinner:
    """Logarithmic binning with O(levels) memory (binning.jl:25-79)."""

    def __init__(self, ncomp: int = 1):
        if ncomp < 1:
            raise ValueError("ncomp must be >= 1")
        self.ncomp = ncomp
        self.count = np.zeros(MAX_LEVELS, np.int64)
        self.sums = np.zeros((MAX_LEVELS, ncomp))
        self.sums2 = np.zeros((MAX_LEVELS, ncomp))
        self.pending = np.zeros((MAX_LEVELS, ncomp))
        self.pending_full = np.zeros(MAX_LEVELS, bool)
        self.n = 0

    def push(self, xs) -> "LogBinner":
        xs = np.atleast_1d(np.asarray(xs, np.float64))
        if xs.size != self.ncomp:
            raise ValueError(f"LogBinner of width {self.ncomp} was pushed {xs.size} values")
        self.n += 1
        for k in range(MAX_LEVELS):
            self.count[k] += 1
            x = xs if k == 0 else self.pending[k - 1].copy()
            self.sums[k] += x
            self.sums2[k] += x * x
            if self.pending_full[k]:
                self.pending[k] = 0.5 * (self.pending[k] + x)
                self.pending_full[k] = False
            else:
                self.pending[k] = x
                self.pending_full[k] = True
                break
        return self

    def push_series(self, series) -> "LogBinner":
        for x in np.asarray(series, np.float64).reshape(-1, self.ncomp):
            self.push(x)
        return self

    def mean(self) -> np.ndarray:
        return np.full(self.ncomp, np.nan) if self.n == 0 else self.sums[0] / self.n

    def level_error(self, k: int, j: int) -> float:
        n = self.count[k]
        if n < 2:
            return math.nan
        var = (self.sums2[k, j] - self.sums[k, j] ** 2 / n) / (n - 1)
        return math.sqrt(max(var, 0.0) / n)

    def plateau_level(self) -> int:
        ks = np.nonzero(self.count >= MIN_BINS)[0]
        return int(ks[-1]) if ks.size else 0

    def std_error(self) -> np.ndarray:
        if self.n < 2:
            return np.full(self.ncomp, np.nan)
        k = self.plateau_level()
        return np.array([self.level_error(k, j) for j in range(self.ncomp)])

    def tau_int(self) -> np.ndarray:
        if self.n < 2:
            return np.full(self.ncomp, np.nan)
        k = self.plateau_level()
        out = []
        for j in range(self.ncomp):
            e1 = self.level_error(0, j)
            out.append(0.5 * ((self.level_error(k, j) / e1) ** 2 - 1) if e1 > 0 else math.nan)
        return np.array(out)


class BinStore:
    """Fixed number of equal-size bin means for jackknifing (binning.jl:138-188)."""

    def __init__(self, ncomp: int, bin_size: int, nbins: int):
        if bin_size < 1 or nbins < 2 or ncomp < 1:
            raise ValueError("invalid BinStore layout")
        self.ncomp, self.bin_size = ncomp, bin_size
        self.means = np.zeros((nbins, ncomp))
        self.nfull = 0
        self.acc = np.zeros(ncomp)
        self.nacc = 0

    def push(self, xs) -> "BinStore":
        xs = np.atleast_1d(np.asarray(xs, np.float64))
        if self.nfull >= self.means.shape[0]:
            return self
        self.acc += xs
        self.nacc += 1
        if self.nacc == self.bin_size:
            self.means[self.nfull] = self.acc / self.bin_size
            self.nfull += 1
            self.acc[:] = 0.0
            self.nacc = 0
        return self

    def bin_means(self) -> np.ndarray:
        return self.means[: self.nfull].copy()


def jackknife(f: Callable, cols: Sequence[np.ndarray]):
    """Leave-one-bin-out jackknife (binning.jl:198-218): (estimate, error)."""
    if not cols:
        raise ValueError("jackknife needs at least one input series")
    nb = len(cols[0])
    if any(len(c) != nb for c in cols):
        raise ValueError("jackknife input series have unequal lengths")
    if nb < 2:
        raise ValueError("jackknife needs >= 2 bins")
    totals = [float(np.sum(c)) for c in cols]
    full = f(*[t / nb for t in totals])
    th = np.array([f(*[(t - c[i]) / (nb - 1) for t, c in zip(totals, cols)]) for i in range(nb)])
    tbar = th.mean()
    est = nb * full - (nb - 1) * tbar
    err = math.sqrt((nb - 1) / nb * float(np.sum((th - tbar) ** 2)))
    return est, err


@dataclass
class ObservableStat:
    name: str
    mean: float
    err: float
    tau_int: float
    count: int


SERIES = ("energy", "energy2", "absm", "m2", "m4")


def standard_series(obs: np.ndarray) -> Dict[str, np.ndarray]:
    """The five series binned per measurement row (E, m_x, m_y, m_z), with the
    device binner's operation order: m^2 = (m_x^2 + m_y^2) + m_z^2."""
    obs = np.asarray(obs, np.float64)
    E = obs[:, 0]
    mx, my, mz = obs[:, 1], obs[:, 2], obs[:, 3]
    m2 = (mx * mx + my * my) + mz * mz
    return {"energy": E, "energy2": E * E, "absm": np.sqrt(m2), "m2": m2, "m4": m2 * m2}


def stats_from_binners(binners: Dict[str, LogBinner], bin_means: Dict[str, np.ndarray], kT: float,
                       n_active: int) -> Dict[str, ObservableStat]:
    """Finalise the standard observables (observables.jl:85-98, 160-166,
    207-249) from the LogBinner and BinStore means of the five series."""
    stats: Dict[str, ObservableStat] = {}
    for name in SERIES:
        b = binners[name]
        stats[name] = ObservableStat(name, float(b.mean()[0]), float(b.std_error()[0]),
                                     float(b.tau_int()[0]), b.n)
    evals = {
        "specific_heat": (("energy", "energy2"), lambda e, e2: (e2 - e * e) / (n_active * kT * kT)),
        "susceptibility": (("m2", "absm"), lambda q2, a: n_active * (q2 - a * a) / kT),
        "binder": (("m2", "m4"), lambda q2, q4: q4 / (q2 * q2)),
    }
    for name, (inputs, f) in evals.items():
        cols = [bin_means[i] for i in inputs]
        if len(cols[0]) < 2:
            stats[name] = ObservableStat(name, math.nan, math.nan, math.nan, 0)
            continue
        est, err = jackknife(f, cols)
        stats[name] = ObservableStat(name, est, err, math.nan, len(cols[0]))
    return stats


def host_binners(obs: np.ndarray, nbins: int = 32):
    """(LogBinner per series, BinStore means per series) of measurement rows,
    BinStore bins of max(1, n // nbins) rows."""
    series = standard_series(obs)
    n = len(series["energy"])
    bin_size = max(1, n // nbins)
    binners, means = {}, {}
    for name in SERIES:
        x = series[name]
        binners[name] = LogBinner(1).push_series(x)
        s = BinStore(1, bin_size, nbins)
        for v in x:
            s.push(v)
        means[name] = s.bin_means()[:, 0]
    return binners, means


def standard_stats(obs: np.ndarray, kT: float, n_active: int, nbins: int = 32) -> Dict[str, ObservableStat]:
    """Finalise the standard observables (observables.jl:85-98, 160-166, 207-249)
    from device measurement rows (E, m_x, m_y, m_z)."""
    binners, means = host_binners(obs, nbins)
    return stats_from_binners(binners, means, kT, n_active)


def tau_int_windowed(series: np.ndarray, c: float = 6.0) -> float:
    """tau_int by summing the normalised autocorrelation up to Sokal's automatic
    window (the first M with M >= c (1/2 + sum_{t<=M} rho(t))), in the
    log-binning convention (1 + 2 tau = variance ratio, so tau = 0 for iid
    data).  A cross-check of the plateau estimate, which scatters widely when
    tau is a few samples; not the reference's definition."""
    x = np.asarray(series, np.float64)
    n = len(x)
    if n < 2:
        return math.nan
    x = x - x.mean()
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    if acf[0] <= 0.0:
        return math.nan
    acf = acf / acf[0]
    t = 0.5
    for m in range(1, n):
        t += acf[m]
        if m >= c * t:
            break
    return t - 0.5


def effective_samples(series: np.ndarray, min_level: int = 1) -> float:
    """n / (1 + 2 tau_int), tau_int from the log-binning plateau (binning.jl:120-128).

    NaN unless the plateau (the deepest level holding >= 32 bins) is at least
    level ``min_level``: with fewer than 32 * 2**min_level samples there is no
    binned level to compare with the raw one, and tau_int would read 0 (every
    sample counted as independent) whatever the correlation."""
    b = LogBinner(1).push_series(series)
    if b.count[min_level] < MIN_BINS:
        return math.nan
    tau = float(b.tau_int()[0])
    if not math.isfinite(tau):
        return math.nan
    return b.n / (1.0 + 2.0 * max(tau, 0.0))
