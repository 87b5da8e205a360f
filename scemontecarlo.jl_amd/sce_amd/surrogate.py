"""Learned delayed acceptance for the device sweep.

A :class:`Surrogate` holds one small MLP per sublattice that predicts a site's
energy from a rotation-invariant descriptor of its spin against its neighbour
spins.  :func:`gpu_da_run_sweeps` runs two-stage (Christen & Fox 2005) sweeps:
the surrogate difference screens each keyed proposal, and only survivors pay
for the exact FP64 walk, whose ``dE - dE_s`` makes the final decision — so the
chain samples exp(-beta E) exactly whatever the network has learned.

Training tuples (phi(e), phi(e'), exact dE) are harvested on the device by plain
Metropolis sweeps (:meth:`Surrogate.harvest`) and fitted by the library's own
MFMA FP32 training kernels with Adam (:meth:`Surrogate.train`).

No reference counterpart: SCEMonteCarlo.jl has plain Metropolis and
overrelaxation only (src/updates.jl, src/gpu/gpu_sweep.jl).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .gpu import GPUChainState, GPUTiledHamiltonian, to_host
from .hamiltonian import ChainState

H = 64  # hidden width of both layers


def param_layout(F: int) -> dict:
    """Slices of the flat FP32 parameter vector of a sublattice with F features."""
    o, out = 0, {}
    for name, n in (("W1", H * F), ("b1", H), ("W2", H * H), ("b2", H), ("w3", H), ("v", F)):
        out[name] = slice(o, o + n)
        o += n
    out["n"] = o
    return out


class Surrogate:
    def __init__(self, gH: GPUTiledHamiltonian, seed: int = 1):
        self.gH = gH
        h = C.c_void_p()
        L.check(L.lib().sce_sur_create(gH.handle, int(seed) & (2**64 - 1), C.byref(h)))
        self._h = h
        self.n_sublattices = gH.n_sublattices
        self.fmax = gH.fmax

    @property
    def handle(self):
        return self._h

    def info(self, sub: int):
        F, n, scale, rows = C.c_int32(), C.c_int64(), C.c_double(), C.c_int64()
        L.check(L.lib().sce_sur_info(self._h, int(sub), C.byref(F), C.byref(n), C.byref(scale),
                                     C.byref(rows)))
        return {"F": F.value, "n_params": n.value, "scale": scale.value, "n_rows": rows.value}

    @property
    def n_rows(self) -> int:
        return self.info(0)["n_rows"]

    def active_sublattices(self):
        return [a for a in range(self.n_sublattices) if self.info(a)["F"] > 0]

    def get_params(self, sub: int) -> np.ndarray:
        out = np.zeros(self.info(sub)["n_params"], np.float32)
        L.check(L.lib().sce_sur_get_params(self._h, int(sub), L.ptr(out)))
        return out

    def set_params(self, sub: int, params, scale: float) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.info(sub)["n_params"]:
            raise L.DimensionMismatch(-2, "parameter vector length")
        L.check(L.lib().sce_sur_set_params(self._h, int(sub), L.ptr(p), float(scale)))

    def set_int8(self, on: bool) -> None:
        """Stage 1 on the int8 networks (matrix-core integer dot products of the
        quantised descriptor and weights) or on the FP32 networks (default)."""
        L.check(L.lib().sce_sur_set_int8(self._h, 1 if on else 0))

    def get_qparams(self, sub: int) -> np.ndarray:
        """The int8 network of sublattice `sub` as a float32 blob: V1 (64 x F),
        r1, b1, V2 (64 x 64), r2, b2, w3 (64 each), v (F), input scales (F)."""
        F = self.info(sub)["F"]
        out = np.zeros(H * F + 5 * H + H * H + 2 * F, np.float32)
        L.check(L.lib().sce_sur_get_qparams(self._h, int(sub), L.ptr(out)))
        return out

    def harvest(self, gst: GPUChainState, beta: float, nsweeps: int) -> int:
        """Plain Metropolis sweeps of gst that record every attempt as a training row."""
        L.check(L.lib().sce_sur_harvest(self._h, gst.handle, float(beta), int(nsweeps)))
        return self.n_rows

    def train(self, sub: int, epochs: int = 4, batch: int = 4096, lr: float = 1e-3,
              seed: int = 7) -> np.ndarray:
        loss = np.zeros(max(epochs, 1))
        L.check(L.lib().sce_sur_train(self._h, int(sub), int(epochs), int(batch), float(lr),
                                      int(seed) & (2**64 - 1), L.ptr(loss)))
        return loss[:epochs]

    def train_all(self, epochs: int = 4, batch: int = 4096, lr: float = 1e-3, seed: int = 7):
        return {a: self.train(a, epochs, batch, lr, seed + a) for a in self.active_sublattices()}

    def grad(self, sub: int, A, B, y):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        y = np.ascontiguousarray(y, np.float64)
        n = len(y)
        if A.shape != (n, self.fmax) or B.shape != (n, self.fmax):
            raise L.DimensionMismatch(-2, "descriptor rows must be n x fmax")
        g = np.zeros(self.info(sub)["n_params"], np.float32)
        loss = C.c_double()
        L.check(L.lib().sce_sur_grad(self._h, int(sub), n, L.ptr(A), L.ptr(B), L.ptr(y), L.ptr(g),
                                     C.byref(loss)))
        return g, loss.value

    def eval(self, sub: int, A, B) -> np.ndarray:
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        n = A.shape[0]
        if A.shape != (n, self.fmax) or B.shape != (n, self.fmax):
            raise L.DimensionMismatch(-2, "descriptor rows must be n x fmax")
        out = np.zeros(max(n, 1))
        L.check(L.lib().sce_sur_eval(self._h, int(sub), n, L.ptr(A), L.ptr(B), L.ptr(out)))
        return out[:n]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().sce_sur_destroy(h)
            finally:
                self._h = C.c_void_p()


def dev_features(gH: GPUTiledHamiltonian, config, sites, e2):
    """Descriptor rows phi(e_site), phi(e2) (n x fmax FP32) computed on the device."""
    cfg = np.ascontiguousarray(config, np.float64)
    s = np.ascontiguousarray(sites, np.int32)
    e = np.ascontiguousarray(e2, np.float64)
    n = len(s)
    A = np.zeros((max(n, 1), gH.fmax), np.float32)
    B = np.zeros((max(n, 1), gH.fmax), np.float32)
    L.check(L.lib().sce_sur_features(gH.handle, L.ptr(cfg), n, L.ptr(s), L.ptr(e), L.ptr(A), L.ptr(B)))
    return A[:n], B[:n]


def gpu_da_run_sweeps(gst: GPUChainState, sur: Surrogate, st: Optional[ChainState], beta: float,
                      nsweeps: int, renorm_interval: int = 1000, measure_interval: int = 0):
    """Delayed-acceptance counterpart of ``gpu_run_sweeps``: the same schedule of
    renormalisation and (E, m) measurement; ``gst.info().acc_stage1`` counts the
    proposals that passed the surrogate stage."""
    nobs = nsweeps // measure_interval if measure_interval > 0 else 0
    obs = np.zeros((max(nobs, 1), 4))
    n = C.c_int64()
    L.check(L.lib().sce_da_run_sweeps(gst.handle, sur.handle, float(beta), int(nsweeps),
                                      int(renorm_interval), int(measure_interval), L.ptr(obs),
                                      int(max(nobs, 1)), C.byref(n)))
    if st is not None:
        to_host(st, gst)
    return obs[: n.value]
