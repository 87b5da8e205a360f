"""Checkpoint and bit-identical resume of device chains.

A checkpoint is a NumPy ``.npz`` of plain data only (no pickled objects; loaded
with ``allow_pickle=False``): the chain payload (config, rows, incremental
energy), its lane-owned keyed-RNG state (seed, sweep index) and counters, the
step, and a fingerprint of the model.  Because every random number of a device
chain is keyed by (seed, site, sweep), that is the whole state: a chain
restored from a checkpoint continues bit-for-bit as if it had not stopped.
Writes go to a temporary file in the same directory and are renamed into place.

Reference: the checkpoint/resume of SCEMonteCarlo.jl (src/checkpoint.jl:54-131,
388-438; plain-data schema of docs/specs/checkpoint-schema.md).  The fingerprint
here is SHA-256 over the flat device tables (every array a TiledHamiltonian
holds, plus dims and lmax) rather than the reference's FNV-style mixer over the
term list; it identifies the same models and detects any change of them.
"""
from __future__ import annotations

import hashlib
import os
import tempfile

import numpy as np

from . import _lib as L
from .gpu import GPUChainState, GPUTiledHamiltonian
from .hamiltonian import ChainState, TiledHamiltonian

FORMAT = "sce_amd-chain-1"
PT_FORMAT = "sce_amd-gpu_pt-1"


def model_fingerprint(H: TiledHamiltonian) -> str:
    h = hashlib.sha256()
    h.update(np.asarray([H.n_cell_atoms, *H.dims, H.lmax, H.nlm, H.n_sites], np.int64).tobytes())
    for name, _, _ in L.ARRAY_FIELDS:
        a = np.ascontiguousarray(getattr(H, name))
        h.update(name.encode())
        h.update(np.asarray(a.shape, np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def _atomic_savez(path: str, data: dict) -> None:
    d = os.path.dirname(os.path.abspath(path)) or "."
    fd, tmp = tempfile.mkstemp(prefix=".ckpt-", suffix=".npz", dir=d)
    os.close(fd)
    try:
        with open(tmp, "wb") as f:
            np.savez(f, **data)
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


def save_chain(path: str, gst: GPUChainState) -> None:
    cfg, z, E = gst.download()
    inf = gst.info()
    data = dict(format=np.asarray(FORMAT), fingerprint=np.asarray(model_fingerprint(gst.gH.host)),
                config=cfg, zrows=z, energy=np.float64(E), seed=np.uint64(inf.seed),
                sweep_index=np.int64(inf.sweep_index), step=np.float64(inf.step),
                counters=np.asarray([inf.acc_metro, inf.att_metro, inf.acc_or, inf.att_or], np.int64),
                max_drift=np.float64(inf.max_drift))
    _atomic_savez(path, data)


def load_chain(path: str, gH: GPUTiledHamiltonian) -> GPUChainState:
    """A device chain on `gH` in the checkpointed state; raises ValueError when the
    file is not a chain checkpoint or was written for another model."""
    with np.load(path, allow_pickle=False) as f:
        if "format" not in f or str(f["format"]) != FORMAT:
            raise ValueError(f"{path}: not a {FORMAT} checkpoint")
        if str(f["fingerprint"]) != model_fingerprint(gH.host):
            raise ValueError(f"{path}: checkpoint was written for a different model")
        H = gH.host
        cfg = np.array(f["config"], np.float64)
        z = np.array(f["zrows"], np.float64)
        if cfg.shape != (H.n_sites, 3) or z.shape != (H.n_sites, H.nlm):
            raise ValueError(f"{path}: array shapes do not match the model")
        st = ChainState(H, cfg, float(f["step"]))
        st.zrows[...] = z
        st.energy = float(f["energy"])
        gst = GPUChainState(gH, st, seed=int(f["seed"]))
        c = [int(x) for x in f["counters"]]
        gst.set_counters(int(f["sweep_index"]), c[0], c[1], c[2], c[3])
        gst.set_drift(float(f["max_drift"]))
    return gst


# ---- parallel tempering ("gpu_pt" kind, src/checkpoint.jl:388-426) ----------
_PLAN_INT = ("exchange_interval", "sweeps_therm", "sweeps_measure", "measure_interval",
             "adapt_interval", "renorm_interval", "nbins", "lanes")
_PLAN_FLOAT = ("step", "adapt_target")


def save_pt(path: str, b, gH: GPUTiledHamiltonian, plan, phase: str, done: int, rows,
            interval: int) -> None:
    """Every rung's payload (config, rows, incremental energy) and lane-owned
    state (keyed-RNG seed and sweep index, step, counters, drift record), the
    exchange schedule's state, the run's plan and progress, and the
    measurements so far.  Reads the device; changes nothing there."""
    R = b.R
    cfgs, zs, es, seeds, sweeps, steps, cnt, drift = [], [], [], [], [], [], [], []
    for k in range(R):
        cfg, z, E = b.download(k)
        inf = b.info(k)
        cfgs.append(cfg)
        zs.append(z)
        es.append(E)
        seeds.append(inf.seed)
        sweeps.append(inf.sweep_index)
        steps.append(inf.step)
        cnt.append([inf.acc_metro, inf.att_metro, inf.acc_or, inf.att_or])
        drift.append(inf.max_drift)
    count, parity = b.exchange_state()
    att, acc = b.swap_stats()
    data = dict(format=np.asarray(PT_FORMAT), kind=np.asarray("gpu_pt"),
                fingerprint=np.asarray(model_fingerprint(gH.host)),
                checkpoint_interval=np.int64(interval), phase=np.asarray(phase), done=np.int64(done),
                plan_kts=np.asarray(plan.kts, np.float64), plan_seed=np.uint64(int(plan.seed) & (2**64 - 1)),
                exchange_count=np.int64(count), exchange_parity=np.int64(parity),
                exchange_seed=np.uint64(b.exchange_seed), swap_att=np.asarray(att, np.int64),
                swap_acc=np.asarray(acc, np.int64), configs=np.stack(cfgs), zrows=np.stack(zs),
                energies=np.asarray(es, np.float64), dev_seeds=np.asarray(seeds, np.uint64),
                sweep_index=np.asarray(sweeps, np.int64), steps=np.asarray(steps, np.float64),
                counters=np.asarray(cnt, np.int64), max_drift=np.asarray(drift, np.float64),
                rows=np.zeros((R, 0, 4)) if rows is None else np.asarray(rows, np.float64))
    for name in _PLAN_INT:
        data["plan_" + name] = np.int64(getattr(plan, name))
    for name in _PLAN_FLOAT:
        data["plan_" + name] = np.float64(getattr(plan, name))
    _atomic_savez(path, data)


def load_pt(path: str, gH: GPUTiledHamiltonian) -> dict:
    """Reads and validates a gpu_pt checkpoint and rebuilds its batch on `gH`;
    raises FileNotFoundError for a missing file and ValueError for a file of
    another kind or model, or an inconsistent one."""
    from .batch import GPUBatch, PTPlan
    if not os.path.isfile(path):
        raise FileNotFoundError(f"no checkpoint file at {path}")
    with np.load(path, allow_pickle=False) as f:
        fmt = str(f["format"]) if "format" in f else ""
        if fmt == FORMAT:
            raise ValueError(f"{path} is a single-chain checkpoint: load it with load_chain(), "
                             "not as a gpu_run_pt checkpoint")
        if fmt != PT_FORMAT:
            raise ValueError(f"{path}: not a {PT_FORMAT} checkpoint")
        if str(f["fingerprint"]) != model_fingerprint(gH.host):
            raise ValueError(f"{path}: checkpoint was written for a different model")
        H = gH.host
        kts = np.array(f["plan_kts"], np.float64)
        R = len(kts)
        plan = PTPlan(kts, **{k: int(f["plan_" + k]) for k in _PLAN_INT if k != "lanes"},
                      **{k: float(f["plan_" + k]) for k in _PLAN_FLOAT},
                      seed=int(f["plan_seed"]), lanes=int(f["plan_lanes"]))
        phase = str(f["phase"])
        done = int(f["done"])
        total = plan.sweeps_therm if phase == "therm" else plan.sweeps_measure
        if phase not in ("therm", "measure") or not 0 <= done <= total:
            raise ValueError(f"{path}: progress {phase}/{done} is not a point of this run")
        cfgs = np.array(f["configs"], np.float64)
        zs = np.array(f["zrows"], np.float64)
        if cfgs.shape != (R, H.n_sites, 3) or zs.shape != (R, H.n_sites, H.nlm):
            raise ValueError(f"{path}: payload shapes do not match {R} rungs of this model")
        rows = np.array(f["rows"], np.float64)
        want = done // plan.measure_interval if phase == "measure" else 0
        if rows.shape != (R, want, 4):
            raise ValueError(f"{path}: holds {rows.shape[1]} measurements per rung, expected {want}")
        energies = np.array(f["energies"], np.float64)
        steps = np.array(f["steps"], np.float64)
        states = []
        for k in range(R):
            st = ChainState(H, cfgs[k], float(steps[k]))
            st.zrows[...] = zs[k]
            st.energy = float(energies[k])
            states.append(st)
        b = GPUBatch([gH] * R, states, [int(x) for x in f["dev_seeds"]], kts, int(f["exchange_seed"]))
        cnt = np.array(f["counters"], np.int64)
        drift = np.array(f["max_drift"], np.float64)
        sweeps = np.array(f["sweep_index"], np.int64)
        for k in range(R):
            b.set_counters(k, int(sweeps[k]), int(cnt[k, 0]), int(cnt[k, 1]), int(cnt[k, 2]),
                           int(cnt[k, 3]), float(drift[k]))
        b.set_exchange(int(f["exchange_count"]), int(f["exchange_parity"]), np.array(f["swap_att"]),
                       np.array(f["swap_acc"]))
        return dict(batch=b, plan=plan, phase=phase, done=done, rows=rows,
                    checkpoint_interval=int(f["checkpoint_interval"]))


def resume(path: str, gH: GPUTiledHamiltonian, **kw):
    """resume(path, gH) of the reference (src/checkpoint.jl:555): continues a
    gpu_run_pt checkpoint (see batch.resume_pt); a single-chain checkpoint is
    reloaded with load_chain instead."""
    from .batch import resume_pt
    return resume_pt(path, gH, **kw)
