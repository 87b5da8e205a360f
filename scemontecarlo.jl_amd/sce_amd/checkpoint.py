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


def model_fingerprint(H: TiledHamiltonian) -> str:
    h = hashlib.sha256()
    h.update(np.asarray([H.n_cell_atoms, *H.dims, H.lmax, H.nlm, H.n_sites], np.int64).tobytes())
    for name, _, _ in L.ARRAY_FIELDS:
        a = np.ascontiguousarray(getattr(H, name))
        h.update(name.encode())
        h.update(np.asarray(a.shape, np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def save_chain(path: str, gst: GPUChainState) -> None:
    cfg, z, E = gst.download()
    inf = gst.info()
    data = dict(format=np.asarray(FORMAT), fingerprint=np.asarray(model_fingerprint(gst.gH.host)),
                config=cfg, zrows=z, energy=np.float64(E), seed=np.uint64(inf.seed),
                sweep_index=np.int64(inf.sweep_index), step=np.float64(inf.step),
                counters=np.asarray([inf.acc_metro, inf.att_metro, inf.acc_or, inf.att_or], np.int64),
                max_drift=np.float64(inf.max_drift))
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
    return gst
