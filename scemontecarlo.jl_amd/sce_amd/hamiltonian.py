"""Host ``TiledHamiltonian`` and ``ChainState`` — the inputs of the device path.

Mirrors the reference's constructor ``TiledHamiltonian(n_cell_atoms, terms; dims,
magmoms, field)`` (src/hamiltonian.jl:267-393) and ``ChainState``
(src/state.jl:17-46): the tables are built by the native builder of
``libsce_hip.so`` and exposed as numpy arrays with the reference's 1-based
conventions, so a parity test can read them field by field.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .models import MultipoleTerm

KB_EV = 1.380649e-23 / 1.602176634e-19        # src/units.jl:17
MU_B_EV_T = 9.2740100783e-24 / 1.602176634e-19  # src/units.jl:29


def flatten_terms(terms: Sequence[MultipoleTerm]):
    body, coef, atoms, shifts, ls, folded = [], [], [], [], [], []
    for t in terms:
        if len(t.atoms) != len(t.shifts) or len(t.atoms) != len(t.ls):
            raise ValueError("atoms/shifts/ls lengths disagree")
        f = np.asarray(t.folded, np.float64)
        if f.shape != tuple(2 * l + 1 for l in t.ls) and not (len(t.ls) == 1 and f.shape == (2 * t.ls[0] + 1,)):
            raise ValueError(f"size(folded) = {f.shape} does not match ls = {t.ls}")
        body.append(len(t.atoms))
        coef.append(float(t.coef))
        atoms.extend(int(a) for a in t.atoms)
        for v in t.shifts:
            shifts.extend(int(x) for x in v)
        ls.extend(int(l) for l in t.ls)
        folded.extend(f.reshape(-1, order="F"))
    return (np.asarray(body, np.int32), np.asarray(coef, np.float64),
            np.asarray(atoms, np.int32), np.asarray(shifts, np.int32),
            np.asarray(ls, np.int32), np.asarray(folded, np.float64))


class TiledHamiltonian:
    """The fitted SCE Hamiltonian tiled onto a ``dims`` supercell (toroidal).

    Attributes mirror the reference struct: ``n_sites``, ``lmax``, ``nlm``,
    ``n_active``, ``n_colors`` and the flat tables (``site_ptr``, ``site_prog``,
    ``sent_w``, ``color_sites`` ...), all 1-based like the reference's.
    """

    def __init__(self, n_cell_atoms: int, terms: Sequence[MultipoleTerm],
                 dims=(1, 1, 1), magmoms: Optional[Sequence[float]] = None,
                 field: Optional[Sequence[float]] = None):
        body, coef, atoms, shifts, ls, folded = flatten_terms(terms)
        d = np.asarray(dims, np.int32)
        mm = None if magmoms is None else np.ascontiguousarray(magmoms, np.float64)
        if mm is not None and len(mm) != n_cell_atoms:
            raise ValueError(f"magmoms has {len(mm)} entries but the cell has {n_cell_atoms} atoms")
        fl = None if field is None else np.ascontiguousarray(field, np.float64)
        h = C.c_void_p()
        L.check(L.lib().sce_tiled_build(n_cell_atoms, len(body), L.ptr(body), L.ptr(coef),
                                        L.ptr(atoms), L.ptr(shifts), L.ptr(ls), L.ptr(folded),
                                        L.ptr(d), L.ptr(mm), L.ptr(fl), C.byref(h)))
        self._h = h
        self.desc = L.TablesDesc()
        L.check(L.lib().sce_tiled_desc(h, C.byref(self.desc)))
        D = self.desc
        self.n_cell_atoms = D.n_cell_atoms
        self.dims = tuple(D.dims)
        self.n_sites = D.n_sites
        self.lmax = D.lmax
        self.nlm = D.nlm
        self.n_active = D.n_active
        self.n_colors = D.n_colors
        self.n_terms = D.n_terms
        self.magmoms = None if mm is None else mm.copy()
        self.field = np.zeros(3) if fl is None else fl.copy()
        sizes = {k: getattr(D, k) for k in ("n_inst", "n_inst_sites", "n_sites", "n_prog",
                                             "n_sent", "n_sfac", "n_terms", "n_eent", "n_efac",
                                             "n_colors", "n_active")}
        for name, expr, dt in L.ARRAY_FIELDS:
            n = int(eval(expr, {}, sizes))
            p = getattr(D, name)
            if n == 0 or not p:
                arr = np.zeros(0, dt)
            else:
                addr = C.cast(p, C.c_void_p).value
                buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(addr)
                arr = np.frombuffer(buf, dtype=dt)
            setattr(self, name, arr)
        self.site_has_l1 = self.site_has_l1.astype(bool)
        self.site_active = self.site_active.astype(bool)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.sce_tiled_free(h)
            self._h = None

    def zrows(self, config: np.ndarray) -> np.ndarray:
        """``_zrows`` (energy.jl:41-50): n_sites x nlm (row s = Z_lm(e_s))."""
        cfg = np.ascontiguousarray(config, np.float64).reshape(self.n_sites, 3)
        z = np.empty((self.n_sites, self.nlm))
        L.check(L.lib().sce_host_zrows(self.lmax, self.n_sites, L.ptr(cfg), L.ptr(z)))
        return z

    def total_energy_rows(self, zrows: np.ndarray) -> float:
        """``_total_energy`` (energy.jl:55-75), serial in instance order."""
        z = np.ascontiguousarray(zrows, np.float64)
        e = C.c_double()
        L.check(L.lib().sce_host_total_energy(C.byref(self.desc), L.ptr(z), C.byref(e)))
        return e.value

    def total_energy(self, config: np.ndarray) -> float:
        return self.total_energy_rows(self.zrows(config))

    def __repr__(self):
        return (f"TiledHamiltonian({self.n_cell_atoms} atoms x {self.dims} = {self.n_sites} sites,"
                f" lmax={self.lmax}, {self.n_terms} terms, {self.n_colors} colours)")


class ChainState:
    """Host chain state (src/state.jl:17-46): config (n x 3), zrows (n x nlm),
    the incremental energy, the proposal step and the acceptance counters."""

    def __init__(self, H: TiledHamiltonian, config: np.ndarray, step: float = 0.6):
        if not step > 0:
            raise ValueError(f"step must be > 0; got {step}")
        cfg = np.array(config, dtype=np.float64, order="C").reshape(H.n_sites, 3)
        self.H = H
        self.config = cfg
        self.zrows = H.zrows(cfg)
        self.energy = H.total_energy_rows(self.zrows)
        self.step = float(step)
        self.acc_metro = self.att_metro = self.acc_or = self.att_or = 0
        self.max_drift = 0.0


def resolve_kt(temperature=None, kT=None) -> np.ndarray:
    """Exactly one of ``temperature`` (K) / ``kT`` (model units) — src/units.jl:40-61."""
    if (temperature is None) == (kT is None):
        raise ValueError("provide exactly one of `temperature` (kelvin) or `kT`")
    if kT is not None:
        vals = np.atleast_1d(np.asarray(kT, np.float64))
    else:
        ts = np.atleast_1d(np.asarray(temperature, np.float64))
        if not np.all(np.isfinite(ts) & (ts > 0)):
            raise ValueError("temperature must be finite and > 0 kelvin")
        vals = KB_EV * ts
    if vals.size == 0:
        raise ValueError("the temperature/kT collection is empty")
    if not np.all(np.isfinite(vals) & (vals > 0)):
        raise ValueError("kT must be finite and > 0")
    return vals
