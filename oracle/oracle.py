"""ctypes binding of the CPU oracle (``libsce_oracle.so``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` — as the checker / CPU baseline, never as
the thing measured or shipped.  The product (``scemontecarlo.jl_amd``) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libsce_oracle.so")

FIELDS = [
    ("inst_term", np.int32), ("inst_ptr", np.int32), ("inst_sites", np.int32),
    ("site_ptr", np.int32), ("site_inst", np.int32), ("site_slot", np.int8),
    ("site_has_l1", np.uint8), ("site_active", np.uint8), ("site_prog", np.int32),
    ("sprog_ptr", np.int32), ("sent_w", np.float64), ("sent_tgt", np.int32),
    ("sfac_ptr", np.int32), ("sfac_row", np.int32), ("sfac_slot", np.int8),
    ("site_col", np.int32), ("site_col2", np.int32), ("pent_row", np.int32),
    ("pent_row2", np.int32), ("eprog_ptr", np.int32), ("eent_w", np.float64),
    ("efac_ptr", np.int32), ("efac_row", np.int32), ("term_coef", np.float64),
    ("color_ptr", np.int32), ("color_sites", np.int32),
]

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        d, i32, u32, u64, i64 = C.c_double, C.c_int32, C.c_uint32, C.c_uint64, C.c_int64
        P = C.c_void_p
        sig = {
            "ora_philox4x32": (None, [P, P, P]),
            "ora_philox_block": (None, [u64, i32, i32, u32, u32, P]),
            "ora_uniform": (d, [u32, u32]),
            "ora_normal2": (None, [P, P]),
            "ora_exp": (d, [d]),
            "ora_log": (d, [d]),
            "ora_sincos": (None, [d, P, P]),
            "ora_zlm_row": (None, [P, C.c_int, P]),
            "ora_grad_zlm_row": (None, [P, C.c_int, P]),
            "ora_energy_gradient": (None, [P, P, P, C.c_int, P]),
            "ora_tiled_build": (C.c_int, [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P,
                                          P, C.c_char_p, C.c_size_t]),
            "ora_tiled_free": (None, [P]),
            "ora_tiled_scalars": (None, [P, P]),
            "ora_tiled_array": (P, [P, C.c_int, P]),
            "ora_zrows": (None, [P, P, P]),
            "ora_total_energy": (d, [P, P]),
            "ora_total_energy_tree": (d, [P, P]),
            "ora_site_coeffs": (None, [P, C.c_int, P, P]),
            "ora_delta_energy": (d, [C.c_int, P, P, P]),
            "ora_reduce_tree": (d, [i64, P]),
            "ora_keyed_proposal": (None, [u64, i32, i32, P, d, P, P]),
            "ora_walk_delta": (d, [P, C.c_int, P, P, C.c_int]),
            "ora_walk_coeffs": (None, [P, C.c_int, P, C.c_int, P]),
            "ora_keyed_sweep": (i64, [P, P, P, P, d, d, u64, i32, C.c_int, P, P]),
            "ora_keyed_or_sweep": (i64, [P, P, P, P, d, u64, i32, i32, C.c_int, P, P, P]),
            "ora_renormalize": (d, [P, P, P, P]),
            "ora_cpu_chain_create": (P, [P, P, u64, d]),
            "ora_cpu_chain_free": (None, [P]),
            "ora_cpu_metropolis_sweep": (i64, [P, d, C.c_int]),
            "ora_cpu_or_sweep": (i64, [P, d, C.c_int]),
            "ora_cpu_renormalize": (d, [P]),
            "ora_cpu_state": (None, [P, P, P, P, P]),
            "ora_cpu_energy": (d, [P]),
            "ora_cpu_mean_spin": (None, [P, P]),
            "ora_cpu_step": (d, [P]),
            "ora_cpu_set_step": (None, [P, d]),
            "ora_cpu_reset_counters": (None, [P]),
            "ora_cpu_counters": (None, [P, P]),
            "ora_sur_features": (C.c_int, [P, P, C.c_int, P, P]),
            "ora_sur_delta": (d, [C.c_int, P, d, P, P]),
            "ora_sur_delta_q": (d, [C.c_int, P, d, P, P]),
            "ora_da_sweep_q": (i64, [P, P, P, P, d, d, u64, i32, P, P, P, P, P]),
            "ora_test_tanhf": (C.c_float, [C.c_float]),
            "ora_da_sweep": (i64, [P, P, P, P, d, d, u64, i32, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---- scalar helpers ---------------------------------------------------------

def philox4x32(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().ora_philox4x32(_p(c), _p(k), _p(o))
    return tuple(int(x) for x in o)


def philox_block(seed, site, sweep, slot, tag=0):
    o = np.zeros(4, np.uint32)
    lib().ora_philox_block(seed, site, sweep, slot, tag, _p(o))
    return o


def uniform(hi, lo):
    return lib().ora_uniform(hi, lo)


def normal2(blk):
    b = np.asarray(blk, dtype=np.uint32)
    o = np.zeros(2)
    lib().ora_normal2(_p(b), _p(o))
    return o


def exp(x):
    return lib().ora_exp(float(x))


def log(x):
    return lib().ora_log(float(x))


def tanhf(x):
    """The surrogate's FP32 tanh of x rounded to single precision."""
    return lib().ora_test_tanhf(float(np.float32(x)))


def sincos(x):
    s, c = C.c_double(), C.c_double()
    lib().ora_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def zlm_row(u, lmax):
    u = np.ascontiguousarray(u, dtype=np.float64)
    z = np.zeros((lmax + 1) ** 2)
    lib().ora_zlm_row(_p(u), lmax, _p(z))
    return z


def grad_zlm_row(u, lmax):
    """Tangent-projected gradients of the Z row at u: nlm x 3."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.zeros((lmax + 1) ** 2 * 3)
    lib().ora_grad_zlm_row(_p(u), lmax, _p(g))
    return g.reshape(-1, 3)


def reduce_tree(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().ora_reduce_tree(len(x), _p(x))


def keyed_proposal(seed, site, sweep, e, step):
    e = np.ascontiguousarray(e, dtype=np.float64)
    e2 = np.zeros(3)
    u = C.c_double()
    lib().ora_keyed_proposal(seed, site, sweep, _p(e), step, _p(e2), C.byref(u))
    return e2, u.value


def sur_delta(F, params, scale, phiA, phiB):
    p = np.ascontiguousarray(params, np.float32)
    a = np.ascontiguousarray(phiA, np.float32)
    b = np.ascontiguousarray(phiB, np.float32)
    return lib().ora_sur_delta(F, _p(p), scale, _p(a), _p(b))


def sur_delta_q(F, qparams, scale, phiA, phiB):
    """The int8 surrogate difference (ora_sur_delta_q) of one site."""
    p = np.ascontiguousarray(qparams, np.float32)
    a = np.ascontiguousarray(phiA, np.float32)
    b = np.ascontiguousarray(phiB, np.float32)
    return lib().ora_sur_delta_q(F, _p(p), scale, _p(a), _p(b))


def delta_energy(c, zold, znew):
    c = np.ascontiguousarray(c, np.float64)
    zo = np.ascontiguousarray(zold, np.float64)
    zn = np.ascontiguousarray(znew, np.float64)
    return lib().ora_delta_energy(len(c), _p(c), _p(zo), _p(zn))


# ---- tiled Hamiltonian ------------------------------------------------------

class Tiled:
    """Oracle restatement of ``TiledHamiltonian`` (src/hamiltonian.jl:267-393)."""

    def __init__(self, n_cell_atoms, terms, dims=(1, 1, 1), magmoms=None, field=None):
        from_terms = _flatten_terms(terms)
        body, coef, atoms, shifts, ls, folded = from_terms
        d = np.asarray(dims, np.int32)
        mm = None if magmoms is None else np.ascontiguousarray(magmoms, np.float64)
        fl = None if field is None else np.ascontiguousarray(field, np.float64)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().ora_tiled_build(n_cell_atoms, len(body), _p(body), _p(coef), _p(atoms),
                                   _p(shifts), _p(ls), _p(folded), _p(d), _p(mm), _p(fl),
                                   C.byref(h), err, 512)
        if rc != 0:
            raise ValueError(err.value.decode())
        self._h = h
        sc = np.zeros(12, np.int64)
        lib().ora_tiled_scalars(h, _p(sc))
        (self.n_cell_atoms, d1, d2, d3, self.n_sites, self.lmax, self.nlm, self.n_active,
         self.n_colors, self.n_terms, self.n_fitted_terms, self.n_inst) = (int(x) for x in sc)
        self.dims = (d1, d2, d3)
        for k, (name, dt) in enumerate(FIELDS):
            n = C.c_int64()
            ptr = lib().ora_tiled_array(h, k, C.byref(n))
            if n.value == 0 or not ptr:
                arr = np.zeros(0, dt)
            else:
                buf = (C.c_char * (n.value * np.dtype(dt).itemsize)).from_address(ptr)
                arr = np.frombuffer(buf, dtype=dt).copy()
            setattr(self, name, arr)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.ora_tiled_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def zrows(self, config):
        config = np.ascontiguousarray(config, np.float64).reshape(self.n_sites, 3)
        z = np.zeros((self.n_sites, self.nlm))
        lib().ora_zrows(self._h, _p(config), _p(z))
        return z

    def total_energy(self, zrows):
        z = np.ascontiguousarray(zrows, np.float64)
        return lib().ora_total_energy(self._h, _p(z))

    def total_energy_tree(self, zrows):
        z = np.ascontiguousarray(zrows, np.float64)
        return lib().ora_total_energy_tree(self._h, _p(z))

    def energy_gradient(self, config, zrows=None, lanes=64):
        """All-site tangent-projected gradient dE/de_s (n_sites x 3), the
        reference's lane definition with `lanes` strided shares."""
        cfg = np.ascontiguousarray(config, np.float64)
        z = np.ascontiguousarray(self.zrows(cfg) if zrows is None else zrows, np.float64)
        G = np.zeros((self.n_sites, 3))
        lib().ora_energy_gradient(self._h, _p(cfg), _p(z), int(lanes), _p(G))
        return G

    def site_coeffs(self, s, zrows):
        z = np.ascontiguousarray(zrows, np.float64)
        c = np.zeros(self.nlm)
        lib().ora_site_coeffs(self._h, s, _p(z), _p(c))
        return c

    def walk_delta(self, s, zrows, znew, lanes):
        z = np.ascontiguousarray(zrows, np.float64)
        zn = np.ascontiguousarray(znew, np.float64)
        return lib().ora_walk_delta(self._h, s, _p(z), _p(zn), lanes)

    def walk_coeffs(self, s, zrows, lanes):
        z = np.ascontiguousarray(zrows, np.float64)
        c = np.zeros(self.nlm)
        lib().ora_walk_coeffs(self._h, s, _p(z), lanes, _p(c))
        return c

    def keyed_sweep(self, config, zrows, energy, beta, step, seed, sweep, lanes):
        """In-place device-contract Metropolis sweep; returns (nacc, energy, dE_q, acc_q)."""
        e = C.c_double(energy)
        dE = np.zeros(max(self.n_active, 1))
        acc = np.zeros(max(self.n_active, 1), np.int32)
        n = lib().ora_keyed_sweep(self._h, _p(config), _p(zrows), C.byref(e), beta, step,
                                  seed, sweep, lanes, _p(dE), _p(acc))
        return int(n), e.value, dE[: self.n_active], acc[: self.n_active]

    def keyed_or_sweep(self, config, zrows, energy, beta, seed, sweep, or_pass, lanes):
        e = C.c_double(energy)
        dE = np.zeros(max(self.n_active, 1))
        acc = np.zeros(max(self.n_active, 1), np.int32)
        natt = C.c_int64()
        n = lib().ora_keyed_or_sweep(self._h, _p(config), _p(zrows), C.byref(e), beta, seed,
                                     sweep, or_pass, lanes, _p(dE), _p(acc), C.byref(natt))
        return int(n), int(natt.value), e.value

    def sur_features(self, config, s, e):
        out = np.zeros(512, np.float32)
        cfg = np.ascontiguousarray(config, np.float64)
        ee = np.ascontiguousarray(e, np.float64)
        F = lib().ora_sur_features(self._h, _p(cfg), s, _p(ee), _p(out))
        return out[:F]

    def da_sweep(self, config, zrows, energy, beta, step, seed, sweep, params, scales, int8=False):
        """params: list (per sublattice) of float32 parameter vectors (with int8:
        the quantised network blobs of ``Surrogate.get_qparams``)."""
        e = C.c_double(energy)
        keep = [np.ascontiguousarray(p, np.float32) for p in params]
        ptrs = (C.c_void_p * len(keep))(*[k.ctypes.data for k in keep])
        sc = np.ascontiguousarray(scales, np.float64)
        dE = np.zeros(max(self.n_active, 1))
        acc = np.zeros(max(self.n_active, 1), np.int32)
        n1 = C.c_int64()
        f = lib().ora_da_sweep_q if int8 else lib().ora_da_sweep
        n = f(self._h, _p(config), _p(zrows), C.byref(e), beta, step, seed, sweep,
              ptrs, _p(sc), _p(dE), _p(acc), C.byref(n1))
        return int(n), int(n1.value), e.value

    def renormalize(self, config, zrows, energy):
        e = C.c_double(energy)
        drift = lib().ora_renormalize(self._h, _p(config), _p(zrows), C.byref(e))
        return drift, e.value


class CpuChain:
    """The reference CPU chain (updates.jl sweeps on per-site streams)."""

    def __init__(self, tiled: Tiled, config, seed, step):
        self.t = tiled
        cfg = np.ascontiguousarray(config, np.float64)
        self._h = lib().ora_cpu_chain_create(tiled.handle, _p(cfg), seed, step)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.ora_cpu_chain_free(h)
            self._h = None

    def metropolis_sweep(self, beta, nthreads=1):
        return lib().ora_cpu_metropolis_sweep(self._h, beta, nthreads)

    def or_sweep(self, beta, nthreads=1):
        return lib().ora_cpu_or_sweep(self._h, beta, nthreads)

    def renormalize(self):
        return lib().ora_cpu_renormalize(self._h)

    @property
    def energy(self):
        return lib().ora_cpu_energy(self._h)

    @property
    def step(self):
        return lib().ora_cpu_step(self._h)

    @step.setter
    def step(self, v):
        lib().ora_cpu_set_step(self._h, v)

    def mean_spin(self):
        m = np.zeros(3)
        lib().ora_cpu_mean_spin(self._h, _p(m))
        return m

    def counters(self):
        c = np.zeros(4, np.int64)
        lib().ora_cpu_counters(self._h, _p(c))
        return c

    def reset_counters(self):
        lib().ora_cpu_reset_counters(self._h)

    def state(self):
        cfg = np.zeros((self.t.n_sites, 3))
        z = np.zeros((self.t.n_sites, self.t.nlm))
        e = C.c_double()
        lib().ora_cpu_state(self._h, _p(cfg), _p(z), C.byref(e), None)
        return cfg, z, e.value


def _flatten_terms(terms):
    """Terms are objects/tuples with (coef, atoms, shifts, ls, folded)."""
    body, coef, atoms, shifts, ls, folded = [], [], [], [], [], []
    for t in terms:
        c, a, sh, l, f = (t.coef, t.atoms, t.shifts, t.ls, t.folded) if hasattr(t, "coef") else t
        body.append(len(a))
        coef.append(float(c))
        atoms.extend(int(x) for x in a)
        for v in sh:
            shifts.extend(int(x) for x in v)
        ls.extend(int(x) for x in l)
        folded.extend(np.asarray(f, np.float64).reshape(-1, order="F"))
    return (np.asarray(body, np.int32), np.asarray(coef, np.float64),
            np.asarray(atoms, np.int32), np.asarray(shifts, np.int32),
            np.asarray(ls, np.int32), np.asarray(folded, np.float64))
