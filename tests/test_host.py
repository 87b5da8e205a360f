"""Host-side product checks that need no GPU: the C ABI loads and exports every
symbol include/sce_hip.h declares, the native TiledHamiltonian builder and host
rows/energies are bit-identical to the oracle, the device layout derives and
verifies for every fixture, error paths raise, and the binning machinery meets
the reference's known answers (test/unit/test_binning.jl)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd import _lib as L
from sce_amd import models as M
from fixtures import build_pair, fixture_cases, rand_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "sce_hip.h")).read()
    declared = set(re.findall(r"\b(sce_[a-z0-9_]+)\s*\(", hdr))
    lib = C.CDLL(L.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(L.EXPORTS) | {"sce_tables_desc"}
    assert lib.sce_version() == 10000


def test_no_device_fails_loudly_without_fallback():
    if L.device_count() > 0:
        pytest.skip("a HIP device is present")
    na, terms, dims = 1, M.chain_terms(0.05), (4, 1, 1)
    H = S.TiledHamiltonian(na, terms, dims)
    with pytest.raises(L.SceError, match="SCE_ENODEV"):
        S.GPUTiledHamiltonian(H)


WORKLOADS = ["sc8", "bcc16", "fcc12"]


@pytest.mark.parametrize("case", fixture_cases() + [(w, *M.config_model(w), None, None) for w in WORKLOADS],
                         ids=lambda c: c[0])
def test_builder_bitwise_equal_to_oracle(case):
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    for f, _ in O.FIELDS:
        a = getattr(H, f)
        a = a.astype(np.uint8) if a.dtype == bool else a
        assert np.array_equal(a, getattr(T, f)), f
    assert (H.n_sites, H.lmax, H.nlm, H.n_active, H.n_colors) == \
        (T.n_sites, T.lmax, T.nlm, T.n_active, T.n_colors)
    cfg = rand_config(H.n_sites, 3)
    zh, zo = H.zrows(cfg), T.zrows(cfg)
    assert np.array_equal(zh.view(np.uint64), zo.view(np.uint64))
    assert H.total_energy_rows(zh) == T.total_energy(zo)


@pytest.mark.parametrize("case", fixture_cases() + [(w, *M.config_model(w), None, None) for w in WORKLOADS + ["bcc32"]],
                         ids=lambda c: c[0])
def test_device_layout_derives_and_verifies(case):
    name, na, terms, dims, mm, fl = case
    H = S.TiledHamiltonian(na, terms, dims, magmoms=mm, field=fl)
    o = np.zeros(5, np.int64)
    L.check(L.lib().sce_layout_check(C.byref(H.desc), L.ptr(o)))
    assert o[0] == H.n_cell_atoms
    # the flattened walk of every active site equals its sublattice's program length
    ent = np.diff(H.sprog_ptr)[H.site_prog - 1]
    per_site = np.add.reduceat(ent, np.minimum(H.site_ptr[:-1] - 1, len(ent) - 1)) if len(ent) else np.zeros(H.n_sites)
    act = np.nonzero(H.site_active)[0]
    if len(act):
        assert per_site[act].max() == o[2]


def test_layout_rejects_non_tiling_tables():
    H = S.TiledHamiltonian(1, M.threebody_terms(0.05), (4, 2, 2))
    d = L.TablesDesc.from_buffer_copy(H.desc)
    col = np.array(H.site_col, copy=True)
    j = int(np.nonzero(col < 0)[0][5])
    col[j] = -((-col[j]) % H.n_sites + 1)  # point one triplet at a wrong neighbour
    d.site_col = col.ctypes.data_as(L.i32p)
    o = np.zeros(5, np.int64)
    rc = L.lib().sce_layout_check(C.byref(d), L.ptr(o))
    assert rc == -6 and "neighbour" in L.last_error()


def test_constructor_guards():
    with pytest.raises(ValueError, match="folds two member sites"):
        S.TiledHamiltonian(1, M.chain_terms(0.05), (1, 1, 1))
    with pytest.raises(ValueError, match="empty"):
        S.TiledHamiltonian(1, [], (2, 2, 2))
    with pytest.raises(ValueError, match="without magmoms"):
        S.TiledHamiltonian(1, M.chain_terms(0.05), (4, 1, 1), field=(0, 0, 1.0))
    bad = M.MultipoleTerm(0.1, [1, 1], [(1, 0, 0), (0, 0, 0)], [1, 1], np.eye(3))
    with pytest.raises(ValueError, match="anchoring"):
        S.TiledHamiltonian(1, [bad], (4, 1, 1))
    big = M.MultipoleTerm(0.1, [1], [(0, 0, 0)], [7], np.ones(15))
    with pytest.raises(ValueError, match="lmax"):
        S.TiledHamiltonian(1, [big], (1, 1, 1))
    # a coef == 0 term is dropped (hamiltonian.jl:280)
    zero = M.MultipoleTerm(0.0, [1, 1], [(0, 0, 0), (1, 0, 0)], [1, 1], np.eye(3))
    H = S.TiledHamiltonian(1, M.chain_terms(0.05) + [zero], (4, 1, 1))
    assert H.n_terms == 2


def test_body5_rejected_by_device_layout():
    f = np.zeros((3,) * 5); f[(1,) * 5] = 1.0
    t = M.MultipoleTerm(0.1, [1] * 5, [(k, 0, 0) for k in range(5)], [1] * 5, f)
    H = S.TiledHamiltonian(1, [t], (6, 1, 1))
    o = np.zeros(5, np.int64)
    rc = L.lib().sce_layout_check(C.byref(H.desc), L.ptr(o))
    assert rc == -1 and "body" in L.last_error()


def test_resolve_kt():
    assert S.resolve_kt(kT=0.01)[0] == 0.01
    assert S.resolve_kt(temperature=100.0)[0] == pytest.approx(100 * 8.617333262e-5, rel=1e-9)
    with pytest.raises(ValueError):
        S.resolve_kt()
    with pytest.raises(ValueError):
        S.resolve_kt(temperature=1.0, kT=1.0)
    with pytest.raises(ValueError):
        S.resolve_kt(kT=-1.0)


# ---- binning (binning.jl) ----------------------------------------------------------

def test_level1_error_is_bessel_sem():
    b = S.LogBinner(1)
    for x in (1.0, 2.0, 3.0, 4.0):
        b.push(x)
    assert b.count[0] == 4
    assert b.level_error(0, 0) == pytest.approx(math.sqrt(5 / 12), rel=1e-14)
    b1 = S.LogBinner(1).push(7.0)
    assert math.isnan(b1.level_error(0, 0))


def test_logbinner_iid_and_ar1():
    rng = np.random.default_rng(42)
    n = 2 ** 14
    b = S.LogBinner(1).push_series(rng.standard_normal(n))
    assert abs(b.mean()[0]) < 5 / math.sqrt(n)
    assert b.std_error()[0] == pytest.approx(1 / math.sqrt(n), rel=0.2)
    assert abs(b.tau_int()[0]) < 0.2
    rho, n = 0.9, 2 ** 16
    x = np.empty(n)
    v = 0.0
    e = np.random.default_rng(1).standard_normal(n)
    for i in range(n):
        v = rho * v + math.sqrt(1 - rho * rho) * e[i]
        x[i] = v
    b = S.LogBinner(1).push_series(x)
    assert b.std_error()[0] == pytest.approx(math.sqrt((1 + rho) / (1 - rho)) / math.sqrt(n), rel=0.3)
    assert b.tau_int()[0] == pytest.approx(rho / (1 - rho), rel=0.5)
    assert b.tau_int()[0] > 5


def test_effective_samples_needs_a_plateau():
    """Too few samples for a binned level with 32 bins: no estimate (NaN), not
    n.  With one: n / (1 + 2 tau) near the AR(1) closed form tau = rho / (1 - rho)."""
    rng = np.random.default_rng(9)
    assert math.isnan(S.effective_samples(rng.standard_normal(20)))
    assert math.isnan(S.effective_samples(rng.standard_normal(63)))
    assert math.isfinite(S.effective_samples(rng.standard_normal(64)))
    rho, n = 0.8, 2 ** 16
    e = np.random.default_rng(2).standard_normal(n)
    x = np.empty(n)
    v = 0.0
    for i in range(n):
        v = rho * v + e[i]
        x[i] = v
    assert S.effective_samples(x) == pytest.approx(n / (1 + 2 * rho / (1 - rho)), rel=0.35)


def test_windowed_tau_on_ar1_and_iid():
    """The windowed estimate against the AR(1) closed form tau = rho / (1 - rho)
    (log-binning convention) for short and long tau, and 0 for iid data."""
    for rho in (0.5, 0.8, 0.9):
        n = 2 ** 17
        e = np.random.default_rng(int(rho * 10)).standard_normal(n)
        x = np.empty(n)
        v = 0.0
        for i in range(n):
            v = rho * v + e[i]
            x[i] = v
        assert S.tau_int_windowed(x) == pytest.approx(rho / (1 - rho), rel=0.1, abs=0.05)
    assert abs(S.tau_int_windowed(np.random.default_rng(3).standard_normal(2 ** 16))) < 0.05


def test_binstore_and_jackknife():
    s = S.BinStore(1, 3, 4)
    for v in range(1, 15):
        s.push(float(v))
    assert np.array_equal(s.bin_means()[:, 0], [2.0, 5.0, 8.0, 11.0])
    cols = [np.array([1.0, 2.0, 3.0, 4.0])]
    est, err = S.jackknife(lambda m: m, cols)
    assert est == pytest.approx(2.5) and err == pytest.approx(math.sqrt(5 / 12), rel=1e-12)
    # nonlinear: jackknife of m^2 removes the O(1/n) bias of the plug-in estimate
    est2, _ = S.jackknife(lambda m: m * m, cols)
    assert est2 == pytest.approx(2.5 ** 2 - 5 / 12, rel=1e-12)


def test_free_spin_closed_forms_from_iid_measurements():
    # free spins: chi = (1 - 8/(3 pi))/kT, U = 5/3 (test_observables.jl:160-172)
    rng = np.random.default_rng(5)
    N, kT, nmeas = 64, 0.7, 20000
    v = rng.standard_normal((nmeas, N, 3))
    v /= np.linalg.norm(v, axis=2, keepdims=True)
    m = v.mean(axis=1)
    obs = np.column_stack([np.zeros(nmeas), m])
    st = S.standard_stats(obs, kT, N, nbins=16)
    chi, U = st["susceptibility"], st["binder"]
    assert abs(chi.mean - (1 - 8 / (3 * math.pi)) / kT) < 5 * chi.err
    assert abs(U.mean - 5 / 3) < 5 * U.err


def test_model_fingerprint_detects_changes():
    from sce_amd import checkpoint as CK
    na, terms = M.lattice_model("bcc", pair_shells=2, pair_lmax=1, seed=3)
    H1 = S.TiledHamiltonian(na, terms, (3, 3, 3))
    H2 = S.TiledHamiltonian(na, terms, (3, 3, 3))
    assert CK.model_fingerprint(H1) == CK.model_fingerprint(H2)
    flipped = list(terms)
    t0 = flipped[0]
    flipped[0] = M.MultipoleTerm(-t0.coef, t0.atoms, t0.shifts, t0.ls, t0.folded)
    assert CK.model_fingerprint(S.TiledHamiltonian(na, flipped, (3, 3, 3))) != CK.model_fingerprint(H1)
    assert CK.model_fingerprint(S.TiledHamiltonian(na, terms, (3, 3, 4))) != CK.model_fingerprint(H1)
    Hf = S.TiledHamiltonian(na, terms, (3, 3, 3), magmoms=[2.0, 2.0], field=(0, 0, 1.0))
    Hg = S.TiledHamiltonian(na, terms, (3, 3, 3), magmoms=[2.0, 2.0], field=(0, 0, -1.0))
    assert CK.model_fingerprint(Hf) != CK.model_fingerprint(Hg)
