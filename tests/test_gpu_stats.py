"""Statistical gates of the device chains (the reference's test_gpu.jl /
test_observables.jl / test_metropolis.jl / test_gpu_pt.jl analogues).

Closed forms: dimer <e1.e2> = L(beta|J|) (test_gpu.jl:298-313), single-site
Langevin <e_z> = -L(beta h) (test_metropolis.jl:31-49), free-spin
chi = (1 - 8/(3 pi))/kT and U = 5/3 (test_observables.jl:160-172).  The drift
gate (test_gpu.jl:290-296) at the full 32^3 size.  The device chain against
the reference CPU path (run_mc's sweeps) on the same model within 2 sigma of
log-binned errors for <E>, C, <|m|>, chi and U.  Replica exchange: rate limits
and the exchange-free composition gate (test_gpu_pt.jl:44-107).  Delayed
acceptance against the plain sweep within 2 sigma.
"""
import math

import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd import models as M
from sce_amd import surrogate as SG
from sce_amd.batch import derive_seeds
from fixtures import bcc_tet_small

pytestmark = pytest.mark.gpu

N1 = math.sqrt(3 / (4 * math.pi))


def _langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


def _chain(H, seed, step=0.6, cfg_seed=0):
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, S.random_config(H.n_sites, np.random.default_rng(cfg_seed)), step)
    return gH, S.GPUChainState(gH, st, seed=seed)


def test_drift_gate_full_size():
    na, terms, dims = M.config_model("bcc32d")
    H = S.TiledHamiltonian(na, terms, dims)
    gH, gst = _chain(H, 11)
    S.gpu_run_sweeps(gst, gH, None, 1 / 0.05, 200, renorm_interval=0)
    cfg, _, E_inc = gst.download()
    T = O.Tiled(na, terms, dims)
    E = T.total_energy(T.zrows(cfg))
    assert abs(E_inc - E) <= 1e-8 * max(1.0, abs(E))


def test_dimer_statistics():
    J = -0.02
    H = S.TiledHamiltonian(4, M.dimer_terms(J), (1, 1, 1))
    beta = 1.5 / abs(J)
    gH, gst = _chain(H, 2026, step=1.2)
    S.gpu_run_sweeps(gst, gH, None, beta, 500, renorm_interval=0)
    obs = S.gpu_run_sweeps(gst, gH, None, beta, 20_000, renorm_interval=0, measure_interval=1)
    x = obs[:, 0] / J                        # E = J e1.e2
    assert abs(x.mean() - _langevin(1.5)) < 0.03


@pytest.mark.parametrize("bh", [1.0, 2.0])
def test_single_site_langevin(bh):
    c0 = 0.3
    h = c0 * N1
    H = S.TiledHamiltonian(1, M.field_term(c0), (1, 1, 1))
    gH, gst = _chain(H, 200 + int(bh), step=1.5)
    beta = bh / h
    S.gpu_run_sweeps(gst, gH, None, beta, 500, renorm_interval=0)
    obs = S.gpu_run_sweeps(gst, gH, None, beta, 40_000, renorm_interval=0, measure_interval=2)
    assert abs(obs[:, 3].mean() + _langevin(bh)) < 0.03


@pytest.mark.parametrize("seed", [12, 13])
def test_free_spin_susceptibility_and_binder(seed):
    J = -0.02
    H = S.TiledHamiltonian(4, M.dimer_terms(J), (4, 4, 4))   # 128 active sites
    kT = 200 * abs(J)
    gH, gst = _chain(H, seed, step=2.5)
    S.gpu_run_sweeps(gst, gH, None, 1 / kT, 2000, renorm_interval=1000)
    obs = S.gpu_run_sweeps(gst, gH, None, 1 / kT, 20_000, renorm_interval=1000, measure_interval=5)
    st = S.standard_stats(obs, kT, H.n_active, nbins=8)
    chi, U = st["susceptibility"], st["binder"]
    assert abs(chi.mean - (1 - 8 / (3 * math.pi)) / kT) < 5 * chi.err
    assert abs(U.mean - 5 / 3) < 5 * U.err
    assert chi.err < 0.1 * chi.mean and U.err < 0.1 * U.mean


def _cpu_run(T, cfg, kT, therm, meas, seed, nthreads=8):
    ch = O.CpuChain(T, cfg, seed, 0.6)
    beta = 1 / kT
    for i in range(1, therm + 1):
        ch.metropolis_sweep(beta, nthreads)
        if i % 1000 == 0:
            ch.renormalize()
    rows = np.zeros((meas, 4))
    for i in range(1, meas + 1):
        ch.metropolis_sweep(beta, nthreads)
        if i % 1000 == 0:
            ch.renormalize()
        rows[i - 1, 0] = ch.energy
        rows[i - 1, 1:] = ch.mean_spin()
    return rows


@pytest.mark.parametrize("kT", [1e-2, 5e-3])
def test_device_chain_matches_cpu_run_mc_within_2_sigma(kT):
    """sc 8^3, body-1 l=2 + body-2 l=1 on two shells (random N(0, 1 meV) couplings),
    2k + 60k sweeps.  At kT = 1 meV this model is a glass (E ~ -5 meV per spin
    against ~12 meV bonds): chains from one start stay in different valleys for
    longer than the run, so the gate is taken where both chains equilibrate
    (asserted through tau_int)."""
    na, terms, dims = M.config_model("sc8")
    H = S.TiledHamiltonian(na, terms, dims)
    T = O.Tiled(na, terms, dims)
    cfg = S.random_config(H.n_sites, np.random.default_rng(5))
    n = 60_000
    cpu = S.standard_stats(_cpu_run(T, cfg, kT, 2000, n, seed=99), kT, H.n_active)
    gH = S.GPUTiledHamiltonian(H)
    gst = S.GPUChainState(gH, S.ChainState(H, cfg, 0.6), seed=98)
    S.gpu_run_sweeps(gst, gH, None, 1 / kT, 2000, renorm_interval=1000)
    dev = S.standard_stats(S.gpu_run_sweeps(gst, gH, None, 1 / kT, n, renorm_interval=1000,
                                            measure_interval=1), kT, H.n_active)
    assert cpu["energy"].tau_int < 100 and dev["energy"].tau_int < 100
    report = {}
    for k in ("energy", "specific_heat", "absm", "susceptibility", "binder"):
        a, b = cpu[k], dev[k]
        sig = math.hypot(a.err, b.err)
        report[k] = (a.mean, b.mean, sig)
        assert abs(a.mean - b.mean) < 2 * sig, (k, report)
    print(report)


def test_pt_exchange_rate_limits():
    H = S.TiledHamiltonian(4, M.dimer_terms(-0.02), (2, 2, 2))
    gH = S.GPUTiledHamiltonian(H)
    r_eq = S.gpu_run_pt(gH, kT=[0.049999, 0.05], sweeps_therm=100, sweeps_measure=400,
                        exchange_interval=5, seed=1)
    assert r_eq.swap_acceptance[0] > 0.95
    H2 = S.TiledHamiltonian(1, M.chain_terms(-0.05), (4, 4, 2))
    gH2 = S.GPUTiledHamiltonian(H2)
    r_far = S.gpu_run_pt(gH2, kT=[0.002, 0.2], sweeps_therm=200, sweeps_measure=1000,
                         exchange_interval=5, seed=2)
    assert r_far.swap_acceptance[0] < 0.2
    Es = [p.stats["energy"].mean for p in r_far.points]
    assert Es == sorted(Es)


def test_pt_exchange_free_ladder_is_independent_chains():
    na, terms, dims = bcc_tet_small(3, dims=(2, 2, 2))
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    kts = [0.5, 0.2, 0.1]
    nt, nm, seed = 40, 60, 77
    r = S.gpu_run_pt(gH, kT=kts, sweeps_therm=nt, sweeps_measure=nm, exchange_interval=nt + nm,
                     adapt_interval=nt + nm, renorm_interval=10 * (nt + nm), nbins=4, seed=seed)
    assert all(math.isnan(x) for x in r.swap_acceptance)
    init, _, dev = derive_seeds(seed, len(kts))
    for i, kt in enumerate(kts):
        st = S.ChainState(H, S.random_config(H.n_sites, np.random.default_rng(init[i])), 0.6)
        gst = S.GPUChainState(gH, st, seed=dev[i])
        for _ in range(nt):
            S.gpu_metropolis_sweep(gst, gH, 1 / kt)
        S.gpu_renormalize(gst)
        for _ in range(nm):
            S.gpu_metropolis_sweep(gst, gH, 1 / kt)
        assert np.array_equal(gst.download()[0], r.final_configs[i])


def test_delayed_acceptance_matches_plain_statistically():
    na, terms, dims = bcc_tet_small(4, dims=(6, 6, 6), isotropic=True)
    H = S.TiledHamiltonian(na, terms, dims)
    kT = 0.02
    gH, gst = _chain(H, 31, step=0.5)
    S.gpu_run_sweeps(gst, gH, None, 1 / kT, 2000, renorm_interval=1000)
    cfg0, _, _ = gst.download()
    sur = SG.Surrogate(gH, seed=5)
    sur.harvest(gst, 1 / kT, 300)
    sur.train_all(epochs=4, batch=1024, lr=1e-3)
    plain = S.GPUChainState(gH, S.ChainState(H, cfg0, 0.5), seed=41)
    da = S.GPUChainState(gH, S.ChainState(H, cfg0, 0.5), seed=42)
    n = 60_000
    sp = S.standard_stats(S.gpu_run_sweeps(plain, gH, None, 1 / kT, n, renorm_interval=1000,
                                           measure_interval=1), kT, H.n_active)
    sd = S.standard_stats(SG.gpu_da_run_sweeps(da, sur, None, 1 / kT, n, renorm_interval=1000,
                                               measure_interval=1), kT, H.n_active)
    # Delayed acceptance is exact whatever the surrogate, so the two chains
    # sample one distribution and each z below is a draw of a unit normal.
    # Five such gates at 2 sigma would fail together by chance about one run
    # in five; the gate is every observable within 3 sigma and at most one of
    # the five beyond 2 sigma (a chance failure rate of about 3.5%, while a
    # bias of 3 sigma in any observable fails).
    report, z = {}, []
    for k in ("energy", "specific_heat", "absm", "susceptibility", "binder"):
        a, b = sp[k], sd[k]
        sig = math.hypot(a.err, b.err)
        report[k] = (a.mean, b.mean, sig)
        z.append(abs(a.mean - b.mean) / sig)
    print(report, z)
    assert max(z) < 3 and sum(x >= 2 for x in z) <= 1, (z, report)
    assert da.info().acc_stage1 < 0.8 * da.info().att_metro
