"""Device gates of the reference's GPU test list not covered elsewhere:
repeated-run identity and seed sensitivity (test_gpu.jl:253-276), inactive
sites bitwise frozen through renormalising runs (:278-288), overrelaxation mixed
with Metropolis against the dimer closed form (test_overrelaxation.jl:6-64),
device-PT marginals on a dimer ladder and payload consistency after exchanges
(test_gpu_pt.jl:13-204), and a checkpoint that does not perturb the run
(test_checkpoint.jl:192-330)."""
import math

import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd import models as M
from sce_amd import checkpoint as CK
from sce_amd.batch import GPUBatch
from fixtures import bcc_tet_small, rand_config

pytestmark = pytest.mark.gpu


def _langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


def test_repeated_run_identity_and_seed_sensitivity():
    na, terms, dims = bcc_tet_small(3, dims=(2, 2, 2))
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 1), 0.6)
    g1 = S.GPUChainState(gH, st, seed=77)
    g2 = S.GPUChainState(gH, st, seed=77)
    g3 = S.GPUChainState(gH, st, seed=78)
    for _ in range(10):
        for g in (g1, g2, g3):
            S.gpu_metropolis_sweep(g, gH, 20.0)
    c1, z1, E1 = g1.download()
    c2, z2, E2 = g2.download()
    c3, _, _ = g3.download()
    assert np.array_equal(c1, c2) and np.array_equal(z1, z2) and E1 == E2
    assert g1.info().acc_metro == g2.info().acc_metro
    assert not np.array_equal(c1, c3)


def test_inactive_sites_bitwise_frozen():
    H = S.TiledHamiltonian(4, M.dimer_terms(-0.02), (2, 1, 1))
    assert H.n_active < H.n_sites
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 4), 0.6)
    cfg0, zr0 = st.config.copy(), st.zrows.copy()
    g = S.GPUChainState(gH, st, seed=5)
    S.gpu_run_sweeps(g, gH, st, 1 / 0.05, 50, renorm_interval=20, or_per_metropolis=1)
    frozen = np.nonzero(H.site_active == 0)[0]
    assert np.array_equal(st.config[frozen], cfg0[frozen])
    assert np.array_equal(st.zrows[frozen], zr0[frozen])
    assert not np.array_equal(st.config[np.nonzero(H.site_active)[0]], cfg0[np.nonzero(H.site_active)[0]])


def test_overrelaxation_mix_keeps_the_dimer_law():
    J = -0.02
    H = S.TiledHamiltonian(4, M.dimer_terms(J), (1, 1, 1))
    gH = S.GPUTiledHamiltonian(H)
    g = S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, 8), 1.2), seed=31)
    beta = 1.5 / abs(J)
    S.gpu_run_sweeps(g, gH, None, beta, 500, renorm_interval=0, or_per_metropolis=2)
    obs = S.gpu_run_sweeps(g, gH, None, beta, 20_000, renorm_interval=0, or_per_metropolis=2,
                           measure_interval=1)
    inf = g.info()
    assert inf.acc_or == inf.att_or > 0          # pure l=1: every reflection accepted
    assert abs((obs[:, 0] / J).mean() - _langevin(1.5)) < 0.03


def test_device_pt_marginals_on_a_dimer_ladder():
    J = -0.02
    H = S.TiledHamiltonian(4, M.dimer_terms(J), (1, 1, 1))
    gH = S.GPUTiledHamiltonian(H)
    bjs = [2.0, 1.0, 0.5]
    kts = [abs(J) / b for b in bjs]
    r = S.gpu_run_pt(gH, kT=kts, exchange_interval=1, sweeps_therm=500, sweeps_measure=20_000,
                     step=1.2, seed=11)
    assert np.all(r.swap_acceptance > 0.05)
    for p, bj in zip(r.points, bjs):
        assert abs(p.stats["energy"].mean / J - _langevin(bj)) < 0.03, (bj, p.stats["energy"].mean / J)


def test_pt_payloads_stay_consistent_through_exchanges():
    na, terms, dims = bcc_tet_small(5, dims=(2, 2, 2))
    H = S.TiledHamiltonian(na, terms, dims)
    T = O.Tiled(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    R = 6
    kts = list(np.geomspace(0.02, 0.2, R))
    sts = [S.ChainState(H, rand_config(H.n_sites, 40 + k), 0.6) for k in range(R)]
    b = GPUBatch([gH] * R, sts, [100 + k for k in range(R)], kts, 9)
    b.run(200, 64, 0, 0, 1, 0, 200)
    att, acc = b.swap_stats()
    assert acc.sum() > 0
    configs = []
    for k in range(R):
        cfg, z, E = b.download(k)
        assert np.array_equal(z, T.zrows(cfg))               # rows travel with their spins
        Ef = T.total_energy(z)
        assert abs(E - Ef) <= 1e-9 * max(1.0, abs(Ef))      # and so does the energy
        configs.append(cfg)
    for i in range(R):
        for j in range(i + 1, R):
            assert not np.array_equal(configs[i], configs[j])


def test_checkpoint_does_not_perturb_the_run(tmp_path):
    na, terms, dims = bcc_tet_small(6, dims=(2, 2, 2))
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 9), 0.6)
    a = S.GPUChainState(gH, st, seed=3)
    b = S.GPUChainState(gH, st, seed=3)
    S.gpu_run_sweeps(a, gH, None, 20.0, 10, renorm_interval=4)
    CK.save_chain(str(tmp_path / "a.npz"), a)
    S.gpu_run_sweeps(a, gH, None, 20.0, 10, renorm_interval=4)
    S.gpu_run_sweeps(b, gH, None, 20.0, 10, renorm_interval=4)
    S.gpu_run_sweeps(b, gH, None, 20.0, 10, renorm_interval=4)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb


def test_buffers_regrow_across_calls():
    """Observation buffers and the harvested dataset grow across calls (old device
    buffers are released) and keep working: measurements, a second harvest,
    training on the merged rows, DA sweeps."""
    from sce_amd import surrogate as SG
    na, terms, dims = bcc_tet_small(7, dims=(2, 2, 2), isotropic=True)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    g = S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, 3), 0.6), seed=8)
    for n in (5, 20, 60):
        obs = S.gpu_run_sweeps(g, gH, None, 1 / 0.05, n, renorm_interval=0, measure_interval=1)
        assert obs.shape == (n, 4)
        assert obs[-1, 0] == g.download()[2]
    sur = SG.Surrogate(gH, seed=2)
    r1 = sur.harvest(g, 1 / 0.05, 30)
    r2 = sur.harvest(g, 1 / 0.05, 90)
    assert r2 == r1 + 90 * H.n_active
    losses = sur.train_all(epochs=2, batch=64, lr=1e-3)
    assert all(np.isfinite(l).all() for l in losses.values())
    for n in (4, 40):
        obs = SG.gpu_da_run_sweeps(g, sur, None, 1 / 0.05, n, renorm_interval=0, measure_interval=1)
        assert obs.shape == (n, 4)
    R = 3
    b = GPUBatch([gH] * R, [S.ChainState(H, rand_config(H.n_sites, 10 + k), 0.6) for k in range(R)],
                 [1, 2, 3], [0.05, 0.06, 0.07])
    for n in (3, 30):
        assert b.run(n, 64, 0, 1).shape == (R, n, 4)
