"""Parity at the benchmark's own size: bcc 32^3 (65 536 spins, the workload of
bench.py) and its dense-tensor twin, device vs the oracle's keyed restatement,
bit for bit — whole graph-replayed sweeps with measurements and the chained
proposals, a delayed-acceptance sweep, and the all-site gradient."""
import numpy as np
import pytest

import sce_amd as S
from sce_amd import models as M
from sce_amd import surrogate as SG
from fixtures import build_pair, rand_config

pytestmark = pytest.mark.gpu


def _model(name):
    na, terms, dims = M.config_model(name, seed=0)
    H, T = build_pair(na, terms, dims)
    return na, H, T, S.GPUTiledHamiltonian(H)


@pytest.mark.parametrize("name,nsw", [("bcc32", 4), ("bcc32d", 2)])
def test_benchmark_size_sweeps_bitwise(name, nsw):
    _, H, T, gH = _model(name)
    st = S.ChainState(H, rand_config(H.n_sites, 2026), 1.3)
    seed, beta = 20261203, 1 / 0.04
    g = S.GPUChainState(gH, st, seed=seed)
    obs = S.gpu_run_sweeps(g, gH, None, beta, nsw, renorm_interval=0, measure_interval=1)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    nacc, Es = 0, []
    for sw in range(1, nsw + 1):
        a, E, _, _ = T.keyed_sweep(c2, z2, E, beta, 1.3, seed, sw, 64)
        nacc += a
        Es.append(E)
    cfg_d, z_d, E_d = g.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E and np.array_equal(obs[:, 0], np.array(Es))
    assert g.info().acc_metro == nacc


def test_benchmark_size_gradient_bitwise():
    _, H, T, gH = _model("bcc32")
    cfg = rand_config(H.n_sites, 7)
    G = S.gpu_energy_gradient(gH, cfg)
    assert np.array_equal(G, T.energy_gradient(cfg, lanes=S.GRADIENT_LANES))


def test_benchmark_size_delayed_acceptance_sweep_bitwise():
    na, H, T, gH = _model("bcc32")
    sur = SG.Surrogate(gH, seed=3)
    params, scales = [], []
    rng = np.random.default_rng(5)
    for a in range(na):
        F = sur.info(a)["F"]
        p = (rng.standard_normal(SG.param_layout(F)["n"]) * 0.1).astype(np.float32)
        sur.set_params(a, p, 0.05)
        params.append(p)
        scales.append(0.05)
    st = S.ChainState(H, rand_config(H.n_sites, 11), 1.3)
    seed, beta = 0xda7a, 1 / 0.04
    g = S.GPUChainState(gH, st, seed=seed)
    SG.gpu_da_run_sweeps(g, sur, None, beta, 1, renorm_interval=0)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    nacc, n1, E = T.da_sweep(c2, z2, E, beta, 1.3, seed, 1, params, np.array(scales))
    cfg_d, z_d, E_d = g.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E
    inf = g.info()
    assert (inf.acc_metro, inf.acc_stage1) == (nacc, n1)


def test_benchmark_size_persistent_sweeps_equal_graph_path():
    """The dependency-driven persistent sweep at bcc 32^3: 40 sweeps measured
    every sweep, with a renormalisation at 25, bit-identical to per-colour
    graph launches (rows, spins, energy, counters, measurement rows)."""
    _, H, T, gH = _model("bcc32")
    st = S.ChainState(H, rand_config(H.n_sites, 99), 1.3)
    a = S.GPUChainState(gH, st, seed=31)
    b = S.GPUChainState(gH, st, seed=31)
    a.set_persistent(True)
    oa = S.gpu_run_sweeps(a, gH, None, 1 / 0.04, 40, renorm_interval=25, measure_interval=1)
    ob = S.gpu_run_sweeps(b, gH, None, 1 / 0.04, 40, renorm_interval=25, measure_interval=1)
    assert np.array_equal(oa.view(np.uint64), ob.view(np.uint64))
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca.view(np.uint64), cb.view(np.uint64))
    assert np.array_equal(za.view(np.uint64), zb.view(np.uint64)) and Ea == Eb
    ia, ib = a.info(), b.info()
    assert (ia.acc_metro, ia.att_metro, ia.sweep_index) == (ib.acc_metro, ib.att_metro, ib.sweep_index)


def test_config2_compound_sweeps_bitwise():
    """The plain form at its own size: bcc 16^3 (8192 spins, body-2 l<=2 to the
    3rd shell + l=1 triangles), compound sweeps of one Metropolis sweep and one
    overrelaxation pass (run.jl:143-150), graph-replayed with the measurement and
    the next proposals fused into the pass's reduction, bit-identical to the
    oracle's composition."""
    _, H, T, gH = _model("bcc16")
    st = S.ChainState(H, rand_config(H.n_sites, 16), 0.9)
    seed, beta = 1616, 1 / 0.02
    g = S.GPUChainState(gH, st, seed=seed)
    obs = S.gpu_run_sweeps(g, gH, None, beta, 3, renorm_interval=0, or_per_metropolis=1, measure_interval=1)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    Es = []
    for sw in range(1, 4):
        _, E, _, _ = T.keyed_sweep(c2, z2, E, beta, 0.9, seed, sw, 64)
        _, _, E = T.keyed_or_sweep(c2, z2, E, beta, seed, sw, 0, 64)
        Es.append(E)
    cfg_d, z_d, E_d = g.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E and np.array_equal(obs[:, 0], np.array(Es))
