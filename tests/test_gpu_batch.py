"""Batched chains: every chain of a batch is bit-identical to the same chain run
alone through the oracle, for a replica ladder on one model and for an uneven
batch of different models and sizes; device replica exchange is bit-identical
to the oracle's restatement of _swap_accepts (src/pt.jl:92-96) with the
alternating-parity schedule of src/gpu/gpu_pt.jl:115-140."""
import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd import models as M
from sce_amd.batch import GPUBatch
from fixtures import bcc_tet_small, build_pair, rand_config

pytestmark = pytest.mark.gpu


def _oracle_exchange(payload, kts, parity, ex, exch_seed):
    R = len(payload)
    att = np.zeros(R - 1, int)
    acc = np.zeros(R - 1, int)
    for i in range(parity, R - 1, 2):
        b = O.philox_block(exch_seed, i + 1, ex, 4, 0)
        u = O.uniform(int(b[0]), int(b[1]))
        Ea, Eb = payload[i][2], payload[i + 1][2]
        logw = (1.0 / kts[i] - 1.0 / kts[i + 1]) * (Ea - Eb)
        att[i] += 1
        if u < O.exp(min(0.0, logw) if logw < 0.0 else 0.0):
            payload[i], payload[i + 1] = payload[i + 1], payload[i]
            acc[i] += 1
    return att, acc


@pytest.mark.parametrize("lanes", [32, 64])
def test_ladder_with_device_exchange_bitwise(lanes):
    na, terms, dims = bcc_tet_small(6)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    R = 5
    kts = np.geomspace(0.001, 0.02, R)
    seeds = [101 + 7 * r for r in range(R)]
    states = [S.ChainState(H, rand_config(H.n_sites, 40 + r), 0.6) for r in range(R)]
    exch_seed = 0xabcdef
    b = GPUBatch([gH] * R, states, seeds, kts, exch_seed)
    nsw = 12
    b.run(nsw, lanes, renorm_interval=5, measure_interval=0, exchange_interval=1, phase_done=0,
          phase_total=nsw)
    payload = [[st.config.copy(), st.zrows.copy(), st.energy] for st in states]
    att = np.zeros(R - 1, int)
    acc = np.zeros(R - 1, int)
    parity, ex = 0, 0
    for j in range(1, nsw + 1):
        for r in range(R):
            c, z, E = payload[r]
            _, E, _, _ = T.keyed_sweep(c, z, E, 1.0 / kts[r], 0.6, seeds[r], j, lanes)
            payload[r][2] = E
        if j % 5 == 0:
            for r in range(R):
                _, payload[r][2] = T.renormalize(payload[r][0], payload[r][1], payload[r][2])
        if j < nsw:
            ex += 1
            a1, a2 = _oracle_exchange(payload, kts, parity, ex, exch_seed)
            att += a1
            acc += a2
            parity = 1 - parity
    for r in range(R):
        cfg, z, E = b.download(r)
        assert np.array_equal(cfg, payload[r][0]), r
        assert np.array_equal(z, payload[r][1]), r
        assert E == payload[r][2], r
        assert b.info(r).sweep_index == nsw
    da, dc = b.swap_stats()
    assert np.array_equal(da, att) and np.array_equal(dc, acc)
    assert acc.sum() > 0  # the ladder actually exchanged


def test_uneven_batch_bitwise_and_equal_to_single_chains():
    cases = [M.lattice_model("bcc", pair_shells=3, pair_lmax=2, triangles=True, seed=3) + ((3, 3, 3),),
             M.lattice_model("fcc", pair_shells=2, pair_lmax=2, seed=5) + ((3, 3, 3),),
             M.lattice_model("sc", pair_shells=2, pair_lmax=1, onsite_l2=True, seed=4) + ((5, 4, 3),),
             M.lattice_model("bcc", pair_shells=2, pair_lmax=2, seed=9) + ((4, 4, 4),)]
    gHs, states, Ts, kts, seeds = [], [], [], [], []
    for k, (na, terms, dims) in enumerate(cases):
        H, T = build_pair(na, terms, dims)
        gHs.append(S.GPUTiledHamiltonian(H))
        states.append(S.ChainState(H, rand_config(H.n_sites, 60 + k), 0.5 + 0.1 * k))
        Ts.append(T)
        kts.append(0.002 * (k + 1))
        seeds.append(1000 + k)
    b = GPUBatch(gHs, states, seeds, kts)
    obs = b.run(9, 64, renorm_interval=4, measure_interval=3)
    assert obs.shape == (4, 3, 4)
    for k, T in enumerate(Ts):
        c, z, E = states[k].config.copy(), states[k].zrows.copy(), states[k].energy
        for j in range(1, 10):
            _, E, _, _ = T.keyed_sweep(c, z, E, 1.0 / kts[k], 0.5 + 0.1 * k, seeds[k], j, 64)
            if j % 4 == 0:
                _, E = T.renormalize(c, z, E)
        cfg, zd, Ed = b.download(k)
        assert np.array_equal(cfg, c) and np.array_equal(zd, z) and Ed == E, k
        # the batch chain equals the same chain run alone on the device
        g = S.GPUChainState(gHs[k], states[k], seed=seeds[k])
        o1 = S.gpu_run_sweeps(g, gHs[k], None, 1.0 / kts[k], 9, renorm_interval=4, measure_interval=3)
        c1, z1, E1 = g.download()
        assert np.array_equal(c1, cfg) and np.array_equal(z1, zd) and E1 == Ed
        assert np.array_equal(o1[:, 0], obs[k, :, 0])


def test_batch_direct_launch_equals_graph():
    na, terms, dims = bcc_tet_small(2)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    R = 3
    states = [S.ChainState(H, rand_config(H.n_sites, r), 0.6) for r in range(R)]
    kts = [0.002, 0.004, 0.008]
    outs = []
    for graph in (True, False):
        b = GPUBatch([gH] * R, states, [5, 6, 7], kts, 3)
        b.set_graph(graph)
        b.run(8, 32, 0, 2, 2, 0, 8)
        outs.append([b.download(r) for r in range(R)] + [b.swap_stats()])
    for r in range(R):
        assert np.array_equal(outs[0][r][0], outs[1][r][0]) and outs[0][r][2] == outs[1][r][2]
    assert np.array_equal(outs[0][R][1], outs[1][R][1])


@pytest.mark.parametrize("renorm, exch", [(40, 7), (0, 0)])
def test_split_runs_equal_one_run(renorm, exch):
    """A phase run as several calls (as checkpointed ladder runs do) gives the
    rows, payloads and exchange state of one call — including a later call
    that takes fewer measurements than an earlier one (the device keeps its
    larger row buffer; every chain's rows are copied from its own offset)."""
    H = S.TiledHamiltonian(*bcc_tet_small(11, dims=(2, 2, 2)))
    gH = S.GPUTiledHamiltonian(H)

    def mk():
        sts = [S.ChainState(H, rand_config(H.n_sites, 60 + k), 0.6) for k in range(3)]
        return GPUBatch([gH] * 3, sts, [5, 6, 7], [0.05, 0.03, 0.02], 99)

    b1, b2 = mk(), mk()
    o1 = b1.run(300, 32, renorm, 3, exch, 0, 300)
    o2 = np.concatenate([b2.run(n, 32, renorm, 3, exch, d, 300)
                         for d, n in [(0, 120), (120, 120), (240, 60)]], axis=1)
    assert o1.shape == o2.shape == (3, 100, 4)
    assert np.array_equal(o1, o2)
    assert b1.exchange_state() == b2.exchange_state()
    for k in range(3):
        assert np.array_equal(b1.download(k)[0], b2.download(k)[0])


def _ladder_vs_oracle(H, T, gH, R, kts, nsw, lanes, renorm, seed0):
    """Run an R-rung ladder with exchange after every sweep on the device and
    through the oracle composition; assert every payload, counter and exchange
    statistic is bit-identical."""
    seeds = [seed0 + 7 * r for r in range(R)]
    states = [S.ChainState(H, rand_config(H.n_sites, seed0 + r), 0.6) for r in range(R)]
    exch_seed = seed0 * 31 + 5
    b = GPUBatch([gH] * R, states, seeds, kts, exch_seed)
    b.run(nsw, lanes, renorm_interval=renorm, measure_interval=0, exchange_interval=1,
          phase_done=0, phase_total=nsw)
    payload = [[st.config.copy(), st.zrows.copy(), st.energy] for st in states]
    att = np.zeros(R - 1, int)
    acc = np.zeros(R - 1, int)
    parity, ex = 0, 0
    for j in range(1, nsw + 1):
        for r in range(R):
            c, z, E = payload[r]
            _, E, _, _ = T.keyed_sweep(c, z, E, 1.0 / kts[r], 0.6, seeds[r], j, lanes)
            payload[r][2] = E
        if renorm and j % renorm == 0:
            for r in range(R):
                _, payload[r][2] = T.renormalize(payload[r][0], payload[r][1], payload[r][2])
        if j < nsw:
            ex += 1
            a1, a2 = _oracle_exchange(payload, kts, parity, ex, exch_seed)
            att += a1
            acc += a2
            parity = 1 - parity
    for r in range(R):
        cfg, z, E = b.download(r)
        assert np.array_equal(cfg, payload[r][0]), r
        assert np.array_equal(z, payload[r][1]), r
        assert E == payload[r][2], r
    da, dc = b.swap_stats()
    assert np.array_equal(da, att) and np.array_equal(dc, acc)
    return acc


@pytest.mark.parametrize("lanes", [32, 64])
def test_c3_fcc_four_shell_ladder_bitwise(lanes):
    """The C3 model form (fcc 4-atom basis, body-2 l <= 2 to the 4th shell) as a
    6-rung geometric ladder on a 3^3 cell, exchanging after every sweep."""
    na, terms = M.lattice_model("fcc", pair_shells=4, pair_lmax=2, seed=12)
    H, T = build_pair(na, terms, (3, 3, 3))
    gH = S.GPUTiledHamiltonian(H)
    R = 6
    kts = np.geomspace(0.002, 0.03, R)
    acc = _ladder_vs_oracle(H, T, gH, R, kts, 10, lanes, 4, 300 + lanes)
    assert acc.sum() > 0


def test_c3_full_size_replicas_two_sweeps_bitwise():
    """Two 6912-site fcc 12^3 replicas of the C3 model: two sweeps with an
    exchange between them, bit-identical to the oracle."""
    na, terms, dims = M.config_model("fcc12", seed=0)
    H, T = build_pair(na, terms, dims)
    assert H.n_sites == 6912
    gH = S.GPUTiledHamiltonian(H)
    _ladder_vs_oracle(H, T, gH, 2, np.array([0.004, 0.006]), 2, 32, 0, 77)


def test_uneven_batch_with_zeeman_fields_bitwise():
    """Uneven batch in the C4 form: per-chain models, moments of 2 mu_B and a
    0-10 T field along a random axis, sizes up to 10^3 = 1000 sites; every chain
    bit-identical to the oracle."""
    rng = np.random.default_rng(404)
    specs = [("sc", 10, 2), ("bcc", 6, 2), ("sc", 7, 1), ("bcc", 5, 2), ("bcc", 7, 2)]
    gHs, states, Ts, kts, seeds, steps = [], [], [], [], [], []
    for k, (lat, N, lmax) in enumerate(specs):
        na, terms = M.lattice_model(lat, pair_shells=2, pair_lmax=lmax, triangles=(lat == "bcc"), seed=500 + k)
        B = rng.uniform(0.0, 10.0) * (lambda v: v / np.linalg.norm(v))(rng.standard_normal(3))
        H, T = build_pair(na, terms, (N, N, N), [2.0] * na, tuple(B))
        gHs.append(S.GPUTiledHamiltonian(H))
        steps.append(0.4 + 0.1 * k)
        states.append(S.ChainState(H, rand_config(H.n_sites, 90 + k), steps[-1]))
        Ts.append(T)
        kts.append(float(rng.uniform(0.005, 0.05)))
        seeds.append(int(rng.integers(1, 2**62)))
    assert max(g.host.n_sites for g in gHs) >= 1000
    # lmax differs between chains of one batch only where the library allows it
    groups = {}
    for k, g in enumerate(gHs):
        groups.setdefault(g.host.lmax, []).append(k)
    for ks in groups.values():
        b = GPUBatch([gHs[k] for k in ks], [states[k] for k in ks], [seeds[k] for k in ks], [kts[k] for k in ks])
        b.run(6, 64, renorm_interval=4, measure_interval=0)
        for i, k in enumerate(ks):
            c, z, E = states[k].config.copy(), states[k].zrows.copy(), states[k].energy
            for j in range(1, 7):
                _, E, _, _ = Ts[k].keyed_sweep(c, z, E, 1.0 / kts[k], steps[k], seeds[k], j, 64)
                if j % 4 == 0:
                    _, E = Ts[k].renormalize(c, z, E)
            cfg, zd, Ed = b.download(i)
            assert np.array_equal(cfg, c) and np.array_equal(zd, z) and Ed == E, k


def test_concurrent_batches_equal_running_them_alone():
    """run_concurrent: three uneven batches (per-chain models and Zeeman
    fields, different lane counts) sweeping at once on their own streams end
    bit-identical to the same batches run one after another, measurements
    included; a second concurrent call continues them identically."""
    from sce_amd.batch import run_concurrent
    groups = [[("sc", 8), ("bcc", 6)], [("bcc", 7), ("sc", 9), ("sc", 6)], [("bcc", 5)]]
    lanes = [16, 32, 64]

    def make():
        out = []
        r = np.random.default_rng(5)
        for g, specs in enumerate(groups):
            gHs, sts, sds, kts = [], [], [], []
            for k, (lat, N) in enumerate(specs):
                na, terms = M.lattice_model(lat, pair_shells=2, pair_lmax=2, triangles=(lat == "bcc"),
                                            seed=700 + 10 * g + k)
                B = r.uniform(0.0, 10.0) * (lambda v: v / np.linalg.norm(v))(r.standard_normal(3))
                H = S.TiledHamiltonian(na, terms, (N, N, N), magmoms=[2.0] * na, field=tuple(B))
                gHs.append(S.GPUTiledHamiltonian(H))
                sts.append(S.ChainState(H, rand_config(H.n_sites, 30 + 10 * g + k), 0.5))
                sds.append(int(r.integers(1, 2**62)))
                kts.append(float(r.uniform(0.005, 0.05)))
            out.append(GPUBatch(gHs, sts, sds, kts))
        return out
    alone = make()
    obs_a = [b.run(9, lanes[i], renorm_interval=4, measure_interval=2) for i, b in enumerate(alone)]
    both = make()
    runs = [(b, (9, lanes[i]), {"renorm_interval": 4, "measure_interval": 2}) for i, b in enumerate(both)]
    obs_c = run_concurrent(runs)
    for i in range(len(groups)):
        assert np.array_equal(obs_a[i], obs_c[i]), i
        for k in range(len(groups[i])):
            ca, za, Ea = alone[i].download(k)
            cc, zc, Ec = both[i].download(k)
            assert np.array_equal(ca, cc) and np.array_equal(za, zc) and Ea == Ec, (i, k)
            assert alone[i].info(k).acc_metro == both[i].info(k).acc_metro
    for i, b in enumerate(alone):
        b.run(5, lanes[i], renorm_interval=4, phase_done=9)
    run_concurrent([(b, (5, lanes[i]), {"renorm_interval": 4, "phase_done": 9}) for i, b in enumerate(both)])
    for i in range(len(groups)):
        for k in range(len(groups[i])):
            assert np.array_equal(alone[i].download(k)[1], both[i].download(k)[1]), (i, k)


def test_concurrent_chains_equal_running_them_alone():
    """run_concurrent over single chains on one model (compound Metropolis +
    overrelaxation sweeps and plain ones, at different kT): the same bits and
    measurements as running each chain alone."""
    from sce_amd.batch import run_concurrent
    na, terms, dims = bcc_tet_small(3, dims=(4, 4, 4))
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    kts = [0.01, 0.02, 0.04]

    def make():
        return [S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, 60 + k), 0.7), seed=800 + k)
                for k in range(len(kts))]
    kw = [{"renorm_interval": 5, "or_per_metropolis": 1, "measure_interval": 1},
          {"renorm_interval": 5, "measure_interval": 2}, {"renorm_interval": 0, "or_per_metropolis": 1}]
    alone = make()
    obs_a = [S.gpu_run_sweeps(g, gH, None, 1.0 / kts[k], 12, **kw[k]) for k, g in enumerate(alone)]
    both = make()
    obs_c = run_concurrent([(S.gpu_run_sweeps, (g, gH, None, 1.0 / kts[k], 12), kw[k]) for k, g in enumerate(both)])
    for k in range(len(kts)):
        assert np.array_equal(obs_a[k], obs_c[k]), k
        ca, za, Ea = alone[k].download()
        cc, zc, Ec = both[k].download()
        assert np.array_equal(ca, cc) and np.array_equal(za, zc) and Ea == Ec, k
        ia, ic = alone[k].info(), both[k].info()
        assert (ia.acc_metro, ia.acc_or, ia.sweep_index) == (ic.acc_metro, ic.acc_or, ic.sweep_index), k


def test_swap_stats_writes_exactly_n_chains_minus_one():
    """sce_batch_swap_stats fills n_chains - 1 entries (pair k, k + 1), the length
    the header documents; nothing past them is written."""
    import ctypes as C
    from sce_amd import _lib as L
    H = S.TiledHamiltonian(*bcc_tet_small(3, dims=(2, 2, 2)))
    gH = S.GPUTiledHamiltonian(H)
    R = 4
    sts = [S.ChainState(H, rand_config(H.n_sites, k), 0.6) for k in range(R)]
    b = GPUBatch([gH] * R, sts, [1, 2, 3, 4], [0.01, 0.02, 0.03, 0.04], 8)
    b.run(6, 32, 0, 0, 1, 0, 6)
    sentinel = -123456789
    att = np.full(R + 3, sentinel, np.int64)
    acc = np.full(R + 3, sentinel, np.int64)
    L.check(L.lib().sce_batch_swap_stats(b._h, L.ptr(att), L.ptr(acc)))
    assert (att[R - 1:] == sentinel).all() and (acc[R - 1:] == sentinel).all()
    # 5 exchanges with parity 0, 1, 0, 1, 0: pairs (0,1), (2,3) on even, (1,2) on odd
    assert att[: R - 1].tolist() == [3, 2, 3]
    assert (acc[: R - 1] >= 0).all() and (acc[: R - 1] <= att[: R - 1]).all()
    a2, c2 = b.swap_stats()
    assert np.array_equal(a2, att[: R - 1]) and np.array_equal(c2, acc[: R - 1])


def test_pt_autotuned_lanes_are_recorded():
    """gpu_run_pt(lanes="auto"): the tuner's choice is pinned in the plan, returned
    with the result, and the run equals the same run with those lanes given."""
    H = S.TiledHamiltonian(*bcc_tet_small(4, dims=(2, 2, 2)))
    gH = S.GPUTiledHamiltonian(H)
    kw = dict(kT=[0.01, 0.02, 0.04], exchange_interval=1, sweeps_therm=20, sweeps_measure=30, seed=17)
    r = S.gpu_run_pt(gH, lanes="auto", **kw)
    assert r.lanes in (16, 32, 64) and r.persistent in (True, False)
    r2 = S.gpu_run_pt(gH, lanes=r.lanes, **kw)
    for a, b in zip(r.final_configs, r2.final_configs):
        assert np.array_equal(a, b)
