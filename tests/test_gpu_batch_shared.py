"""Shared walks: a ladder on one model with 2 or 4 chains per wavefront
walking each site's sublattice walk once (GPUBatch.set_shared_walk), and any
batch with two same-sublattice sites of one chain per wavefront
(GPUBatch.set_site_pairs), are bit-identical to the oracle at 64 lanes per
site and to the unshared launches — payloads, energies, counters,
measurements and exchange statistics, with chain counts and classes that
leave a partial last group or a lone site, in graph replay and direct
launches."""
import numpy as np
import pytest

import sce_amd as S
from sce_amd import _lib as L
from sce_amd import models as M
from sce_amd.batch import GPUBatch
from fixtures import bcc_tet_small, build_pair, rand_config
from test_gpu_batch import _oracle_exchange

pytestmark = pytest.mark.gpu


def _batch(H, gH, R, kts, seed0):
    seeds = [seed0 + 7 * r for r in range(R)]
    states = [S.ChainState(H, rand_config(H.n_sites, seed0 + r), 0.6) for r in range(R)]
    return GPUBatch([gH] * R, states, seeds, kts, seed0 * 31 + 5), seeds, states


def _shared_vs_oracle(H, T, gH, R, kts, nsw, renorm, seed0, share, graph):
    b, seeds, states = _batch(H, gH, R, kts, seed0)
    b.set_shared_walk(share)
    b.set_graph(graph)
    b.run(nsw, 64, renorm_interval=renorm, measure_interval=0, exchange_interval=1,
          phase_done=0, phase_total=nsw)
    payload = [[st.config.copy(), st.zrows.copy(), st.energy] for st in states]
    att = np.zeros(R - 1, int)
    acc = np.zeros(R - 1, int)
    parity, ex = 0, 0
    nacc = np.zeros(R, int)
    for j in range(1, nsw + 1):
        for r in range(R):
            c, z, E = payload[r]
            a, E, _, _ = T.keyed_sweep(c, z, E, 1.0 / kts[r], 0.6, seeds[r], j, 64)
            payload[r][2] = E
            nacc[r] += a
        if renorm and j % renorm == 0:
            for r in range(R):
                _, payload[r][2] = T.renormalize(payload[r][0], payload[r][1], payload[r][2])
        if j < nsw:
            ex += 1
            a1, a2 = _oracle_exchange(payload, kts, parity, ex, seed0 * 31 + 5)
            att += a1
            acc += a2
            parity = 1 - parity
    for r in range(R):
        cfg, z, E = b.download(r)
        assert np.array_equal(cfg.view(np.uint64), payload[r][0].view(np.uint64)), r
        assert np.array_equal(z.view(np.uint64), payload[r][1].view(np.uint64)), r
        assert E == payload[r][2], r
        # counters stay with the rung while the payloads move between rungs
        assert (b.info(r).acc_metro, b.info(r).att_metro) == (nacc[r], nsw * H.n_active), r
    da, dc = b.swap_stats()
    assert np.array_equal(da, att) and np.array_equal(dc, acc)
    return acc


@pytest.mark.parametrize("share, R, graph", [(2, 5, True), (4, 6, True), (4, 4, False), (2, 2, False),
                                               (2, 9, False), (4, 3, True)])
def test_fcc_four_shell_ladder_shared_walk_bitwise(share, R, graph):
    """The fcc 4-atom basis with body-2 l <= 2 to the 4th shell (pair walks
    only) on a 3^3 cell, exchanging after every sweep."""
    na, terms = M.lattice_model("fcc", pair_shells=4, pair_lmax=2, seed=12)
    H, T = build_pair(na, terms, (3, 3, 3))
    gH = S.GPUTiledHamiltonian(H)
    kts = np.geomspace(0.002, 0.03, R)
    acc = _shared_vs_oracle(H, T, gH, R, kts, 8, 4, 500 + 10 * share + R, share, graph)
    if R >= 5:  # narrow enough kT spacing for accepted exchanges
        assert acc.sum() > 0


@pytest.mark.parametrize("share", [2, 4])
def test_cluster_walk_ladder_shared_walk_bitwise(share):
    """Triangles and tetrahedra (walks with one-, two- and three-factor
    entries) on bcc, 7 rungs (a partial last group)."""
    na, terms, dims = bcc_tet_small(6)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    kts = np.geomspace(0.001, 0.02, 7)
    _shared_vs_oracle(H, T, gH, 7, kts, 6, 5, 900 + share, share, True)


def test_full_size_ladder_shared_equals_unshared_with_measurements():
    """Eight fcc 12^3 replicas: shared walks of 2 and 4 chains and the default give the same
    bits as 64 lanes per site unshared — payloads, measurements, counters and
    exchange statistics — over sweeps with renormalisation and exchange."""
    na, terms, dims = M.config_model("fcc12", seed=0)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    R = 8
    kts = np.geomspace(0.001, 0.008, R)
    out = []
    for share in (1, 2, 4, 0):
        b, _, _ = _batch(H, gH, R, kts, 61)
        b.set_shared_walk(share)
        obs = b.run(6, 64, renorm_interval=4, measure_interval=2, exchange_interval=1,
                    phase_done=0, phase_total=6)
        out.append((obs, [b.download(r) for r in range(R)], [b.info(r).acc_metro for r in range(R)],
                    b.swap_stats()))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0])
        for (c1, z1, e1), (c2, z2, e2) in zip(o[1], out[0][1]):
            assert np.array_equal(c1, c2) and np.array_equal(z1, z2) and e1 == e2
        assert o[2] == out[0][2]
        assert all(np.array_equal(x, y) for x, y in zip(o[3], out[0][3]))


def test_shared_walk_guards():
    """Chains per shared walk must be 0 (default), 1, 2 or 4, and a batch of
    different tables cannot share walks; other lane counts ignore the
    setting."""
    na, terms, dims = bcc_tet_small(6)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    H2, _ = build_pair(na, terms, dims)
    gH2 = S.GPUTiledHamiltonian(H2)
    st = [S.ChainState(H, rand_config(H.n_sites, 3 + r), 0.6) for r in range(2)]
    b = GPUBatch([gH, gH2], st, [1, 2], [0.01, 0.02])
    with pytest.raises(L.SceError):
        b.set_shared_walk(2)
    b.set_shared_walk(1)
    h, _, _ = _batch(H, gH, 3, np.array([0.01, 0.02, 0.03]), 5)
    for bad in (-1, 3, 8):
        with pytest.raises(L.SceError):
            h.set_shared_walk(bad)
    # at 32 lanes the setting is unused: the results are those of 32 lanes
    h.set_shared_walk(2)
    h.run(3, 32, 0)
    u, _, _ = _batch(H, gH, 3, np.array([0.01, 0.02, 0.03]), 5)
    u.run(3, 32, 0)
    for r in range(3):
        assert np.array_equal(h.download(r)[1], u.download(r)[1])


def _uneven(seed, specs):
    rng = np.random.default_rng(seed)
    gHs, states, Ts, kts, seeds = [], [], [], [], []
    for k, (lat, N) in enumerate(specs):
        na, terms = M.lattice_model(lat, pair_shells=2, pair_lmax=2, triangles=(lat == "bcc"), seed=seed + k)
        B = rng.uniform(0.0, 10.0) * (lambda v: v / np.linalg.norm(v))(rng.standard_normal(3))
        H, T = build_pair(na, terms, (N, N, N), [2.0] * na, tuple(B))
        gHs.append(S.GPUTiledHamiltonian(H))
        states.append(S.ChainState(H, rand_config(H.n_sites, seed + 50 + k), 0.5))
        Ts.append(T)
        kts.append(float(rng.uniform(0.005, 0.05)))
        seeds.append(int(rng.integers(1, 2**62)))
    return gHs, states, Ts, kts, seeds


@pytest.mark.parametrize("graph", [True, False])
def test_uneven_batch_site_pairs_bitwise(graph):
    """Site pairs on an uneven batch (sc and bcc chains of different sizes,
    per-chain models and Zeeman fields; odd class sizes leave lone sites):
    every chain bit-identical to the oracle at 64 lanes per site."""
    specs = [("sc", 5), ("bcc", 4), ("sc", 3), ("bcc", 5), ("bcc", 3)]
    gHs, states, Ts, kts, seeds = _uneven(610, specs)
    b = GPUBatch(gHs, states, seeds, kts)
    b.set_site_pairs(True)
    b.set_graph(graph)
    obs = b.run(7, 64, renorm_interval=3, measure_interval=2)
    for k in range(len(specs)):
        c, z, E = states[k].config.copy(), states[k].zrows.copy(), states[k].energy
        nacc = 0
        for j in range(1, 8):
            a, E, _, _ = Ts[k].keyed_sweep(c, z, E, 1.0 / kts[k], 0.5, seeds[k], j, 64)
            nacc += a
            if j % 3 == 0:
                _, E = Ts[k].renormalize(c, z, E)
        cfg, zd, Ed = b.download(k)
        assert np.array_equal(cfg.view(np.uint64), c.view(np.uint64)), k
        assert np.array_equal(zd.view(np.uint64), z.view(np.uint64)), k
        assert Ed == E, k
        assert b.info(k).acc_metro == nacc, k
    assert obs.shape[1] == 3


def test_site_pairs_equal_unpaired_at_full_size():
    """Ten chains of the uneven benchmark's sizes (sc / bcc up to 20^3): site
    pairs give the bits of 64 lanes per site unpaired, measurements included;
    switching the pairs off and on between calls continues identically; on a
    ladder they take the place of the replica sharing only when it is off."""
    specs = [("sc", 20), ("bcc", 17), ("sc", 13), ("bcc", 11), ("sc", 9), ("bcc", 20), ("sc", 7), ("bcc", 8),
             ("sc", 16), ("bcc", 6)]
    gHs, states, _, kts, seeds = _uneven(720, specs)
    runs = []
    for mode in ("plain", "pairs", "switch"):
        b = GPUBatch(gHs, states, seeds, kts)
        b.set_site_pairs(mode != "plain")
        obs = [b.run(4, 64, renorm_interval=3, measure_interval=1)]
        if mode == "switch":
            b.set_site_pairs(False)
        obs.append(b.run(3, 64, renorm_interval=3, measure_interval=1, phase_done=4))
        runs.append((obs, [b.download(k) for k in range(len(specs))]))
    for obs, dl in runs[1:]:
        for x, y in zip(obs, runs[0][0]):
            assert np.array_equal(x, y)
        for (c1, z1, e1), (c2, z2, e2) in zip(dl, runs[0][1]):
            assert np.array_equal(c1, c2) and np.array_equal(z1, z2) and e1 == e2
    na, terms, dims = bcc_tet_small(6)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    kts = np.geomspace(0.001, 0.02, 5)
    b, seeds, sts = _batch(H, gH, 5, kts, 930)
    b.set_shared_walk(1)
    b.set_site_pairs(True)
    b.run(4, 64, renorm_interval=0, exchange_interval=1, phase_done=0, phase_total=4)
    u, _, _ = _batch(H, gH, 5, kts, 930)
    u.set_shared_walk(1)
    u.run(4, 64, renorm_interval=0, exchange_interval=1, phase_done=0, phase_total=4)
    for r in range(5):
        assert np.array_equal(b.download(r)[1], u.download(r)[1])
