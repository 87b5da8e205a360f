"""The device binner (LogBinner of src/binning.jl:25-79 and BinStore of
:138-188 over E, E^2, |m|, m^2, m^4) against the host classes fed with the
downloaded measurement rows: bit-identical state, identical statistics."""
import numpy as np
import pytest

import sce_amd as S
from sce_amd import observables as OB
from sce_amd._lib import SceError
from fixtures import bcc_small, build_pair, rand_config

pytestmark = pytest.mark.gpu


def _chain(seed=3):
    na, terms, dims = bcc_small(4)
    H, _ = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    g = S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, seed), 0.9), seed=seed + 10)
    return H, gH, g


def _assert_same(g, obs, nbins):
    hb, hm = OB.host_binners(obs, nbins)
    db, dm = g.binner_state()
    for name in OB.SERIES:
        assert db[name].n == hb[name].n == len(obs)
        assert np.array_equal(db[name].count, hb[name].count)
        assert np.array_equal(db[name].sums, hb[name].sums), name
        assert np.array_equal(db[name].sums2, hb[name].sums2), name
        assert np.array_equal(dm[name], hm[name]), name


@pytest.mark.parametrize("graph", [True, False])
def test_device_binner_matches_host_binners(graph):
    H, gH, g = _chain()
    g.set_graph(graph)
    S.gpu_run_sweeps(g, gH, None, 1 / 0.05, 20, renorm_interval=0)
    n1, n2, nbins = 150, 173, 32
    g.binner_reset((n1 + n2) // nbins, nbins)
    o1 = S.gpu_run_sweeps(g, gH, None, 1 / 0.05, n1, renorm_interval=40, measure_interval=1)
    o2 = S.gpu_run_sweeps(g, gH, None, 1 / 0.05, n2, renorm_interval=40, measure_interval=1)
    obs = np.vstack([o1, o2])
    _assert_same(g, obs, nbins)
    ds = g.device_stats(0.05)
    hs = S.standard_stats(obs, 0.05, H.n_active, nbins)
    for k, v in hs.items():
        d = ds[k]
        assert (d.mean, d.count) == (v.mean, v.count) or (np.isnan(d.mean) and np.isnan(v.mean)), k
        assert d.err == v.err or (np.isnan(d.err) and np.isnan(v.err)), k


def test_device_binner_interval_and_off():
    H, gH, g = _chain(7)
    g.binner_reset(5, 8)
    o = S.gpu_run_sweeps(g, gH, None, 1 / 0.08, 120, renorm_interval=0, measure_interval=3)
    _assert_same(g, o, 8)
    g.binner_off()
    S.gpu_run_sweeps(g, gH, None, 1 / 0.08, 30, renorm_interval=0, measure_interval=1)
    _assert_same(g, o, 8)          # no pushes while off
    g.binner_reset()               # LogBinner alone
    o = S.gpu_run_sweeps(g, gH, None, 1 / 0.08, 64, renorm_interval=0, measure_interval=1)
    db, dm = g.binner_state()
    hb, _ = OB.host_binners(o)
    assert np.array_equal(db["energy"].sums, hb["energy"].sums)
    assert all(len(v) == 0 for v in dm.values())


def test_device_binner_fed_by_delayed_acceptance():
    from sce_amd import surrogate as SG
    H, gH, g = _chain(9)
    sur = SG.Surrogate(gH, seed=4)
    g.binner_reset(4, 10)
    o = SG.gpu_da_run_sweeps(g, sur, None, 1 / 0.05, 44, renorm_interval=0, measure_interval=1)
    _assert_same(g, o, 10)


def test_device_binner_rejects_bad_layouts():
    _, _, g = _chain()
    for bs, nb in ((0, 4), (3, 0), (3, 1), (3, 2000), (-1, 4)):
        with pytest.raises(SceError):
            g.binner_reset(bs, nb)
