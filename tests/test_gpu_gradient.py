"""The device all-site energy gradient (gpu_energy_gradient!,
src/gpu/gpu_gradient.jl) against the oracle's restatement of the reference's
lane contract (_gradient_lane_ref!): bit-identical for 64, 128 and 256 strided
shares on every fixture model, for the configuration entry point (rows rebuilt
on the device) and the chain entry point (the chain's rows)."""
import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd._lib import SceError
from fixtures import build_pair, fixture_cases, rand_config, bcc_tet_small

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_gradient_bitwise_vs_oracle(case):
    _, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    cfg = rand_config(H.n_sites, 23)
    z = T.zrows(cfg)
    for lanes in (64, 128, 256):
        G = S.gpu_energy_gradient(gH, cfg, lanes)
        R = T.energy_gradient(cfg, z, lanes)
        assert np.array_equal(G, R), (case[0], lanes, np.abs(G - R).max())
    inactive = np.nonzero(H.site_active == 0)[0]
    assert np.all(G[inactive] == 0.0)


def test_chain_gradient_reads_the_chain_rows():
    na, terms, dims = bcc_tet_small(3, dims=(3, 3, 3))
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    g = S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, 4), 0.8), seed=6)
    S.gpu_run_sweeps(g, gH, None, 1 / 0.05, 12, renorm_interval=5)
    cfg, z, _ = g.download()
    for lanes in (64, 128):
        Gc = S.gpu_chain_energy_gradient(g, lanes)
        assert np.array_equal(Gc, T.energy_gradient(cfg, z, lanes))
        assert np.array_equal(Gc, S.gpu_energy_gradient(gH, cfg, lanes))
    assert S.gpu_time_gradient(g, 128, 3) > 0.0
    assert np.array_equal(g.download()[0], cfg)          # read-only pass


def test_gradient_at_the_benchmark_size_is_tangent_and_matches_the_oracle():
    from sce_amd import models as M
    na, terms, dims = M.config_model("bcc16", seed=0)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    cfg = rand_config(H.n_sites, 31)
    G = S.gpu_energy_gradient(gH, cfg)
    assert np.array_equal(G, T.energy_gradient(cfg, lanes=S.GRADIENT_LANES))
    assert np.abs(np.sum(G * cfg, axis=1)).max() <= 1e-12 * np.abs(G).max()


def test_gradient_rejects_bad_arguments():
    na, terms, dims = bcc_tet_small(1, dims=(2, 2, 2))
    H, _ = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    cfg = rand_config(H.n_sites, 1)
    with pytest.raises(SceError):
        S.gpu_energy_gradient(gH, cfg, 32)
    with pytest.raises(ValueError):
        S.gpu_energy_gradient(gH, cfg[:-1], 128)


def test_gradient_both_kernels_at_lmax_6():
    """A body-1 l = 6 term raises the rows to 49 entries: with the 26-site bcc
    stencil the walk-layout kernel's LDS (rows, gradient rows and share
    partials of four sites) fits for 64 and 128 shares but not for 256, where
    the adjacency-table kernel runs; both equal the oracle bit for bit."""
    from sce_amd import models as M
    na, terms = M.lattice_model("bcc", pair_shells=3, pair_lmax=1, seed=8)
    rng = np.random.default_rng(8)
    for a in range(1, na + 1):
        terms.append(M.MultipoleTerm(1e-3 * rng.normal(), [a], [(0, 0, 0)], [6], rng.normal(size=13)))
    H, T = build_pair(na, terms, (3, 3, 3))
    assert H.lmax == 6
    gH = S.GPUTiledHamiltonian(H)
    assert gH.nb_max == 26
    for lanes in (64, 128, 256):
        lds = 4 * (26 * 49 + 1 + 3 * 49 + 3 * lanes) * 8
        assert (lds <= 65536) == (lanes < 256)
        cfg = rand_config(H.n_sites, 40 + lanes)
        assert np.array_equal(S.gpu_energy_gradient(gH, cfg, lanes), T.energy_gradient(cfg, lanes=lanes)), lanes
