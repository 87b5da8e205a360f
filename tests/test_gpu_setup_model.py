"""Device sweeps on models read from a setup file (the Nd2Fe14B cell of the
reference's bench fixtures, bench/fixtures.jl:66-88): walk rows, keyed sweeps,
overrelaxation and chained sweeps bitwise against the oracle, on a 64-colour
all-pairs l <= 4 cell with a Zeeman field (a 63-site stencil that fits LDS at
64 lanes only — narrower lanes must fail with an error) and on isotropic and
anisotropic pair + triplet models at 2x2x2 cells."""
import os

import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd import gpu as G
from sce_amd._lib import SceError
from fixtures import ZB, build_pair, rand_config

pytestmark = pytest.mark.gpu

SETUP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nd2fe14b.toml")
L2 = [2] * 8 + [0]

# (name, setup_model keywords, dims, field, lanes that fit)
CASES = [
    ("pairs-l4+field", dict(), (1, 1, 1), ZB, (64,)),
    ("3body-iso", dict(nbody=3, cutoff=3.5, lmax=L2), (2, 2, 2), None, (16, 32, 64)),
    ("3body-aniso", dict(nbody=3, cutoff=3.0, lmax=L2, isotropy=False), (2, 2, 2), None, (16, 32, 64)),
]
_BUILT = {}


def _case(c):
    name, kw, dims, field, lanes = c
    if name not in _BUILT:
        st = S.read_setup(SETUP)
        na, terms = S.setup_model(st, **kw)
        mm = S.setup_magmoms(st) if field is not None else None
        H, T = build_pair(na, terms, dims, mm, field)
        _BUILT[name] = (H, T, S.GPUTiledHamiltonian(H))
    return _BUILT[name] + (lanes,)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_walk_rows_bitwise(case):
    H, T, gH, lanes = _case(case)
    z = H.zrows(rand_config(H.n_sites, 31))
    sites = np.array(T.color_sites[:64], np.int32)
    znew = np.array([O.zlm_row(v, H.lmax) for v in S.random_config(len(sites), np.random.default_rng(7))])
    for ln in lanes:
        dev = G.dev_walk_delta(gH, z, sites, znew, ln)
        cdev = G.dev_walk_coeffs(gH, z, sites, ln)
        for i, s in enumerate(sites):
            assert dev[i] == T.walk_delta(int(s), z, znew[i], ln), (ln, s)
            assert np.array_equal(cdev[i], T.walk_coeffs(int(s), z, ln)), (ln, s)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_sweeps_and_overrelaxation_bitwise(case):
    H, T, gH, lanes = _case(case)
    st = S.ChainState(H, rand_config(H.n_sites, 7), 0.6)
    for ln in lanes:
        gst = S.GPUChainState(gH, st, seed=404)
        c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
        acc = 0
        for sw in range(1, 4):
            n = S.gpu_metropolis_sweep(gst, gH, 20.0, lanes=ln)
            m, E, _, _ = T.keyed_sweep(c2, z2, E, 20.0, 0.6, 404, sw, ln)
            assert n == m
            acc += n
            S.gpu_overrelaxation_sweep(gst, gH, 20.0, or_pass=0, lanes=ln)
            _, _, E = T.keyed_or_sweep(c2, z2, E, 20.0, 404, sw, 0, ln)
        c, z, Ed = gst.download()
        assert np.array_equal(c.view(np.uint64), c2.view(np.uint64)), ln
        assert np.array_equal(z.view(np.uint64), z2.view(np.uint64)), ln
        assert Ed == E
        assert gst.info().acc_metro == acc


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_chained_sweeps_equal_single_sweeps(case):
    H, T, gH, lanes = _case(case)
    ln = lanes[-1]
    st = S.ChainState(H, rand_config(H.n_sites, 21), 0.6)
    ref = S.GPUChainState(gH, st, seed=9)
    for _ in range(6):
        S.gpu_metropolis_sweep(ref, gH, 30.0, lanes=ln)
    cr, zr, Er = ref.download()
    g = S.GPUChainState(gH, st, seed=9)
    S.gpu_run_sweeps(g, gH, None, 30.0, 6, renorm_interval=0, measure_interval=2, lanes=ln)
    c, z, E = g.download()
    assert np.array_equal(c.view(np.uint64), cr.view(np.uint64))
    assert np.array_equal(z.view(np.uint64), zr.view(np.uint64)) and E == Er
    assert abs(E - g.total_energy()) <= 1e-10 * max(1.0, abs(E))


@pytest.mark.parametrize("case", CASES[1:], ids=lambda c: c[0])
@pytest.mark.parametrize("lanes", [128, 256])
def test_wide_lanes_bitwise(case, lanes):
    """Two and four wavefronts per site on the heavy triplet walks (thousands
    of entries per site, so every count of trailing pads occurs): walk rows
    and chained Metropolis sweeps with renormalisation against the oracle at
    that width."""
    H, T, gH, _ = _case(case)
    z = H.zrows(rand_config(H.n_sites, 41))
    sites = np.array(T.color_sites[:64], np.int32)
    znew = np.array([O.zlm_row(v, H.lmax) for v in S.random_config(len(sites), np.random.default_rng(8))])
    dev = G.dev_walk_delta(gH, z, sites, znew, lanes)
    for i, s in enumerate(sites):
        assert dev[i] == T.walk_delta(int(s), z, znew[i], lanes), (lanes, s)
    st = S.ChainState(H, rand_config(H.n_sites, 17), 0.6)
    g = S.GPUChainState(gH, st, seed=515)
    S.gpu_run_sweeps(g, gH, None, 25.0, 4, renorm_interval=2, measure_interval=1, lanes=lanes)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    acc = 0
    for sw in range(1, 5):
        a, E, _, _ = T.keyed_sweep(c2, z2, E, 25.0, 0.6, 515, sw, lanes)
        acc += a
        if sw % 2 == 0:
            _, E = T.renormalize(c2, z2, E)
    c, zd, Ed = g.download()
    assert np.array_equal(c.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(zd.view(np.uint64), z2.view(np.uint64)) and Ed == E
    assert g.info().acc_metro == acc


def test_wide_stencil_rejects_narrow_lanes():
    H, T, gH, _ = _case(CASES[0])
    gst = S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, 1), 0.6), seed=1)
    with pytest.raises(SceError, match="use more lanes"):
        S.gpu_metropolis_sweep(gst, gH, 20.0, lanes=16)
    # the chain is untouched and still sweeps at 64 lanes
    assert gst.info().sweep_index == 0
    S.gpu_metropolis_sweep(gst, gH, 20.0, lanes=64)
    assert gst.info().sweep_index == 1
