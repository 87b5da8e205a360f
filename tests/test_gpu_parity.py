"""Device parity: every piece of device arithmetic against the CPU oracle.

Bit-exact: Philox blocks, deterministic exp/log/sincos, Z rows, keyed
proposals, the walk ΔE / coefficients, the tree reduction, whole sweeps
(config, rows, energy, counters), overrelaxation sweeps, renormalisation.
Tolerance: walk ΔE vs the reference's site_coeffs! + delta_energy (1e-12
scaled, test_gpu.jl:135-178).
"""
import numpy as np
import pytest

import oracle as O
import sce_amd as S
from sce_amd import gpu as G
from sce_amd._lib import SceError
from fixtures import build_pair, fixture_cases, rand_config, bcc_small, bcc_tet_small

pytestmark = pytest.mark.gpu

LANES = (16, 32, 64)


def _dirs(n, seed):
    rng = np.random.default_rng(seed)
    d = [(0, 0, 1.0), (0, 0, -1.0), (1, 0, 0.0), (-1, 0, 0.0), (0, 1, 0.0), (0, -1, 0.0)]
    d += [(np.cos(k * np.pi / 6), np.sin(k * np.pi / 6), 0.0) for k in range(12)]
    d = np.array(d, float)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return np.vstack([d, v])


def test_philox_kat_and_streams_on_device():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    seeds = [k[0] | (k[1] << 32) for _, k, _ in kat]
    out = G.dev_philox(seeds, [c for c, _, _ in kat])
    for row, (_, _, want) in zip(out, kat):
        assert tuple(int(x) for x in row) == want
    rng = np.random.default_rng(1)
    seeds = rng.integers(0, 2**63, 500, dtype=np.uint64)
    ctrs = rng.integers(0, 2**32, (500, 4), dtype=np.uint64).astype(np.uint32)
    dev = G.dev_philox(seeds, ctrs)
    for s, c, d in zip(seeds, ctrs, dev):
        assert O.philox4x32(c, (int(s) & 0xffffffff, int(s) >> 32)) == tuple(int(x) for x in d)


@pytest.mark.parametrize("op", [0, 1, 2, 3, 4])
def test_detmath_bitwise(op):
    rng = np.random.default_rng(op)
    if op == 4:   # the surrogate's FP32 tanh, compared as single-precision bits
        x = np.concatenate([rng.normal(0, 3, 5000), rng.uniform(-0.3, 0.3, 2000), [0.0, -0.0, 0.25, -10.0, 10.5]])
        x = x.astype(np.float32).astype(np.float64)
        dev = G.dev_detmath(op, x).astype(np.float32)
        ref = np.array([O.tanhf(v) for v in x], np.float32)
        assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))
        return
    if op == 0:
        x = np.concatenate([-rng.exponential(5, 5000), rng.uniform(-700, 700, 500), [0.0, -745.0, 709.0]])
        ref = [O.exp(v) for v in x]
    elif op == 1:
        x = np.concatenate([rng.uniform(0, 1, 5000), [2.0**-53, 1 - 2.0**-53, 5e-324, 1.0, 3.5e7]])
        ref = [O.log(v) for v in x]
    else:
        x = np.concatenate([rng.uniform(0, 2 * np.pi, 3000), rng.normal(0, 5, 3000), [0.0, np.pi / 4]])
        ref = [O.sincos(v)[op - 2] for v in x]
    dev = G.dev_detmath(op, x)
    assert np.array_equal(dev.view(np.uint64), np.asarray(ref).view(np.uint64))


@pytest.mark.parametrize("lmax", range(7))
def test_zlm_row_bitwise(lmax):
    d = _dirs(2000, 2026 + lmax)
    dev = G.dev_zlm(d, lmax)
    ref = np.array([O.zlm_row(u, lmax) for u in d])
    assert np.array_equal(dev.view(np.uint64), ref.view(np.uint64))


def test_keyed_proposal_bitwise():
    rng = np.random.default_rng(5)
    n = 3000
    sites = rng.integers(1, 70000, n).astype(np.int32)
    sweeps = rng.integers(1, 10000, n).astype(np.int32)
    e = _dirs(n - 18, 3)
    seed = 0xbeef1234cafe
    e2, u = G.dev_proposal(seed, sites, sweeps, e, 0.6)
    nflip = 0
    for i in range(n):
        r2, ru = O.keyed_proposal(seed, int(sites[i]), int(sweeps[i]), e[i], 0.6)
        assert np.array_equal(r2.view(np.uint64), e2[i].view(np.uint64)) and ru == u[i]
        nflip += int(np.array_equal(r2, -e[i]))
    assert 0.1 * n < nflip < 0.3 * n


@pytest.mark.parametrize("scale", [1e-300, 1e-310])
def test_overrelaxation_row_sums_in_the_subnormal_range(scale):
    """The overrelaxation walk's row sums (leave-one-out coefficients) where
    the terms and partial sums are tiny (1e-300) or subnormal (1e-310): the
    device's row sums, formed by LDS adds into lane-private slots, equal the
    oracle's register sums bit for bit at every lane count."""
    import dataclasses
    na, terms, dims = bcc_tet_small(5)
    terms = [dataclasses.replace(t, coef=t.coef * scale) for t in terms]
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    z = H.zrows(rand_config(H.n_sites, 41))
    sites = np.array(T.color_sites[: min(64, T.n_active)], np.int32)
    for lanes in LANES:
        cdev = G.dev_walk_coeffs(gH, z, sites, lanes)
        for i, s in enumerate(sites):
            cref = T.walk_coeffs(int(s), z, lanes)
            assert np.array_equal(cdev[i].view(np.uint64), cref.view(np.uint64)), (scale, lanes, s)
        if scale < 2.3e-308:
            nz = cdev[cdev != 0.0]
            assert nz.size and np.abs(nz).max() < 2.3e-308  # subnormal sums were formed


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("lanes", [128, 256])
def test_walk_delta_at_wide_lanes(case, lanes):
    """Two and four wavefronts per site: the walk's ΔE at 128 and 256 lanes
    (entries dealt round-robin over the lanes, the fold's +128 and +64 steps
    across the wavefronts) equals the oracle's contract at that width bit for
    bit."""
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    z = H.zrows(rand_config(H.n_sites, 33))
    rng = np.random.default_rng(9)
    sites = np.array(T.color_sites[: min(64, T.n_active)], np.int32)
    znew = np.array([O.zlm_row(v, H.lmax) for v in S.random_config(len(sites), rng)])
    dev = G.dev_walk_delta(gH, z, sites, znew, lanes)
    for i, s in enumerate(sites):
        assert dev[i] == T.walk_delta(int(s), z, znew[i], lanes), (name, s)


@pytest.mark.parametrize("case", fixture_cases()[:6], ids=lambda c: c[0])
@pytest.mark.parametrize("graph, lanes", [(True, 128), (False, 128), (True, 256), (False, 256)])
def test_sweeps_at_wide_lanes_bitwise(case, graph, lanes):
    """Chained 128- and 256-lane sweeps with measurements and
    renormalisation, graph replay and direct launches, against the oracle's
    keyed sweeps at that width; overrelaxation and persistent sweeps refuse
    the width."""
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 21), 0.7)
    g = S.GPUChainState(gH, st, seed=4242)
    g.set_graph(graph)
    beta = 1 / 0.02
    obs = S.gpu_run_sweeps(g, gH, None, beta, 7, renorm_interval=3, lanes=lanes, measure_interval=2)
    c, z, E = st.config.copy(), st.zrows.copy(), st.energy
    nacc = 0
    for j in range(1, 8):
        a, E, _, _ = T.keyed_sweep(c, z, E, beta, 0.7, 4242, j, lanes)
        nacc += a
        if j % 3 == 0:
            _, E = T.renormalize(c, z, E)
    cfg, zd, Ed = g.download()
    assert np.array_equal(cfg.view(np.uint64), c.view(np.uint64)), name
    assert np.array_equal(zd.view(np.uint64), z.view(np.uint64)), name
    assert Ed == E and g.info().acc_metro == nacc and obs.shape[0] == 3
    with pytest.raises(SceError):
        S.gpu_run_sweeps(g, gH, None, beta, 1, lanes=lanes, or_per_metropolis=1)
    g.set_persistent(True)
    with pytest.raises(SceError):
        S.gpu_run_sweeps(g, gH, None, beta, 1, lanes=lanes)


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_walk_delta_and_coeffs(case):
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    cfg = rand_config(H.n_sites, 31)
    z = H.zrows(cfg)
    rng = np.random.default_rng(7)
    sites = np.array(T.color_sites[: min(64, T.n_active)], np.int32)
    znew = np.array([O.zlm_row(v, H.lmax) for v in S.random_config(len(sites), rng)])
    for lanes in LANES:
        dev = G.dev_walk_delta(gH, z, sites, znew, lanes)
        cdev = G.dev_walk_coeffs(gH, z, sites, lanes)
        for i, s in enumerate(sites):
            ref = T.walk_delta(int(s), z, znew[i], lanes)
            assert dev[i] == ref, (name, lanes, s)
            cref = T.walk_coeffs(int(s), z, lanes)
            assert np.array_equal(cdev[i], cref)
            c = T.site_coeffs(int(s), z)
            dE_cpu = O.delta_energy(c, z[s - 1], znew[i])
            scale = np.abs(c).sum() * (np.abs(znew[i]).sum() + np.abs(z[s - 1]).sum())
            assert abs(dev[i] - dE_cpu) <= 1e-12 * max(scale, 1e-30)
            assert np.all(np.abs(cdev[i] - c) <= 1e-12 * max(np.abs(c).sum(), 1e-30))


@pytest.mark.parametrize("n", [1, 7, 511, 512, 513, 100_000, 300_000])
def test_reduce_tree_bitwise(n):
    x = np.random.default_rng(n).standard_normal(n) * 10.0 ** np.random.default_rng(n + 1).integers(-8, 3, n)
    assert G.dev_reduce(x) == O.reduce_tree(x)


def _sweeps_case(case, lanes, nsweeps=5, beta=1 / 0.05, seed=0xc0ffee, step=0.6):
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    cfg = rand_config(H.n_sites, 7)
    st = S.ChainState(H, cfg, step)
    gst = S.GPUChainState(gH, st, seed=seed)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    naccs, naccs_ref = [], []
    for sw in range(1, nsweeps + 1):
        naccs.append(S.gpu_metropolis_sweep(gst, gH, beta, lanes=lanes))
        n, E, _, _ = T.keyed_sweep(c2, z2, E, beta, step, seed, sw, lanes)
        naccs_ref.append(n)
    cfg_d, z_d, E_d = gst.download()
    return (H, T, gH, gst, st), (cfg_d, z_d, E_d, naccs), (c2, z2, E, naccs_ref)


@pytest.mark.parametrize("lanes", LANES)
@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_full_sweep_bitwise(case, lanes):
    (H, T, gH, gst, st), dev, ref = _sweeps_case(case, lanes)
    cfg_d, z_d, E_d, naccs = dev
    c2, z2, E, naccs_ref = ref
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E
    assert naccs == naccs_ref
    inf = gst.info()
    assert inf.acc_metro == sum(naccs_ref) and inf.att_metro == 5 * H.n_active
    assert inf.sweep_index == 5


@pytest.mark.parametrize("case", [c for c in fixture_cases() if c[0] in ("dimer", "3body", "bcc-l2-tri", "bcc-tet", "dimer+field")],
                         ids=lambda c: c[0])
def test_overrelaxation_bitwise(case):
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 9), 0.6)
    seed, beta, lanes = 77, 1 / 0.05, 32
    gst = S.GPUChainState(gH, st, seed=seed)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    acc_or = att_or = 0
    for sw in range(1, 4):
        S.gpu_metropolis_sweep(gst, gH, beta, lanes=lanes)
        _, E, _, _ = T.keyed_sweep(c2, z2, E, beta, 0.6, seed, sw, lanes)
        for p in range(2):
            S.gpu_overrelaxation_sweep(gst, gH, beta, or_pass=p, lanes=lanes)
            a, n, E = T.keyed_or_sweep(c2, z2, E, beta, seed, sw, p, lanes)
            acc_or += a
            att_or += n
    cfg_d, z_d, E_d = gst.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E
    inf = gst.info()
    assert (inf.acc_or, inf.att_or) == (acc_or, att_or)


@pytest.mark.parametrize("case", [c for c in fixture_cases() if c[0] in ("dimer", "3body", "bcc-l2-tri", "bcc-tet", "dimer+field", "sc-onsite")],
                         ids=lambda c: c[0])
def test_compound_sweeps_graph_and_chained_bitwise(case):
    """Compound sweeps (run.jl:143-150: one Metropolis sweep then two
    overrelaxation passes) through sce_run_sweeps — graph-replayed and direct,
    the next sweep's proposals made in the last pass's reduction, the
    measurement riding on it — equal the oracle's composition sweep by sweep,
    measurement rows included, across renormalisations."""
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 19), 0.6)
    seed, beta, lanes, nsw = 505, 1 / 0.05, 64, 11
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    rows = []
    acc_or = att_or = 0
    act = H.site_active.astype(bool)
    for sw in range(1, nsw + 1):
        _, E, _, _ = T.keyed_sweep(c2, z2, E, beta, 0.6, seed, sw, lanes)
        for p in range(2):
            a, n, E = T.keyed_or_sweep(c2, z2, E, beta, seed, sw, p, lanes)
            acc_or += a
            att_or += n
        if sw % 5 == 0:
            _, E = T.renormalize(c2, z2, E)
        if sw % 2 == 0:
            rows.append([E, *np.mean(c2[act], axis=0)])
    rows = np.array(rows)
    outs = []
    for graph in (True, False):
        g = S.GPUChainState(gH, st, seed=seed)
        g.set_graph(graph)
        obs = S.gpu_run_sweeps(g, gH, None, beta, nsw, renorm_interval=5, lanes=lanes, or_per_metropolis=2,
                               measure_interval=2)
        cfg_d, z_d, E_d = g.download()
        assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64)), graph
        assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64)), graph
        assert E_d == E, graph
        assert np.array_equal(obs[:, 0], rows[:, 0]) and np.allclose(obs[:, 1:], rows[:, 1:], rtol=0, atol=1e-15)
        inf = g.info()
        assert (inf.acc_or, inf.att_or, inf.sweep_index) == (acc_or, att_or, nsw)
        outs.append(obs)
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))


def test_renormalize_and_total_energy_bitwise():
    na, terms, dims = bcc_small(3)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 4), 0.6)
    gst = S.GPUChainState(gH, st, seed=5)
    S.gpu_run_sweeps(gst, gH, None, 1 / 0.05, 30, renorm_interval=0)
    cfg, z, E = gst.download()
    assert gst.total_energy() == T.total_energy_tree(z)
    drift = S.gpu_renormalize(gst)
    c2, z2 = cfg.copy(), z.copy()
    d_ref, E_ref = T.renormalize(c2, z2, E)
    cfg_d, z_d, E_d = gst.download()
    assert np.array_equal(cfg_d, c2) and np.array_equal(z_d, z2)
    assert E_d == E_ref and drift == d_ref
    assert abs(E_ref - T.total_energy(z2)) <= 1e-12 * max(1.0, abs(E_ref))


def test_run_sweeps_matches_single_sweeps():
    na, terms, dims = bcc_small(4)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 8), 0.6)
    a = S.GPUChainState(gH, st, seed=11)
    b = S.GPUChainState(gH, st, seed=11)
    obs = S.gpu_run_sweeps(a, gH, None, 20.0, 12, renorm_interval=5, measure_interval=3)
    assert obs.shape == (4, 4)
    for i in range(1, 13):
        S.gpu_metropolis_sweep(b, gH, 20.0)
        if i % 5 == 0:
            S.gpu_renormalize(b)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    m = np.mean(ca, axis=0)
    assert np.allclose(obs[-1, 1:], m, atol=1e-14) and obs[-1, 0] == Ea


@pytest.mark.parametrize("lanes", LANES)
@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_chained_sweeps_equal_single_sweeps(case, lanes):
    """Sweeps chained within one call (the next sweep's proposals made during
    the walks, the final reduction deferred into the next sweep's
    smallest-class launch, every colour count and class-size order of the
    fixture models) equal the same sweeps made one call each, which are
    bitwise the oracle's; graph replay and direct launches alike, with the
    measured energies of the deferred finals in the right rows."""
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 21), 0.6)
    ref = S.GPUChainState(gH, st, seed=404)
    energies = []
    for _ in range(7):
        S.gpu_metropolis_sweep(ref, gH, 20.0, lanes=lanes)
        energies.append(ref.download()[2])
    cr, zr, Er = ref.download()
    ir = ref.info()
    for graph in (True, False):
        g = S.GPUChainState(gH, st, seed=404)
        g.set_graph(graph)
        obs = S.gpu_run_sweeps(g, gH, None, 20.0, 7, renorm_interval=0, measure_interval=2, lanes=lanes)
        c, z, E = g.download()
        assert np.array_equal(c.view(np.uint64), cr.view(np.uint64)), graph
        assert np.array_equal(z.view(np.uint64), zr.view(np.uint64)), graph
        assert E == Er
        i = g.info()
        assert (i.sweep_index, i.acc_metro, i.att_metro) == (ir.sweep_index, ir.acc_metro, ir.att_metro)
        assert obs.shape == (3, 4) and list(obs[:, 0]) == [energies[1], energies[3], energies[5]]


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_persistent_equals_graph_path_and_oracle(case):
    """The dependency-driven persistent sweep (each site waits only on its
    neighbours' sweep counters) against per-colour graph launches and the
    oracle: the same bits for config, rows, energy, counters and every
    measurement row, across renormalisations and measurement schedules."""
    lanes = 64
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 12), 0.6)
    a = S.GPUChainState(gH, st, seed=99)
    b = S.GPUChainState(gH, st, seed=99)
    a.set_persistent(True)
    beta = 1 / 0.05
    oa = S.gpu_run_sweeps(a, gH, None, beta, 23, renorm_interval=10, lanes=lanes, measure_interval=2)
    ob = S.gpu_run_sweeps(b, gH, None, beta, 23, renorm_interval=10, lanes=lanes, measure_interval=2)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    assert np.array_equal(oa.view(np.uint64), ob.view(np.uint64))
    # oracle: the same sweeps with device-contract renormalisations
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    for sw in range(1, 24):
        _, E, _, _ = T.keyed_sweep(c2, z2, E, beta, 0.6, 99, sw, lanes)
        if sw % 10 == 0:
            _, E = T.renormalize(c2, z2, E)
    assert np.array_equal(ca, c2) and np.array_equal(za, z2) and Ea == E
    ia, ib = a.info(), b.info()
    assert (ia.acc_metro, ia.att_metro, ia.sweep_index) == (ib.acc_metro, ib.att_metro, ib.sweep_index)


def test_persistent_long_segments_and_lane_guard():
    """Calls longer than one segment buffer (128 sweeps), every sweep measured,
    no renormalisation inside: persistent == graph path, row by row; the
    dependency-driven walk takes 64 lanes per site and says so otherwise."""
    na, terms, dims = bcc_small(13)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 5), 0.6)
    a = S.GPUChainState(gH, st, seed=7)
    b = S.GPUChainState(gH, st, seed=7)
    a.set_persistent(True)
    for n in (300, 17):
        oa = S.gpu_run_sweeps(a, gH, None, 30.0, n, renorm_interval=0, measure_interval=1)
        ob = S.gpu_run_sweeps(b, gH, None, 30.0, n, renorm_interval=0, measure_interval=1)
        assert np.array_equal(oa.view(np.uint64), ob.view(np.uint64)), n
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    with pytest.raises(SceError, match="64 lanes"):
        S.gpu_run_sweeps(a, gH, None, 30.0, 3, renorm_interval=0, lanes=32)


def test_resume_from_downloaded_state_and_counters_is_bitwise():
    """A chain rebuilt from to_host! output plus its counters continues the same
    keyed stream (the lane-owned state of gpu_state.jl:19-47)."""
    na, terms, dims = bcc_small(6)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 3), 0.6)
    a = S.GPUChainState(gH, st, seed=4242)
    S.gpu_run_sweeps(a, gH, None, 20.0, 10, renorm_interval=0)
    cfg, z, E = a.download()
    inf = a.info()
    st2 = S.ChainState(H, cfg, 0.6)
    st2.zrows[...] = z
    st2.energy = E
    b = S.GPUChainState(gH, st2, seed=4242)
    b.set_counters(inf.sweep_index, inf.acc_metro, inf.att_metro)
    S.gpu_run_sweeps(a, gH, None, 20.0, 7, renorm_interval=0)
    S.gpu_run_sweeps(b, gH, None, 20.0, 7, renorm_interval=0)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    ia, ib = a.info(), b.info()
    assert (ia.sweep_index, ia.acc_metro, ia.att_metro) == (ib.sweep_index, ib.acc_metro, ib.att_metro)
    with pytest.raises(SceError):
        b.set_counters(2**31, 0, 0)


def test_walk_timing_leaves_the_chain_unchanged():
    na, terms, dims = bcc_small(7)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 5), 0.6)
    a = S.GPUChainState(gH, st, seed=99)
    b = S.GPUChainState(gH, st, seed=99)
    S.gpu_run_sweeps(a, gH, None, 20.0, 3, renorm_interval=0)
    S.gpu_run_sweeps(b, gH, None, 20.0, 3, renorm_interval=0)
    ms, n = a.time_walk(64, 3)
    assert ms > 0 and n == 3 * sum(1 for k in range(H.n_colors) if H.color_ptr[k + 1] > H.color_ptr[k])
    S.gpu_run_sweeps(a, gH, None, 20.0, 4, renorm_interval=0)
    S.gpu_run_sweeps(b, gH, None, 20.0, 4, renorm_interval=0)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    assert (a.info().acc_metro, a.info().sweep_index) == (b.info().acc_metro, b.info().sweep_index)


def test_direct_launches_equal_graph_replay():
    na, terms, dims = bcc_small(8)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 6), 0.6)
    a = S.GPUChainState(gH, st, seed=5)
    b = S.GPUChainState(gH, st, seed=5)
    b.set_graph(False)
    oa = S.gpu_run_sweeps(a, gH, None, 20.0, 9, renorm_interval=4, measure_interval=2)
    ob = S.gpu_run_sweeps(b, gH, None, 20.0, 9, renorm_interval=4, measure_interval=2)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    assert np.array_equal(oa, ob)


def test_checkpoint_resume_is_bitwise(tmp_path):
    from sce_amd import checkpoint as CK
    na, terms, dims = bcc_small(9)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    st = S.ChainState(H, rand_config(H.n_sites, 2), 0.7)
    a = S.GPUChainState(gH, st, seed=2718)
    S.gpu_run_sweeps(a, gH, None, 20.0, 11, renorm_interval=5)
    path = str(tmp_path / "chain.npz")
    CK.save_chain(path, a)
    b = CK.load_chain(path, gH)
    S.gpu_run_sweeps(a, gH, None, 20.0, 9, renorm_interval=5)
    S.gpu_run_sweeps(b, gH, None, 20.0, 9, renorm_interval=5)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    ia, ib = a.info(), b.info()
    assert (ia.sweep_index, ia.acc_metro, ia.att_metro, ia.step) == (ib.sweep_index, ib.acc_metro, ib.att_metro, ib.step)
    assert ia.max_drift == ib.max_drift             # the drift record travels too
    c = CK.load_chain(path, gH)
    assert c.info().max_drift == np.load(path)["max_drift"] > 0.0
    other = S.GPUTiledHamiltonian(S.TiledHamiltonian(*bcc_small(10)))
    with pytest.raises(ValueError, match="different model"):
        CK.load_chain(path, other)


def test_swap_payload_survives_destroying_the_partner():
    """_swap_payload!(a, b) (gpu_pt.jl:46-51) moves the buffers with ownership:
    after the swap, destroying b (and allocating a new chain into the memory it
    frees) leaves a sweeping b's former payload with a's own keyed stream,
    bit-identical to the oracle."""
    import gc
    na, terms, dims = bcc_small(8)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    beta = 1.0 / 0.01
    sa = S.ChainState(H, rand_config(H.n_sites, 21), 0.6)
    sb = S.ChainState(H, rand_config(H.n_sites, 22), 0.6)
    a = S.GPUChainState(gH, sa, seed=1111)
    b = S.GPUChainState(gH, sb, seed=2222)
    S.gpu_metropolis_sweep(a, gH, beta)
    S.gpu_metropolis_sweep(b, gH, beta)
    cb, zb, Eb = sb.config.copy(), sb.zrows.copy(), sb.energy
    _, Eb, _, _ = T.keyed_sweep(cb, zb, Eb, beta, 0.6, 2222, 1, S.DEFAULT_LANES)
    S.swap_payload(a, b)
    del b
    gc.collect()
    filler = [S.GPUChainState(gH, S.ChainState(H, rand_config(H.n_sites, 30 + k), 0.6), seed=k)
              for k in range(3)]
    for f in filler:
        S.gpu_metropolis_sweep(f, gH, beta)
    for sw in range(2, 5):   # a's own counter continues: sweeps 2, 3, 4
        n_dev = S.gpu_metropolis_sweep(a, gH, beta)
        n_ref, Eb, _, _ = T.keyed_sweep(cb, zb, Eb, beta, 0.6, 1111, sw, S.DEFAULT_LANES)
        assert n_dev == n_ref
    ca, za, Ea = a.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    del filler
