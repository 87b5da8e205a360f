"""Device parity of the learned delayed-acceptance path.

Bit-exact against the oracle: descriptor rows, the FP32 inference path, whole
two-stage sweeps (config, rows, energy, both acceptance counters).  Tolerance:
the MFMA FP32 training gradient against the FP64 numpy backprop (itself pinned
by finite differences).  Behaviour: training lowers the loss, and delayed
acceptance with the trained surrogate against the plain sweep walks the exact
dE for a fraction of the proposals at the plain acceptance and samples the
same distribution.
"""
import numpy as np
import pytest

import oracle as O
import sur_mlp as R
import sce_amd as S
from sce_amd import surrogate as SG
from fixtures import build_pair, fixture_cases, rand_config, bcc_tet_small

pytestmark = pytest.mark.gpu

DA_CASES = [c for c in fixture_cases() if c[0] in ("dimer+field", "3body", "4body", "sc-onsite", "bcc-l2-tri",
                                                    "bcc-tet", "fcc-l2")]


def _rand_params(F, seed, scale_w=1.0):
    rng = np.random.default_rng(seed)
    p = R.glorot(F, rng).astype(np.float64)
    lay = np.cumsum([0, R.H * F, R.H, R.H * R.H, R.H, R.H, F])
    p[lay[1]:lay[2]] = rng.uniform(-0.2, 0.2, R.H)
    p[lay[3]:lay[4]] = rng.uniform(-0.2, 0.2, R.H)
    p[lay[5]:lay[6]] = rng.uniform(-0.3, 0.3, F)
    return (p * scale_w).astype(np.float32)


@pytest.mark.parametrize("case", DA_CASES, ids=lambda c: c[0])
def test_descriptor_rows_bitwise(case):
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    cfg = rand_config(H.n_sites, 3)
    rng = np.random.default_rng(4)
    act = np.nonzero(H.site_active)[0] + 1
    sites = rng.choice(act, min(len(act), 200), replace=False).astype(np.int32)
    e2 = rand_config(len(sites), 5)
    A, B = SG.dev_features(gH, cfg, sites, e2)
    assert A.shape == (len(sites), gH.fmax)
    for k, s in enumerate(sites):
        ra = T.sur_features(cfg, int(s), cfg[s - 1])
        rb = T.sur_features(cfg, int(s), e2[k])
        F = len(ra)
        assert np.array_equal(A[k, :F].view(np.uint32), ra.view(np.uint32))
        assert np.array_equal(B[k, :F].view(np.uint32), rb.view(np.uint32))
        assert not A[k, F:].any() and not B[k, F:].any()


def test_inference_path_bitwise():
    na, terms, dims = bcc_tet_small(2)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    sur = SG.Surrogate(gH, seed=9)
    cfg = rand_config(H.n_sites, 6)
    sites = np.arange(1, H.n_sites + 1, dtype=np.int32)[:300]
    e2 = rand_config(len(sites), 7)
    A, B = SG.dev_features(gH, cfg, sites, e2)
    for variant in range(2):
        for a in sur.active_sublattices():
            F = sur.info(a)["F"]
            if variant == 0:
                p, sc = sur.get_params(a), sur.info(a)["scale"]   # the library's own initialisation
            else:
                p, sc = _rand_params(F, 30 + a), 0.37
                sur.set_params(a, p, sc)
            rows = np.nonzero((sites - 1) % na == a)[0]
            got = sur.eval(a, A[rows], B[rows])
            want = np.array([O.sur_delta(F, p, sc, A[r, :F], B[r, :F]) for r in rows])
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("n", [250, 1500])   # 1500 pairs: K = 3000, the split-K weight gradients
def test_training_gradient_matches_fp64_backprop(n):
    na, terms, dims = bcc_tet_small(2)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    sur = SG.Surrogate(gH, seed=3)
    a = sur.active_sublattices()[0]
    F = sur.info(a)["F"]
    p = _rand_params(F, 5)
    sur.set_params(a, p, 0.05)
    cfg = rand_config(H.n_sites, 2)
    sites = np.arange(a + 1, H.n_sites + 1, na, dtype=np.int32)
    sites = np.concatenate([sites] * (n // len(sites) + 1))[:n].astype(np.int32)
    e2 = rand_config(len(sites), 3)
    A, B = SG.dev_features(gH, cfg, sites, e2)
    y = np.random.default_rng(1).normal(0, 0.05, len(sites))
    g, loss = sur.grad(a, A, B, y)
    loss_ref, g_ref = R.loss_and_grad(p, F, A, B, y, 0.05)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    tol = 2e-4 * np.abs(g_ref).max()
    np.testing.assert_allclose(g, g_ref, rtol=2e-3, atol=tol)


@pytest.mark.parametrize("case", [c for c in DA_CASES if c[0] in ("dimer+field", "4body", "bcc-tet", "fcc-l2", "sc-onsite")],
                         ids=lambda c: c[0])
def test_da_sweeps_bitwise(case):
    name, na, terms, dims, mm, fl = case
    H, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H)
    sur = SG.Surrogate(gH, seed=4)
    cfg = rand_config(H.n_sites, 12)
    params, scales = [], []
    for a in range(na):
        F = sur.info(a)["F"]
        if F:
            p = _rand_params(F, 50 + a)
            sur.set_params(a, p, 0.04)
            params.append(p)
            scales.append(0.04)
        else:
            params.append(np.zeros(1, np.float32))
            scales.append(1.0)
    seed, step, beta = 0xda7a, 0.6, 1 / 0.05
    st = S.ChainState(H, cfg, step)
    gst = S.GPUChainState(gH, st, seed=seed)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    nacc = n1 = 0
    obs = SG.gpu_da_run_sweeps(gst, sur, None, beta, 6, renorm_interval=0, measure_interval=2)
    for sw in range(1, 7):
        a_, b_, E = T.da_sweep(c2, z2, E, beta, step, seed, sw, params, np.array(scales))
        nacc += a_
        n1 += b_
    cfg_d, z_d, E_d = gst.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E
    inf = gst.info()
    assert (inf.acc_metro, inf.att_metro, inf.acc_stage1) == (nacc, 6 * H.n_active, n1)
    assert inf.sweep_index == 6
    assert obs.shape == (3, 4) and obs[-1, 0] == E_d
    assert 0 < n1 < 6 * H.n_active


def test_harvest_is_the_plain_chain():
    na, terms, dims = bcc_tet_small(2)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    sur = SG.Surrogate(gH, seed=1)
    st = S.ChainState(H, rand_config(H.n_sites, 13), 0.6)
    gst = S.GPUChainState(gH, st, seed=21)
    beta = 1 / 0.05
    rows = sur.harvest(gst, beta, 4)
    assert rows == 4 * H.n_active
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    for sw in range(1, 5):
        _, E, _, _ = T.keyed_sweep(c2, z2, E, beta, 0.6, 21, sw, 64)
    cfg_d, z_d, E_d = gst.download()
    assert np.array_equal(cfg_d, c2) and np.array_equal(z_d, z2) and E_d == E


def test_trained_delayed_acceptance_against_the_plain_sweep():
    """What the trained network gains over the method without it: the same
    chain, from the same state, sampled by the plain sweep and by delayed
    acceptance with the trained surrogate.  The DA chain walks the exact dE
    only for the proposals its surrogate passes (a fraction well below one),
    keeps the plain acceptance (stage 2 rarely overturns stage 1), and samples
    the same distribution (<E> and <|m|> within 3 sigma).  An untrained network
    of the same shape, as a control, wastes far more exact walks on moves that
    stage 2 then rejects."""
    # isotropic tensors: the site energy is a function of the rotation-invariant
    # descriptor, so a trained surrogate can follow the exact dE closely
    # kT = 0.04 eV: above the random-sign couplings' freezing (tau_int(E) of a few
    # sweeps), where two chains from one state sample the same equilibrium
    na, terms, dims = bcc_tet_small(2, dims=(6, 6, 6), isotropic=True)
    H, _ = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    kT = 0.04
    beta = 1 / kT
    st = S.ChainState(H, rand_config(H.n_sites, 1), 0.5)
    gst = S.GPUChainState(gH, st, seed=5)
    S.gpu_run_sweeps(gst, gH, None, beta, 2000, renorm_interval=100)   # equilibrate
    cfg0, _, _ = gst.download()

    trained = SG.Surrogate(gH, seed=2)
    untrained = SG.Surrogate(gH, seed=2)
    trained.harvest(gst, beta, 400)
    losses = trained.train_all(epochs=6, batch=1024, lr=1e-3)
    for a, l in losses.items():
        assert l[-1] < 0.5 * l[0] and l[-1] < 0.01, (a, l)   # loss in units of the rms dE^2
        untrained.set_params(a, untrained.get_params(a), trained.info(a)["scale"])

    n = 8000
    p = S.GPUChainState(gH, S.ChainState(H, cfg0, 0.5), seed=77)
    obs_p = S.gpu_run_sweeps(p, gH, None, beta, n, renorm_interval=1000, measure_interval=1)
    ip = p.info()
    plain_acc = ip.acc_metro / ip.att_metro
    res = {}
    for name, sur in (("trained", trained), ("untrained", untrained)):
        g = S.GPUChainState(gH, S.ChainState(H, cfg0, 0.5), seed=77)
        obs = SG.gpu_da_run_sweeps(g, sur, None, beta, n, renorm_interval=1000, measure_interval=1)
        i = g.info()
        res[name] = (i.acc_metro / i.att_metro, i.acc_stage1 / i.att_metro, i.acc_metro / max(i.acc_stage1, 1), obs)
    acc_t, walks_t, s2_t, obs_t = res["trained"]
    acc_u, walks_u, s2_u, _ = res["untrained"]
    print("plain acceptance", plain_acc, "trained (acc, exact-walk fraction, stage 2)", acc_t, walks_t, s2_t,
          "untrained", acc_u, walks_u, s2_u)
    # exact walks only for passed proposals, nearly all of which stage 2 keeps
    assert walks_t < 0.9 and s2_t > 0.9 and acc_t > 0.9 * plain_acc
    # an untrained surrogate spends its exact walks on moves stage 2 rejects
    assert s2_u < s2_t and acc_u < acc_t
    # the same distribution as the plain sweep
    sd = S.standard_stats(obs_t, kT, H.n_active)
    sp = S.standard_stats(obs_p, kT, H.n_active)
    for k in ("energy", "absm"):
        sig = np.hypot(sd[k].err, sp[k].err)
        assert abs(sd[k].mean - sp[k].mean) < 3 * sig, (k, sd[k].mean, sp[k].mean, sig)


def test_da_direct_launches_equal_graph_replay():
    na, terms, dims = bcc_tet_small(2)
    H, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    sur = SG.Surrogate(gH, seed=4)
    st = S.ChainState(H, rand_config(H.n_sites, 12), 0.6)
    a = S.GPUChainState(gH, st, seed=9)
    b = S.GPUChainState(gH, st, seed=9)
    b.set_graph(False)
    oa = SG.gpu_da_run_sweeps(a, sur, None, 20.0, 7, renorm_interval=3, measure_interval=2)
    ob = SG.gpu_da_run_sweeps(b, sur, None, 20.0, 7, renorm_interval=3, measure_interval=2)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    assert np.array_equal(oa, ob)
    assert a.info().acc_stage1 == b.info().acc_stage1
