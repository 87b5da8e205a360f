"""Delayed acceptance with stage 1 on the int8 networks (Surrogate.set_int8):
the quantised network against its restatement, whole two-stage sweeps bit for
bit against the oracle (ora_da_sweep_q), graph replay against direct launches,
the benchmark's own size, and what the trained int8 screen keeps of the plain
chain's statistics."""
import numpy as np
import pytest

import sce_amd as S
from sce_amd import models as M
from sce_amd import surrogate as SG
from fixtures import build_pair, fixture_cases, rand_config, bcc_tet_small
from test_gpu_da import DA_CASES, _rand_params
from test_oracle_da import quantise_np

pytestmark = pytest.mark.gpu

H = SG.H


def _blob_parts(blob, F):
    o = np.cumsum([0, H * F, H, H, H * H, H, H, H, F, F])
    return [blob[o[k]:o[k + 1]] for k in range(9)]


def _set_random(sur, na, seed, scale=0.04):
    qps, scales = [], []
    for a in range(na):
        F = sur.info(a)["F"]
        if F:
            sur.set_params(a, _rand_params(F, seed + a), scale)
            qps.append(sur.get_qparams(a))
            scales.append(scale)
        else:
            qps.append(np.zeros(1, np.float32))
            scales.append(1.0)
    return qps, np.array(scales)


def test_int8_network_is_the_restated_quantisation():
    """get_qparams == the quantisation restated in numpy, for the library's
    initialisation and for set parameters; biases and head are the FP32 ones;
    input scales are 127 or 42 and keep every descriptor feature in int8."""
    na, terms, dims = bcc_tet_small(2)
    H_, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H_)
    sur = SG.Surrogate(gH, seed=9)
    cfg = rand_config(H_.n_sites, 6)
    sites = np.arange(1, H_.n_sites + 1, dtype=np.int32)
    A, B = SG.dev_features(gH, cfg, sites, rand_config(len(sites), 7))
    for variant in range(2):
        for a in sur.active_sublattices():
            F = sur.info(a)["F"]
            if variant:
                sur.set_params(a, _rand_params(F, 70 + a), 0.2)
            p = sur.get_params(a)
            V1, r1, b1, V2, r2, b2, w3, v, qin = _blob_parts(sur.get_qparams(a), F)
            assert set(np.unique(qin)) <= {np.float32(127.0), np.float32(42.0)}
            W1, r1n, W2, r2n = quantise_np(p, F, qin)
            assert np.array_equal(V1, W1.ravel()) and np.array_equal(r1.view(np.uint32), r1n.view(np.uint32))
            assert np.array_equal(V2, W2.ravel()) and np.array_equal(r2.view(np.uint32), r2n.view(np.uint32))
            lay = np.cumsum([0, H * F, H, H * H, H, H, F])
            assert np.array_equal(b1, p[lay[1]:lay[2]]) and np.array_equal(b2, p[lay[3]:lay[4]])
            assert np.array_equal(w3, p[lay[4]:lay[5]]) and np.array_equal(v, p[lay[5]:lay[6]])
            rows = np.nonzero((sites - 1) % na == a)[0]
            for X in (A[rows, :F], B[rows, :F]):
                assert np.abs(X * qin[None, :]).max() <= 127.5


def test_int8_inference_bitwise():
    """The device's int8 surrogate differences on descriptor rows against the
    oracle's (ora_sur_delta_q), for the library's initialisation and for set
    parameters, on every active sublattice."""
    import oracle as O
    na, terms, dims = bcc_tet_small(2)
    H_, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H_)
    sur = SG.Surrogate(gH, seed=9)
    sur.set_int8(True)
    cfg = rand_config(H_.n_sites, 6)
    sites = np.arange(1, H_.n_sites + 1, dtype=np.int32)[:300]
    A, B = SG.dev_features(gH, cfg, sites, rand_config(len(sites), 7))
    for variant in range(2):
        for a in sur.active_sublattices():
            F = sur.info(a)["F"]
            sc = sur.info(a)["scale"]
            if variant:
                sc = 0.37
                sur.set_params(a, _rand_params(F, 30 + a), sc)
            qp = sur.get_qparams(a)
            rows = np.nonzero((sites - 1) % na == a)[0]
            got = sur.eval(a, A[rows], B[rows])
            want = np.array([O.sur_delta_q(F, qp, sc, A[r, :F], B[r, :F]) for r in rows])
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (variant, a, got[:4], want[:4])


@pytest.mark.parametrize("case", [c for c in DA_CASES if c[0] in ("dimer+field", "4body", "bcc-tet", "fcc-l2", "sc-onsite")],
                         ids=lambda c: c[0])
def test_int8_da_sweeps_bitwise(case):
    name, na, terms, dims, mm, fl = case
    H_, T = build_pair(na, terms, dims, mm, fl)
    gH = S.GPUTiledHamiltonian(H_)
    sur = SG.Surrogate(gH, seed=4)
    sur.set_int8(True)
    qps, scales = _set_random(sur, na, 50)
    seed, step, beta = 0xda7a, 0.6, 1 / 0.05
    st = S.ChainState(H_, rand_config(H_.n_sites, 12), step)
    gst = S.GPUChainState(gH, st, seed=seed)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    nacc = n1 = 0
    obs = SG.gpu_da_run_sweeps(gst, sur, None, beta, 6, renorm_interval=0, measure_interval=2)
    for sw in range(1, 7):
        a_, b_, E = T.da_sweep(c2, z2, E, beta, step, seed, sw, qps, scales, int8=True)
        nacc += a_
        n1 += b_
    cfg_d, z_d, E_d = gst.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E
    inf = gst.info()
    assert (inf.acc_metro, inf.att_metro, inf.acc_stage1) == (nacc, 6 * H_.n_active, n1)
    assert obs.shape == (3, 4) and obs[-1, 0] == E_d
    assert 0 < n1 < 6 * H_.n_active


def test_int8_da_graph_equals_direct_and_mode_switch():
    """Graph-replayed and direct int8 DA sweeps are the same bits, and switching
    the mode between calls runs each call on its own networks."""
    na, terms, dims = bcc_tet_small(2)
    H_, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H_)
    sur = SG.Surrogate(gH, seed=4)
    qps, scales = _set_random(sur, na, 80)
    params = [sur.get_params(a) if sur.info(a)["F"] else np.zeros(1, np.float32) for a in range(na)]
    st = S.ChainState(H_, rand_config(H_.n_sites, 12), 0.6)
    a = S.GPUChainState(gH, st, seed=9)
    b = S.GPUChainState(gH, st, seed=9)
    b.set_graph(False)
    sur.set_int8(True)
    oa = SG.gpu_da_run_sweeps(a, sur, None, 20.0, 7, renorm_interval=3, measure_interval=2)
    ob = SG.gpu_da_run_sweeps(b, sur, None, 20.0, 7, renorm_interval=3, measure_interval=2)
    assert np.array_equal(oa, ob)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    assert np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb
    # then two FP32 sweeps, then two int8 sweeps: against the oracle composition
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    for sw in range(1, 8):
        _, _, E = T.da_sweep(c2, z2, E, 20.0, 0.6, 9, sw, qps, scales, int8=True)
        if sw % 3 == 0:
            _, E = T.renormalize(c2, z2, E)
    assert np.array_equal(ca, c2) and Ea == E
    sur.set_int8(False)
    SG.gpu_da_run_sweeps(a, sur, None, 20.0, 2, renorm_interval=0)
    sur.set_int8(True)
    SG.gpu_da_run_sweeps(a, sur, None, 20.0, 2, renorm_interval=0)
    for sw in (8, 9):
        _, _, E = T.da_sweep(c2, z2, E, 20.0, 0.6, 9, sw, params, scales)
    for sw in (10, 11):
        _, _, E = T.da_sweep(c2, z2, E, 20.0, 0.6, 9, sw, qps, scales, int8=True)
    ca, za, Ea = a.download()
    assert np.array_equal(ca, c2) and np.array_equal(za, z2) and Ea == E


def test_benchmark_size_int8_da_sweep_bitwise():
    na, terms, dims = M.config_model("bcc32", seed=0)
    H_, T = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H_)
    sur = SG.Surrogate(gH, seed=3)
    sur.set_int8(True)
    qps, scales = _set_random(sur, na, 5, scale=0.05)
    st = S.ChainState(H_, rand_config(H_.n_sites, 11), 1.3)
    seed, beta = 0xda7a, 1 / 0.04
    g = S.GPUChainState(gH, st, seed=seed)
    SG.gpu_da_run_sweeps(g, sur, None, beta, 1, renorm_interval=0)
    c2, z2, E = st.config.copy(), st.zrows.copy(), st.energy
    nacc, n1, E = T.da_sweep(c2, z2, E, beta, 1.3, seed, 1, qps, scales, int8=True)
    cfg_d, z_d, E_d = g.download()
    assert np.array_equal(cfg_d.view(np.uint64), c2.view(np.uint64))
    assert np.array_equal(z_d.view(np.uint64), z2.view(np.uint64))
    assert E_d == E
    inf = g.info()
    assert (inf.acc_metro, inf.acc_stage1) == (nacc, n1)


def test_trained_int8_screen_against_the_plain_sweep():
    """A trained surrogate's int8 screen: most exact walks skipped, stage 2
    keeps nearly all it passes, the plain acceptance kept, and <E>, <|m|>
    within 3 sigma of the plain chain from the same state."""
    na, terms, dims = bcc_tet_small(2, dims=(6, 6, 6), isotropic=True)
    H_, _ = build_pair(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H_)
    kT = 0.04
    beta = 1 / kT
    gst = S.GPUChainState(gH, S.ChainState(H_, rand_config(H_.n_sites, 1), 0.5), seed=5)
    S.gpu_run_sweeps(gst, gH, None, beta, 2000, renorm_interval=100)
    cfg0, _, _ = gst.download()
    sur = SG.Surrogate(gH, seed=2)
    sur.harvest(gst, beta, 400)
    sur.train_all(epochs=6, batch=1024, lr=1e-3)
    sur.set_int8(True)
    n = 8000
    p = S.GPUChainState(gH, S.ChainState(H_, cfg0, 0.5), seed=77)
    obs_p = S.gpu_run_sweeps(p, gH, None, beta, n, renorm_interval=1000, measure_interval=1)
    ip = p.info()
    g = S.GPUChainState(gH, S.ChainState(H_, cfg0, 0.5), seed=77)
    obs = SG.gpu_da_run_sweeps(g, sur, None, beta, n, renorm_interval=1000, measure_interval=1)
    i = g.info()
    acc, walks, s2 = i.acc_metro / i.att_metro, i.acc_stage1 / i.att_metro, i.acc_metro / max(i.acc_stage1, 1)
    print("plain acceptance", ip.acc_metro / ip.att_metro, "int8 (acc, exact-walk fraction, stage 2)", acc, walks, s2)
    assert walks < 0.9 and s2 > 0.85 and acc > 0.85 * ip.acc_metro / ip.att_metro
    sd = S.standard_stats(obs, kT, H_.n_active)
    sp = S.standard_stats(obs_p, kT, H_.n_active)
    for k in ("energy", "absm"):
        sig = np.hypot(sd[k].err, sp[k].err)
        assert abs(sd[k].mean - sp[k].mean) < 3 * sig, (k, sd[k].mean, sp[k].mean, sig)
