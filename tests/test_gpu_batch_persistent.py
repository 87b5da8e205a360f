"""The batched persistent sweep (sce_batch_set_persistent): every chain of a
batch in one cooperative launch per renormalisation segment, with the replica
exchange decided on the device between sweeps.  Every payload, energy,
counter, measurement row and exchange statistic is bit-identical to the
per-colour launches of the same batch at 64 lanes per site, which the other
batch tests pin to the oracle composition of keyed sweeps and
_swap_accepts (src/pt.jl:92-96) on the schedule of src/gpu/gpu_pt.jl:115-140."""
import numpy as np
import pytest

import sce_amd as S
from sce_amd import models as M
from sce_amd.batch import GPUBatch
from sce_amd._lib import SceError
from fixtures import bcc_tet_small, fixture_cases, rand_config

pytestmark = pytest.mark.gpu


def _snapshot(b):
    out = []
    for k in range(b.R):
        cfg, z, E = b.download(k)
        inf = b.info(k)
        out.append((cfg, z, E, inf.sweep_index, inf.acc_metro, inf.att_metro, inf.max_drift))
    return out, b.swap_stats(), b.exchange_state()


def _assert_same(a, b):
    (ca, sa, ea), (cb, sb, eb) = a, b
    for k, (x, y) in enumerate(zip(ca, cb)):
        assert np.array_equal(x[0], y[0]), k
        assert np.array_equal(x[1], y[1]), k
        assert x[2] == y[2], k
        assert x[3:] == y[3:], k
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])
    assert ea == eb


def _both(make, calls):
    """Run the same calls (kwargs of GPUBatch.run) on a per-colour and on a
    persistent batch; returns (rows, snapshot) of each."""
    res = []
    for pers in (False, True):
        b = make()
        b.set_persistent(pers)
        rows = [b.run(**kw) for kw in calls]
        res.append((rows, _snapshot(b)))
    return res


def _ladder(H, R, kts, step=0.6, seed0=11):
    gH = S.GPUTiledHamiltonian(H)
    seeds = [seed0 + 7 * r for r in range(R)]

    def make():
        sts = [S.ChainState(H, rand_config(H.n_sites, seed0 + r), step) for r in range(R)]
        return GPUBatch([gH] * R, sts, seeds, kts, seed0 * 31 + 5)
    return make


@pytest.mark.parametrize("nsw, renorm, meas, exch, lanes", [
    (12, 5, 1, 1, 64),     # segments 5, 5, 2; the renormalised sweeps exchange after renormalising
    (12, 5, 1, 1, 32),
    (12, 5, 1, 1, 16),
    (20, 7, 3, 2, 64),
    (140, 0, 2, 3, 64),    # a 128-sweep segment (both words of the exchange mask), then 12
    (140, 0, 2, 3, 16),
    (9, 0, 0, 0, 32),      # independent chains
])
def test_ladder_persistent_equals_per_colour(nsw, renorm, meas, exch, lanes):
    H = S.TiledHamiltonian(*bcc_tet_small(6))
    R = 5
    make = _ladder(H, R, np.geomspace(0.001, 0.02, R))
    kw = dict(nsweeps=nsw, lanes=lanes, renorm_interval=renorm, measure_interval=meas, exchange_interval=exch,
              phase_done=0, phase_total=nsw)
    (r0, s0), (r1, s1) = _both(make, [kw])
    assert np.array_equal(r0[0], r1[0])
    _assert_same(s0, s1)
    if exch:
        assert s0[1][0].sum() > 0  # exchanges were attempted
    if nsw == 12:
        assert s0[1][1].sum() > 0  # and here accepted (payloads moved in-kernel)


def test_split_calls_and_switching_paths():
    """A phase split over calls, alternating persistent and per-colour calls,
    equals the phase run per-colour in one call."""
    H = S.TiledHamiltonian(*bcc_tet_small(4, dims=(3, 3, 3)))
    R = 4
    make = _ladder(H, R, np.geomspace(0.002, 0.03, R), seed0=23)
    ref = make()
    o_ref = ref.run(30, 64, 8, 2, 1, 0, 30)
    b = make()
    rows = []
    for d, n, pers in [(0, 11, True), (11, 6, False), (17, 13, True)]:
        b.set_persistent(pers)
        rows.append(b.run(n, 64, 8, 2, 1, d, 30))
    assert np.array_equal(o_ref, np.concatenate(rows, axis=1))
    _assert_same(_snapshot(ref), _snapshot(b))


@pytest.mark.parametrize("lanes", [16, 32, 64])
def test_c3_fcc_four_shell_ladder_persistent(lanes):
    """The C3 model form (fcc, body-2 l <= 2 to the 4th shell, 54 neighbour
    slots: two or four per lane below 64 lanes) as a 6-rung ladder, exchanging
    after every sweep."""
    na, terms = M.lattice_model("fcc", pair_shells=4, pair_lmax=2, seed=12)
    H = S.TiledHamiltonian(na, terms, (3, 3, 3))
    R = 6
    make = _ladder(H, R, np.geomspace(0.002, 0.03, R), seed0=41)
    kw = dict(nsweeps=10, lanes=lanes, renorm_interval=4, measure_interval=1, exchange_interval=1, phase_done=0,
              phase_total=10)
    (r0, s0), (r1, s1) = _both(make, [kw])
    assert np.array_equal(r0[0], r1[0])
    _assert_same(s0, s1)


@pytest.mark.parametrize("lanes", [32, 64])
def test_c3_full_size_ladder_persistent(lanes):
    """Four 6912-site fcc 12^3 replicas of the C3 model (14 reduction chunks
    per sweep, so the in-kernel energy runs both tree levels), three sweeps
    with exchanges, every sweep measured."""
    na, terms, dims = M.config_model("fcc12", seed=0)
    H = S.TiledHamiltonian(na, terms, dims)
    assert H.n_sites == 6912
    make = _ladder(H, 4, np.array([0.004, 0.005, 0.006, 0.008]), seed0=77)
    kw = dict(nsweeps=3, lanes=lanes, renorm_interval=0, measure_interval=1, exchange_interval=1, phase_done=0,
              phase_total=3)
    (r0, s0), (r1, s1) = _both(make, [kw])
    assert np.array_equal(r0[0], r1[0])
    _assert_same(s0, s1)


@pytest.mark.parametrize("lanes", [16, 32, 64])
def test_fixture_models_as_batches_persistent(lanes):
    """The ten fixture models (pairs, triplets, bodies of 4, Zeeman-only,
    inactive sites, lattices) as independent chains, one batch per lmax."""
    groups = {}
    for k, (name, na, terms, dims, mm, fl) in enumerate(fixture_cases()):
        H = S.TiledHamiltonian(na, terms, dims, magmoms=mm, field=fl)
        groups.setdefault(H.lmax, []).append((k, H))
    for chains in groups.values():
        gHs = [S.GPUTiledHamiltonian(H) for _, H in chains]
        sts = [S.ChainState(H, rand_config(H.n_sites, 7 + k), 0.5) for k, H in chains]

        def make():
            return GPUBatch(gHs, sts, [100 + k for k, _ in chains], [0.01 + 0.002 * k for k, _ in chains])
        kw = dict(nsweeps=7, lanes=lanes, renorm_interval=3, measure_interval=2, exchange_interval=0,
                  phase_done=0, phase_total=7)
        (r0, s0), (r1, s1) = _both(make, [kw])
        assert np.array_equal(r0[0], r1[0])
        _assert_same(s0, s1)


@pytest.mark.parametrize("lanes", [32, 64])
def test_uneven_zeeman_batch_persistent(lanes):
    """Uneven chains (sizes 125-1000 sites, per-chain fields and moments), no
    exchange: each chain advances on its own waves."""
    rng = np.random.default_rng(404)
    specs = [("sc", 10, 2), ("bcc", 6, 2), ("bcc", 5, 2), ("bcc", 7, 2), ("sc", 4, 2)]
    gHs, states, kts, seeds = [], [], [], []
    for k, (lat, N, lmax) in enumerate(specs):
        na, terms = M.lattice_model(lat, pair_shells=2, pair_lmax=lmax, triangles=(lat == "bcc"), seed=500 + k)
        B = rng.uniform(0.0, 10.0) * (lambda v: v / np.linalg.norm(v))(rng.standard_normal(3))
        H = S.TiledHamiltonian(na, terms, (N, N, N), magmoms=[2.0] * na, field=tuple(B))
        gHs.append(S.GPUTiledHamiltonian(H))
        states.append(S.ChainState(H, rand_config(H.n_sites, 90 + k), 0.4 + 0.1 * k))
        kts.append(float(rng.uniform(0.005, 0.05)))
        seeds.append(int(rng.integers(1, 2**62)))

    def make():
        return GPUBatch(gHs, states, seeds, kts)
    kw = dict(nsweeps=14, lanes=lanes, renorm_interval=6, measure_interval=3, exchange_interval=0, phase_done=0,
              phase_total=14)
    (r0, s0), (r1, s1) = _both(make, [kw])
    assert np.array_equal(r0[0], r1[0])
    _assert_same(s0, s1)


def test_persistent_batch_lane_guard():
    H = S.TiledHamiltonian(*bcc_tet_small(2, dims=(2, 2, 2)))
    b = _ladder(H, 2, [0.01, 0.02])()
    b.set_persistent(True)
    with pytest.raises(SceError, match="lanes"):
        b.run(2, 8, 0, 0, 1, 0, 2)


def test_gpu_run_pt_persistent_equals_per_colour():
    """gpu_run_pt with persistent=True (thermalisation with step adaptation,
    measurement phase, renormalisations) returns the per-colour run's result."""
    H = S.TiledHamiltonian(*bcc_tet_small(4, dims=(2, 2, 2)))
    gH = S.GPUTiledHamiltonian(H)
    kw = dict(kT=[0.01, 0.02, 0.04], exchange_interval=2, sweeps_therm=60, sweeps_measure=40, renorm_interval=25,
              adapt_interval=20, seed=5, lanes=64)
    a = S.gpu_run_pt(gH, persistent=False, **kw)
    b = S.gpu_run_pt(gH, persistent=True, **kw)
    assert b.persistent and not a.persistent
    for x, y in zip(a.final_configs, b.final_configs):
        assert np.array_equal(x, y)
    assert np.array_equal(a.swap_acceptance, b.swap_acceptance, equal_nan=True)
    for p, q in zip(a.points, b.points):
        assert p.final_step == q.final_step and p.acceptance_metropolis == q.acceptance_metropolis
        for name, st in p.stats.items():
            assert (st.mean, st.err, st.tau_int, st.count) == (q.stats[name].mean, q.stats[name].err,
                                                                q.stats[name].tau_int, q.stats[name].count)
