"""Pins of the oracle's delayed-acceptance restatement (CPU only).

The descriptor, the FP32 network and the two-stage sweep have no counterpart in
the reference, so they are pinned by what the mathematics fixes: the descriptor
against the instance table read directly (multisets, rotation invariance), the
network against an FP64 numpy forward whose backprop is itself checked by
finite differences, the sweep against the plain keyed sweep where both must
agree (beta = 0) and against the exact single-spin law (Langevin) under a
random, untrained surrogate — the law holds only if stage 2 corrects stage 1.
"""
import math

import numpy as np
import pytest

import oracle as O
import sur_mlp as R
from fixtures import build_pair, fixture_cases, rand_config, bcc_tet_small

CASES = {c[0]: c for c in fixture_cases()}


def _members(T, s):
    """Instances touching site s (1-based): list of (members in slot order, own slot)."""
    out = []
    for j in range(T.site_ptr[s - 1], T.site_ptr[s]):
        inst = T.site_inst[j - 1]
        mem = [int(x) for x in T.inst_sites[T.inst_ptr[inst - 1] - 1:T.inst_ptr[inst] - 1]]
        out.append((mem, int(T.site_slot[j - 1])))
    return out


def _expected_features(T, cfg, s, e):
    nbs, tri, tet, body1 = [], [], [], False
    for mem, own in _members(T, s):
        others = [m for k, m in enumerate(mem, start=1) if k != own]
        for m in others:
            if m not in nbs:
                nbs.append(m)
        E = [cfg[m - 1] for m in others]
        if len(others) == 2:
            tri.append(float(np.dot(e, np.cross(E[0], E[1]))))
        elif len(others) == 3:
            x, y, z = E
            tet.append(np.dot(e, x) * np.dot(y, z) + np.dot(e, y) * np.dot(x, z) + np.dot(e, z) * np.dot(x, y))
        elif len(others) == 0:
            body1 = True
    d = np.array([np.dot(e, cfg[m - 1]) for m in nbs])
    return d, np.array(tri), np.array(tet), body1


@pytest.mark.parametrize("name", ["dimer+field", "3body", "4body", "sc-onsite", "bcc-l2-tri", "bcc-tet", "fcc-l2"])
def test_descriptor_matches_instance_table(name):
    _, na, terms, dims, mm, fl = CASES[name]
    H, T = build_pair(na, terms, dims, mm, fl)
    cfg = rand_config(T.n_sites, 11)
    rng = np.random.default_rng(3)
    for s in np.nonzero(T.site_active)[0][:40] + 1:
        e = rng.standard_normal(3)
        e /= np.linalg.norm(e)
        phi = T.sur_features(cfg, int(s), e)
        d, tri, tet, body1 = _expected_features(T, cfg, s, e)
        nb = len(d)
        assert len(phi) == 2 * nb + len(tri) + len(tet) + (3 if body1 else 0)
        np.testing.assert_allclose(np.sort(phi[:nb]), np.sort(d.astype(np.float32)), atol=2e-7)
        np.testing.assert_allclose(np.sort(phi[nb:2 * nb]), np.sort((d * d).astype(np.float32)), atol=2e-7)
        k = 2 * nb
        np.testing.assert_allclose(np.sort(phi[k:k + len(tri)]), np.sort(tri), atol=4e-7)
        k += len(tri)
        np.testing.assert_allclose(np.sort(phi[k:k + len(tet)]), np.sort(tet), atol=8e-7)
        k += len(tet)
        if body1:
            np.testing.assert_array_equal(phi[k:], e.astype(np.float32))
        # the features pair with the neighbours in one consistent order
        np.testing.assert_allclose(phi[nb:2 * nb], phi[:nb].astype(np.float64) ** 2, atol=4e-7)


def test_descriptor_rotation_invariance_without_field():
    _, na, terms, dims, mm, fl = CASES["bcc-tet"]
    _, T = build_pair(na, terms, dims, mm, fl)
    cfg = rand_config(T.n_sites, 5)
    q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]  # proper rotation: triple products keep their sign
    for s in (1, 7, 30):
        e = cfg[s - 1]
        a = T.sur_features(cfg, s, e)
        b = T.sur_features(cfg @ q.T, s, q @ e)
        np.testing.assert_allclose(a, b, atol=1e-6)
    # an improper rotation flips the triple products only
    r = -np.eye(3)
    a = T.sur_features(cfg, 7, cfg[6])
    b = T.sur_features(cfg @ r.T, 7, r @ cfg[6])
    d, tri, tet, _ = _expected_features(T, cfg, 7, cfg[6])
    nb = len(d)
    np.testing.assert_array_equal(b[:2 * nb], a[:2 * nb])
    np.testing.assert_array_equal(b[2 * nb:2 * nb + len(tri)], -a[2 * nb:2 * nb + len(tri)])
    np.testing.assert_array_equal(b[2 * nb + len(tri):], a[2 * nb + len(tri):])


def _random_params(F, seed, scale_w=1.0):
    rng = np.random.default_rng(seed)
    p = R.glorot(F, rng).astype(np.float64)
    lay = np.cumsum([0, R.H * F, R.H, R.H * R.H, R.H, R.H, F])
    p[lay[1]:lay[2]] = rng.uniform(-0.2, 0.2, R.H)      # b1
    p[lay[3]:lay[4]] = rng.uniform(-0.2, 0.2, R.H)      # b2
    p[lay[5]:lay[6]] = rng.uniform(-0.3, 0.3, F)        # v
    return (p * scale_w).astype(np.float32)


@pytest.mark.parametrize("F", [3, 17, 64, 130])
def test_network_forward_matches_fp64(F):
    rng = np.random.default_rng(F)
    p = _random_params(F, F)
    for _ in range(20):
        a = rng.uniform(-1, 1, F).astype(np.float32)
        b = rng.uniform(-1, 1, F).astype(np.float32)
        got = O.sur_delta(F, p, 2.5, a, b)
        want = 2.5 * (R.forward(p, F, b[None])[0] - R.forward(p, F, a[None])[0])
        assert abs(got - want) <= 2e-5 * (1 + abs(want))
        assert O.sur_delta(F, p, 2.5, b, a) == -got       # exactly antisymmetric
        assert O.sur_delta(F, p, 2.5, a, a) == 0.0


def test_surrogate_tanh_pinned():
    """The surrogate's FP32 tanh against tanh in double: within 2.5 ulp of the
    single-precision result over [-12, 12] and near zero, odd, monotone,
    saturated to exactly 1 beyond |x| = 10, both zeros to +0, and continuous
    across the polynomial/exponential switch at |x| = 1/4."""
    xs = np.concatenate([np.linspace(-12, 12, 60001), np.geomspace(1e-9, 0.6, 4000),
                         -np.geomspace(1e-9, 0.6, 400)]).astype(np.float32)
    got = np.array([O.tanhf(x) for x in xs])
    want = np.tanh(xs.astype(np.float64))
    ulp = np.spacing(np.abs(want.astype(np.float32))).astype(np.float64)
    assert np.max(np.abs(got - want) / ulp) <= 2.5
    for x in xs[::97]:
        if x != 0:
            assert O.tanhf(-x) == -O.tanhf(x)
    grid = np.array([O.tanhf(x) for x in np.linspace(-11, 11, 20001).astype(np.float32)])
    assert np.all(np.diff(grid) >= 0) and np.all(np.abs(grid) <= 1.0)
    assert O.tanhf(10.001) == 1.0 and O.tanhf(-40.0) == -1.0
    for z in (0.0, -0.0):
        assert O.tanhf(z) == 0.0 and math.copysign(1.0, O.tanhf(z)) == 1.0
    q = np.float32(0.25)
    below = O.tanhf(np.nextafter(q, np.float32(0)))
    assert 0 < O.tanhf(q) - below <= 3 * np.spacing(np.float32(0.245))


def test_network_tanh_saturates_and_is_odd():
    F = 1
    p = np.zeros(R.H * F + R.H + R.H * R.H + 3 * R.H + F, np.float32)
    p[0] = 1.0                                   # W1[0, 0]
    p[R.H * F + R.H] = 1.0                       # W2[0, 0]
    p[R.H * F + 2 * R.H + R.H * R.H] = 1.0       # w3[0]
    for x in (-30.0, -2.0, -0.3, 1e-4, 0.7, 5.0, 40.0):
        d = O.sur_delta(F, p, 1.0, np.zeros(1, np.float32), np.array([x], np.float32))
        assert abs(d - np.tanh(np.tanh(x))) <= 3e-7
        d2 = O.sur_delta(F, p, 1.0, np.zeros(1, np.float32), np.array([-x], np.float32))
        assert d2 == -d


def test_backprop_matches_finite_differences():
    F, n = 9, 12
    rng = np.random.default_rng(2)
    p = _random_params(F, 4).astype(np.float64)
    A = rng.uniform(-1, 1, (n, F))
    B = rng.uniform(-1, 1, (n, F))
    y = rng.normal(0, 0.3, n)
    loss, g = R.loss_and_grad(p, F, A, B, y, 0.7)
    idx = np.concatenate([rng.choice(len(p), 40, replace=False), np.arange(len(p) - F - R.H, len(p))])
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_and_grad(pp, F, A, B, y, 0.7)[0] - R.loss_and_grad(pm, F, A, B, y, 0.7)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i])
    assert loss == pytest.approx(np.mean(((R.forward(p, F, B) - R.forward(p, F, A)) - y / 0.7) ** 2))


def _sub_params(T, cfg, seed, scale_w=1.0):
    params, scales = [], []
    for a in range(T.n_cell_atoms):
        s = a + 1
        F = len(T.sur_features(cfg, s, cfg[s - 1])) if T.site_active[s - 1] else 0
        params.append(_random_params(F, seed + a, scale_w) if F else np.zeros(1, np.float32))
        scales.append(1.0)
    return params, np.array(scales)


def test_da_sweep_at_beta_zero_is_the_plain_sweep():
    na, terms, dims = bcc_tet_small(2)
    _, T = build_pair(na, terms, dims)
    cfg = rand_config(T.n_sites, 8)
    params, scales = _sub_params(T, cfg, 3)
    c1, c2 = cfg.copy(), cfg.copy()
    z1, z2 = T.zrows(c1), T.zrows(c2)
    E1 = E2 = T.total_energy_tree(z1)
    for sw in range(1, 4):
        n1, n1s, E1 = T.da_sweep(c1, z1, E1, 0.0, 0.6, 77, sw, params, scales)
        n2, E2, _, _ = T.keyed_sweep(c2, z2, E2, 0.0, 0.6, 77, sw, 64)
        assert n1 == n1s == n2 == T.n_active
        assert E1 == E2
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(z1, z2)


def test_da_sweep_decisions_restated_from_parts():
    """Zero network: stage 1 passes everything (dE_s = 0), stage 2 is Metropolis
    on the exact walk dE with the slot-3 uniform."""
    na, terms, dims = bcc_tet_small(2)
    _, T = build_pair(na, terms, dims)
    cfg = rand_config(T.n_sites, 9)
    params = [np.zeros(R.H * len(T.sur_features(cfg, a + 1, cfg[a])) + 3 * R.H + R.H * R.H + R.H
                       + len(T.sur_features(cfg, a + 1, cfg[a])), np.float32) for a in range(na)]
    z = T.zrows(cfg)
    E0 = T.total_energy_tree(z)
    beta, seed, sweep = 1.0 / 2e-3, 5, 3
    c2, z2 = cfg.copy(), z.copy()
    n, n1, E = T.da_sweep(c2, z2, E0, beta, 0.4, seed, sweep, params, np.ones(na))
    assert n1 == T.n_active
    want = 0
    c, zz = cfg.copy(), z.copy()
    for q in range(T.n_active):
        s = int(T.color_sites[q])
        e2, _ = O.keyed_proposal(seed, s, sweep, c[s - 1], 0.4)
        zn = O.zlm_row(e2, T.lmax)
        dE = T.walk_delta(s, zz, zn, 64)
        b = O.philox_block(seed, s, sweep, 3, 0)
        if dE <= 0.0 or O.uniform(int(b[0]), int(b[1])) < O.exp(-beta * dE):
            c[s - 1], zz[s - 1] = e2, zn
            want += 1
    assert n == want
    np.testing.assert_array_equal(c2, c)
    assert abs(E - T.total_energy_tree(zz)) <= 1e-12 * max(1.0, abs(E)) * T.n_active


def _langevin(x):
    return 1.0 / np.tanh(x) - 1.0 / x


def test_da_chain_samples_the_exact_single_spin_law():
    """Independent moments in a field: <e . h_hat> = L(beta |h|) per sublattice,
    with an untrained random surrogate screening every proposal."""
    _, na, terms, dims, mm, fl = CASES["zeeman-only"]
    dims = (6, 6, 4)
    _, T = build_pair(na, terms, dims, mm, fl)
    cfg = rand_config(T.n_sites, 1)
    z = T.zrows(cfg)
    # field strength per sublattice from the exact energy: E(-h_hat) - E(h_hat) = 2|h|
    hhat = np.asarray(fl) / np.linalg.norm(fl)
    hs = []
    for a in range(na):
        c = cfg.copy()
        c[a] = hhat
        Ep = T.total_energy(T.zrows(c))
        c[a] = -hhat
        Em = T.total_energy(T.zrows(c))
        hs.append((Em - Ep) / 2)
    kT = 0.8 * max(hs)
    beta = 1.0 / kT
    params, _ = _sub_params(T, cfg, 21, scale_w=2.0)
    scales = np.array([2.0 * h for h in hs])
    E = T.total_energy_tree(z)
    n_sw, burn = 3000, 200
    proj = [[] for _ in range(na)]
    n1tot = 0
    for sw in range(1, n_sw + 1):
        _, n1, E = T.da_sweep(cfg, z, E, beta, 1.5, 99, sw, params, scales)
        n1tot += n1
        if sw > burn:
            m = cfg @ hhat
            for a in range(na):
                proj[a].append(m[a::na].mean())
    assert 0.2 * T.n_active * n_sw < n1tot < 0.98 * T.n_active * n_sw   # stage 1 really screens
    for a in range(na):
        x = np.array(proj[a])
        blocks = x[: len(x) // 20 * 20].reshape(20, -1).mean(1)
        err = blocks.std(ddof=1) / np.sqrt(len(blocks))
        want = _langevin(beta * hs[a])
        assert abs(x.mean() - want) < 4.5 * err + 1e-4, (a, x.mean(), want, err)


# ---- int8 stage-1 inference -------------------------------------------------

def _f32(q):
    """A rational rounded to the nearest float32 (ties to even), exactly."""
    from fractions import Fraction
    q = Fraction(q)
    if q == 0:
        return np.float32(0.0)
    e = math.floor(math.log2(abs(q)))
    while Fraction(2) ** e > abs(q):
        e -= 1
    while Fraction(2) ** (e + 1) <= abs(q):
        e += 1
    ulp = Fraction(2) ** (max(e, -126) - 23)
    m = q / ulp
    n = math.floor(m)
    rem = m - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(float(n * ulp))


def _fmaf(x, y, z):
    from fractions import Fraction
    return _f32(Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z)))


def _delta_q_restated(F, qp, scale, phiA, phiB):
    """ora_sur_delta_q from its definition, with exact rationals for every
    rounding step and Python integers for the dot products."""
    H = R.H
    o = np.cumsum([0, H * F, H, H, H * H, H, H, H, F, F])
    V1 = qp[o[0]:o[1]].reshape(H, F).astype(np.int64)
    r1, b1 = qp[o[1]:o[2]], qp[o[2]:o[3]]
    V2 = qp[o[3]:o[4]].reshape(H, H).astype(np.int64)
    r2, b2, w3, v, qin = qp[o[4]:o[5]], qp[o[5]:o[6]], qp[o[6]:o[7]], qp[o[7]:o[8]], qp[o[8]:o[9]]

    def quant(phi):
        x = np.rint(np.float32(phi) * qin)          # one float32 product, then RNE
        return np.clip(x, -127, 127).astype(np.int64)

    def head(phi, u):
        acc1 = V1 @ u
        h1 = np.array([int(np.rint(np.float32(127.0) * np.float32(O.tanhf(_fmaf(float(a), r1[j], b1[j])))))
                       for j, a in enumerate(acc1)], np.int64)
        acc2 = V2 @ h1
        t = np.zeros(H, np.float32)
        for j in range(H):
            o2 = np.float32(w3[j] * np.float32(O.tanhf(_fmaf(float(acc2[j]), r2[j], b2[j]))))
            s = np.float32(0.0)
            for f in range(j, F, H):
                s = np.float32(s + np.float32(v[f] * np.float32(phi[f])))
            t[j] = np.float32(o2 + s)
        off = H // 2
        while off >= 1:
            t[:off] = (t[:off] + t[off:2 * off]).astype(np.float32)
            off //= 2
        return t[0]

    tA = head(phiA, quant(phiA))
    tB = head(phiB, quant(phiB))
    return (float(tB) - float(tA)) * scale


def quantise_np(params, F, qin):
    """The library's quantisation of an FP32 network (float32 steps in its order)."""
    H = R.H
    p = np.asarray(params, np.float32)
    W1 = p[:H * F].reshape(H, F)
    W2 = p[H * F + H:H * F + H + H * H].reshape(H, H)

    def layer(W, sin):
        Wd = (W / sin[None, :]).astype(np.float32)
        t = np.abs(Wd).max(axis=1)
        V = np.zeros_like(Wd)
        r = np.zeros(W.shape[0], np.float32)
        for j in range(W.shape[0]):
            if t[j] == 0:
                continue
            m = np.float32(127.0) / t[j]
            V[j] = np.clip(np.rint((Wd[j] * m).astype(np.float32)), -127, 127)
            r[j] = t[j] / np.float32(127.0)
        return V, r

    V1, r1 = layer(W1, np.asarray(qin, np.float32))
    V2, r2 = layer(W2, np.full(H, 127.0, np.float32))
    return V1, r1, V2, r2


def _rand_qblob(F, rng, qin):
    p = R.glorot(F, rng).astype(np.float32)
    H = R.H
    V1, r1, V2, r2 = quantise_np(p, F, qin)
    lay = np.cumsum([0, H * F, H, H * H, H, H, F])
    b1 = rng.uniform(-0.2, 0.2, H).astype(np.float32)
    b2 = rng.uniform(-0.2, 0.2, H).astype(np.float32)
    w3, v = p[lay[4]:lay[5]], rng.uniform(-0.3, 0.3, F).astype(np.float32)
    return p, np.concatenate([V1.ravel(), r1, b1, V2.ravel(), r2, b2, w3, v, qin]).astype(np.float32)


@pytest.mark.parametrize("F", [7, 20, 112])
def test_int8_surrogate_matches_exact_restatement(F):
    """ora_sur_delta_q against its definition computed with exact integers and
    correctly rounded rationals, on descriptor-like inputs (|phi| <= 1, a few
    four-body pairing sums up to 3 in magnitude with input scale 42)."""
    rng = np.random.default_rng(F)
    qin = np.full(F, 127.0, np.float32)
    qin[-(F // 5):] = 42.0
    _, qp = _rand_qblob(F, rng, qin)
    for k in range(3):
        a = rng.uniform(-1, 1, F).astype(np.float32)
        b = rng.uniform(-1, 1, F).astype(np.float32)
        a[-(F // 5):] *= 3
        b[-(F // 5):] *= 3
        want = _delta_q_restated(F, qp, 0.37, a, b)
        got = O.sur_delta_q(F, qp, 0.37, a, b)
        assert got == want, (k, got, want)


def test_int8_quantisation_and_network_agree_with_fp32():
    """The quantised weights are int8 with per-unit factors that reproduce the
    FP32 weights to half a step, and the int8 surrogate differences follow the
    FP32 network's (a screen, not an exact value)."""
    F = 40
    rng = np.random.default_rng(5)
    qin = np.full(F, 127.0, np.float32)
    p, qp = _rand_qblob(F, rng, qin)
    H = R.H
    V1, r1, V2, r2 = quantise_np(p, F, qin)
    assert np.abs(V1).max() <= 127 and np.abs(V2).max() <= 127
    W1 = p[:H * F].reshape(H, F)
    assert np.all(np.abs(V1 * r1[:, None] - W1 / 127.0) <= 0.5 * r1[:, None] * 1.0001 + 1e-12)
    # the FP32 network with the blob's biases and head
    o = np.cumsum([0, H * F, H, H, H * H, H, H, H, F, F])
    pf = np.concatenate([W1.ravel(), qp[o[2]:o[3]], p[H * F + H:H * F + H + H * H], qp[o[5]:o[6]],
                         qp[o[6]:o[7]], qp[o[7]:o[8]]]).astype(np.float32)
    d8, d32 = [], []
    for _ in range(200):
        a = rng.uniform(-1, 1, F).astype(np.float32)
        b = rng.uniform(-1, 1, F).astype(np.float32)
        d8.append(O.sur_delta_q(F, qp, 1.0, a, b))
        d32.append(O.sur_delta(F, pf, 1.0, a, b))
    d8, d32 = np.array(d8), np.array(d32)
    assert np.corrcoef(d8, d32)[0, 1] > 0.99
    assert np.sqrt(np.mean((d8 - d32) ** 2)) < 0.05 * np.sqrt(np.mean(d32 ** 2))
