"""Pins of the CPU oracle against what the reference's tests hold and what the
mathematics fixes — every oracle function has at least one check here that a
plausible mistake (dropped term, wrong sign/index, transposed operand) fails.

Sources: Philox Random123 known answers (test/unit/test_gpu.jl:49-84);
closed forms of Z_lm, the addition theorem and mpmath values; the reference's
hand-built fixtures (test/unit/fixtures.jl:94-131) and their energies
(test_hamiltonian.jl:92-122); ΔE ≡ total-energy difference across body orders
(test_energy.jl:19-53); colouring properness (test_parallel.jl:7-24).
"""
import itertools
import math

import mpmath as mp
import numpy as np
import pytest

import oracle as O
from sce_amd import models as M
from fixtures import fixture_cases, rand_config

mp.mp.prec = 120


# ---- Philox4x32-10 (philox.jl) ------------------------------------------------

def test_philox_random123_known_answers():
    assert O.philox4x32((0, 0, 0, 0), (0, 0)) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert O.philox4x32((0xffffffff,) * 4, (0xffffffff,) * 2) == \
        (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)
    assert O.philox4x32((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0)) == \
        (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_philox_counter_layout_and_stream_separation():
    seed = 0x299f31d0a4093822
    # block (seed, site, sweep, slot) == philox(ctr=(site, sweep, slot, 0), key=(lo, hi))
    b = O.philox_block(seed, 0x243f6a88 - 2**32, -0x7a5cf72d, 0x13198a2e, 0x03707344)
    assert tuple(int(x) for x in b) == (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)
    ref = tuple(O.philox_block(1, 3, 5, 0))
    for args in ((2, 3, 5, 0), (1, 4, 5, 0), (1, 3, 6, 0), (1, 3, 5, 1), (1, 3, 5, 0, 1)):
        assert tuple(O.philox_block(*args)) != ref
    assert tuple(O.philox_block(1, 3, 5, 0)) == ref


def test_proposal_reads_its_named_slots_value_for_value():
    """The proposal restated from the slot constants outside the proposal function
    (test_gpu.jl:219-251): slot 0 = (flip uniform, accept uniform), slot 1 = (n1, n2),
    slot 2 = (n3, angle n4); flip fraction 0.2; e' = cos t e + sin t (a x e) +
    (1 - cos t)(a.e) a with a = normalize(n1, n2, n3), t = step n4.  Collapsing two
    slots inside the function moves device and oracle together and only this pin
    sees it."""
    seed, step = 0xBEEF, 0.6
    e = np.array([0.0, 0.0, 1.0])
    nflip = 0
    for site in range(1, 25):
        for sw in range(1, 5):
            b0 = O.philox_block(seed, site, sw, 0)
            u_flip, u_acc = O.uniform(int(b0[0]), int(b0[1])), O.uniform(int(b0[2]), int(b0[3]))
            e2, ua = O.keyed_proposal(seed, site, sw, e, step)
            assert ua == u_acc
            if u_flip < 0.2:
                nflip += 1
                assert np.array_equal(e2, -e)
            else:
                n1, n2 = O.normal2(O.philox_block(seed, site, sw, 1))
                n3, n4 = O.normal2(O.philox_block(seed, site, sw, 2))
                ax = np.array([n1, n2, n3]) / math.sqrt(n1 * n1 + n2 * n2 + n3 * n3)
                t = step * n4
                want = math.cos(t) * e + math.sin(t) * np.cross(ax, e) + (1 - math.cos(t)) * np.dot(ax, e) * ax
                np.testing.assert_allclose(e2, want, atol=1e-14)
    assert 0 < nflip < 96


def test_reflection_is_an_involution():
    rng = np.random.default_rng(2)
    for _ in range(5):
        e, a = rng.standard_normal(3), rng.standard_normal(3)
        e /= np.linalg.norm(e)
        a /= np.linalg.norm(a)
        S = lambda x: 2 * np.dot(x, a) * a - x
        np.testing.assert_allclose(S(S(e)), e, atol=1e-15)
        assert abs(np.linalg.norm(S(e)) - 1.0) < 1e-15


def test_uniform_open_interval_exact_endpoints():
    assert O.uniform(0, 0) == 2.0 ** -53
    assert O.uniform(0xffffffff, 0xffffffff) == 1.0 - 2.0 ** -53
    # top 52 bits of hi<<32|lo
    assert O.uniform(0x80000000, 0) == 0.5 + 2.0 ** -53


def test_box_muller_pair():
    blk = O.philox_block(7, 1, 1, 1)
    u1 = O.uniform(int(blk[0]), int(blk[1]))
    u2 = O.uniform(int(blk[2]), int(blk[3]))
    r = math.sqrt(-2 * math.log(u1))
    n = O.normal2(blk)
    assert n[0] == pytest.approx(r * math.cos(2 * math.pi * u2), rel=1e-14)
    assert n[1] == pytest.approx(r * math.sin(2 * math.pi * u2), rel=1e-14)


# ---- deterministic elementary functions -----------------------------------------

def _ulps(a, b):
    b = float(b)
    return abs(a - b) / np.spacing(abs(b)) if b != 0 else abs(a)


def test_exp_within_one_ulp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([-rng.exponential(5, 3000), rng.uniform(-700, 700, 500),
                         [0.0, -1e-300, 1e-20, -745.0, -740.0, 709.7, 1.0, -1.0]])
    assert max(_ulps(O.exp(x), mp.exp(x)) for x in xs) <= 1.0
    assert O.exp(0.0) == 1.0 and O.exp(-800.0) == 0.0 and math.isinf(O.exp(710.0))


def test_log_within_two_ulps():
    rng = np.random.default_rng(2)
    ys = np.concatenate([rng.uniform(0, 1, 3000), rng.uniform(1, 1e6, 300),
                         [2.0 ** -53, 1 - 2.0 ** -53, 1e-300, 5e-324, 0.5, 2.0]])
    assert max(_ulps(O.log(y), mp.log(y)) for y in ys) <= 2.0
    assert O.log(1.0) == 0.0


def test_sincos_accuracy_and_quadrants():
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.uniform(0, 2 * math.pi, 3000), rng.normal(0, 8, 1000)]):
        s, c = O.sincos(x)
        assert abs(s - float(mp.sin(x))) <= np.spacing(1.0)
        assert abs(c - float(mp.cos(x))) <= np.spacing(1.0)
    assert O.sincos(0.0) == (0.0, 1.0)


# ---- tesseral row (zlm_device.jl) -----------------------------------------------

def test_zlm_closed_forms():
    rng = np.random.default_rng(4)
    N1 = math.sqrt(3 / (4 * math.pi))
    for u in rng.standard_normal((200, 3)):
        u = u / np.linalg.norm(u)
        z = O.zlm_row(u, 2)
        assert z[0] == pytest.approx(1 / math.sqrt(4 * math.pi), rel=1e-15)
        # Z_1m = N1 (y, z, x)
        assert np.allclose(z[1:4], N1 * np.array([u[1], u[2], u[0]]), rtol=0, atol=4e-16)
        # l = 2 real harmonics
        c2 = math.sqrt(15 / (4 * math.pi))
        x, y, zz = u
        ref2 = [c2 * x * y, c2 * y * zz, math.sqrt(5 / (16 * math.pi)) * (3 * zz * zz - 1),
                c2 * x * zz, c2 / 2 * (x * x - y * y)]
        assert np.allclose(z[4:9], ref2, rtol=0, atol=2e-15)


@pytest.mark.parametrize("lmax", [3, 4, 5, 6])
def test_zlm_addition_theorem(lmax):
    rng = np.random.default_rng(lmax)
    for _ in range(40):
        a = rng.standard_normal(3); a /= np.linalg.norm(a)
        b = rng.standard_normal(3); b /= np.linalg.norm(b)
        za, zb = O.zlm_row(a, lmax), O.zlm_row(b, lmax)
        for l in range(lmax + 1):
            s = float(np.dot(za[l * l:(l + 1) ** 2], zb[l * l:(l + 1) ** 2]))
            pl = float(mp.legendre(l, float(np.dot(a, b))))
            assert s == pytest.approx((2 * l + 1) / (4 * math.pi) * pl, abs=1e-13)


def test_zlm_poles_and_signs():
    # at the north pole only m = 0 survives; Z_l0(±z) = ±^l sqrt((2l+1)/4π)
    for sgn in (1.0, -1.0):
        z = O.zlm_row(np.array([0, 0, sgn]), 6)
        for l in range(7):
            row = z[l * l:(l + 1) ** 2]
            assert row[l] == pytest.approx(sgn ** l * math.sqrt((2 * l + 1) / (4 * math.pi)), rel=1e-14)
            assert np.all(np.abs(np.delete(row, l)) < 1e-15)
    # Z_{l,±l} on the equator: tesseral sign convention of Harmonics (parity^2 = +1)
    z = O.zlm_row(np.array([1.0, 0, 0]), 3)
    assert z[3] > 0 and z[8] > 0 and z[15] > 0


def test_zlm_orthonormal_by_quadrature():
    # Gauss–Legendre in cos(theta) x uniform in phi integrates degree <= 12 exactly
    xs, ws = np.polynomial.legendre.leggauss(16)
    phis = np.arange(32) * 2 * math.pi / 32
    G = np.zeros((49, 49))
    for x, w in zip(xs, ws):
        st = math.sqrt(1 - x * x)
        for p in phis:
            z = O.zlm_row(np.array([st * math.cos(p), st * math.sin(p), x]), 6)
            G += w * (2 * math.pi / 32) * np.outer(z, z)
    assert np.allclose(G, np.eye(49), atol=1e-13)


# ---- tiled Hamiltonian (hamiltonian.jl) -----------------------------------------

def test_self_image_chain_energies_and_instances():
    J = -0.05
    with pytest.raises(ValueError):
        O.Tiled(1, M.chain_terms(J), (1, 1, 1))
    T = O.Tiled(1, M.chain_terms(J), (4, 1, 1))
    assert T.n_sites == 4 and T.n_inst == 8
    up = np.tile([0.0, 0.0, 1.0], (4, 1))
    assert T.total_energy(T.zrows(up)) == pytest.approx(4 * J, abs=1e-12)
    flipped = up.copy(); flipped[2] = -flipped[2]
    assert T.total_energy(T.zrows(flipped)) == pytest.approx(0.0, abs=1e-12)


def test_scale_once_field_term():
    c0 = 0.3
    T = O.Tiled(1, M.field_term(c0), (1, 1, 1))
    e = np.array([0.3, -0.4, 0.85]); e /= np.linalg.norm(e)
    N1 = math.sqrt(3 / (4 * math.pi))
    assert T.total_energy(T.zrows(e[None])) == pytest.approx(c0 * N1 * e[2], abs=1e-14)


def test_zeeman_closed_form():
    mu_b = 9.2740100783e-24 / 1.602176634e-19
    B = np.array([0.3, -1.2, 2.0]); mm = [1.0, 2.0]
    T = O.Tiled(2, [], (2, 2, 1), magmoms=mm, field=B)
    cfg = rand_config(T.n_sites, 5)
    E = T.total_energy(T.zrows(cfg))
    ref = -mu_b * sum(mm[s % 2] * float(np.dot(cfg[s], B)) for s in range(T.n_sites))
    assert E == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_csr_self_consistent_and_colouring_proper(case):
    name, na, terms, dims, mm, fl = case
    T = O.Tiled(na, terms, dims, magmoms=mm, field=fl)
    assert T.inst_ptr[0] == 1 and T.inst_ptr[-1] == len(T.inst_sites) + 1
    assert T.site_ptr[0] == 1 and T.site_ptr[-1] == len(T.inst_sites) + 1
    counted = np.zeros(T.n_inst, int)
    for s in range(1, T.n_sites + 1):
        for j in range(T.site_ptr[s - 1], T.site_ptr[s]):
            i, slot = T.site_inst[j - 1], T.site_slot[j - 1]
            assert T.inst_sites[T.inst_ptr[i - 1] + slot - 2] == s
            counted[i - 1] += 1
    assert np.array_equal(counted, np.diff(T.inst_ptr))
    colour = np.zeros(T.n_sites + 1, int)
    for c in range(T.n_colors):
        colour[T.color_sites[T.color_ptr[c] - 1:T.color_ptr[c + 1] - 1]] = c + 1
    for i in range(T.n_inst):
        mem = T.inst_sites[T.inst_ptr[i] - 1:T.inst_ptr[i + 1] - 1]
        assert len(set(colour[mem])) == len(mem)
    assert np.array_equal(np.sort(T.color_sites), np.nonzero(T.site_active)[0] + 1)
    assert T.color_ptr[-1] - 1 == T.n_active


def _brute_greedy(T):
    conflict = [set() for _ in range(T.n_sites + 1)]
    for i in range(T.n_inst):
        mem = list(T.inst_sites[T.inst_ptr[i] - 1:T.inst_ptr[i + 1] - 1])
        for a, b in itertools.permutations(mem, 2):
            conflict[a].add(b)
    col = {}
    for s in range(1, T.n_sites + 1):
        if not T.site_active[s - 1]:
            continue
        used = {col[x] for x in conflict[s] if x in col}
        c = 1
        while c in used:
            c += 1
        col[s] = c
    return col


@pytest.mark.parametrize("case", fixture_cases()[:7], ids=lambda c: c[0])
def test_colouring_equals_brute_force_greedy(case):
    name, na, terms, dims, mm, fl = case
    T = O.Tiled(na, terms, dims, magmoms=mm, field=fl)
    col = _brute_greedy(T)
    for c in range(T.n_colors):
        for s in T.color_sites[T.color_ptr[c] - 1:T.color_ptr[c + 1] - 1]:
            assert col[int(s)] == c + 1


def test_periodic_replication_is_n_times_cell_energy():
    na, terms = M.lattice_model("bcc", pair_shells=2, pair_lmax=2, triangles=True, seed=3)
    T1 = O.Tiled(na, terms, (3, 3, 3))
    T2 = O.Tiled(na, terms, (6, 3, 3))
    cell = rand_config(na, 1)
    c1 = np.tile(cell, (27, 1))
    c2 = np.tile(cell, (54, 1))
    assert T2.total_energy(T2.zrows(c2)) == pytest.approx(2 * T1.total_energy(T1.zrows(c1)), rel=1e-12)


@pytest.mark.parametrize("lattice,shells,expected", [("sc", 2, [3, 6]), ("bcc", 3, [8, 6, 12]),
                                                      ("fcc", 4, [24, 12, 48, 24])])
def test_neighbour_shell_counts(lattice, shells, expected):
    # bonds per cubic cell = n_atoms * coordination / 2 (sc: 6, 12; bcc: 8, 6, 12; fcc: 12, 6, 24, 12)
    assert [len(M.bonds(lattice, s)) for s in range(1, shells + 1)] == expected


def test_bcc_triangles_and_tetrahedra_counts():
    assert len(M.bcc_triangles()) == 24     # 6 2NN bonds x 4 common 1NN neighbours
    assert len(M.bcc_tetrahedra()) == 12    # 6 Delaunay tetrahedra per atom x 2 atoms


# ---- energy contract (energy.jl) -------------------------------------------------

@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_delta_energy_is_total_energy_difference(case):
    name, na, terms, dims, mm, fl = case
    T = O.Tiled(na, terms, dims, magmoms=mm, field=fl)
    cfg = rand_config(T.n_sites, 3)
    z = T.zrows(cfg)
    E0 = T.total_energy(z)
    rng = np.random.default_rng(11)
    for s in T.color_sites[:6]:
        e2 = rng.standard_normal(3); e2 /= np.linalg.norm(e2)
        zn = O.zlm_row(e2, T.lmax)
        c = T.site_coeffs(int(s), z)
        dE = O.delta_energy(c, z[s - 1], zn)
        cfg2 = cfg.copy(); cfg2[s - 1] = e2
        exact = T.total_energy(T.zrows(cfg2)) - E0
        assert abs(dE - exact) < 1e-13 * max(1.0, abs(E0))
        # c is independent of the site's own spin
        z3 = z.copy(); z3[s - 1] = zn
        assert np.array_equal(T.site_coeffs(int(s), z3), c)


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_walk_contract_within_1e12_of_site_coeffs(case):
    name, na, terms, dims, mm, fl = case
    T = O.Tiled(na, terms, dims, magmoms=mm, field=fl)
    cfg = rand_config(T.n_sites, 31)
    z = T.zrows(cfg)
    rng = np.random.default_rng(7)
    for s in T.color_sites[:16]:
        e2 = rng.standard_normal(3); e2 /= np.linalg.norm(e2)
        zn = O.zlm_row(e2, T.lmax)
        c = T.site_coeffs(int(s), z)
        ref = O.delta_energy(c, z[s - 1], zn)
        scale = np.abs(c).sum() * (np.abs(zn).sum() + np.abs(z[s - 1]).sum())
        for lanes in (16, 32, 64, 128, 256):
            assert abs(T.walk_delta(int(s), z, zn, lanes) - ref) <= 1e-12 * max(scale, 1e-30)
            assert np.all(np.abs(T.walk_coeffs(int(s), z, lanes) - c) <= 1e-12 * max(np.abs(c).sum(), 1e-30))


def test_walk_single_lane_is_pid_ordered_serial_sum():
    # with one lane the walk is the serial fma sum over pid-sorted entries: check against
    # an independent numpy restatement built from the same tables
    na, terms, dims = 1, M.threebody_terms(0.05), (4, 2, 2)
    T = O.Tiled(na, terms, dims)
    cfg = rand_config(T.n_sites, 2)
    z = T.zrows(cfg)
    s = int(T.color_sites[3])
    e2 = np.array([0.0, 0.6, 0.8])
    zn = O.zlm_row(e2, T.lmax)
    dz = zn - z[s - 1]
    js = sorted(range(T.site_ptr[s - 1], T.site_ptr[s]), key=lambda j: T.site_prog[j - 1])
    a = 0.0
    for j in js:
        pid = T.site_prog[j - 1]
        col, col2 = T.site_col[j - 1], T.site_col2[j - 1]
        for e in range(T.sprog_ptr[pid - 1], T.sprog_ptr[pid]):
            p = z[-col - 1, T.pent_row[e - 1] - 1] * z[col2 - 1, T.pent_row2[e - 1] - 1]
            if p != 0.0:
                a = _fma(T.sent_w[e - 1] * p, dz[T.sent_tgt[e - 1] - 1], a)
    assert O.lib().ora_walk_delta(T.handle, s, z.ctypes.data_as(O.C.c_void_p),
                                  zn.ctypes.data_as(O.C.c_void_p), 1) == a


def _fma(x, y, z):
    from fractions import Fraction
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def test_reduce_tree_shape():
    x = np.arange(1, 1301, dtype=float) * 1e-3
    ref = math.fsum(x)
    assert O.reduce_tree(x) == pytest.approx(ref, rel=1e-15)
    # the shape is a pairwise tree (offsets 256..1), not a serial sum:
    # serial ((1e16 + 1) - 1e16) + 1 = 1, the tree (1e16 - 1e16) + (1 + 1) = 2
    assert O.reduce_tree(np.array([1e16, 1.0, -1e16, 1.0])) == 2.0
    y = np.ones(1025)
    assert O.reduce_tree(y) == 1025.0


def test_keyed_sweep_energy_stays_exact():
    na, terms = M.lattice_model("bcc", pair_shells=2, pair_lmax=2, triangles=True, tetrahedra=True, seed=8)
    T = O.Tiled(na, terms, (3, 3, 3))
    cfg = rand_config(T.n_sites, 4)
    z = T.zrows(cfg)
    E = T.total_energy(z)
    for sw in range(1, 41):
        _, E, _, _ = T.keyed_sweep(cfg, z, E, 1 / 0.003, 0.6, 99, sw, 64)
    assert abs(E - T.total_energy(T.zrows(cfg))) <= 1e-10 * max(1.0, abs(E))
    assert np.allclose(np.linalg.norm(cfg, axis=1), 1.0, atol=1e-12)


def test_or_pure_l1_conserves_energy():
    # pure l=1 isotropic dimer: the reflection has ΔE == 0 and is always accepted
    T = O.Tiled(4, M.dimer_terms(-0.02), (1, 1, 1))
    cfg = rand_config(4, 9)
    z = T.zrows(cfg)
    E0 = T.total_energy(z)
    E = E0
    for sw in range(1, 21):
        acc, att, E = T.keyed_or_sweep(cfg, z, E, 100.0, 5, sw, 0, 64)
        assert acc == att == 2
    assert E == pytest.approx(E0, abs=1e-12)
    assert T.total_energy(T.zrows(cfg)) == pytest.approx(E0, abs=1e-12)


def test_cpu_chain_serial_equals_threaded():
    na, terms = M.lattice_model("bcc", pair_shells=2, pair_lmax=2, triangles=True, seed=2)
    T = O.Tiled(na, terms, (4, 4, 4))
    cfg = rand_config(T.n_sites, 1)
    runs = []
    for nt in (1, 3, 7):
        ch = O.CpuChain(T, cfg, 17, 0.6)
        for _ in range(5):
            ch.metropolis_sweep(1 / 0.002, nt)
            ch.or_sweep(1 / 0.002, nt)
        runs.append((ch.state(), ch.counters()))
    for (st, cnt) in runs[1:]:
        assert np.array_equal(st[0], runs[0][0][0]) and st[2] == runs[0][0][2]
        assert np.array_equal(cnt, runs[0][1])


# ---- contraction order of asymmetric tensors (closed forms) -------------------

def _z_closed(u, l):
    """Real tesseral row of degree l from its closed form (lm order m = -l..l)."""
    x, y, z = u
    if l == 1:
        return math.sqrt(3 / (4 * math.pi)) * np.array([y, z, x])
    c2 = math.sqrt(15 / (4 * math.pi))
    return np.array([c2 * x * y, c2 * y * z, math.sqrt(5 / (16 * math.pi)) * (3 * z * z - 1),
                     c2 * x * z, c2 / 2 * (x * x - y * y)])


def _closed_form_energy(na, terms, dims, cfg):
    """E = sum over cells and terms of (4 pi)^(body/2) coef sum_idx folded[idx]
    prod_k Z_{l_k, idx_k}(e of member k), member k at atom a_k, cell c + shift_k
    (site = (a - 1) + na (c1 + d1 (c2 + d2 c3)), hamiltonian.jl:99-181)."""
    d1, d2, d3 = dims
    E = 0.0
    for c3 in range(d3):
        for c2 in range(d2):
            for c1 in range(d1):
                for t in terms:
                    rows = []
                    for a, sh, l in zip(t.atoms, t.shifts, t.ls):
                        s = (a - 1) + na * ((c1 + sh[0]) % d1 + d1 * ((c2 + sh[1]) % d2 + d2 * ((c3 + sh[2]) % d3)))
                        rows.append(_z_closed(cfg[s], l))
                    acc = 0.0
                    for idx in itertools.product(*[range(2 * l + 1) for l in t.ls]):
                        f = t.folded[idx]
                        if f != 0.0:
                            acc += f * math.prod(rows[k][i] for k, i in enumerate(idx))
                    E += (4 * math.pi) ** (len(t.atoms) / 2) * t.coef * acc
    return E


def _pair12_terms():
    """A (1, 2) bond along +y with two off-diagonal entries of folded (3 x 5):
    member 1 carries l = 1, member 2 l = 2."""
    f = np.zeros((3, 5))
    f[0, 3] = 1.3    # Z_{1,-1}(e_0) Z_{2,1}(e_+y)
    f[2, 0] = -0.5   # Z_{1,1}(e_0) Z_{2,-2}(e_+y)
    return [M.MultipoleTerm(0.02, [1, 1], [(0, 0, 0), (0, 1, 0)], [1, 2], f)]


@pytest.mark.parametrize("name, terms, dims", [
    ("3body", M.threebody_terms(0.05), (4, 2, 2)),
    ("4body", M.fourbody_terms(0.03), (4, 2, 2)),
    ("pair12", _pair12_terms(), (2, 3, 2)),
], ids=lambda v: v if isinstance(v, str) else "")
def test_asymmetric_fixture_energies_closed_form(name, terms, dims):
    """The reference's asymmetric 3- and 4-body fixtures (fixtures.jl:109-131)
    and an off-diagonal (1,2) pair: the oracle's and the product builder's total
    energies equal the closed-form contraction, and differ from it with the
    tensor's member axes reversed (so a transposed folded or swapped member
    slots fail)."""
    import sce_amd as S
    n = int(np.prod(dims))
    for seed in range(3):
        cfg = rand_config(n, 100 + seed)
        ref = _closed_form_energy(1, terms, dims, cfg)
        T = O.Tiled(1, terms, dims)
        assert T.total_energy(T.zrows(cfg)) == pytest.approx(ref, rel=1e-13, abs=1e-15)
        H = S.TiledHamiltonian(1, terms, dims)
        assert H.total_energy_rows(H.zrows(cfg)) == pytest.approx(ref, rel=1e-13, abs=1e-15)
        # the plausible slip: member axes of folded reversed (and, for the
        # mixed-l pair, the members' l swapped with them)
        rev = [M.MultipoleTerm(t.coef, t.atoms, t.shifts, t.ls[::-1], np.ascontiguousarray(np.transpose(t.folded)))
               for t in terms]
        wrong = _closed_form_energy(1, rev, dims, cfg)
        assert abs(wrong - ref) > 1e-6 * max(abs(ref), 1e-3)
