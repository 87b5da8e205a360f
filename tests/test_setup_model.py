"""Model input from a setup file (the step before the path, src/hamiltonian.jl:396-400;
bench/fixtures.jl:66-88): the loader against the Nd2Fe14B setup data the
reference's bench holds (tests/golden/nd2fe14b.toml), the cluster enumeration
against brute force over explicit supercells, and the isotropic tensors against
what the mathematics fixes — rotation invariance of the energy, the l = 0
reduction of the real Gaunt tensor, the closed form of its (1, 1, 2) coupling
and its parity / triangle zeros."""
import itertools
import math
import os

import numpy as np
import pytest

import sce_amd as S
from sce_amd import _lib as L
from sce_amd import models as M

HERE = os.path.dirname(os.path.abspath(__file__))
SETUP = os.path.join(HERE, "golden", "nd2fe14b.toml")
L2 = [2] * 8 + [0]


@pytest.fixture(scope="module")
def setup():
    return S.read_setup(SETUP)


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_read_setup(setup):
    assert setup.n_atoms == 68 and len(setup.species_labels) == 9
    assert setup.species_labels[0] == "Nd1" and setup.species_labels[-1] == "B"
    assert [setup.species.count(k) for k in range(1, 10)] == [4, 4, 16, 16, 8, 8, 4, 4, 4]
    assert np.allclose(setup.lattice, np.diag([8.8, 8.8, 12.2]))
    assert setup.nbody == 2 and math.isinf(setup.cutoff) and setup.isotropy
    assert setup.lmax == [4, 4, 2, 2, 2, 2, 2, 2, 0]
    assert np.all((setup.positions >= 0) & (setup.positions < 1))
    mm = S.setup_magmoms(setup, 2.5)
    assert [i for i, m in enumerate(mm) if m == 0] == list(range(64, 68))
    assert set(mm[:64]) == {2.5}


@pytest.mark.parametrize("old, new, match", [
    ("lmax = [4, 4, 2, 2, 2, 2, 2, 2, 0]", "lmax = [4, 4, 2]", "one entry"),
    ("nbody = 2", "nbody = 4", "nbody must be"),
    ("cutoff = inf", "cutoff = -1.0", "cutoff must be"),
])
def test_read_setup_rejects_malformed(tmp_path, old, new, match):
    text = open(SETUP).read()
    p = tmp_path / "bad.toml"
    p.write_text(text.replace(old, new))
    with pytest.raises(ValueError, match=match):
        S.read_setup(str(p))
    p.write_text(text.replace("nbody = 2", "nbody = 3"))
    with pytest.raises(ValueError, match="finite cutoff"):
        S.read_setup(str(p))


def _supercell(setup, n=2):
    """Cartesian positions and (atom, shift) labels of every magnetic site in the
    (2n+1)^3 block of cells around the home cell."""
    P = setup.cartesian()
    mag = [setup.lmax[s - 1] >= 1 for s in setup.species]
    pts, lab = [], []
    for c in itertools.product(range(-n, n + 1), repeat=3):
        for a in range(setup.n_atoms):
            if mag[a]:
                pts.append(P[a] + np.asarray(c, float) @ setup.lattice)
                lab.append((a + 1, c))
    return np.asarray(pts), lab, mag


def _dist(setup, key):
    P = setup.cartesian()
    p = [P[a - 1] + np.asarray(c, float) @ setup.lattice for a, c in key]
    return [float(np.linalg.norm(p[i] - p[j])) for i in range(len(p)) for j in range(i + 1, len(p))]


@pytest.mark.parametrize("cutoff", [2.6, 3.5])
def test_pairs_within_cutoff_against_brute_force(setup, cutoff):
    pts, lab, mag = _supercell(setup)
    home = [i for i, (a, c) in enumerate(lab) if c == (0, 0, 0)]
    d = np.linalg.norm(pts[home][:, None, :] - pts[None, :, :], axis=-1)
    coord = int(np.sum((d > 1e-9) & (d <= cutoff + setup.tol)))
    pairs = M.setup_pairs(setup, cutoff, mag)
    # each translation-unique pair is seen once from either member's home cell
    assert 2 * len(pairs) == coord
    assert len(set(pairs)) == len(pairs)
    assert all(max(_dist(setup, k)) <= cutoff + setup.tol for k in pairs)
    assert all(k[0][1] == (0, 0, 0) for k in pairs)


def test_minimum_image_pairs(setup):
    mag = [setup.lmax[s - 1] >= 1 for s in setup.species]
    pairs = M.setup_pairs(setup, math.inf, mag)
    assert len(pairs) == 64 * 63 // 2
    assert {tuple(sorted((k[0][0], k[1][0]))) for k in pairs} == set(itertools.combinations(range(1, 65), 2))
    P = setup.cartesian()
    for k in pairs[::37]:
        (a, _), (b, _) = k
        best = min(float(np.linalg.norm(P[b - 1] - P[a - 1] + np.asarray(c, float) @ setup.lattice))
                   for c in itertools.product(range(-2, 3), repeat=3))
        assert abs(_dist(setup, k)[0] - best) < 1e-9


def test_minimum_image_pairs_on_a_skewed_cell():
    """An unreduced, strongly sheared cell (b = (3, 1, 0) in units of a): its
    minimum images lie beyond the {-1, 0, 1}^3 shifts; every pair's distance
    equals brute force over +-6 cells, and the chosen shift attains it."""
    lat = np.array([[1.0, 0.0, 0.0], [3.0, 1.0, 0.0], [0.7, 0.4, 1.1]])
    pos = np.array([[0.0, 0.0, 0.0], [0.43, 0.61, 0.27], [0.9, 0.15, 0.8]])
    st = M.Setup(lattice=lat, positions=pos, species=[1, 1, 1], species_labels=["X"], lmax=[1])
    pairs = M.setup_pairs(st, math.inf, [True] * 3)
    assert len(pairs) == 3
    P = st.cartesian()
    far = False
    for (a, ca), (b, cb) in pairs:
        sh = np.asarray(cb, float) - np.asarray(ca, float)
        got = float(np.linalg.norm(P[b - 1] + sh @ lat - P[a - 1]))
        best = min(float(np.linalg.norm(P[b - 1] - P[a - 1] + np.asarray(c, float) @ lat))
                   for c in itertools.product(range(-6, 7), repeat=3))
        assert abs(got - best) < 1e-9, ((a, ca), (b, cb))
        far |= max(abs(x) for x in sh) > 1
    assert far   # at least one minimum image lies outside {-1, 0, 1}^3


def test_triplets_against_brute_force(setup):
    cutoff = 3.0
    pts, lab, mag = _supercell(setup, 1)
    home = [i for i, (a, c) in enumerate(lab) if c == (0, 0, 0)]
    D = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=-1)
    near = (D > 1e-9) & (D <= cutoff + setup.tol)
    count = 0
    for i in home:
        nb = np.flatnonzero(near[i])
        sub = near[np.ix_(nb, nb)]
        count += int(np.triu(sub, 1).sum())
    tri = M.setup_triplets(setup, cutoff, mag)
    # each translation-unique triplet is seen once from each member's home cell
    assert 3 * len(tri) == count
    assert all(max(_dist(setup, k)) <= cutoff + setup.tol for k in tri)


def test_real_gaunt_closed_forms():
    rng = np.random.default_rng(5)
    # l = 0 member: the integral of Z_00 Z_lm Z_lm' is delta_mm' / sqrt(4 pi)
    for l in (1, 2, 3, 4):
        t = M.real_gaunt((0, l, l))
        assert np.allclose(t[0], np.eye(2 * l + 1) / math.sqrt(2 * l + 1), atol=1e-13)
    # parity and triangle zeros
    for ls in ((1, 1, 1), (1, 2, 2), (1, 1, 3), (2, 2, 1)):
        assert not np.any(M.real_gaunt(ls))
    # (1, 1, 2): the invariant is proportional to 3 (u.w)(v.w) - u.v
    t = M.real_gaunt((1, 1, 2))
    ratios = []
    for _ in range(6):
        u, v, w = (x / np.linalg.norm(x) for x in rng.standard_normal((3, 3)))
        z = np.empty((3, 9))
        dirs = np.ascontiguousarray(np.stack([u, v, w]))
        L.check(L.lib().sce_host_zrows(2, 3, L.ptr(dirs), L.ptr(z)))
        val = np.einsum("abc,a,b,c->", t, z[0, 1:4], z[1, 1:4], z[2, 4:9])
        ratios.append(val / (3 * (u @ w) * (v @ w) - u @ v))
    assert np.ptp(ratios) < 1e-12 * max(1.0, abs(ratios[0])) and abs(ratios[0]) > 1e-6


@pytest.mark.parametrize("kw, dims", [
    (dict(), (1, 1, 1)),                                      # l <= 4 pairs, minimum image
    (dict(nbody=3, cutoff=3.5, lmax=L2), (2, 2, 2)),          # pairs + Gaunt triplets
])
def test_isotropic_model_is_rotation_invariant(setup, kw, dims):
    na, terms = S.setup_model(setup, **kw)
    H = S.TiledHamiltonian(na, terms, dims)
    cfg = S.random_config(H.n_sites, np.random.default_rng(3))
    E0 = H.total_energy(cfg)
    for seed in (1, 2):
        E1 = H.total_energy(cfg @ _rotation(seed).T)
        assert abs(E1 - E0) <= 1e-12 * max(1.0, sum(abs(t.coef) for t in terms))
    # an anisotropic model of the same cell is not
    na, terms = S.setup_model(setup, **kw, isotropy=False)
    Ha = S.TiledHamiltonian(na, terms, dims)
    assert abs(Ha.total_energy(cfg @ _rotation(1).T) - Ha.total_energy(cfg)) > 1e-6


def test_model_shape_and_determinism(setup):
    na, t1 = S.setup_model(setup, nbody=3, cutoff=3.5, lmax=L2, seed=13)
    _, t2 = S.setup_model(setup, nbody=3, cutoff=3.5, lmax=L2, seed=13)
    assert [t.coef for t in t1] == [t.coef for t in t2]
    assert na == 68 and all(a <= 64 for t in t1 for a in t.atoms)   # boron carries no term
    assert {tuple(t.ls) for t in t1 if t.body == 3} == {(1, 1, 2), (1, 2, 1), (2, 1, 1), (2, 2, 2)}
    assert {tuple(t.ls) for t in t1 if t.body == 2} == {(1, 1), (2, 2)}
    # default lmax: Nd pairs reach l = 4, Fe ones stop at 2
    _, t3 = S.setup_model(setup)
    nd = {tuple(t.ls) for t in t3 if max(t.atoms) <= 8}
    assert (4, 4) in nd and all(max(t.ls) <= 2 for t in t3 if min(t.atoms) > 8)
